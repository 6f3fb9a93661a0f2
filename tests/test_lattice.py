"""Beam-lattice generator: closed-form beam results and structural invariants."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

E, G = 2.1e11, 8.1e10


def cantilever_tip(K, free=slice(6, 12)):
    """Clamp node 0, return the compliance of the free end node."""
    return np.linalg.inv(K[free, free])


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_cantilever_closed_forms(hbj, direction):
    A, Iy, Iz, J, L = 4e-3, 3e-6, 5e-6, 7e-6, 2.5
    K = hbj.beam_element_stiffness(E, G, A, Iy, Iz, J, L, direction)
    C = cantilever_tip(K)  # 6x6 tip compliance in global DOFs
    ax = direction
    # axial load along the beam: u = P L / (E A)
    assert np.isclose(C[ax, ax], L / (E * A), rtol=1e-12)
    # torsion about the beam axis: theta = T L / (G J)
    assert np.isclose(C[3 + ax, 3 + ax], L / (G * J), rtol=1e-12)
    # the two transverse directions: w = P L^3 / (3 E I), theta = P L^2 / (2 E I)
    trans = [d for d in range(3) if d != ax]
    Is = sorted([C[d, d] for d in trans])
    expect = sorted([L ** 3 / (3 * E * Iy), L ** 3 / (3 * E * Iz)])
    assert np.allclose(Is, expect, rtol=1e-12)
    for d in trans:
        rot = [r for r in range(3) if r not in (ax, d)][0] + 3
        assert np.isclose(abs(C[d, rot]), L ** 2 / (2 * E * (Iz if C[d, d] == pytest.approx(L ** 3 / (3 * E * Iz)) else Iy)), rtol=1e-9)


@pytest.mark.parametrize("direction", [0, 1, 2])
def test_element_rigid_body_modes(hbj, direction):
    K = hbj.beam_element_stiffness(E, G, 5e-3, 2e-6, 3e-6, 4e-6, 1.0, direction)
    assert np.allclose(K, K.T, rtol=0, atol=1e-6 * np.abs(K).max())
    # node positions: node 0 at origin, node 1 at e_dir
    p1 = np.zeros(3)
    p1[direction] = 1.0
    modes = []
    for t in range(3):  # translations
        u = np.zeros(12)
        u[t] = u[6 + t] = 1
        modes.append(u)
    for r in range(3):  # small rotations: u = w x p, theta = w
        w = np.zeros(3)
        w[r] = 1
        u = np.zeros(12)
        u[3:6] = w
        u[6:9] = np.cross(w, p1)
        u[9:12] = w
        modes.append(u)
    for u in modes:
        assert np.abs(K @ u).max() <= 1e-6 * np.abs(K).max()
    ev = np.linalg.eigvalsh(K)
    assert (ev > 1e-6 * ev.max()).sum() == 6  # exactly six deformation modes


@pytest.mark.parametrize("N", [1, 2, 3, 5, 10])
def test_lattice_structure(hbj, N):
    Ls = hbj.lattice(N, seed=0)
    n, nnz = Ls["n"], Ls["nnz"]
    assert n == 6 * N ** 3
    free_nodes = N ** 3 - N ** 2
    free_edges = 3 * N * N * (N - 1) - (2 * N * (N - 1) + (N * N if N > 1 else 0))
    assert nnz == 36 * (free_nodes + 2 * free_edges) + 6 * N * N
    A = sp.csr_matrix((Ls["values"], Ls["col_idx"], Ls["row_ptr"]), shape=(n, n))
    assert abs(A - A.T).max() == 0
    assert np.array_equal(Ls["block_starts"], np.arange(0, n + 1, 6, dtype=np.uint32))
    cl = np.arange(6 * N * N)
    assert np.all(A[cl][:, cl].toarray() == np.eye(cl.size))
    assert np.all(Ls["rhs"][cl] == 0)
    rows = np.diff(Ls["row_ptr"])
    assert rows.max() <= 42
    # columns sorted within rows
    for i in range(0, n, max(1, n // 50)):
        c = Ls["col_idx"][Ls["row_ptr"][i]:Ls["row_ptr"][i + 1]]
        assert np.all(np.diff(c.astype(np.int64)) > 0)
    free = Ls["rhs"][6 * N * N:]
    if free.size:
        assert free.min() >= -1e3 and free.max() <= 1e3 and np.abs(free).max() > 100


def test_lattice_spd_and_deterministic(hbj):
    a = hbj.lattice(4, seed=7, e_log10_half_range=1.5)
    b = hbj.lattice(4, seed=7, e_log10_half_range=1.5)
    c = hbj.lattice(4, seed=8, e_log10_half_range=1.5)
    for k in ("row_ptr", "col_idx", "values", "rhs"):
        assert np.array_equal(a[k], b[k])
    assert not np.array_equal(a["values"], c["values"])
    n = a["n"]
    A = sp.csr_matrix((a["values"].astype(np.float64), a["col_idx"], a["row_ptr"]), shape=(n, n))
    ev = np.linalg.eigvalsh(A.toarray())
    assert ev.min() > 0


def test_lattice_gravity_column(hbj):
    # N=2 lattice: a unit cell.  Loading all four top nodes with -P in z gives
    # a total reaction of 4P spread over the 4 columns: each vertical column
    # (E A / L axial stiffness, fixed at the base) carries its own node's load
    # if the top beams do not transfer load: check the solve against the
    # closed form of the axial column within the shear/bending coupling bound.
    Ls = hbj.lattice(2, seed=1)
    n = Ls["n"]
    A = sp.csr_matrix((Ls["values"].astype(np.float64), Ls["col_idx"], Ls["row_ptr"]), shape=(n, n))
    b = np.zeros(n)
    top = [4, 5, 6, 7]
    for t in top:
        b[6 * t + 2] = -1e4
    x = spla.spsolve(A.tocsc(), b)
    # vertical displacement of each top node is -P L / (E A_col) with the
    # column's own A: recover A_col and check it lies in the sampled range
    for t in top:
        w = x[6 * t + 2]
        A_col = 1e4 * 1.0 / (E * -w)
        assert 1e-3 * 0.999 <= A_col <= 1e-2 * 1.001
