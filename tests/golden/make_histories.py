"""Write tests/golden/histories.npz: the oracle's FP32 residual histories.

    python tests/golden/make_histories.py

For the three reference datasets (solve parameters of the reference's
commands.txt: rel 1e-4, abs 1e-7, 2000 iterations, x0 from the dataset) and
the 10^3 beam lattice of the CPU-runnable workload (seed 0, rel 1e-6, 5000
iterations, x0 = 0): per iteration the FP32 r.r that the stop test sees, the
iteration count and the status, plus a SHA-256 of the generated lattice's
CSR, right-hand side and block starts.  Regenerated only when the oracle or
the generator is meant to change; tests/test_oracle.py checks the oracle
against it bit for bit.
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)


def lattice_digest(L):
    h = hashlib.sha256()
    for k in ("row_ptr", "col_idx", "values", "rhs", "block_starts"):
        h.update(np.ascontiguousarray(L[k]).tobytes())
    return h.hexdigest()


def main():
    import oracle_lib as O
    import wgpu_solver_backend_amd as H
    O.build()
    out = {}
    cases = []
    for k in (1, 2, 3):
        d = O.load_dataset(k)
        cases.append((f"ds{k}", d["row_ptr"], d["col_idx"], d["values"], d["block_starts"], d["b"],
                      d["x0"], 2000, 1e-4, 1e-7))
    L = H.lattice(10, seed=0)
    cases.append(("lattice10", L["row_ptr"], L["col_idx"], L["values"], L["block_starts"], L["rhs"],
                  np.zeros(L["n"], np.float32), 5000, 1e-6, 0.0))
    out["lattice10_sha256"] = np.array(lattice_digest(L))
    for name, rp, ci, va, bs, b, x0, mi, rel, ab in cases:
        st, lu = O.build_lu_blocks_6(len(rp) - 1, rp, ci, va, bs)
        assert st == 0
        st, it, x, hist = O.pcg(rp, ci, va, lu, bs, b, x0, mi, rel, ab, history=True)
        out[f"{name}_rr"] = hist
        out[f"{name}_status_iterations"] = np.array([st, it], np.int64)
        print(name, st, it, hist[-1])
    np.savez_compressed(os.path.join(HERE, "histories.npz"), **out)


if __name__ == "__main__":
    main()
