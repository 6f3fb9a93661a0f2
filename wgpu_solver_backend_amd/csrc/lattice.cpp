// lattice.cpp -- synthetic SPD systems: a 3D space frame of Euler-Bernoulli
// beams on an N x N x N node lattice (6 DOF per node: u, v, w, rx, ry, rz).
//
// The reference ships no generator (its inputs are FEA exports, datasets/*.txt);
// this is the benchmark workload generator, documented here in full:
//   node id        (k*N + j)*N + i, DOF index 6*node + d
//   elements       one per axis-neighbour pair (node, node + step(dir)),
//                  dir = x, y, z; spacing L = 1 m; 3 N^2 (N-1) elements
//   properties     E = 2.1e11, G = 8.1e10; per element, in the order
//                  k, j, i, dir: A ~ U[1e-3,1e-2], Iy, Iz, J ~ U[1e-6,1e-5],
//                  then (if h > 0) f = 10^U[-h,h] scaling both E and G
//   RNG            xoshiro256** seeded by splitmix64(seed); U = (next>>11)*2^-53
//   assembly       full 12x12 element matrices scatter-added in FP64 (explicit
//                  zeros kept: every coupled node pair is a dense 6x6 block),
//                  rounded once to FP32
//   clamping       nodes with k = 0: rows and columns removed, diagonal 1,
//                  RHS 0 (n stays 6 N^3)
//   RHS            U[-1e3, 1e3] per free DOF in DOF order, after all elements
//   blocks         block_starts = 0, 6, 12, ..., n (one block per node)
// nnz = 36 (free nodes + 2 free edges) + 6 clamped nodes.
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "hbj_internal.hpp"

namespace {

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) v = splitmix(x);
  }
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform(double a, double b) { return a + (b - a) * ((double)(next() >> 11) * 0x1.0p-53); }
};

struct Props {
  double E, G, A, Iy, Iz, J;
};

// Local 12x12 stiffness (x along the beam, bending about local z uses Iz,
// about local y uses Iy), then rotation to global axes.  Rows of R are the
// local axes in global coordinates; for the three lattice directions R is a
// signed permutation with det +1.
void element_global(const Props& p, double L, int dir, double* K) {
  double k[12][12];
  std::memset(k, 0, sizeof k);
  const double EA = p.E * p.A / L, GJ = p.G * p.J / L;
  const double z12 = 12 * p.E * p.Iz / (L * L * L), z6 = 6 * p.E * p.Iz / (L * L);
  const double z4 = 4 * p.E * p.Iz / L, z2 = 2 * p.E * p.Iz / L;
  const double y12 = 12 * p.E * p.Iy / (L * L * L), y6 = 6 * p.E * p.Iy / (L * L);
  const double y4 = 4 * p.E * p.Iy / L, y2 = 2 * p.E * p.Iy / L;
  auto set = [&](int i, int j, double v) {
    k[i][j] = v;
    k[j][i] = v;
  };
  set(0, 0, EA); set(0, 6, -EA); set(6, 6, EA);
  set(3, 3, GJ); set(3, 9, -GJ); set(9, 9, GJ);
  // v, rz (bending in the local x-y plane)
  set(1, 1, z12); set(1, 5, z6); set(1, 7, -z12); set(1, 11, z6);
  set(5, 5, z4); set(5, 7, -z6); set(5, 11, z2);
  set(7, 7, z12); set(7, 11, -z6); set(11, 11, z4);
  // w, ry (bending in the local x-z plane)
  set(2, 2, y12); set(2, 4, -y6); set(2, 8, -y12); set(2, 10, -y6);
  set(4, 4, y4); set(4, 8, y6); set(4, 10, y2);
  set(8, 8, y12); set(8, 10, y6); set(10, 10, y4);
  double R[3][3];
  std::memset(R, 0, sizeof R);
  if (dir == 0) {
    R[0][0] = 1; R[1][1] = 1; R[2][2] = 1;
  } else if (dir == 1) {
    R[0][1] = 1; R[1][0] = -1; R[2][2] = 1;
  } else {
    R[0][2] = 1; R[1][1] = 1; R[2][0] = -1;
  }
  double T[12][12];
  std::memset(T, 0, sizeof T);
  for (int b = 0; b < 4; ++b)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) T[3 * b + i][3 * b + j] = R[i][j];
  // K_global = T^T k T
  double tmp[12][12];
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j < 12; ++j) {
      double s = 0;
      for (int l = 0; l < 12; ++l) s += k[i][l] * T[l][j];
      tmp[i][j] = s;
    }
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j < 12; ++j) {
      double s = 0;
      for (int l = 0; l < 12; ++l) s += T[l][i] * tmp[l][j];
      K[i * 12 + j] = s;
    }
}

}  // namespace

extern "C" {

hbj_status hbj_beam_element_stiffness(double E, double G, double A, double Iy, double Iz,
                                      double J, double L, int dir, double* k144) {
  if (!k144 || dir < 0 || dir > 2 || !(L > 0)) {
    hbj::set_error("hbj_beam_element_stiffness: bad argument");
    return HBJ_E_INVALID;
  }
  element_global(Props{E, G, A, Iy, Iz, J}, L, dir, k144);
  return HBJ_OK;
}

hbj_status hbj_lattice_size(uint32_t N, uint32_t* n, uint64_t* nnz) {
  if (N < 1 || N > 1300) {
    hbj::set_error("hbj_lattice_size: N must be in [1, 1300], got %u", N);
    return HBJ_E_INVALID;
  }
  const uint64_t nodes = (uint64_t)N * N * N, clamped = (uint64_t)N * N;
  const uint64_t free_nodes = nodes - clamped;
  const uint64_t edges = 3ull * N * N * (N - 1);
  const uint64_t clamped_edges = 2ull * N * (N - 1) + (N > 1 ? (uint64_t)N * N : 0ull);
  const uint64_t free_edges = edges - clamped_edges;
  if (n) *n = (uint32_t)(6 * nodes);
  if (nnz) *nnz = 36ull * (free_nodes + 2 * free_edges) + 6ull * clamped;
  return HBJ_OK;
}

hbj_status hbj_lattice_generate(uint32_t N, uint64_t seed, double h, uint32_t* row_ptr,
                                uint32_t* col_idx, float* values, float* rhs,
                                uint32_t* block_starts) {
  uint32_t n;
  uint64_t nnz;
  if (hbj_lattice_size(N, &n, &nnz) != HBJ_OK) return HBJ_E_INVALID;
  if (nnz >= (1ull << 32)) {
    hbj::set_error("hbj_lattice_generate: nnz %llu exceeds 32-bit indices",
                   (unsigned long long)nnz);
    return HBJ_E_INVALID;
  }
  if (!row_ptr || !col_idx || !values || !rhs || !block_starts) {
    hbj::set_error("hbj_lattice_generate: null output array");
    return HBJ_E_INVALID;
  }
  const uint64_t nodes = (uint64_t)N * N * N;
  const double E0 = 2.1e11, G0 = 8.1e10, L = 1.0;
  // element properties, drawn sequentially; index 3*node + dir
  std::vector<Props> props(3 * nodes);
  std::vector<uint8_t> has(3 * nodes, 0);
  Rng rng(seed);
  for (uint32_t k = 0; k < N; ++k)
    for (uint32_t j = 0; j < N; ++j)
      for (uint32_t i = 0; i < N; ++i) {
        const uint64_t a = ((uint64_t)k * N + j) * N + i;
        const bool nb[3] = {i + 1 < N, j + 1 < N, k + 1 < N};
        for (int d = 0; d < 3; ++d) {
          if (!nb[d]) continue;
          Props p;
          p.A = rng.uniform(1e-3, 1e-2);
          p.Iy = rng.uniform(1e-6, 1e-5);
          p.Iz = rng.uniform(1e-6, 1e-5);
          p.J = rng.uniform(1e-6, 1e-5);
          double f = 1.0;
          if (h > 0) f = std::pow(10.0, rng.uniform(-h, h));
          p.E = E0 * f;
          p.G = G0 * f;
          props[3 * a + d] = p;
          has[3 * a + d] = 1;
        }
      }
  for (uint64_t a = 0; a < nodes; ++a) {
    const bool clamped = a < (uint64_t)N * N;
    for (int d = 0; d < 6; ++d) rhs[6 * a + d] = clamped ? 0.0f : (float)rng.uniform(-1e3, 1e3);
  }
  for (uint64_t a = 0; a <= nodes; ++a) block_starts[a] = (uint32_t)(6 * a);

  // row_ptr from the per-node neighbour counts
  const int64_t step[3] = {1, (int64_t)N, (int64_t)N * N};
  auto coord = [&](uint64_t a, uint32_t c[3]) {
    c[0] = a % N;
    c[1] = (a / N) % N;
    c[2] = a / ((uint64_t)N * N);
  };
  auto is_clamped = [&](uint64_t a) { return a < (uint64_t)N * N; };
  // neighbours of a in increasing node order: -z, -y, -x, self, +x, +y, +z
  auto neighbours = [&](uint64_t a, int64_t out[7], int out_dir[7], int out_side[7]) {
    uint32_t c[3];
    coord(a, c);
    int cnt = 0;
    for (int d = 2; d >= 0; --d)
      if (c[d] > 0) {
        out[cnt] = (int64_t)a - step[d];
        out_dir[cnt] = d;
        out_side[cnt] = 1;  // a is the element's second node
        ++cnt;
      }
    out[cnt] = (int64_t)a;
    out_dir[cnt] = -1;
    out_side[cnt] = 0;
    ++cnt;
    for (int d = 0; d < 3; ++d)
      if (c[d] + 1 < N) {
        out[cnt] = (int64_t)a + step[d];
        out_dir[cnt] = d;
        out_side[cnt] = 0;  // a is the element's first node
        ++cnt;
      }
    return cnt;
  };
  row_ptr[0] = 0;
  {
    uint64_t acc = 0;
    for (uint64_t a = 0; a < nodes; ++a) {
      uint32_t per_row;
      if (is_clamped(a)) {
        per_row = 1;
      } else {
        int64_t nbr[7];
        int nd[7], ns[7];
        const int cnt = neighbours(a, nbr, nd, ns);
        uint32_t free_cnt = 0;
        for (int q = 0; q < cnt; ++q) free_cnt += is_clamped((uint64_t)nbr[q]) ? 0 : 1;
        per_row = 6 * free_cnt;
      }
      for (int d = 0; d < 6; ++d) {
        acc += per_row;
        row_ptr[6 * a + d + 1] = (uint32_t)acc;
      }
    }
    if (acc != nnz) {
      hbj::set_error("hbj_lattice_generate: internal nnz mismatch %llu vs %llu",
                     (unsigned long long)acc, (unsigned long long)nnz);
      return HBJ_E_INVALID;
    }
  }

  auto assemble = [&](uint64_t a0, uint64_t a1) {
    double K[144];
    for (uint64_t a = a0; a < a1; ++a) {
      if (is_clamped(a)) {
        for (int d = 0; d < 6; ++d) {
          const uint32_t e = row_ptr[6 * a + d];
          col_idx[e] = (uint32_t)(6 * a + d);
          values[e] = 1.0f;
        }
        continue;
      }
      int64_t nbr[7];
      int nd[7], ns[7];
      const int cnt = neighbours(a, nbr, nd, ns);
      double blocks[7][36];
      std::memset(blocks, 0, sizeof blocks);
      int self = -1;
      for (int q = 0; q < cnt; ++q)
        if (nbr[q] == (int64_t)a) self = q;
      // every incident element adds to the diagonal block and (if the
      // neighbour is free) to the coupling block; summed in neighbour order
      for (int q = 0; q < cnt; ++q) {
        if (q == self) continue;
        const int d = nd[q];
        const uint64_t first = ns[q] ? (uint64_t)nbr[q] : a;  // element's first node
        element_global(props[3 * first + d], L, d, K);
        const int ra = ns[q] ? 6 : 0;   // a's rows in the element matrix
        const int rb = ns[q] ? 0 : 6;   // neighbour's columns
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 6; ++c) {
            blocks[self][r * 6 + c] += K[(ra + r) * 12 + (ra + c)];
            blocks[q][r * 6 + c] += K[(ra + r) * 12 + (rb + c)];
          }
      }
      for (int r = 0; r < 6; ++r) {
        uint32_t e = row_ptr[6 * a + r];
        for (int q = 0; q < cnt; ++q) {
          if (is_clamped((uint64_t)nbr[q])) continue;
          for (int c = 0; c < 6; ++c) {
            col_idx[e] = (uint32_t)(6 * nbr[q] + c);
            values[e] = (float)blocks[q][r * 6 + c];
            ++e;
          }
        }
      }
    }
  };
  unsigned nt = nodes >= 8192 ? std::thread::hardware_concurrency() : 1u;
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  if (nt == 1) {
    assemble(0, nodes);
  } else {
    std::vector<std::thread> th;
    const uint64_t per = (nodes + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
      const uint64_t a0 = t * per, a1 = a0 + per < nodes ? a0 + per : nodes;
      if (a0 < a1) th.emplace_back(assemble, a0, a1);
    }
    for (auto& x : th) x.join();
  }
  return HBJ_OK;
}

}  // extern "C"
