// Host-side tables of the boundary term of rhs_3x3 (library and CPU emulator).
#pragma once
#include <cmath>
#include <cstdint>
#include <vector>

namespace sfx {

// Quadratic edge element with three Gauss points (SOURCES/GAUSS_POINTS/GP_2D_p2.F90:176-212):
// shape values w[3*ls + ns] and derivatives d[3*ls + ns] of the end nodes
// (x = -1, +1) and the midpoint (x = 0) at x = -sqrt(3/5), 0, sqrt(3/5);
// weights p = 5/9, 8/9, 5/9.
inline void p2_edge_tables(double* w, double* d, double* p) {
  const double xx[3] = {-std::sqrt(3.0 / 5.0), 0.0, std::sqrt(3.0 / 5.0)};
  for (int ls = 0; ls < 3; ++ls) {
    const double x = xx[ls];
    w[3 * ls + 0] = (x - 1) * x / 2;  d[3 * ls + 0] = (2 * x - 1) / 2;
    w[3 * ls + 1] = (x + 1) * x / 2;  d[3 * ls + 1] = (2 * x + 1) / 2;
    w[3 * ls + 2] = (x + 1) * (1 - x); d[3 * ls + 2] = -2 * x;
  }
  p[0] = p[2] = 5.0 / 9.0;
  p[1] = 8.0 / 9.0;
}

// Boundary sides jjs(3, mes) (1-based: two ends, then the midpoint) ->
// the boundary nodes in order of first appearance, each node's
// (side, local node) entries in side order (CSR), and per side Gauss point
// the weight ray * rjs, rjs = |d rr/ds| p (gauss_points_2d.F90:182-190), or
// 0 where ray < 1e-10 (assembling_rhs.F90:156-157 skips those points).
struct BdyTables {
  std::vector<int64_t> ptr, node;
  std::vector<int32_t> ent;
  std::vector<double> bw;
};

inline void build_boundary_tables(const int32_t* jjs, int64_t mes, const double* rr, int64_t np, BdyTables& t) {
  double w[9], d[9], p[3];
  p2_edge_tables(w, d, p);
  std::vector<int64_t> idx(np, -1);
  std::vector<int64_t> cnt;
  for (int64_t s = 0; s < mes; ++s)
    for (int ns = 0; ns < 3; ++ns) {
      const int64_t n = jjs[3 * s + ns] - 1;
      if (idx[n] < 0) {
        idx[n] = (int64_t)t.node.size();
        t.node.push_back(n);
        cnt.push_back(0);
      }
      cnt[idx[n]] += 1;
    }
  const int64_t nb = (int64_t)t.node.size();
  t.ptr.assign(nb + 1, 0);
  for (int64_t b = 0; b < nb; ++b) t.ptr[b + 1] = t.ptr[b] + cnt[b];
  t.ent.assign(t.ptr[nb], 0);
  std::vector<int64_t> fill(t.ptr.begin(), t.ptr.end() - 1);
  for (int64_t s = 0; s < mes; ++s)
    for (int ns = 0; ns < 3; ++ns) t.ent[fill[idx[jjs[3 * s + ns] - 1]]++] = (int32_t)(3 * s + ns);
  t.bw.assign(3 * mes, 0.0);
  for (int64_t s = 0; s < mes; ++s)
    for (int ls = 0; ls < 3; ++ls) {
      double ray = 0, dr = 0, dz = 0;
      for (int ns = 0; ns < 3; ++ns) {
        const int64_t n = jjs[3 * s + ns] - 1;
        ray += rr[2 * n] * w[3 * ls + ns];
        dr += rr[2 * n] * d[3 * ls + ns];
        dz += rr[2 * n + 1] * d[3 * ls + ns];
      }
      t.bw[3 * s + ls] = ray < 1e-10 ? 0.0 : ray * std::sqrt(dr * dr + dz * dz) * p[ls];
    }
}

}  // namespace sfx
