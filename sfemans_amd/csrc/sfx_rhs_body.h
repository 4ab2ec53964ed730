// Right-hand-side assembly from Gauss-point values (body for gfx950 and the
// CPU emulator).
//
// rhs_3x3 (SOURCES/SOLVER/assembling_rhs.F90:5-126, without the tensor
// options) and qs_00_gauss (SOURCES/SOLVER/fem_rhs_axi.F90:940-987) scatter
// per-element contributions to the nodes with VecSetValues(ADD_VALUES), one
// element after the other.  Here each output entry is gathered by one thread
// from the node's (element, local node) list in element order, so the sums
// are formed in the reference's order, deterministically and without atomics.
//
//   wgt(e,l)     = rj(l,e) * ray(l,e)                 (assembling_rhs.F90:73)
//   vb(n, ki, m) = sum_(e,i) sum_l ww(i,l) wgt(e,l) rhs(7e+l, comp(ki), m)
// with comp = (1,4,5) for vb_145 and (2,3,6) for vb_236 (:75-95).
#pragma once
#include "sfx_portable.h"

namespace sfx {

struct RhsArgs {
  const int64_t* adj_ptr;   // (np+1)
  const int32_t* adj_ent;   // e*6 + i
  const double* wgt;        // (7, me): rj * ray
  const double* tab;        // ww(6,7), ...
  int64_t me, np, G;
  int M;                    // modes (or columns for the scalar assembly)
  double scale;             // multiplies the result (-1 for the -rotv_v share)
  // rhs_3x3
  const double* rhs;        // (G, 6, M)
  double* v145;             // (np, 3, M)
  double* v236;             // (np, 3, M)
  // qs_00_gauss (ncols independent right-hand sides)
  const int32_t* jj;        // (6, me)
  const double* heat_capa;  // (me)
  const double* ff;         // (np, M) nodal source
  const double* ffg;        // (G, M) Gauss source
  double* vect;             // (np, M)
  int64_t n_rhs, n_ff, n_ffg, n_out, n_adj;
};

// One thread per (node, mode): all six components.
SFX_DEV void rhs3x3_body(const RhsArgs& a, int64_t gid) {
  const int64_t n = gid % a.np;
  const int64_t m = gid / a.np;
  if (m >= a.M) return;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int64_t b = a.adj_ptr[n], e_ = a.adj_ptr[n + 1];
  for (int64_t q = b; q < e_; ++q) {
    SFX_CHECK("adj", q, a.n_adj);
    const int32_t ent = a.adj_ent[q];
    const int64_t e = ent / 6;
    const int i = ent % 6;
    double loc[6] = {0, 0, 0, 0, 0, 0};
    for (int l = 0; l < 7; ++l) {
      const double c = a.tab[i + 6 * l] * a.wgt[7 * e + l];
      const int64_t gi = 7 * e + l + a.G * 6 * m;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        SFX_CHECK("rhs load", gi + a.G * k, a.n_rhs);
        loc[k] = fma(c, a.rhs[gi + a.G * k], loc[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) acc[k] += loc[k];
  }
  const int64_t o = n + a.np * 3 * m;
  SFX_CHECK("v145 store", o + 2 * a.np, a.n_out);
  a.v145[o] = a.scale * acc[0];
  a.v145[o + a.np] = a.scale * acc[3];
  a.v145[o + 2 * a.np] = a.scale * acc[4];
  a.v236[o] = a.scale * acc[1];
  a.v236[o + a.np] = a.scale * acc[2];
  a.v236[o + 2 * a.np] = a.scale * acc[5];
}

// One thread per (node, column): fl = (hc(e) SUM(ff_loc*ww(:,l)) + ffg(7e+l)) * rj * ray.
SFX_DEV void qs00_body(const RhsArgs& a, int64_t gid) {
  const int64_t n = gid % a.np;
  const int64_t c = gid / a.np;
  if (c >= a.M) return;
  double acc = 0;
  const int64_t b = a.adj_ptr[n], e_ = a.adj_ptr[n + 1];
  for (int64_t q = b; q < e_; ++q) {
    SFX_CHECK("adj", q, a.n_adj);
    const int32_t ent = a.adj_ent[q];
    const int64_t e = ent / 6;
    const int i = ent % 6;
    double ffl[6];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const int64_t fi = (int64_t)a.jj[6 * e + j] - 1 + a.np * c;
      SFX_CHECK("ff load", fi, a.n_ff);
      ffl[j] = a.ff[fi];
    }
    const double hc = a.heat_capa[e];
    double loc = 0;
    for (int l = 0; l < 7; ++l) {
      double s = 0;
#pragma unroll
      for (int j = 0; j < 6; ++j) s = fma(ffl[j], a.tab[j + 6 * l], s);
      const int64_t gi = 7 * e + l + a.G * c;
      SFX_CHECK("ffg load", gi, a.n_ffg);
      const double fl = (hc * s + a.ffg[gi]) * a.wgt[7 * e + l];
      loc = fma(a.tab[i + 6 * l], fl, loc);
    }
    acc += loc;
  }
  SFX_CHECK("vect store", n + a.np * c, a.n_out);
  a.vect[n + a.np * c] = a.scale * acc;
}

}  // namespace sfx
