// Right-hand-side assembly from Gauss-point values (body for gfx950 and the
// CPU emulator).
//
// rhs_3x3 (SOURCES/SOLVER/assembling_rhs.F90:5-126, without the tensor
// options) and qs_00_gauss (SOURCES/SOLVER/fem_rhs_axi.F90:940-987) scatter
// per-element contributions to the nodes with VecSetValues(ADD_VALUES), one
// element after the other.  Here each output entry is gathered by one thread
// from the node's (element, local node) list in element order, so the sums
// are formed in the reference's order, deterministically and without atomics.
// For rhs_3x3 the element contributions are formed first, one thread per
// (element, mode) reading the Gauss array once, into a work array.
//
//   wgt(e,l)     = rj(l,e) * ray(l,e)                 (assembling_rhs.F90:73)
//   vb(n, ki, m) = sum_(e,i) sum_l ww(i,l) wgt(e,l) rhs(7e+l, comp(ki), m)
// with comp = (1,4,5) for vb_145 and (2,3,6) for vb_236 (:75-95).
#pragma once
#include "sfx_gather.h"
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
  double* elo;              // per-element nodal contributions, index ((6e+i)*6 + k)*M + m
  int64_t n_rhs, n_ff, n_ffg, n_out, n_adj, n_elo;
};

// Work item gid of the element kernels -> (element e, mode or column m).  A
// wave takes ELEM_EB consecutive elements x ELEM_MB consecutive modes, modes
// fastest: a load of one Gauss plane then touches the 7 * ELEM_EB
// consecutive points of those elements (two cache lines) for ELEM_MB planes
// instead of 64 planes, and the stores still write ELEM_MB consecutive modes
// (one whole line) of each entry.
constexpr int ELEM_EB = 4, ELEM_MB = 16;
SFX_DEV void elem_item(int64_t gid, int M, int64_t& e, int& m) {
  const int64_t w = gid / 64;
  const int lane = (int)(gid % 64);
  const int nmb = (M + ELEM_MB - 1) / ELEM_MB;
  e = (w / nmb) * ELEM_EB + lane / ELEM_MB;
  m = (int)(w % nmb) * ELEM_MB + lane % ELEM_MB;
}
inline int64_t elem_items(int64_t me, int M) {
  return ((me + ELEM_EB - 1) / ELEM_EB) * (int64_t)((M + ELEM_MB - 1) / ELEM_MB) * 64;
}

// Element share of rhs_3x3, one thread per (element, mode) (elem_item):
//   elo(6e+i, k, m) = sum_l ww(i,l) wgt(e,l) rhs(7e+l, comp k, m)
// (the values the reference adds with VecSetValues, assembling_rhs.F90:73-95).
// A thread reads the 7 consecutive Gauss values of its element per
// component; the stores of a wave cover consecutive modes.  rhs_gather_block
// then sums the entries of each node in element order.
SFX_DEV void rhs_elem_body(const RhsArgs& a, int64_t gid) {
  int64_t e;
  int m;
  elem_item(gid, a.M, e, m);
  if (e >= a.me || m >= a.M) return;
  double w[7], r[7][6];
#pragma unroll
  for (int l = 0; l < 7; ++l) {
    w[l] = a.wgt[7 * e + l];
    const int64_t gi = 7 * e + l + a.G * 6 * m;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      SFX_CHECK("rhs load", gi + a.G * k, a.n_rhs);
      r[l][k] = a.rhs[gi + a.G * k];
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double loc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int l = 0; l < 7; ++l) {
      const double c = a.tab[i + 6 * l] * w[l];
#pragma unroll
      for (int k = 0; k < 6; ++k) loc[k] = fma(c, r[l][k], loc[k]);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int64_t o = ((6 * e + i) * 6 + k) * a.M + m;
      SFX_CHECK("elo store", o, a.n_elo);
      a.elo[o] = loc[k];
    }
  }
}

// vb_145 / vb_236 of work-group bid: the nodes' element contributions summed
// in element order (sfx_gather.h), components (1,4,5) and (2,3,6), times scale.
SFX_DEV void rhs_gather_block(const RhsArgs& a, int64_t bid, int tid, double* tile) {
  gather_body<6>(a.adj_ptr, a.adj_ent, a.n_adj, a.elo, a.n_elo, a.np, a.M, bid, tid, tile,
                 [&](int k, int m, int64_t n, double v) {
                   // component k -> (array, column): 1,4,5 -> vb_145; 2,3,6 -> vb_236
                   const bool v145 = (k == 0 || k == 3 || k == 4);
                   const int col = (k == 0 || k == 1) ? 0 : (k == 2 || k == 3) ? 1 : 2;
                   const int64_t o = n + a.np * (col + 3 * (int64_t)m);
                   SFX_CHECK("vb store", o, a.n_out);
                   (v145 ? a.v145 : a.v236)[o] = a.scale * v;
                 });
}

// Element share of qs_00_gauss, one thread per (element, column) (elem_item):
//   fl(l)       = (hc(e) SUM(ff_loc * ww(:,l)) + ff_gauss(7e+l)) * rj * ray
//   elo(6e+i,c) = sum_l ww(i,l) fl(l)                    (fem_rhs_axi.F90:958-975)
SFX_DEV void qs00_elem_body(const RhsArgs& a, int64_t gid) {
  int64_t e;
  int c;
  elem_item(gid, a.M, e, c);
  if (e >= a.me || c >= a.M) return;
  double ffl[6];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const int64_t fi = (int64_t)a.jj[6 * e + j] - 1 + a.np * (int64_t)c;
    SFX_CHECK("ff load", fi, a.n_ff);
    ffl[j] = a.ff[fi];
  }
  const double hc = a.heat_capa[e];
  double fl[7];
#pragma unroll
  for (int l = 0; l < 7; ++l) {
    double s = 0;
#pragma unroll
    for (int j = 0; j < 6; ++j) s = fma(ffl[j], a.tab[j + 6 * l], s);
    const int64_t gi = 7 * e + l + a.G * (int64_t)c;
    SFX_CHECK("ffg load", gi, a.n_ffg);
    fl[l] = (hc * s + a.ffg[gi]) * a.wgt[7 * e + l];
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double loc = 0;
#pragma unroll
    for (int l = 0; l < 7; ++l) loc = fma(a.tab[i + 6 * l], fl[l], loc);
    const int64_t o = (6 * e + i) * a.M + c;
    SFX_CHECK("elo store", o, a.n_elo);
    a.elo[o] = loc;
  }
}

// vect(n, c) of work-group bid: the nodes' element contributions summed in
// element order (sfx_gather.h).
SFX_DEV void qs00_gather_block(const RhsArgs& a, int64_t bid, int tid, double* tile) {
  gather_body<1>(a.adj_ptr, a.adj_ent, a.n_adj, a.elo, a.n_elo, a.np, a.M, bid, tid, tile,
                 [&](int, int c, int64_t n, double v) {
                   SFX_CHECK("vect store", n + a.np * c, a.n_out);
                   a.vect[n + a.np * c] = a.scale * v;
                 });
}

}  // namespace sfx
