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
#include "sfx_geom.h"
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
  // full NS right-hand side (rhs_ns_elem_body); a null field counts as zero
  const double* rr;         // (2, np) node coordinates
  const int32_t* lmode;     // (M) mode numbers
  const double* v1m;        // (np, 6, M) time-derivative term
  const double* fsrc;       // (np, 6, M) nodal momentum source
  const double* pres;       // (np, 2, M) pressure on the velocity mesh
  const double* pinc;       // (np, 2, M) pressure increment on the velocity mesh
  const double* tens;       // (3, G, 6, M) opt_tensor of rhs_3x3
  double rsign;             // sign of the Gauss field rhs: -1 for -rotv_v, +1 for rhs_3x3's rhs_in
  // velocity imposed in solids by penalty (computing_rhs_gauss.F90:103-121):
  // isign * SUM(imp*ww) added to the mode slot imp_slot (-1: none)
  const double* imp;        // (np, 6) 3 u_imposed / (2 dt)
  int imp_slot;
  double isign;
  double* gout;             // (G, 6, M): with the Gauss output, rhs without fp goes here
  int64_t n_v1m, n_fsrc, n_pres, n_pinc, n_tens, n_imp, n_gout;
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

// Full explicit NS right-hand side of rhs_ns_gauss_3x3 without the penalty
// options (SOURCES/SOLVER/computing_rhs_gauss.F90:7-146) fused with the
// element share of rhs_3x3, one thread per (element, mode) (elem_item):
//   rhs(l, :) = (ft + fp_inc + fs - rotv_v(7e+l, :)) + fp         (:100, :145)
// (rsign = -1; with rsign = +1 and the other fields absent it is rhs_3x3's
// own rhs_in, so the kernel also serves rhs_3x3 with opt_tensor)
//   ft = SUM(V1m*ww), fs = SUM(ff*ww),
//   fp = (-d_r P_c, -d_r P_s, -P_s/ray*m, P_c/ray*m, -d_z P_c, -d_z P_s) (:69-97)
// with P and P_inc the pressure injected onto the velocity mesh; then
//   elo(6e+i, k, m) = sum_l ww(i,l) wgt(e,l) rhs(l, k)   as rhs_elem_body.
// With TENSOR, rhs_3x3's opt_tensor T(3, G, 6, M) (assembling_rhs.F90:56-121):
// node i of component k adds, per Gauss point,
//   (ww(i,l) f_tensor(k) + d_r phi_i T(k/2, k%2) + d_z phi_i T(k/2, 4 + k%2)) wgt(e,l)
// with f_tensor the 1/r and mode terms (row k/2 of T, its cos columns for
// even k, sin for odd).
// The Gauss array rhs_gauss is never stored.
// With GOUT the kernel writes the Gauss values (ft + fp_inc + fs - rotv_v,
// minus the imposed velocity on its slot) to gout instead of assembling them:
// the first step of the penalty branch, whose product with the penalty
// function mixes the modes at each Gauss point (computing_rhs_gauss.F90:124-129).
template <bool TENSOR, bool GOUT = false>
SFX_DEV void rhs_ns_elem_body(const RhsArgs& a, int64_t gid) {
  int64_t e;
  int m;
  elem_item(gid, a.M, e, m);
  if (e >= a.me || m >= a.M) return;
  int64_t j[6];
  double rn[6], zn[6];
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    j[n] = (int64_t)a.jj[6 * e + n] - 1;
    SFX_CHECK("rr load", 2 * j[n] + 1, 2 * a.np);
    rn[n] = a.rr[2 * j[n]];
    zn[n] = a.rr[2 * j[n] + 1];
  }
  const double mv = (double)a.lmode[m];
  double V[6][6], F[6][6], P[6][2], Q[6][2];
#pragma unroll
  for (int n = 0; n < 6; ++n) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int64_t i6 = j[n] + a.np * (k + 6 * (int64_t)m);
      SFX_CHECK("V1m load", i6, a.v1m ? a.n_v1m : 1 << 30);
      SFX_CHECK("ff load", i6, a.fsrc ? a.n_fsrc : 1 << 30);
      V[n][k] = a.v1m ? a.v1m[i6] : 0.0;
      F[n][k] = a.fsrc ? a.fsrc[i6] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int64_t i2 = j[n] + a.np * (k + 2 * (int64_t)m);
      SFX_CHECK("P load", i2, a.pres ? a.n_pres : 1 << 30);
      SFX_CHECK("P_inc load", i2, a.pinc ? a.n_pinc : 1 << 30);
      P[n][k] = a.pres ? a.pres[i2] : 0.0;
      Q[n][k] = a.pinc ? a.pinc[i2] : 0.0;
    }
  }
  const bool imposed = a.imp && m == a.imp_slot;
  double loc[6][6];
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int k = 0; k < 6; ++k) loc[i][k] = 0.0;
#pragma unroll
  for (int l = 0; l < 7; ++l) {
    const double* w = a.tab + 6 * l;
    const double* d = a.tab + 42 + 12 * l;
    // isoparametric Jacobian at this Gauss point: d/dr = a11 d/dx + a12 d/dy,
    // d/dz = a21 d/dx + a22 d/dy (exact for curved elements)
    double a11, a12, a21, a22;
    p2_inverse_jacobian(rn, zn, d, a11, a12, a21, a22);
    double ray = 0, ft[6] = {0, 0, 0, 0, 0, 0}, fs[6] = {0, 0, 0, 0, 0, 0};
    double pv[2] = {0, 0}, px[2] = {0, 0}, py[2] = {0, 0}, qv[2] = {0, 0}, qx[2] = {0, 0}, qy[2] = {0, 0};
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      ray = fma(rn[n], w[n], ray);
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        ft[k] = fma(V[n][k], w[n], ft[k]);
        fs[k] = fma(F[n][k], w[n], fs[k]);
      }
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        pv[k] = fma(P[n][k], w[n], pv[k]);
        px[k] = fma(P[n][k], d[2 * n], px[k]);
        py[k] = fma(P[n][k], d[2 * n + 1], py[k]);
        qv[k] = fma(Q[n][k], w[n], qv[k]);
        qx[k] = fma(Q[n][k], d[2 * n], qx[k]);
        qy[k] = fma(Q[n][k], d[2 * n + 1], qy[k]);
      }
    }
    double fp[6], fq[6];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      fp[k] = -fma(a11, px[k], a12 * py[k]);
      fp[4 + k] = -fma(a21, px[k], a22 * py[k]);
      fq[k] = -fma(a11, qx[k], a12 * qy[k]);
      fq[4 + k] = -fma(a21, qx[k], a22 * qy[k]);
    }
    fp[2] = -pv[1] / ray * mv;
    fp[3] = pv[0] / ray * mv;
    fq[2] = -qv[1] / ray * mv;
    fq[3] = qv[0] / ray * mv;
    const double wl = a.wgt[7 * e + l];
    const int64_t gi = 7 * e + l + a.G * 6 * (int64_t)m;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      SFX_CHECK("rotv load", gi + a.G * k, a.rhs ? a.n_rhs : 1 << 30);
      const double rv = a.rhs ? a.rhs[gi + a.G * k] : 0.0;
      double base = ft[k] + fq[k] + fs[k] + a.rsign * rv;
      if (imposed) {
        double iv = 0;
#pragma unroll
        for (int n = 0; n < 6; ++n) {
          const int64_t jn = (int64_t)a.jj[6 * e + n] - 1;
          SFX_CHECK("imposed velocity load", jn + a.np * k, a.n_imp);
          iv = fma(a.imp[jn + a.np * k], w[n], iv);
        }
        base = base + a.isign * iv;
      }
      if constexpr (GOUT) {
        SFX_CHECK("gauss store", gi + a.G * k, a.n_gout);
        a.gout[gi + a.G * k] = base;
      } else {
        const double r = (base + fp[k]) * wl;
#pragma unroll
        for (int i = 0; i < 6; ++i) loc[i][k] = fma(w[i], r, loc[i][k]);
      }
    }
  }
  if constexpr (GOUT) return;
  // The tensor terms in a second pass over the Gauss points, once the nodal
  // fields are dead: the two passes never hold both in registers.
  if constexpr (TENSOR) {
    // the node coordinates again (from L2), so that they are not held over
    // the first pass: the opaque pointer keeps these loads from being merged
    // with the first ones
    const double* rr2 = a.rr;
    const int32_t* jj2 = a.jj;
    SFX_OPAQUE_S(rr2);
    SFX_OPAQUE_S(jj2);
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      const int64_t jn = (int64_t)jj2[6 * e + n] - 1;
      rn[n] = rr2[2 * jn];
      zn[n] = rr2[2 * jn + 1];
    }
#pragma unroll 1
    for (int l = 0; l < 7; ++l) {
      const double* w = a.tab + 6 * l;
      const double* d = a.tab + 42 + 12 * l;
      double a11, a12, a21, a22;
      p2_inverse_jacobian(rn, zn, d, a11, a12, a21, a22);
      double ray = 0;
#pragma unroll
      for (int n = 0; n < 6; ++n) ray = fma(rn[n], w[n], ray);
      const double wl = a.wgt[7 * e + l];
      const int64_t gi = 7 * e + l + a.G * 6 * (int64_t)m;
      double T[3][6];
#pragma unroll
      for (int t = 0; t < 6; ++t)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const int64_t it = k + 3 * (gi + a.G * t);
          SFX_CHECK("tensor load", it, a.n_tens);
          T[k][t] = a.tens[it];
        }
      double fT[6];
      fT[0] = (-mv * T[0][3] + T[1][2]) / ray;
      fT[1] = (mv * T[0][2] + T[1][3]) / ray;
      fT[2] = (-T[0][2] - mv * T[1][3]) / ray;
      fT[3] = (-T[0][3] + mv * T[1][2]) / ray;
      fT[4] = -T[2][3] * mv / ray;
      fT[5] = T[2][2] * mv / ray;
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        const double dr = fma(a11, d[2 * i], a12 * d[2 * i + 1]) * wl;
        const double dz = fma(a21, d[2 * i], a22 * d[2 * i + 1]) * wl;
        const double wi = w[i] * wl;
#pragma unroll
        for (int k = 0; k < 6; ++k)
          loc[i][k] = fma(wi, fT[k], fma(dr, T[k / 2][k % 2], fma(dz, T[k / 2][4 + k % 2], loc[i][k])));
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const int64_t o = ((6 * e + i) * 6 + k) * a.M + m;
      SFX_CHECK("elo store", o, a.n_elo);
      a.elo[o] = loc[i][k];
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
    SFX_CHECK("ff load", fi, a.ff ? a.n_ff : 1 << 30);
    ffl[j] = a.ff ? a.ff[fi] : 0.0;
  }
  // a null heat capacity is 1; a null nodal or Gauss source is zero (the mass
  // matrix applied to ff, or the Galerkin right-hand side of ff_gauss alone)
  const double hc = a.heat_capa ? a.heat_capa[e] : 1.0;
  double fl[7];
#pragma unroll
  for (int l = 0; l < 7; ++l) {
    double s = 0;
#pragma unroll
    for (int j = 0; j < 6; ++j) s = fma(ffl[j], a.tab[j + 6 * l], s);
    const int64_t gi = 7 * e + l + a.G * (int64_t)c;
    SFX_CHECK("ffg load", gi, a.ffg ? a.n_ffg : 1 << 30);
    fl[l] = (hc * s + (a.ffg ? a.ffg[gi] : 0.0)) * a.wgt[7 * e + l];
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

// ---- boundary term of rhs_3x3 (opt_tensor_scaln_bdy, assembling_rhs.F90:129-185)
// For every boundary side ms (mesh%jjs: two ends, then the midpoint) and its
// three Gauss points, smb = T_bdy(3 ms + ls, :) * ray * rjs (points on the
// axis, ray < 1e-10, skipped), added to the side's nodes with the edge shape
// values wws and the (1,4,5)/(2,3,6) packing of the volume term.  Each
// boundary node is gathered by one thread from its (side, local node) list
// in side order, after the volume sums (the reference's VecSetValues order:
// elements first, then sides).
struct BdyArgs {
  const int64_t* bptr;      // (nb+1) offsets of each boundary node's entries
  const int32_t* bent;      // 3*ms + local node, side order
  const int64_t* bnode;     // (nb) node index (0-based)
  const double* bw;         // (3, mes): ray * rjs per side Gauss point, 0 where skipped
  double wws[9];            // wws(ns, ls) at 3*ls + ns
  int64_t nb, mes, np;
  int M;
  double scale;
  const double* tb;         // (3*mes, 6, M)
  double* v145;             // (np, 3, M)
  double* v236;             // (np, 3, M)
  int64_t n_tb, n_out, n_ent;
};

// One thread per (boundary node, mode), nodes fastest.
SFX_DEV void rhs_bdy_body(const BdyArgs& a, int64_t gid) {
  if (a.nb <= 0) return;
  const int64_t b = gid % a.nb, m = gid / a.nb;
  if (m >= a.M) return;
  const int64_t n = a.bnode[b];
  // component k -> (array, column): 1,4,5 -> vb_145 columns 1,2,3; 2,3,6 -> vb_236 columns 1,2,3
  int64_t o[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int col = (k == 0 || k == 1) ? 0 : (k == 2 || k == 3) ? 1 : 2;
    o[k] = n + a.np * (col + 3 * m);
    SFX_CHECK("vb bdy", o[k], a.n_out);
  }
  double acc[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const bool v145 = (k == 0 || k == 3 || k == 4);
    acc[k] = (v145 ? a.v145 : a.v236)[o[k]];
  }
  for (int64_t t = a.bptr[b]; t < a.bptr[b + 1]; ++t) {
    SFX_CHECK("bdy entry", t, a.n_ent);
    const int32_t en = a.bent[t];
    const int64_t ms = en / 3;
    const int ns = en % 3;
    double side[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int ls = 0; ls < 3; ++ls) {
      const double wl = a.bw[3 * ms + ls];
      if (wl == 0.0) continue;                      // a point on the axis (ray < 1e-10)
      const double c = a.wws[3 * ls + ns] * wl;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const int64_t it = 3 * ms + ls + 3 * a.mes * (k + 6 * m);
        SFX_CHECK("T_bdy load", it, a.n_tb);
        side[k] = fma(c, a.tb[it], side[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) acc[k] = fma(a.scale, side[k], acc[k]);
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const bool v145 = (k == 0 || k == 3 || k == 4);
    (v145 ? a.v145 : a.v236)[o[k]] = acc[k];
  }
}

}  // namespace sfx
