// The inner product of SFEMaNS fields, the r-weighted mass matrix and its
// solve, column by column (body for gfx950 and the CPU emulator).
//
// dot_product_SF (SOURCES/SOLVER/tn_axi.F90:9-37, fem_tn_ns_mhd.F90:92-135,
// fem_tn_axi.F90:8-48): 2 pi SUM_k w_k SUM_c SUM_(e,l) v_g w_g rj ray with
// v_g = SUM(v(jj(:,e),c,k) ww(:,l)), w_k = 1 for mode 0 and 1/2 otherwise.
// The nodal arrays (np, nc, M) are read as ncols = nc*M columns; column c
// belongs to mode slot c / nc.
#pragma once
#include "sfx_portable.h"
#include "sfx_rhs_body.h"

namespace sfx {

constexpr int INNER_THREADS = 256;

struct InnerArgs {
  const int32_t* jj;        // (6, me)
  const double* tab;        // ww(6,7), dd(2,6,7)
  const double* wgt;        // (7, me): rj * ray
  const int32_t* lmode;     // (M) mode numbers
  int nc;                   // columns per mode slot
  int ncols;                // nc * M
  int64_t me, np, G;
  double w0, wm;            // column scale: w0 for mode 0, wm otherwise
  int weighted;             // gauss values times rj * ray
  const double* z;          // (np, ncols) nodal input
  double* out;              // (G, ncols) Gauss output, or (np, ncols) nodal output
  const double* v;          // dot product: (np, ncols) each
  const double* w;
  double* partial;          // per-work-group partial sums
  int64_t n_z, n_out, n_part;
};

SFX_DEV double col_scale(const InnerArgs& a, int c) { return a.lmode[c / a.nc] == 0 ? a.w0 : a.wm; }

// Gauss values of nodal columns, one thread per (element, column) (elem_item):
//   out(7e+l, c) = scale(c) * [rj ray](e,l) * SUM_n ww(n,l) z(jj(n,e), c)
// (R^T z when weighted; the plain Gauss values otherwise).
SFX_DEV void gauss_values_body(const InnerArgs& a, int64_t gid) {
  int64_t e;
  int c;
  elem_item(gid, a.ncols, e, c);
  if (e >= a.me || c >= a.ncols) return;
  double zn[6];
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    const int64_t i = (int64_t)a.jj[6 * e + n] - 1 + a.np * (int64_t)c;
    SFX_CHECK("z load", i, a.n_z);
    zn[n] = a.z[i];
  }
  const double sc = col_scale(a, c);
#pragma unroll
  for (int l = 0; l < 7; ++l) {
    double s = 0;
#pragma unroll
    for (int n = 0; n < 6; ++n) s = fma(a.tab[n + 6 * l], zn[n], s);
    const double wl = a.weighted ? a.wgt[7 * e + l] : 1.0;
    const int64_t o = 7 * e + l + a.G * (int64_t)c;
    SFX_CHECK("gauss store", o, a.n_out);
    a.out[o] = sc * (s * wl);
  }
}

// Tree sum of the INNER_THREADS values in sh (fixed order: deterministic).
SFX_DEV double block_sum(double* sh, int tid) {
  for (int s = INNER_THREADS / 2; s > 0; s >>= 1) {
    SFX_SYNC();
    if (tid < s) sh[tid] += sh[tid + s];
  }
  SFX_SYNC();
  return sh[0];
}

// At most this many work-groups (and partial sums) per dot product.
constexpr int DOTSF_MAX_BLOCKS = 2048;

// Work-group bid of nblk: its threads take the (element, column) items
// (bid + j*nblk)*INNER_THREADS + tid for j = 0, 1, ..., each adding
// SUM_l v_g w_g rj ray times the mode weight in that order; the group's sum
// goes to partial[bid].  The split depends only on the sizes, so the sum is
// formed in the same order on every call.
SFX_DEV double dotsf_item(const InnerArgs& a, int64_t gid);
SFX_DEV void dotsf_block(const InnerArgs& a, int64_t bid, int64_t nblk, int64_t nitems, int tid, double* sh) {
  double t = 0;
  for (int64_t g = bid * INNER_THREADS + tid; g < nitems; g += nblk * INNER_THREADS) t += dotsf_item(a, g);
  sh[tid] = t;
  const double s = block_sum(sh, tid);
  if (tid == 0) {
    SFX_CHECK("partial store", bid, a.n_part);
    a.partial[bid] = s;
  }
}

SFX_DEV double dotsf_item(const InnerArgs& a, int64_t gid) {
  int64_t e;
  int c;
  elem_item(gid, a.ncols, e, c);
  double t = 0;
  if (e < a.me && c < a.ncols) {
    double vn[6], wn[6];
#pragma unroll
    for (int n = 0; n < 6; ++n) {
      const int64_t i = (int64_t)a.jj[6 * e + n] - 1 + a.np * (int64_t)c;
      SFX_CHECK("v load", i, a.n_z);
      vn[n] = a.v[i];
      wn[n] = a.w[i];
    }
#pragma unroll
    for (int l = 0; l < 7; ++l) {
      double vg = 0, wg = 0;
#pragma unroll
      for (int n = 0; n < 6; ++n) {
        vg = fma(a.tab[n + 6 * l], vn[n], vg);
        wg = fma(a.tab[n + 6 * l], wn[n], wg);
      }
      t = fma(vg * wg, a.wgt[7 * e + l], t);
    }
    t *= col_scale(a, c);
  }
  return t;
}

// One work-group: out[0] = scale * SUM of the n partials, summed in a fixed order.
SFX_DEV void sum_partials_block(const double* partial, int64_t n, double scale, double* out, int tid, double* sh) {
  double t = 0;
  for (int64_t i = tid; i < n; i += INNER_THREADS) t += partial[i];
  sh[tid] = t;
  const double s = block_sum(sh, tid);
  if (tid == 0) out[0] = scale * s;
}

// Column dot products over the np rows, one work-group per column:
// out[c] = SUM_i A(i,c) B(i,c).
SFX_DEV void coldot_block(const double* A, const double* B, int64_t np, double* out, int64_t c, int tid,
                          double* sh) {
  double t = 0;
  const double* a = A + np * c;
  const double* b = B + np * c;
  for (int64_t i = tid; i < np; i += INNER_THREADS) t = fma(a[i], b[i], t);
  sh[tid] = t;
  const double s = block_sum(sh, tid);
  if (tid == 0) out[c] = s;
}

// ---- Jacobi-preconditioned conjugate gradients on every column at once
struct PcgArgs {
  int64_t np;
  int ncols;
  const double* dinv;       // (np) inverse of the mass matrix diagonal
  double *x, *r, *z, *p;    // (np, ncols)
  const double* q;          // (np, ncols): M p
  const double* b;          // (np, ncols) right-hand side
  const double* pq;         // (ncols) (p, q)
  const double* rz;         // (ncols) (r, z) of the current iterate
  const double* rz_new;     // (ncols) (r, z) after the update
};

// x = 0, r = b, z = D^{-1} r, p = z
SFX_DEV void pcg_init_item(const PcgArgs& a, int64_t i) {
  if (i >= a.np * a.ncols) return;
  const double ri = a.b[i];
  const double zi = ri * a.dinv[i % a.np];
  a.x[i] = 0.0;
  a.r[i] = ri;
  a.z[i] = zi;
  a.p[i] = zi;
}

// alpha = (r, z) / (p, q); x += alpha p; r -= alpha q; z = D^{-1} r
SFX_DEV void pcg_step_item(const PcgArgs& a, int64_t i) {
  if (i >= a.np * a.ncols) return;
  const int64_t c = i / a.np;
  const double alpha = a.pq[c] != 0.0 ? a.rz[c] / a.pq[c] : 0.0;
  a.x[i] = fma(alpha, a.p[i], a.x[i]);
  const double ri = fma(-alpha, a.q[i], a.r[i]);
  a.r[i] = ri;
  a.z[i] = ri * a.dinv[i % a.np];
}

// beta = (r, z)_new / (r, z); p = z + beta p
SFX_DEV void pcg_dir_item(const PcgArgs& a, int64_t i) {
  if (i >= a.np * a.ncols) return;
  const int64_t c = i / a.np;
  const double beta = a.rz[c] != 0.0 ? a.rz_new[c] / a.rz[c] : 0.0;
  a.p[i] = fma(beta, a.p[i], a.z[i]);
}

// y(i, c) = scale(c) x(i, c), nodal columns (x and y may alias).
SFX_DEV void scale_cols_item(const InnerArgs& a, int64_t i) {
  if (i >= a.np * a.ncols) return;
  a.out[i] = col_scale(a, (int)(i / a.np)) * a.z[i];
}

}  // namespace sfx
