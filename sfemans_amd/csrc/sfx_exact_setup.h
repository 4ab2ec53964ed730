// Host-side set-up of the exact-grid products, shared by the library and the
// CPU emulator: DFT tables and the field descriptors of each product.
#pragma once
#include <math.h>
#include <string.h>

#include <vector>

#include "sfx_exact_body.h"

namespace sfx {

inline int exact_nj(int K) { return (K + 1 + EXACT_NJ_ALIGN - 1) / EXACT_NJ_ALIGN * EXACT_NJ_ALIGN; }
inline int exact_mr(int M) { return (M + EXACT_MC - 1) / EXACT_MC * EXACT_MC; }
inline size_t exact_ctab_size(int M, int K) { return (size_t)2 * M * exact_nj(K); }
inline size_t exact_rtab_size(int M, int K) { return (size_t)2 * (K + 1) * exact_mr(M); }

// ctab then rtab (layouts in ExactArgs), angles 2 pi (m j mod N)/N in long double.
inline void exact_host_tables(int M, int K, std::vector<double>& h) {
  const int N = 2 * K + 1, nj = exact_nj(K), MR = exact_mr(M);
  const size_t nc = exact_ctab_size(M, K);
  h.assign(nc + exact_rtab_size(M, K), 0.0);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < M; ++m)
    for (int j = 0; j <= K; ++j) {
      const long double t = two_pi * (long double)(((int64_t)m * j) % N) / (long double)N;
      const long double c = cosl(t), s = sinl(t);
      h[(size_t)m * nj + j] = (double)c;
      h[(size_t)(M + m) * nj + j] = (double)s;
      h[nc + (size_t)j * MR + m] = (double)((m == 0 ? 1.0L : 2.0L) / (long double)N * c);
      h[nc + (size_t)(K + 1 + j) * MR + m] = m == 0 ? 0.0 : (double)(2.0L / (long double)N * s);
    }
}

inline void exact_setup_sizes(ExactArgs& a, int64_t np, int M, int m_max_pad) {
  a.np = np; a.M = M; a.K = m_max_pad - 1;
  a.nj = exact_nj(a.K);
  a.MR = exact_mr(M);
}

inline ExactField exact_field(const double* base, int64_t sp, int64_t sm, int64_t off) {
  ExactField f;
  f.base = base; f.sp = sp; f.sm = sm; f.off = off;
  return f;
}
inline ExactOut exact_outf(double* base, int64_t sp, int64_t sm, int64_t off) {
  ExactOut f;
  f.base = base; f.sp = sp; f.sm = sm; f.off = off;
  return f;
}

// Field k of an (np, nb, M) array: cos column 2k, sin column 2k+1.
inline ExactField exact_col(const double* V, int64_t np, int nb, int k) {
  return exact_field(V + 2 * k * np, 1, (int64_t)nb * np, np);
}
inline ExactOut exact_colo(double* V, int64_t np, int nb, int k) {
  return exact_outf(V + 2 * k * np, 1, (int64_t)nb * np, np);
}

// FFT_PAR_DIV_DCL: c1, c2, c (np, 2, M).
inline void exact_setup_div(ExactArgs& a, const double* c1, const double* c2, double* c, int64_t np, int M) {
  memset(&a, 0, sizeof(a));
  a.nin = 2; a.nout = 1;
  a.in[0] = exact_col(c1, np, 2, 0);
  a.in[1] = exact_col(c2, np, 2, 0);
  a.out[0] = exact_colo(c, np, 2, 0);
  a.n_in[0] = a.n_in[1] = a.n_out[0] = np * 2 * (int64_t)M;
}

// FFT_SCALAR_VECT_DCL pb = 2: V (np, 6, M) / s (np, 2, M) -> out (np, 6, M).
inline void exact_setup_sv2(ExactArgs& a, const double* V, const double* s, double* out, int64_t np, int M) {
  memset(&a, 0, sizeof(a));
  a.nin = 4; a.nout = 3;
  for (int k = 0; k < 3; ++k) {
    a.in[k] = exact_col(V, np, 6, k);
    a.out[k] = exact_colo(out, np, 6, k);
    a.n_in[k] = a.n_out[k] = np * 6 * (int64_t)M - 2 * k * np;
  }
  a.in[3] = exact_col(s, np, 2, 0);
  a.n_in[3] = np * 2 * (int64_t)M;
}

// FFT_SCALAR_VECT_NO_OVERSHOOT: as pb = 1 / 2 with the scalar clamped to [smin, smax].
inline void exact_setup_svclamp(ExactArgs& a, const double* V, const double* s, double* out, int64_t np, int M,
                                int pb, double smin, double smax) {
  exact_setup_sv2(a, V, s, out, np, M);
  a.pb = pb; a.smin = smin; a.smax = smax;
}

// FFT_PAR_ENTRO_VISC_DCL: V (np, 6, M), c (np, 2, M), c_in_real (N, np) ->
// V_out (np, 6, M) = (coeff - c_in_real) V, c_out (np, 2, M) = MAX(0, c (1 - c)).
inline void exact_setup_entro(ExactArgs& a, const double* V, const double* c, const double* creal, double coeff,
                              double* Vout, double* cout, int64_t np, int M, int m_max_pad) {
  memset(&a, 0, sizeof(a));
  a.nin = 4; a.nout = 4;
  for (int k = 0; k < 3; ++k) {
    a.in[k] = exact_col(V, np, 6, k);
    a.out[k] = exact_colo(Vout, np, 6, k);
    a.n_in[k] = a.n_out[k] = np * 6 * (int64_t)M - 2 * k * np;
  }
  a.in[3] = exact_col(c, np, 2, 0);
  a.out[3] = exact_colo(cout, np, 2, 0);
  a.n_in[3] = a.n_out[3] = np * 2 * (int64_t)M;
  a.creal = creal;
  a.n_creal = (2 * (int64_t)m_max_pad - 1) * np;
  a.coeff = coeff;
}

// FFT_PAR_COMPR_ENTRO_VISC_DCL: grad(phi) V1, grad(phi_reg) V2 (np, 6, M), phi c (np, 2, M),
// c_in_real (N, np) -> V_out (np, 6, M), c_out (np, 2, M).
inline void exact_setup_compr(ExactArgs& a, const double* V1, const double* V2, const double* c, const double* creal,
                              double coeff, double comp, double* Vout, double* cout, int64_t np, int M,
                              int m_max_pad) {
  exact_setup_entro(a, V1, c, creal, coeff, Vout, cout, np, M, m_max_pad);
  a.nin = 7;
  a.in[6] = a.in[3];
  a.n_in[6] = a.n_in[3];
  for (int k = 0; k < 3; ++k) {
    a.in[3 + k] = exact_col(V2, np, 6, k);
    a.n_in[3 + k] = np * 6 * (int64_t)M - 2 * k * np;
  }
  a.coeff2 = comp;
}

// FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM: V1, V2, V3 (np, 6, M), c_in_real (N, np) ->
// out (3, np, 6, M), out(a, :, 2f:2f+1, :) = coefficients of (c_in_real - coeff) V_a,f.
inline void exact_setup_gradmom(ExactArgs& a, const double* V1, const double* V2, const double* V3,
                                const double* creal, double coeff, double* out, int64_t np, int M, int m_max_pad) {
  memset(&a, 0, sizeof(a));
  a.nin = 9; a.nout = 9;
  const double* V[3] = {V1, V2, V3};
  for (int r = 0; r < 3; ++r)
    for (int f = 0; f < 3; ++f) {
      a.in[3 * r + f] = exact_col(V[r], np, 6, f);
      a.n_in[3 * r + f] = np * 6 * (int64_t)M - 2 * f * np;
      a.out[3 * r + f] = exact_outf(out + r + 3 * np * 2 * f, 3, 3 * np * 6, 3 * np);
      a.n_out[3 * r + f] = 3 * np * 6 * (int64_t)M - r - 3 * np * 2 * f;
    }
  a.creal = creal;
  a.n_creal = (2 * (int64_t)m_max_pad - 1) * np;
  a.coeff = coeff;
}

// FFT_COMPUTE_ENTROPY_VISC: velocity V1, residual V2, tensor rows V3, V4, V5
// (np, 6, M), hloc (np), lg points per element -> O1, O2, O3 (np, 6, M).  The
// per-element maxima array (me * N doubles) is set by the caller in a.nint.
inline void exact_setup_evisc(ExactArgs& a, const double* const (&V)[5], const double* hloc, double* const (&O)[3],
                              int64_t np, int M, int lg, double c1, double c2, double c4, double res) {
  memset(&a, 0, sizeof(a));
  a.nin = 15; a.nout = 9;
  for (int v = 0; v < 5; ++v)
    for (int k = 0; k < 3; ++k) {
      a.in[3 * v + k] = exact_col(V[v], np, 6, k);
      a.n_in[3 * v + k] = np * 6 * (int64_t)M - 2 * k * np;
    }
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k) {
      a.out[3 * r + k] = exact_colo(O[r], np, 6, k);
      a.n_out[3 * r + k] = np * 6 * (int64_t)M - 2 * k * np;
    }
  a.hloc = hloc;
  a.lg = lg;
  a.coeff = c1; a.coeff2 = c2; a.coeff3 = c4; a.coeff4 = res;
}

// FFT_COMPUTE_ENTROPY_VISC in two passes from its full argument set `a`: `v`
// writes the viscosity on the grid into vreal (N, np) from the velocity and
// residual; `w` multiplies the three tensor rows by it (EX_GRADMOM with no
// offset) and carries the r2c outputs.
inline void exact_evisc_split(const ExactArgs& a, double* vreal, ExactArgs& v, ExactArgs& w) {
  v = a;
  v.nin = 6; v.nout = 0;
  v.rout1 = vreal;
  v.n_nint2 = (2 * (int64_t)a.K + 1) * a.np;
  w = a;
  w.nin = 9; w.nout = 9;
  for (int f = 0; f < 9; ++f) {
    w.in[f] = a.in[6 + f];
    w.n_in[f] = a.n_in[6 + f];
  }
  w.creal = vreal;
  w.n_creal = v.n_nint2;
  w.coeff = 0.0;
}

// FFT_COMPUTE_ENTROPY_VISC_MOM: u V1, momentum V2, residual V3 (np, 6, M), mass
// residual c (np, 2, M), hloc (np); real-space outputs (N, np).  The caller sets
// the per-element arrays, pmax and the sizes.
inline void exact_setup_emom(ExactArgs& a, const double* V1, const double* V2, const double* V3, const double* c,
                             const double* hloc, double* out1, double* out2, int64_t np, int M, int lg, double c2,
                             double c4) {
  memset(&a, 0, sizeof(a));
  a.nin = 10; a.nout = 0;
  const double* V[3] = {V1, V2, V3};
  for (int v = 0; v < 3; ++v)
    for (int k = 0; k < 3; ++k) {
      a.in[3 * v + k] = exact_col(V[v], np, 6, k);
      a.n_in[3 * v + k] = np * 6 * (int64_t)M - 2 * k * np;
    }
  a.in[9] = exact_col(c, np, 2, 0);
  a.n_in[9] = np * 2 * (int64_t)M;
  a.hloc = hloc;
  a.rout1 = out1; a.rout2 = out2;
  a.lg = lg;
  a.coeff2 = c2; a.coeff3 = c4;
}

// FFT_COMPRESSION_LEVEL_SET_DCL, product pass: grad(phi) V1 (np, 6, M), phi c
// (np, 2, M) -> c_out (np, 2, M); dmax (device) the maximum of |velocity|.
inline void exact_setup_cls(ExactArgs& a, const double* V1, const double* c, double* cout, const double* dmax,
                            int64_t np, int M, double comp, double les_coeff4, double h_min_distance) {
  memset(&a, 0, sizeof(a));
  a.nin = 4; a.nout = 1;
  for (int k = 0; k < 3; ++k) {
    a.in[k] = exact_col(V1, np, 6, k);
    a.n_in[k] = np * 6 * (int64_t)M - 2 * k * np;
  }
  a.in[3] = exact_col(c, np, 2, 0);
  a.n_in[3] = np * 2 * (int64_t)M;
  a.out[0] = exact_colo(cout, np, 2, 0);
  a.n_out[0] = np * 2 * (int64_t)M;
  a.dmax = dmax;
  a.coeff = comp * 4.0 * les_coeff4;
  a.coeff2 = h_min_distance;
}

// FFT_PAR_REAL, field f of V (np, nb, M) -> its values on the N = 2K+1 grid
// angles, out(N, nb/2, np) column-major: out[j + N*(f + nb/2*p)].
inline void exact_setup_real(ExactArgs& a, const double* V, int f, int nb, double* out, int64_t np, int M, int N) {
  memset(&a, 0, sizeof(a));
  a.nin = 1; a.nout = 0;
  a.in[0] = exact_col(V, np, nb, f);
  a.n_in[0] = np * nb * (int64_t)M - 2 * (int64_t)f * np;
  a.rout1 = out + (int64_t)N * f;
  a.rstride = (int64_t)N * (nb / 2);
  a.n_nint2 = a.rstride * np - (int64_t)N * f;
}

// FFT_SQUAREROOT_SCALAR and FFT_NO_OVERSHOOT_LEVEL_SET: first scalar of c (np, nb, M)
// -> first scalar of out (np, nb, M).
inline void exact_setup_sqrt(ExactArgs& a, const double* c, double* out, int64_t np, int nb, int M) {
  memset(&a, 0, sizeof(a));
  a.nin = 1; a.nout = 1;
  a.in[0] = exact_col(c, np, nb, 0);
  a.out[0] = exact_colo(out, np, nb, 0);
  a.n_in[0] = a.n_out[0] = np * nb * (int64_t)M;
}

// FFT_HEAVISIDE_DCL: V1 (n_int, np, 2, M), values (n_int + 1) -> out (np, 2, M).
inline void exact_setup_heavi(ExactArgs& a, const double* V1, const double* values, int n_values, double* out,
                              int64_t np, int M, double eps) {
  memset(&a, 0, sizeof(a));
  const int n_int = n_values - 1;
  a.nin = n_int; a.nout = 1; a.n_int = n_int; a.eps = eps;
  for (int i = 0; i < n_values; ++i) a.vals[i] = values[i];
  for (int i = 0; i < n_int; ++i) {
    a.in[i] = exact_field(V1 + i, n_int, (int64_t)n_int * np * 2, (int64_t)n_int * np);
    a.n_in[i] = (int64_t)n_int * np * 2 * M - i;
  }
  a.out[0] = exact_colo(out, np, 2, 0);
  a.n_out[0] = np * 2 * (int64_t)M;
}

// FFT_TENSOR_DCL with opt_tension: V1, V2 (np, 6, M) -> out (3, np, 6, M), out(r, :, 2c:2c+1, :).
inline void exact_setup_tension(ExactArgs& a, const double* V1, const double* V2, double* out, int64_t np,
                                int M) {
  memset(&a, 0, sizeof(a));
  a.nin = 6; a.nout = 9;
  for (int k = 0; k < 3; ++k) {
    a.in[k] = exact_col(V1, np, 6, k);
    a.in[3 + k] = exact_col(V2, np, 6, k);
    a.n_in[k] = a.n_in[3 + k] = np * 6 * (int64_t)M - 2 * k * np;
  }
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      a.out[3 * r + c] = exact_outf(out + r + 3 * np * 2 * c, 3, 3 * np * 6, 3 * np);
      a.n_out[3 * r + c] = 3 * np * 6 * (int64_t)M - r - 3 * np * 2 * c;
    }
}

// FFT_MAX_VEL_DCL: V1 (np, nb, M), nb = 6 (|v|) or 2 (|c|).
inline void exact_setup_maxvel(ExactArgs& a, const double* V1, int64_t np, int nb, int M) {
  memset(&a, 0, sizeof(a));
  a.nin = nb == 6 ? 3 : 1; a.nout = 0;
  for (int k = 0; k < a.nin; ++k) {
    a.in[k] = exact_col(V1, np, nb, k);
    a.n_in[k] = np * nb * (int64_t)M - 2 * k * np;
  }
}

// FFT_PAR_VAR_ETA_PROD_T_DCL / _GAUSS_DCL: eta (N, np) on the grid angles,
// c (np, nb, M), nb = 2 or 6 -> out (np, nb, M), the coefficients of eta * c.
inline void exact_setup_eta(ExactArgs& a, const double* eta, const double* c, double* out, int64_t np, int nb,
                            int M, int m_max_pad) {
  memset(&a, 0, sizeof(a));
  a.nin = a.nout = nb / 2;
  for (int k = 0; k < a.nin; ++k) {
    a.in[k] = exact_col(c, np, nb, k);
    a.out[k] = exact_colo(out, np, nb, k);
    a.n_in[k] = a.n_out[k] = np * nb * (int64_t)M - 2 * k * np;
  }
  a.creal = eta;
  a.n_creal = (2 * (int64_t)m_max_pad - 1) * np;
}

// Analysis of values on the grid: real (N, np) -> c (np, 2, M), the
// truncated, scaled real DFT the products end with.
inline void exact_setup_fromreal(ExactArgs& a, const double* real, double* c, int64_t np, int M, int m_max_pad) {
  memset(&a, 0, sizeof(a));
  a.nin = 1; a.nout = 1;
  a.out[0] = exact_colo(c, np, 2, 0);
  a.n_out[0] = np * 2 * (int64_t)M;
  a.creal = real;
  a.n_creal = (2 * (int64_t)m_max_pad - 1) * np;
}

// FFT_CHECK_INTERFACE: V1 (n_int, np, 2, M) level sets, as for the Heaviside.
inline void exact_setup_checkif(ExactArgs& a, const double* V1, int n_int, int64_t np, int M) {
  memset(&a, 0, sizeof(a));
  a.nin = n_int; a.nout = 0; a.n_int = n_int;
  for (int i = 0; i < n_int; ++i) {
    a.in[i] = exact_field(V1 + i, n_int, (int64_t)n_int * np * 2, (int64_t)n_int * np);
    a.n_in[i] = (int64_t)n_int * np * 2 * M - i;
  }
}

}  // namespace sfx
