// Non-polynomial FFT_PAR_* products on the reference's own azimuthal grid.
//
// The reference evaluates every product on N = 2*m_max_pad - 1 angles
// (SOURCES/POST_PROCESSING/fft_parallel.F90:185, :254-288).  For bilinear
// products any alias-free grid gives the same coefficients, but a division,
// a square root, a smoothed Heaviside or a norm depends on the grid, so these
// run here on exactly that odd grid, as real DFTs written as sums:
//
//   c2r:  u(t_j) = E_j + O_j,  u(t_{N-j}) = E_j - O_j,  j = 0..K, K = m_max_pad-1,
//         E_j = sum_m a_m cos(m t_j),  O_j = sum_{m>=1} b_m sin(m t_j);
//   r2c:  with S_j = w(t_j) + w(t_{N-j}), D_j = w(t_j) - w(t_{N-j}) (S_0 = w(t_0), D_0 = 0),
//         a_0 = (1/N) sum_j S_j,  a_m = (2/N) sum_j S_j cos(m t_j),
//         b_m = (2/N) sum_j D_j sin(m t_j)      (the reference's 1/N scaling and
//         2*CONJG of modes m >= 1, :291-299; mode-0 sine input dropped, :241).
//
// exact_grid_body: one thread per (point, block of JC angles) accumulates E
// and O of every input field over the modes (the cos/sin tables are uniform
// across the wave), applies the pointwise operation at t_j and t_{N-j} and
// writes S and D of every output field.  exact_r2c_body: one thread per
// (point, output field, block of MC modes) sums S and D against the r2c
// tables.  Points are the fastest index everywhere (coalesced loads).
#pragma once
#include "sfx_portable.h"

namespace sfx {

constexpr int EXACT_MAX_IN = 15, EXACT_MAX_OUT = 9, EXACT_MAX_INT = 4;
constexpr int EXACT_NJ_ALIGN = 32;   // c2r table rows padded to a multiple of every JC
constexpr int EXACT_MC = 16;         // modes per r2c thread

enum ExactOp { EX_DIV = 0, EX_SV2 = 1, EX_SQRT = 2, EX_HEAVI = 3, EX_TENSION = 4, EX_MAXVEL = 5,
               EX_SVCLAMP = 6, EX_CLAMP01 = 7, EX_ENTRO = 8, EX_GRADMOM = 9, EX_COMPR = 10,
               EX_EVISC = 11, EX_EVISCV = 12, EX_CLS = 13, EX_REAL = 14, EX_VELSTAT = 15, EX_CHECKIF = 16,
               EX_ETA = 17, EX_ETA3 = 18, EX_FROMREAL = 19 };

// Operations reduced to one value over all points and angles (a maximum per
// work-group, then one reduction block) instead of transformed back.
constexpr bool exact_is_reduction(int op) { return op == EX_MAXVEL || op == EX_CHECKIF; }

// Angles per thread of the angle pass (2*NI*JC accumulators; the tension
// tensor, 96 of them, re-reads its six inputs fewest times at 8).
constexpr int exact_jc(int op, int nin) {
  return op == EX_SQRT ? 32 : nin <= 2 ? 16 : nin <= 6 ? 8 : nin <= 9 ? 4 : 2;
}
// Inputs and outputs of each operation (EX_HEAVI: nin = interfaces; EX_MAXVEL: 3 or 1).
constexpr int exact_nout(int op) {
  return (op == EX_SV2 || op == EX_SVCLAMP) ? 3 : op == EX_TENSION ? 9 : op == EX_ETA3 ? 3 : (exact_is_reduction(op) || op == EX_EVISCV || op == EX_REAL) ? 0
         : (op == EX_ENTRO || op == EX_COMPR) ? 4 : (op == EX_GRADMOM || op == EX_EVISC) ? 9 : 1;
}

// Fourier field: cos coefficient of mode m at point p at base[p*sp + m*sm],
// sin coefficient at base[p*sp + off + m*sm].
struct ExactField {
  const double* base;
  int64_t sp, sm, off;
};
struct ExactOut {
  double* base;
  int64_t sp, sm, off;
};

struct ExactArgs {
  int64_t np;
  int M;          // modes kept (m_max)
  int K;          // m_max_pad - 1; N = 2K + 1
  int nj;         // c2r table row length: K + 1 rounded up to EXACT_NJ_ALIGN
  int MR;         // r2c table row length: M rounded up to EXACT_MC
  const double* ctab;   // cos(2 pi m j/N) at [m*nj + j], sin at [(M + m)*nj + j], zero for j > K
  const double* rtab;   // (c_m/N) cos(2 pi j m/N) at [j*MR + m], (2/N) sin at [(K+1 + j)*MR + m],
                        // c_0 = 1, c_m = 2; zero for m >= M and for the sine of m = 0
  int nin, nout;
  ExactField in[EXACT_MAX_IN];
  ExactOut out[EXACT_MAX_OUT];
  double* W;            // S of output g, angle j, point p at W[(2g*(K+1) + j)*np + p]; D at +(K+1)*np
  int64_t n_W;
  double* pmax;         // reductions: one maximum per work-group
  int n_int;            // EX_HEAVI: interfaces
  double vals[EXACT_MAX_INT + 1];
  double eps;           // EX_HEAVI: regul half-width (coeff_tanh)
  int pb;               // EX_SVCLAMP: 1 product, 2 quotient
  double smin, smax;    // EX_SVCLAMP: bounds of the scalar
  const double* creal;  // EX_ENTRO: real-space field c_in_real(N, np) at [p*N + j]
  int64_t n_creal;
  double coeff;         // EX_ENTRO, EX_COMPR: coeff1_in_level; EX_GRADMOM: LES_coeff1_mom
  double coeff2;        // EX_COMPR: level_set_comp_factor; EX_EVISC: LES_coeff2
  double coeff3, coeff4;  // EX_EVISC: LES_coeff4, residual_normalization
  const double* hloc;   // EX_EVISC: mesh size per point (np)
  double* nint;         // EX_EVISC: per element and angle, the maximum of |v| over its points, [e*N + j]
  int lg;               // EX_EVISC: points per element
  int64_t n_nint;
  // FFT_COMPUTE_ENTROPY_VISC_MOM (exact_emom_*): per element and angle maxima of
  // |v| (nint), |m| (nint2) and of the raw viscosity (nint3); real-space outputs
  double* nint2;
  double* nint3;
  double* rout1;        // c1_real_out (N, np) at [p*N + j]
  double* rout2;        // opt_c2_real_out or null
  const double* dmax;   // device scalar: the maximum of |v| over points and angles
  int64_t n_nint2;      // EX_EVISCV, EX_REAL: size of rout1
  int64_t rstride;      // EX_REAL: doubles between the points of rout1
  int64_t n_in[EXACT_MAX_IN], n_out[EXACT_MAX_OUT];   // bounds, emulator checks
};

// regul of fft_parallel.F90:2234-2246.
SFX_DEV double exact_regul(double phi, double eps) {
  const double x = phi - 0.5;
  if (x <= -eps) return 0.0;
  if (x <= eps) return (1.0 + x * (x * x - 3.0 * eps * eps) / (-2.0 * eps * eps * eps)) / 2.0;
  return 1.0;
}

// The pointwise operation at one angle: u (NI fields) -> w (NO fields); x is
// the real-space input at that angle (EX_ENTRO only).
template <int OP, int NI, int NO>
SFX_DEV void exact_op(const ExactArgs& a, const double (&u)[NI], double (&w)[NO > 0 ? NO : 1], double x,
                      double y) {
  if constexpr (OP == EX_EVISC || OP == EX_EVISCV) {
    // x: |v| maximised over the element and the neighbouring angles, y: the
    // point's mesh size; v (0-2), residual (3-5), three tensor rows (6-14)   :2601-2620
    const double dot = u[0] * u[3] + u[1] * u[4] + u[2] * u[5];
    const double ve = -a.coeff + fmin(a.coeff3 * x, a.coeff2 * y * fabs(dot) / (a.coeff4 + 1.e-14));
    if constexpr (OP == EX_EVISCV) {
      w[0] = ve;                      // the viscosity alone, on the grid
    } else {
#pragma unroll
      for (int k = 0; k < 9; ++k) w[k] = ve * u[6 + k];
    }
  } else if constexpr (OP == EX_CLS) {
    // grad(phi) (0-2), phi (3); coeff = level_set_comp_factor * 4 * LES_coeff4,
    // coeff2 = h_min_distance, dmax = max |v| over points and angles   :3997-4006
    const double vm = fmin(1.0, a.dmax[0]);
    const double ng = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double t = 2.0 * u[3] - 1.0;
    w[0] = a.coeff * vm * (2.0 * exact_regul(u[3], 0.01) - 1.0) * (fmax(1.0 - t * t, 0.0) / (2.0 * a.coeff2) - ng);
  } else if constexpr (OP == EX_REAL) {
    w[0] = u[0];
  } else if constexpr (OP == EX_ETA || OP == EX_ETA3) {
    // x: eta at this point and angle                              :1954-1962
#pragma unroll
    for (int k = 0; k < NI; ++k) w[k] = x * u[k];
  } else if constexpr (OP == EX_FROMREAL) {
    w[0] = x;
  } else if constexpr (OP == EX_ENTRO) {
    const double v = a.coeff - x;                                 // -visc_entro, :4212-4219
#pragma unroll
    for (int k = 0; k < 3; ++k) w[k] = v * u[k];
    w[3] = fmax(0.0, u[3] * (1.0 - u[3]));                        // :4209
  } else if constexpr (OP == EX_COMPR) {
    // inputs grad(phi) (0-2), grad(phi_reg) (3-5), phi (6)            :3316-3330
    const double c2 = fmax(0.0, u[6] * (1.0 - u[6]));
    const double ng = sqrt(u[3] * u[3] + u[4] * u[4] + u[5] * u[5]) + 1.e-14;
    const double ve = -a.coeff + x;
    const double vc = x * a.coeff2 * c2 / ng;
#pragma unroll
    for (int k = 0; k < 3; ++k) w[k] = -ve * u[k] + vc * u[3 + k];
    w[3] = c2;
  } else if constexpr (OP == EX_GRADMOM) {
    const double v = x - a.coeff;                                 // visco_entro, :3552-3568
#pragma unroll
    for (int k = 0; k < 9; ++k) w[k] = v * u[k];
  } else if constexpr (OP == EX_DIV) {
    w[0] = u[0] / u[1];                                           // :782
  } else if constexpr (OP == EX_SV2) {
#pragma unroll
    for (int k = 0; k < 3; ++k) w[k] = u[k] / u[3];               // :1155-1157
  } else if constexpr (OP == EX_SVCLAMP) {
    const double sc = fmax(a.smin, fmin(u[3], a.smax));         // :1352-1361
#pragma unroll
    for (int k = 0; k < 3; ++k) w[k] = a.pb == 1 ? sc * u[k] : u[k] / sc;
  } else if constexpr (OP == EX_CLAMP01) {
    w[0] = fmax(0.0, fmin(1.0, u[0]));                            // :3771
  } else if constexpr (OP == EX_SQRT) {
    w[0] = sqrt(fmax(0.0, fmin(1.0, u[0])));                      // :4723
  } else if constexpr (OP == EX_HEAVI) {
    double rho = a.vals[0];                                       // :959-968
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const double h = exact_regul(u[i], a.eps);
      rho = rho * (1.0 - h) + a.vals[i + 1] * h;
    }
    w[0] = rho;
  } else if constexpr (OP == EX_CHECKIF) {
    // 1 where a level set in [0.1, 0.9] lies below the next one   :2394-2402
    double bad = 0.0;
#pragma unroll
    for (int i = 0; i + 1 < NI; ++i)
      if (0.1 <= u[i] && u[i] <= 0.9 && u[i] < u[i + 1]) bad = 1.0;
    w[0] = bad;
  } else if constexpr (OP == EX_TENSION) {
    const double nrm = sqrt(u[3] * u[3] + u[4] * u[4] + u[5] * u[5]);   // :1690-1706
    const double inv = nrm + 1.e-14;
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) w[3 * r + c] = u[r] * u[3 + c] / inv - (r == c ? nrm : 0.0);
  } else {
    if constexpr (NI == 3)                                        // :1532-1536
      w[0] = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    else
      w[0] = sqrt(u[0] * u[0]);
  }
}

// c2r of the NI input fields on angles j0..j0+JC-1 (and their mirrors), the
// operation, S and D of the NO outputs (or the running maximum for EX_MAXVEL).
// E and O of the NI input fields at point p on angles j0..j0+JC-1.
template <int NI, int JC>
SFX_DEV void exact_c2r(const ExactArgs& a, int64_t p, int j0, double (&E)[NI][JC], double (&O)[NI][JC]) {
#pragma unroll
  for (int f = 0; f < NI; ++f)
#pragma unroll
    for (int q = 0; q < JC; ++q) E[f][q] = O[f][q] = 0.0;
#pragma unroll 2
  for (int m = 0; m < a.M; ++m) {
    const double* tc = a.ctab + (int64_t)m * a.nj + j0;
    const double* ts = a.ctab + (int64_t)(a.M + m) * a.nj + j0;
    double c[NI], s[NI];
#pragma unroll
    for (int f = 0; f < NI; ++f) {
      const ExactField& F = a.in[f];
      const int64_t ic = p * F.sp + m * F.sm;
      SFX_CHECK("exact in", ic + (m ? F.off : 0), a.n_in[f]);
      c[f] = F.base[ic];
      s[f] = m ? F.base[ic + F.off] : 0.0;     // mode-0 sine dropped (:241)
    }
#pragma unroll
    for (int q = 0; q < JC; ++q) {
      const double cq = tc[q], sq = ts[q];
#pragma unroll
      for (int f = 0; f < NI; ++f) {
        E[f][q] = fma(c[f], cq, E[f][q]);
        O[f][q] = fma(s[f], sq, O[f][q]);
      }
    }
  }
}

// Maximum of the element's values at angles j-1, j, j+1 (periodic) of the
// per-element maxima (:2572-2591; one angle: the value itself).
SFX_DEV double exact_evisc_norm(const ExactArgs& a, int64_t e, int j) {
  const int N = 2 * a.K + 1;
  const double* r = a.nint + e * N;
  SFX_CHECK("exact nint", e * N + N - 1, a.n_nint);
  if (N < 3) return r[j];
  const int jm = j == 0 ? N - 1 : j - 1, jp = j == N - 1 ? 0 : j + 1;
  return fmax(fmax(r[jm], r[j]), r[jp]);
}

template <int OP, int NI, int NO = exact_nout(OP), int JC = exact_jc(OP, NI)>
SFX_DEV void exact_grid_body(const ExactArgs& a, int64_t p, int jc, double& vmax) {
  if (p >= a.np) return;
  const int j0 = jc * JC;
  double E[NI][JC], O[NI][JC];
  if constexpr (OP == EX_FROMREAL) {
#pragma unroll
    for (int f = 0; f < NI; ++f)
#pragma unroll
      for (int q = 0; q < JC; ++q) E[f][q] = O[f][q] = 0.0;   // no Fourier input
  } else {
    exact_c2r<NI, JC>(a, p, j0, E, O);
  }
  const int64_t plane = (int64_t)(a.K + 1) * a.np;
#pragma unroll
  for (int q = 0; q < JC; ++q) {
    const int j = j0 + q;
    if (j > a.K) break;
    double up[NI], um[NI];
#pragma unroll
    for (int f = 0; f < NI; ++f) {
      up[f] = E[f][q] + O[f][q];
      um[f] = E[f][q] - O[f][q];
    }
    double xp = 0.0, xm = 0.0, y = 0.0;
    if constexpr (OP == EX_ENTRO || OP == EX_GRADMOM || OP == EX_COMPR || OP == EX_ETA || OP == EX_ETA3 ||
                  OP == EX_FROMREAL) {
      const int64_t N = 2 * a.K + 1;
      SFX_CHECK("exact creal", p * N + (j > 0 ? N - j : 0), a.n_creal);
      xp = a.creal[p * N + j];
      xm = j > 0 ? a.creal[p * N + (N - j)] : 0.0;
    } else if constexpr (OP == EX_EVISC || OP == EX_EVISCV) {
      const int64_t e = p / a.lg;
      xp = exact_evisc_norm(a, e, j);
      xm = j > 0 ? exact_evisc_norm(a, e, 2 * a.K + 1 - j) : 0.0;
      y = a.hloc[p];
    }
    double wp[NO > 0 ? NO : 1], wm[NO > 0 ? NO : 1];
    exact_op<OP, NI, NO>(a, up, wp, xp, y);
    if (j > 0) exact_op<OP, NI, NO>(a, um, wm, xm, y);
    if constexpr (OP == EX_MAXVEL || OP == EX_CHECKIF) {
      vmax = fmax(vmax, wp[0]);
      if (j > 0) vmax = fmax(vmax, wm[0]);
    } else if constexpr (OP == EX_REAL) {
      // real planes: the value at angle j of point p at rout1[p*rstride + j]
      const int64_t N = 2 * a.K + 1;
      SFX_CHECK("exact real", p * a.rstride + (j > 0 ? N - j : 0), a.n_nint2);
      a.rout1[p * a.rstride + j] = wp[0];
      if (j > 0) a.rout1[p * a.rstride + (N - j)] = wm[0];
    } else if constexpr (OP == EX_EVISCV) {
      // real-space output (N, np) at [p*N + angle]
      const int64_t N = 2 * a.K + 1;
      SFX_CHECK("exact rout", p * N + (j > 0 ? N - j : 0), a.n_nint2);
      a.rout1[p * N + j] = wp[0];
      if (j > 0) a.rout1[p * N + (N - j)] = wm[0];
    } else {
#pragma unroll
      for (int g = 0; g < NO; ++g) {
        const int64_t o = (2 * g * (int64_t)(a.K + 1) + j) * a.np + p;
        SFX_CHECK("exact W", o + plane, a.n_W);
        a.W[o] = j > 0 ? wp[g] + wm[g] : wp[g];
        a.W[o + plane] = j > 0 ? wp[g] - wm[g] : 0.0;
      }
    }
  }
}

// FFT_MAX_MIN_VEL_DCL in one pass: at point p, angle block jc (JC angles and
// their mirrors), the running maximum of |v| (r[0]), of -|v| over |v| > 0
// (r[1], starts at -inf) and whether |v| > 0 at the last angle N-1 (r[2]:
// 1 or 0: the last angle is the mirror of j = 1, or j = 0 itself when
// N = 1); |v| from three fields (nin = 3) or the first one.
template <int NI, int JC = exact_jc(EX_MAXVEL, NI)>
SFX_DEV void exact_velstat_body(const ExactArgs& a, int64_t p, int jc, double (&r)[3]) {
  if (p >= a.np) return;
  const int j0 = jc * JC;
  double E[NI][JC], O[NI][JC];
  exact_c2r<NI, JC>(a, p, j0, E, O);
#pragma unroll
  for (int q = 0; q < JC; ++q) {
    const int j = j0 + q;
    if (j > a.K) break;
    double up[NI], um[NI], wp[1], wm[1];
#pragma unroll
    for (int f = 0; f < NI; ++f) {
      up[f] = E[f][q] + O[f][q];
      um[f] = E[f][q] - O[f][q];
    }
    exact_op<EX_MAXVEL, NI, 0>(a, up, wp, 0.0, 0.0);
    exact_op<EX_MAXVEL, NI, 0>(a, um, wm, 0.0, 0.0);
    r[0] = fmax(r[0], wp[0]);
    if (wp[0] > 0.0) r[1] = fmax(r[1], -wp[0]);
    if (j > 0) {
      r[0] = fmax(r[0], wm[0]);
      if (wm[0] > 0.0) r[1] = fmax(r[1], -wm[0]);
    }
    if (a.K == 0 ? wp[0] > 0.0 : (j == 1 && wm[0] > 0.0)) r[2] = 1.0;
  }
}

// |v| of the first three input fields at point p, angles j0..j0+15 and their
// mirrors: nrm[q] at t_j, nrm[16 + q] at t_{N-j} (0 for j = 0 or j > K).
constexpr int EXACT_NORM_JC = 16;
SFX_DEV void exact_norm_body(const ExactArgs& a, int64_t p, int j0, double (&nrm)[2 * EXACT_NORM_JC]) {
  double E[3][EXACT_NORM_JC], O[3][EXACT_NORM_JC];
  exact_c2r<3, EXACT_NORM_JC>(a, p, j0, E, O);
#pragma unroll
  for (int q = 0; q < EXACT_NORM_JC; ++q) {
    const int j = j0 + q;
    double sp = 0.0, sm = 0.0;
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const double up = E[f][q] + O[f][q], um = E[f][q] - O[f][q];
      sp = fma(up, up, sp);
      sm = fma(um, um, sm);
    }
    nrm[q] = j <= a.K ? sqrt(sp) : 0.0;                              // :2571
    nrm[EXACT_NORM_JC + q] = (j > 0 && j <= a.K) ? sqrt(sm) : 0.0;
  }
}

// FFT_COMPUTE_ENTROPY_VISC_MOM, inputs u (0-2), m (3-5), residual (6-8) and the
// mass residual c (9): at point p, angles j0..j0+EXACT_EMOM_JC-1 and their
// mirrors (slot q and EXACT_EMOM_JC + q), |u|, |m| (:2839-2840) and the raw
// viscosity LES_coeff2 h MAX(|u . res|, |c |u|^2|) (:2890-2893).
constexpr int EXACT_EMOM_JC = 4;
SFX_DEV void exact_emom_body(const ExactArgs& a, int64_t p, int j0, double (&nv)[2 * EXACT_EMOM_JC],
                             double (&nm)[2 * EXACT_EMOM_JC], double (&vr)[2 * EXACT_EMOM_JC]) {
  double E[10][EXACT_EMOM_JC], O[10][EXACT_EMOM_JC];
  exact_c2r<10, EXACT_EMOM_JC>(a, p, j0, E, O);
  const double h = a.hloc[p];
#pragma unroll
  for (int q = 0; q < EXACT_EMOM_JC; ++q) {
    const int j = j0 + q;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      double u[10];
#pragma unroll
      for (int f = 0; f < 10; ++f) u[f] = side == 0 ? E[f][q] + O[f][q] : E[f][q] - O[f][q];
      const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
      const double mm = u[3] * u[3] + u[4] * u[4] + u[5] * u[5];
      const double ur = u[0] * u[6] + u[1] * u[7] + u[2] * u[8];
      const bool ok = side == 0 ? j <= a.K : (j > 0 && j <= a.K);
      const int s = side * EXACT_EMOM_JC + q;
      nv[s] = ok ? sqrt(uu) : 0.0;
      nm[s] = ok ? sqrt(mm) : 0.0;
      vr[s] = ok ? a.coeff2 * h * fmax(fabs(ur), fabs(u[9] * uu)) : 0.0;
    }
  }
}

// Periodic maximum over angles j-1, j, j+1 of row e of a per-element array.
SFX_DEV double exact_smooth3(const double* arr, int64_t e, int j, int N) {
  const double* r = arr + e * N;
  if (N < 3) return r[j];
  const int jm = j == 0 ? N - 1 : j - 1, jp = j == N - 1 ? 0 : j + 1;
  return fmax(fmax(r[jm], r[j]), r[jp]);
}

// Outputs at point p and angle j (:2914-2931): with nv, nm, vis the element and
// angle maxima of |u|, |m| and the raw viscosity and vmax the maximum of |u|,
// l_G = 3: MIN(4 c4 vmax, 16 vis/(nv nm + 1e-30)); otherwise MIN(c4 nv, vis/(nv nm
// + 1e-30)) and, when asked for, MIN(c4 vmax, vis/(nv nm + 1e-30)).
SFX_DEV void exact_emom_out_body(const ExactArgs& a, int64_t p, int j) {
  const int N = 2 * a.K + 1;
  const int64_t e = p / a.lg;
  SFX_CHECK("exact emom", e * N + N - 1, a.n_nint);
  const double nv = exact_smooth3(a.nint, e, j, N), nm = exact_smooth3(a.nint2, e, j, N);
  const double vis = exact_smooth3(a.nint3, e, j, N);
  const double vmax = a.dmax[0], c4 = a.coeff3;
  const double q = vis / (nv * nm + 1.e-30);
  const int64_t o = p * N + j;
  if (a.lg == 3) {
    a.rout1[o] = fmin(4.0 * c4 * vmax, 16.0 * q);
  } else {
    a.rout1[o] = fmin(c4 * nv, q);
    if (a.rout2) a.rout2[o] = fmin(c4 * vmax, q);
  }
}

// r2c of output g, modes mc*EXACT_MC .. +EXACT_MC-1, at point p.
SFX_DEV void exact_r2c_body(const ExactArgs& a, int64_t p, int g, int mc) {
  if (p >= a.np) return;
  const int m0 = mc * EXACT_MC;
  double A[EXACT_MC], B[EXACT_MC];
#pragma unroll
  for (int q = 0; q < EXACT_MC; ++q) A[q] = B[q] = 0.0;
  const int64_t plane = (int64_t)(a.K + 1) * a.np;
  const double* S = a.W + 2 * g * plane + p;
#pragma unroll 4
  for (int j = 0; j <= a.K; ++j) {
    SFX_CHECK("exact W", 2 * g * plane + p + plane + (int64_t)j * a.np, a.n_W);
    const double sj = S[(int64_t)j * a.np], dj = S[plane + (int64_t)j * a.np];
    const double* tc = a.rtab + (int64_t)j * a.MR + m0;
    const double* ts = a.rtab + (int64_t)(a.K + 1 + j) * a.MR + m0;
#pragma unroll
    for (int q = 0; q < EXACT_MC; ++q) {
      A[q] = fma(sj, tc[q], A[q]);
      B[q] = fma(dj, ts[q], B[q]);
    }
  }
  const ExactOut& F = a.out[g];
#pragma unroll
  for (int q = 0; q < EXACT_MC; ++q) {
    const int m = m0 + q;
    if (m >= a.M) break;
    const int64_t o = p * F.sp + m * F.sm;
    SFX_CHECK("exact out", o + F.off, a.n_out[g]);
    F.base[o] = A[q];
    F.base[o + F.off] = B[q];
  }
}

}  // namespace sfx
