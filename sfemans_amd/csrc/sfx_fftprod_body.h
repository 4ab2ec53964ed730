// Body of the FFT_PAR_*_DCL kernels (compiled for gfx950 and for the CPU emulator).
//
// Replaces, for nb_procs = 1, the per-point pipeline of
// SOURCES/POST_PROCESSING/fft_parallel.F90: FFT_PAR_CROSS_PROD_DCL (:147-328),
// FFT_PAR_DOT_PROD_DCL (:330-500), FFT_PAR_PROD_DCL (:502-665),
// FFT_SCALAR_VECT_DCL pb=1 (:1021-1220) and FFT_TENSOR_DCL without
// opt_tension (:1543-1796).  Inputs and outputs keep the reference layout
// V(np, nb_field, m_max) column-major (point index fastest), so a wave's
// loads and stores for one (column, mode) touch PW consecutive points.
#pragma once
#include "sfx_engine.h"
#include "sfx_kernels.h"

namespace sfx {

template <int OP> struct OpT;
// NF: real physical fields after the inverse transform; NO: real outputs,
// packed two per complex output signal (the tensor's 9 outputs, row-major
// (row g, column c) -> 3g + c, in 5 signals).
template <> struct OpT<OP_CROSS>  { static constexpr int NF = 6, NO = 3; };
template <> struct OpT<OP_DOT>    { static constexpr int NF = 6, NO = 1; };
template <> struct OpT<OP_PROD>   { static constexpr int NF = 2, NO = 1; };
template <> struct OpT<OP_SVECT1> { static constexpr int NF = 4, NO = 3; };
template <> struct OpT<OP_TENSOR> { static constexpr int NF = 6, NO = 9; };

// Field f -> (array, first column, columns of that array, element count).
template <int OP>
SFX_DEV const double* field_base(const FftProdArgs& a, int f, int& col, int& nb, int64_t& n) {
  const bool first = (OP == OP_PROD) ? (f == 0) : (f < 3);
  if (OP == OP_PROD) col = 0;
  else if (OP == OP_SVECT1) col = first ? 2 * f : 0;
  else col = first ? 2 * f : 2 * (f - 3);
  nb = first ? a.nb1 : a.nb2;
  n = first ? a.n1 : a.n2;
  return first ? a.V1 : a.V2;
}

// Real output j at one physical point from the real fields u.
template <int OP>
SFX_DEV double apply_op(const double (&u)[6], int j) {
  if (OP == OP_CROSS) {
    return j == 0 ? u[1] * u[5] - u[2] * u[4] : j == 1 ? u[2] * u[3] - u[0] * u[5] : u[0] * u[4] - u[1] * u[3];
  } else if (OP == OP_DOT) {
    return u[0] * u[3] + u[1] * u[4] + u[2] * u[5];
  } else if (OP == OP_PROD) {
    return u[0] * u[1];
  } else if (OP == OP_SVECT1) {
    return u[3] * u[j];
  } else {  // tensor (g, c) = V1_g * V2_c
    return u[j / 3] * u[3 + j % 3];
  }
}

// Packed spectra of the complex signal made of fields (f1, f2) at point p:
// Z at own modes m = l + L r, B (partner part) from mode MP - m.  Each lane
// loads its own modes only; the partner modes come from the lane that
// loaded them, through the wave's LDS region xf (64 x 16 doubles).
template <int MP, int OP>
SFX_DEV void load_signal(const FftProdArgs& a, int64_t p, bool valid, int pt, int l, int f1, int f2,
                         double* xf, double2 (&Z)[4], double2 (&B)[4]) {
  constexpr int L = Shape<MP>::L;
  int c1, c2, nb1, nb2;
  int64_t n1, n2;
  const double* b1 = field_base<OP>(a, f1, c1, nb1, n1);
  const double* b2 = field_base<OP>(a, f2, c2, nb2, n2);
  const int64_t s1 = a.np * nb1, s2 = a.np * nb2;   // mode strides
  // Loads are unconditional at clamped indices (selected to zero afterwards)
  // so that all of them issue back to back instead of one branch at a time.
  const int64_t pc = valid ? p : 0;
  const int64_t o1 = pc + a.np * c1, o2 = pc + a.np * c2;
  double v[4][4], w[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = l + L * r;
    const int mc = (m < a.M) ? m : 0;
    const int64_t i1 = o1 + s1 * mc, i2 = o2 + s2 * mc;
    SFX_CHECK("V1 load", i1 + a.np, n1);
    SFX_CHECK("V2 load", i2 + a.np, n2);
    v[r][0] = b1[i1]; v[r][1] = b1[i1 + a.np]; v[r][2] = b2[i2]; v[r][3] = b2[i2 + a.np];
  }
  // mode MP - m of slot r is held by lane (L - l) % L in slot 3 - r (l >= 1)
  // or 4 - r (l = 0, r >= 1); layout xf[(4 slot + j) * 64 + lane]
  const int lane = pt * L + l, plane = pt * L + (L - l) % L;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      SFX_CHECK("signal swap", (4 * r + j) * 64 + lane, 64 * 16);
      xf[(4 * r + j) * 64 + lane] = v[r][j];
    }
  SFX_WAVE_SYNC();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int rp = (l != 0) ? 3 - r : ((4 - r) & 3);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      SFX_CHECK("signal swap", (4 * rp + j) * 64 + plane, 64 * 16);
      w[r][j] = xf[(4 * rp + j) * 64 + plane];
    }
  }
  SFX_WAVE_SYNC();
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = l + L * r;
    const int k = MP - m;
    const bool own = valid && m < a.M;
    const bool par = valid && m > 0 && k < a.M;
    const double a1 = own ? v[r][0] : 0.0, bb1 = own ? v[r][1] : 0.0;
    const double a2 = own ? v[r][2] : 0.0, bb2 = own ? v[r][3] : 0.0;
    const double ka1 = par ? w[r][0] : 0.0, kb1 = par ? w[r][1] : 0.0;
    const double ka2 = par ? w[r][2] : 0.0, kb2 = par ? w[r][3] : 0.0;
    if (m == 0) Z[r] = make_double2(a1, a2);
    else Z[r] = make_double2(0.5 * (a1 + bb2), 0.5 * (a2 - bb1));
    B[r] = make_double2(0.5 * (ka1 - kb2), 0.5 * (kb1 + ka2));
  }
}

// Pair Q[m] (lo) with Q[N-m] (hi of the partner lane) and store the (cos,sin)
// coefficients of the two real outputs packed in each complex output signal.
// fft_parallel.F90:291-324: scale 1/N, mode 0 as is, 2*conj for m >= 1.
template <int MP, int NOC, int G = NOC>
SFX_DEV void pair_modes(const double2 (&lo)[NOC][4], const double2 (&hi)[NOC][4], int pt, int l,
                        double2* lds, double (&oa)[NOC][2][4], double (&ob)[NOC][2][4]) {
  constexpr int L = Shape<MP>::L, N = Shape<MP>::N, WC = Shape<MP>::WAVE_CPLX;
  // exchanged G signals at a time: the wave region holds G * WC complex slots
  double2 qn[NOC][4];
#pragma unroll
  for (int t0 = 0; t0 < NOC; t0 += G) {
#pragma unroll
    for (int t = t0; t < NOC && t < t0 + G; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = (t - t0) * WC + lds_slot<MP>(4, pt, l + L * r);
        SFX_CHECK("lds pair write", i, G * WC);
        lds[i] = hi[t][r];
      }
    SFX_WAVE_SYNC();
#pragma unroll
    for (int t = t0; t < NOC && t < t0 + G; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = (t - t0) * WC + lds_slot<MP>(4, pt, (MP - (l + L * r)) & (MP - 1));
        SFX_CHECK("lds pair read", i, G * WC);
        qn[t][r] = lds[i];
      }
    SFX_WAVE_SYNC();
  }
  const double invN = 1.0 / (double)N;
#pragma unroll
  for (int t = 0; t < NOC; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const bool m0 = (l + L * r) == 0;
      const double2 Qm = lo[t][r];
      const double2 Qn = sel2(m0, Qm, qn[t][r]);
      oa[t][0][r] = m0 ? Qm.x * invN : (Qm.x + Qn.x) * invN;
      ob[t][0][r] = m0 ? 0.0 : (Qn.y - Qm.y) * invN;
      oa[t][1][r] = m0 ? Qm.y * invN : (Qm.y + Qn.y) * invN;
      ob[t][1][r] = m0 ? 0.0 : (Qm.x - Qn.x) * invN;
    }
}

// Output of the work-group's PW*wpb consecutive points through LDS, one
// output field (cos and sin planes of all modes) at a time: each store
// instruction then writes runs of consecutive points of one plane instead of
// the 4-point pieces the transform layout holds.  `stage` is the exchange
// region of the work-group (free once the modes are paired); the rows of
// PTS points are padded by one double.  Tensor output 3g + c goes to row g,
// column c of the (3, np, 6, M) output: the points of a plane are 3 doubles
// apart, so a store instruction writes a third of each line it touches (the
// other rows fill the rest), 16 lines per instruction against 64 from the
// transform layout.  The arrays hold output signals t0 .. t0 + NOC - 1.
template <int MP, int OP, int NOC, int NO>
SFX_DEV void stage_store(const FftProdArgs& a, const double (&oa)[NOC][2][4], const double (&ob)[NOC][2][4],
                         int tid, int wave, int pt, int l, int wpb, int64_t bid, double2* stage2, int t0 = 0) {
  constexpr int L = Shape<MP>::L, PW = Shape<MP>::PW;
  double* stage = reinterpret_cast<double*>(stage2);
  const int PTS = PW * wpb, ROW = PTS + 1;
  const int q = wave * PW + pt;
  const int64_t p0 = bid * PTS;
  const int nst = 2 * a.M * PTS;
#pragma unroll
  for (int t = 0; t < NOC; ++t)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int o = 2 * (t0 + t) + h;
      if (o >= NO) continue;
      SFX_SYNC();
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = l + L * r;
        if (m >= a.M) continue;
        SFX_CHECK("stage write", (2 * m + 1) * ROW + q, 2 * a.M * ROW);
        stage[(2 * m) * ROW + q] = oa[t][h][r];
        stage[(2 * m + 1) * ROW + q] = ob[t][h][r];
      }
      SFX_SYNC();
      for (int i = tid; i < nst; i += 64 * wpb) {
        const int plane = i / PTS, qq = i % PTS;
        const int64_t pp = p0 + qq;
        if (pp >= a.np) continue;
        const int c = plane & 1, m = plane >> 1;
        const int64_t idx = (OP == OP_TENSOR) ? o / 3 + 3 * (pp + a.np * (int64_t)(2 * (o % 3) + c + 6 * m))
                                              : pp + a.np * (int64_t)(2 * o + c + 2 * NO * m);
        SFX_CHECK("Vout store", idx, a.nout);
        a.Vout[idx] = stage[plane * ROW + qq];
      }
    }
}

// LDS per work-group: twiddles (4 MP complex) + per wave 3 signals x 256 complex.
template <int MP>
constexpr size_t fftprod_smem_cplx(int wpb) { return 4 * MP + (size_t)wpb * 3 * Shape<MP>::WAVE_CPLX; }

template <int MP, int OP>
SFX_DEV void fftprod_body(const FftProdArgs& a, int tid, int64_t bid, int wpb, double2* smem) {
  using S = Shape<MP>;
  constexpr int L = S::L, PW = S::PW;
  constexpr int NF = OpT<OP>::NF, NS = NF / 2, NO = OpT<OP>::NO;
  constexpr int NOC = (NO + 1) / 2;
  constexpr int GX = 3;                           // signals per LDS exchange
  const int wave = tid >> 6, lane = tid & 63;
  const int pt = lane / L, l = lane % L;
  // twiddle tables staged in LDS once per work-group: wN (3 MP) then wMP (MP)
  double2* tw = smem;
  for (int i = tid; i < 4 * MP; i += 64 * wpb) tw[i] = (i < 3 * MP) ? a.wN[i] : a.wMP[i - 3 * MP];
  SFX_SYNC();
  const double2* wN = tw;
  const double2* wMP = tw + 3 * MP;
  double2* lds = smem + 4 * MP + wave * (GX * S::WAVE_CPLX);
  const int64_t p = (bid * wpb + wave) * PW + pt;
  const bool valid = p < a.np;

  // inverse: NS complex signals x 3 branches
  double2 y[3 * NS][4];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    double2 Z[4], B[4];
    load_signal<MP, OP>(a, p, valid, pt, l, 2 * s, 2 * s + 1, reinterpret_cast<double*>(lds), Z, B);
    prestep<MP>(Z, B, l, y[3 * s + 0], y[3 * s + 1], y[3 * s + 2], wN);
  }
  inverse_fft<MP, 3 * NS, GX>(y, pt, l, lds, wMP);

  // The tensor product transforms its 5 output signals one at a time (three
  // branches each): its 9 input branches stay live through all of them, and
  // with more output signals at once the kernel needs more than the 256
  // registers of 2 waves per SIMD.
  constexpr int NT1 = (OP == OP_TENSOR) ? 1 : NOC;
#pragma unroll
  for (int t0 = 0; t0 < NOC; t0 += NT1) {
    // pointwise op on the 3 branches x 4 positions held by this lane
    double2 q[3 * NT1][4];
#pragma unroll
    for (int j1 = 0; j1 < 3; ++j1)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double u[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < NS; ++s) { u[2 * s] = y[3 * s + j1][r].x; u[2 * s + 1] = y[3 * s + j1][r].y; }
#pragma unroll
        for (int t = 0; t < NT1; ++t) {
          const int ot = 2 * (t0 + t);
          q[3 * t + j1][r] = make_double2(apply_op<OP>(u, ot), (ot + 1 < NO) ? apply_op<OP>(u, ot + 1) : 0.0);
        }
      }
    forward_fft<MP, 3 * NT1, GX>(q, pt, l, lds, wMP);

    double2 lo[NT1][4], hi[NT1][4];
#pragma unroll
    for (int t = 0; t < NT1; ++t) poststep<MP>(q[3 * t], q[3 * t + 1], q[3 * t + 2], l, lo[t], hi[t], wN);
    double oa[NT1][2][4], ob[NT1][2][4];
    pair_modes<MP, NT1>(lo, hi, pt, l, lds, oa, ob);
    stage_store<MP, OP, NT1, NO>(a, oa, ob, tid, wave, pt, l, wpb, bid, smem + 4 * MP, t0);
    // the next exchanges reuse the stage area other waves may still read
    if (t0 + NT1 < NOC) SFX_SYNC();
  }
}

}  // namespace sfx
