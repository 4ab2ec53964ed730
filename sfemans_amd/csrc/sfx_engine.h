// Azimuthal FFT engine for the de-aliased pseudo-spectral products (gfx950, FP64).
//
// Grid.  A product of two trigonometric polynomials of degree M-1 is exactly
// de-aliased on any grid with N >= 3M-2 points.  The reference uses
// N = 2*floor(3M/2)-1 (SOURCES/POST_PROCESSING/fft_parallel.F90:185 with
// m_max_pad set by every caller, e.g. subroutine_ns_with_velocity.F90:626);
// here N = 3*MP with MP = next power of two >= max(M,4), which gives the same
// bilinear result up to rounding.
//
// Per point, two real fields f1, f2 with coefficients (a_m, b_m) are packed
// in one complex signal with spectrum Z[m] = c1[m] + i c2[m] (m < MP) and
// Z[N-k] = conj(c1[k]) + i conj(c2[k]) (0 < k < MP), where c[0] = a_0 and
// c[m] = (a_m - i b_m)/2 (fft_parallel.F90:236-241).  With j = 3*j2 + j1,
//     h[3 j2 + j1] = sum_m' e^{2 pi i m' j2/MP} Y_j1[m'],
//     Y_j1[m']     = w^{m' j1} Z[m'] + w^{j1 (m' + 2 MP)} B[m'],
// with w = e^{2 pi i/N} and B[m'] = Z[m' + 2MP]: the length-N transform is three
// length-MP "branch" transforms.  Re h = f1, Im h = f2 on the grid.
//
// The length-MP transforms are radix-4 decimation-in-frequency stages (with a
// final radix-2 stage when log2 MP is odd).  A point is handled by L = MP/4
// lanes holding 4 values each, so a 64-lane wave carries PW = 256/MP points.
// Between stages the values are exchanged through a wave-private LDS region.
// The forward transform applies the conjugate transposes of the inverse
// stages in reverse order; it starts from the inverse's digit-reversed output
// positions (a pointwise product commutes with that permutation) and ends
// with lane l holding modes l + L*r, r = 0..3, the same layout as the input.
#pragma once
#include "sfx_portable.h"

namespace sfx {

SFX_DEV double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
SFX_DEV double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
SFX_DEV double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// a * w
SFX_DEV double2 cmul(double2 a, double2 w) {
  return make_double2(fma(a.x, w.x, -a.y * w.y), fma(a.x, w.y, a.y * w.x));
}
// a * conj(w)
SFX_DEV double2 cmulc(double2 a, double2 w) {
  return make_double2(fma(a.x, w.x, a.y * w.y), fma(a.y, w.x, -a.x * w.y));
}
SFX_DEV double2 mul_i(double2 a) { return make_double2(-a.y, a.x); }
// c ? a : b, component by component.  A conditional expression whose operands
// are double2 lvalues (array elements) is lowered as a select of their
// addresses, which keeps the whole array in scratch memory; scalar selects
// keep it in registers.
SFX_DEV double2 sel2(bool c, double2 a, double2 b) { return make_double2(c ? a.x : b.x, c ? a.y : b.y); }
SFX_DEV double2 mul_mi(double2 a) { return make_double2(a.y, -a.x); }

template <int MP>
struct Shape {
  static_assert(MP >= 4 && MP <= 256 && (MP & (MP - 1)) == 0, "MP must be a power of two in [4,256]");
  static constexpr int L = MP / 4;          // lanes per point
  static constexpr int PW = 64 / L;         // points per wave
  static constexpr int LOG2 = (MP == 4) ? 2 : (MP == 8) ? 3 : (MP == 16) ? 4 : (MP == 32) ? 5
                              : (MP == 64) ? 6 : (MP == 128) ? 7 : 8;
  static constexpr int N4 = LOG2 / 2;       // radix-4 stages
  static constexpr bool R2 = (LOG2 & 1) != 0;
  static constexpr int NST = N4 + (R2 ? 1 : 0);
  static constexpr int N = 3 * MP;
  static constexpr int WAVE_CPLX = PW * MP;  // complex slots per signal per wave (= 256)
};

// Position (0..MP-1) of value r held by lane l at stage s.
template <int MP>
SFX_DEV int stage_pos(int s, int l, int r) {
  if (s < Shape<MP>::N4) {
    const int lb = Shape<MP>::LOG2 - 2 * s;   // log2 block size
    const int ld = lb - 2;                    // log2 stride
    const int b = l >> ld;
    const int o = l & ((1 << ld) - 1);
    return (b << lb) + o + (r << ld);
  }
  return 4 * l + r;
}

// LDS slot of (point, position) inside a wave region for exchange `ex`
// (ex = s for the exchange between stages s and s+1, ex = 4 for the mode
// pairing).  q = pt*MP + pos indexes the 256 slots of one signal; the low
// four bits are XORed with a GF(2)-linear function of bits 2..7 of q.  The
// per-exchange masks were found by exhaustive search so that every 8- and
// 16-lane group of every access pattern hits distinct 16-byte bank groups;
// masks on bits 2 and 3 only touch bits 0 and 1, so the map is a bijection.
constexpr unsigned char SWZ[7][5][6] = {
  {{0,0,0,0,0,0}, {0,0,0,0,0,0}, {0,0,0,0,0,0}, {0,0,0,0,0,0}, {0,0,0,0,0,0}},  // MP=4
  {{1,3,15,2,6,6}, {0,0,0,0,0,0}, {0,0,0,0,0,0}, {0,0,0,0,0,0}, {3,2,5,7,3,7}},  // MP=8
  {{0,3,14,6,9,4}, {0,0,0,0,0,0}, {0,0,0,0,0,0}, {0,0,0,0,0,0}, {3,0,15,7,11,10}},  // MP=16
  {{0,2,14,12,15,10}, {2,3,7,8,13,7}, {0,0,0,0,0,0}, {0,0,0,0,0,0}, {3,0,8,8,0,4}},  // MP=32
  {{2,0,4,14,1,5}, {3,2,12,5,15,3}, {0,0,0,0,0,0}, {0,0,0,0,0,0}, {0,0,0,0,12,3}},  // MP=64
  {{0,2,15,13,15,14}, {1,2,14,8,11,11}, {0,2,15,6,0,1}, {0,0,0,0,0,0}, {3,0,0,0,0,15}},  // MP=128
  {{3,3,9,2,13,3}, {0,2,7,12,8,14}, {3,1,12,8,0,14}, {0,0,0,0,0,0}, {0,0,0,0,0,0}},  // MP=256
};

template <int MP>
SFX_DEV int lds_slot(int ex, int pt, int pos) {
  const int q = pt * MP + pos;
  const unsigned char* A = SWZ[Shape<MP>::LOG2 - 2][ex];
  int x = 0;
#pragma unroll
  for (int i = 0; i < 6; ++i) x ^= ((q >> (2 + i)) & 1) * A[i];
  return q ^ x;
}

// Work-groups are dealt round-robin to the 8 XCDs (each with its own L2).
// Remap the hardware block id so that each XCD walks a contiguous range of
// logical blocks: neighbouring points, which share cache lines, then meet in
// the same L2.  The map is a bijection on [0, nblk).
SFX_DEV int64_t xcd_block(int64_t b, int64_t nblk) {
  const int64_t per = nblk / 8, rem = nblk % 8;
  const int64_t x = b % 8, local = b / 8;
  return x < rem ? x * (per + 1) + local : rem * (per + 1) + (x - rem) * per + local;
}

// Inverse (e^{+i}) radix stage s on the 4 values of lane l, in place.
template <int MP>
SFX_DEV void inv_stage(int s, int l, double2 (&v)[4], const double2* __restrict__ wMP) {
  if (s < Shape<MP>::N4) {
    double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    double2 a2 = cadd(v[1], v[3]), a3 = mul_i(csub(v[1], v[3]));
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
    const int ld = Shape<MP>::LOG2 - 2 * s - 2;
    if (ld > 0) {
      const int o = l & ((1 << ld) - 1);
      const int e = o << (2 * s);
      v[1] = cmul(v[1], wMP[e]);
      v[2] = cmul(v[2], wMP[2 * e]);
      v[3] = cmul(v[3], wMP[3 * e]);
    }
  } else {
    double2 t0 = cadd(v[0], v[1]), t1 = csub(v[0], v[1]);
    double2 t2 = cadd(v[2], v[3]), t3 = csub(v[2], v[3]);
    v[0] = t0; v[1] = t1; v[2] = t2; v[3] = t3;
  }
}

// Forward (e^{-i}) transposed stage s, in place.
template <int MP>
SFX_DEV void fwd_stage(int s, int l, double2 (&v)[4], const double2* __restrict__ wMP) {
  if (s < Shape<MP>::N4) {
    const int ld = Shape<MP>::LOG2 - 2 * s - 2;
    if (ld > 0) {
      const int o = l & ((1 << ld) - 1);
      const int e = o << (2 * s);
      v[1] = cmulc(v[1], wMP[e]);
      v[2] = cmulc(v[2], wMP[2 * e]);
      v[3] = cmulc(v[3], wMP[3 * e]);
    }
    double2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    double2 a2 = cadd(v[1], v[3]), a3 = mul_mi(csub(v[1], v[3]));
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
  } else {
    double2 t0 = cadd(v[0], v[1]), t1 = csub(v[0], v[1]);
    double2 t2 = cadd(v[2], v[3]), t3 = csub(v[2], v[3]);
    v[0] = t0; v[1] = t1; v[2] = t2; v[3] = t3;
  }
}

// Exchange a group of NS signals from stage layout `sa` to stage layout `sb`.
// `lds` is this wave's region: NS * WAVE_CPLX complex slots.
template <int MP, int NS>
SFX_DEV void exchange(double2 (*v)[4], int sa, int sb, int pt, int l, double2* lds) {
  constexpr int WC = Shape<MP>::WAVE_CPLX;
#pragma unroll
  for (int g = 0; g < NS; ++g)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = g * WC + lds_slot<MP>(sa < sb ? sa : sb, pt, stage_pos<MP>(sa, l, r));
      SFX_CHECK("lds exchange write", i, NS * WC);
      lds[i] = v[g][r];
    }
  SFX_WAVE_SYNC();
#pragma unroll
  for (int g = 0; g < NS; ++g)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = g * WC + lds_slot<MP>(sa < sb ? sa : sb, pt, stage_pos<MP>(sb, l, r));
      SFX_CHECK("lds exchange read", i, NS * WC);
      v[g][r] = lds[i];
    }
  SFX_WAVE_SYNC();
}

// Full inverse length-MP transform of NS signals (register arrays v[NS][4]),
// exchanging in groups of at most G signals.
template <int MP, int NS, int G>
SFX_DEV void inverse_fft(double2 (&v)[NS][4], int pt, int l, double2* lds,
                                            const double2* __restrict__ wMP) {
  constexpr int NST = Shape<MP>::NST;
#pragma unroll
  for (int s = 0; s < NST; ++s) {
#pragma unroll
    for (int g = 0; g < NS; ++g) inv_stage<MP>(s, l, v[g], wMP);
    if (s + 1 < NST) {
#pragma unroll
      for (int g0 = 0; g0 < NS; g0 += G) {
        if (g0 + G <= NS) exchange<MP, G>(&v[g0], s, s + 1, pt, l, lds);
        else exchange<MP, (NS % G == 0 ? G : NS % G)>(&v[g0], s, s + 1, pt, l, lds);
      }
    }
  }
}

template <int MP, int NS, int G>
SFX_DEV void forward_fft(double2 (&v)[NS][4], int pt, int l, double2* lds,
                                            const double2* __restrict__ wMP) {
  constexpr int NST = Shape<MP>::NST;
#pragma unroll
  for (int s = NST - 1; s >= 0; --s) {
#pragma unroll
    for (int g = 0; g < NS; ++g) fwd_stage<MP>(s, l, v[g], wMP);
    if (s > 0) {
#pragma unroll
      for (int g0 = 0; g0 < NS; g0 += G) {
        if (g0 + G <= NS) exchange<MP, G>(&v[g0], s, s - 1, pt, l, lds);
        else exchange<MP, (NS % G == 0 ? G : NS % G)>(&v[g0], s, s - 1, pt, l, lds);
      }
    }
  }
}

// Pre-step: from the packed spectra Z (own modes m' = l + L*r) and B
// (partner part of the same slots) build the three branch inputs.
//   Y_j1 = w^{m' j1} Z + w^{j1 (m' + 2MP)} B
template <int MP>
SFX_DEV void prestep(const double2 (&Z)[4], const double2 (&B)[4], int l,
                                        double2 (&Y0)[4], double2 (&Y1)[4], double2 (&Y2)[4],
                                        const double2* __restrict__ wN) {
  constexpr int L = Shape<MP>::L, N = Shape<MP>::N;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = l + L * r;
    Y0[r] = cadd(Z[r], B[r]);
    int e1 = m, e1b = m + 2 * MP;                 // < N
    int e2 = 2 * m, e2b = 2 * (m + 2 * MP);
    e2 -= (e2 >= N) ? N : 0;
    e2b -= (e2b >= N) ? N : 0;
    e2b -= (e2b >= N) ? N : 0;
    SFX_CHECK("wN", e2b, N);
    SFX_CHECK("wN", e2, N);
    Y1[r] = cadd(cmul(Z[r], wN[e1]), cmul(B[r], wN[e1b]));
    Y2[r] = cadd(cmul(Z[r], wN[e2]), cmul(B[r], wN[e2b]));
  }
}

// Post-step: combine the three forward branch transforms G_j1 at own modes
// into Q[m'] (lo) and Q[m' + 2MP] (hi).
//   T_j1 = conj(w)^{m' j1} G_j1,  lo = T0+T1+T2,  hi = T0 + w3 T1 + w3^2 T2
template <int MP>
SFX_DEV void poststep(const double2 (&G0)[4], const double2 (&G1)[4], const double2 (&G2)[4],
                                         int l, double2 (&lo)[4], double2 (&hi)[4],
                                         const double2* __restrict__ wN) {
  constexpr int L = Shape<MP>::L, N = Shape<MP>::N;
  const double c = -0.5;
  const double s = 0.86602540378443864676372317075294;  // sqrt(3)/2
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = l + L * r;
    int e2 = 2 * m;
    e2 -= (e2 >= N) ? N : 0;
    double2 T1 = cmulc(G1[r], wN[m]);
    double2 T2 = cmulc(G2[r], wN[e2]);
    lo[r] = cadd(G0[r], cadd(T1, T2));
    // w3 = c + i s, w3^2 = c - i s
    double2 sum = cadd(T1, T2), dif = csub(T1, T2);
    hi[r] = make_double2(G0[r].x + c * sum.x - s * dif.y, G0[r].y + c * sum.y + s * dif.x);
  }
}

}  // namespace sfx
