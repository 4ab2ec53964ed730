// Adjoint (exact transpose, Euclidean inner product on the coefficient
// arrays) of the tangent of the convective term, body for gfx950 and the CPU
// emulator.  No counterpart in the reference: its LightKrylov operator binds
// rmatvec to the forward matvec (SOURCES/LIGHTKRYLOV_OBJECTS/exponential_propagator.F90:61-62).
//
// Tangent around U, per Gauss point p (W_p: interpolation, R_p: curl in
// cos/sin form, Psi: synthesis on the grid, Phi: analysis and truncation):
//     y_p = Phi( r_U x Psi(W_p d) + Psi(R_p d) x w_U )
// so, with pbar = Phi^T ybar_p,
//     dbar = sum_p  W_p^T Psi^T (pbar x r_U)  +  R_p^T Psi^T (w_U x pbar).
// Phi^T is a synthesis of the coefficients scaled by (1/N, 2/N, 2/N) (mode 0,
// cos, sin; the mode-0 sine is dropped); Psi^T is an analysis scaled by
// (N, N/2, N/2) with a zero mode-0 sine.  R_p^T and W_p^T are applied per
// mode and Gauss point; the 7 Gauss points of an element are summed through
// LDS into per-element nodal contributions elo(6 me, 6, M), which a node-
// centred gather (adjoint_gather_block) sums in element order.
#pragma once
#include "sfx_gather.h"
#include "sfx_gauss_body.h"

namespace sfx {

struct AdjArgs {
  GaussArgs g;              // mesh, base field A (g.U: its curl enters), list_mode, twiddles, tables
  const double* UB;         // split form: base field B (its values enter), (np, 6, M)
  int64_t nUB;
  double coef;              // multiplies both products
  const double* ybar;       // (G, 6, M)
  int64_t nybar;
  double* elo;              // per-element nodal contributions, index ((6e+i)*6 + k)*M + m
                            // (W^T part in the split form)
  double* elo2;             // split form: R^T part
  int64_t nelo;
};

// Doubles per (Gauss point, mode) in the element reduction: value, d/dx and
// d/dy weights of the six components; the split form keeps the value
// weights of the two outputs apart.
template <bool SPLIT> struct AdjRed { static constexpr int K = SPLIT ? 24 : 18; };

// Waves per work-group of the adjoint kernel (see gauss_waves): 7 (256
// registers, the parking form) for the convective form at 16 <= MP <= 128,
// 4 (512 registers) elsewhere.  Measured at 200,704 triangles, 7 against 4
// waves: M = 16 2.23 / 2.56 ms, M = 32 4.92 / 5.97, M = 128 25.9 / 29.6
// (M = 64 was the first size moved to 7).
template <int MP, bool SPLIT>
constexpr int adjoint_waves() { return (!SPLIT && MP >= 16 && MP <= 128) ? 7 : 4; }

template <int MP, bool SPLIT, int NW>
struct AdjLds {
  using LY = GaussLds<MP, NW>;
  static constexpr int NT = 64 * NW;
  // Parking (convective form on 7 waves, 256 registers per lane): the base
  // spectra wait in LDS while pbar is formed, and pbar waits in LDS during
  // the base-field transforms, so that neither is spilled to scratch.
  static constexpr bool PARK = !SPLIT && NW > 4;
  // signals exchanged per LDS round of the transforms; the per-wave region
  // also holds the partner swap of the base fields (64 x 12 doubles) and,
  // with 4 waves, of ybar (64 x 24)
  static constexpr int GX = PARK ? 2 : 3;
  static constexpr int RW0 = 2 * GX * Shape<MP>::WAVE_CPLX;
  static constexpr int RWS = 64 * (NW <= 4 ? 24 : 12);
  static constexpr int RW = RW0 > RWS ? RW0 : RWS;
  static_assert(RW >= 2 * GX * Shape<MP>::WAVE_CPLX, "exchange region too small");
  // nodal staging: two alternating chunk buffers of the base field(s)
  static constexpr int NODAL = (SPLIT ? 4 : 2) * LY::NOD6;                 // doubles from BIG
  static constexpr int XCH = NODAL;                                       // first wave region
  // Element sums by component: work item (element, component, mode offset)
  // for the six local nodes.  red[i + k * SK] for entry i = local point * L
  // + mode offset and weight k, weight-major: a wave writes 64 consecutive
  // entries of one weight and reads 16 consecutive entries of each of 4
  // weights SK apart; the 16-double pad puts those groups in different bank
  // halves.  (By local node, entry-major, each item reads all 18 weights of
  // its 7 entries, six times the LDS reads: 11.18 against 10.93 ms for the
  // 7-wave form at M = 64, and slower on 4 waves too.)
  static constexpr int RPL = 7 * LY::E * (MP / 4);
  static constexpr int SK = RPL + 16;
  static constexpr int RED = SK * AdjRed<SPLIT>::K;
  // the reduction area aliases the staging and wave regions (used after them)
  static constexpr int BIGD = (NODAL + NW * RW > RED) ? NODAL + NW * RW : RED;
  // parked doubles per lane: PK1 in the staging area (idle once the base
  // fields are interpolated; slot f of lane t at f * NT + t), PK2 past BIGD
  // in what is left of the 160 KB
  static constexpr int LDS_DOUBLES = 160 * 1024 / 8;
  static constexpr int PK1 = PARK ? NODAL / NT : 0;
  static constexpr int PK2 = PARK ? (LDS_DOUBLES - 2 * LY::BIG - BIGD) / NT : 0;
  static constexpr int PK = PK1 + PK2;
  static constexpr int TOTAL = LY::BIG + (BIGD + PK2 * NT + 1) / 2;   // double2 units
  static_assert(!PARK || (PK2 >= 0 && 2 * TOTAL <= LDS_DOUBLES), "parking exceeds the LDS");
};

// Parked slot f (0 <= f < PK) of lane t.
template <class AL>
SFX_DEV double& park_slot(double* big, int f, int t) {
  SFX_CHECK("park slot", f, AL::PK);
  return f < AL::PK1 ? big[f * AL::NT + t] : big[AL::BIGD + (f - AL::PK1) * AL::NT + t];
}

// Load ybar at the own and partner modes (MP - m) of Gauss point p, scale
// for Phi^T and pack the three real output fields as signals (pr, pth) and
// (pz, 0).  Each lane loads its own modes only and takes the partner modes
// from the lanes that loaded them, through the wave's LDS region xf: all
// four slots at once (64 x 24 doubles, the 512-register form), or with
// SWAP2 in two rounds of two mode slots, packing as the partners arrive, so
// that it fits the 256-register form (its wave region holds 12 doubles per
// lane).  Loading the partner modes as well would double the load
// instructions, each of which touches 16 cache lines: the work-group's
// loads queue in the vector memory path (adjoint at M = 64: 10.93 ms with
// own and partner loads two slots at a time, 10.56 ms with SWAP2).
template <int MP, bool SWAP2>
SFX_DEV void load_ybar(const AdjArgs& A, int64_t p, bool valid, int pt, int l, double* xf,
                       double2 (&Z)[2][4], double2 (&B)[2][4]) {
  constexpr int L = MP / 4, N = 3 * MP;
  const GaussArgs& a = A.g;
  const int64_t pc = valid ? p : 0;
  const double two_n = 2.0 / N, one_n = 1.0 / N;
  // the plane offsets are the same for every group: kept live between calls
  // they would be spilled, so they are recomputed each time
  int lo = l;
  SFX_OPAQUE(lo);
  // own (v) and partner (w) values [slot][a_r, b_r, a_th, b_th, a_z, b_z]
  double v[4][6], w[4][6];
  auto pack = [&](int r) {
    const int m = l + L * r, k = MP - m;
    const bool own = valid && m < a.M;
    const bool par = valid && m > 0 && k < a.M;
    const double so = own ? (m == 0 ? one_n : two_n) : 0.0;
    const double sp = par ? two_n : 0.0;
    double o[6], q[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) { o[c] = so * v[r][c]; q[c] = sp * w[r][c]; }
    if (m == 0) { o[1] = 0; o[3] = 0; o[5] = 0; }
    // signal 0: fields (r, theta); signal 1: (z, 0)
    Z[0][r] = make_double2(m == 0 ? o[0] : 0.5 * (o[0] + o[3]), m == 0 ? o[2] : 0.5 * (o[2] - o[1]));
    B[0][r] = make_double2(0.5 * (q[0] - q[3]), 0.5 * (q[1] + q[2]));
    Z[1][r] = make_double2(m == 0 ? o[4] : 0.5 * o[4], m == 0 ? 0.0 : -0.5 * o[5]);
    B[1][r] = make_double2(0.5 * q[4], 0.5 * q[5]);
  };
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int m = lo + L * r;
    const int mc = (m < a.M) ? m : 0;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const int64_t i1 = pc + a.G * (c + 6 * (int64_t)mc);
      SFX_CHECK("ybar load", i1, A.nybar);
      v[r][c] = A.ybar[i1];
    }
  }
  if constexpr (SWAP2) {
    // own modes only, in one batch; the partner modes through xf in two
    // rounds of two slots (xf[(6 j + c) * 64 + lane], 64 x 12 doubles):
    // round A slots 0 and 3, round B slots 1 and 2.  The partner of slot r
    // is slot 3 - r at lane (L - l) % L for l >= 1, slot (4 - r) % 4 of
    // lane 0 for l = 0 (whose slot-0 partner, mode MP, is dropped by pack).
    const int lane = pt * L + l, plane = pt * L + (L - l) % L;
    const bool l0 = (l == 0);
    double a0[6], a1[6], b0[6], b1[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      SFX_CHECK("ybar swap", (6 + c) * 64 + lane, 64 * 12);
      xf[c * 64 + lane] = v[0][c];
      xf[(6 + c) * 64 + lane] = v[3][c];
    }
    SFX_WAVE_SYNC();
    // l >= 1: a0 = slot 3 (partner of r = 0), a1 = slot 0 (partner of r = 3);
    // l = 0:  a0 = slot 0 (partner of r = 0), a1 = slot 3 (partner of r = 1)
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      SFX_CHECK("ybar swap", (6 + c) * 64 + plane, 64 * 12);
      a0[c] = xf[((l0 ? 0 : 6) + c) * 64 + plane];
      a1[c] = xf[((l0 ? 6 : 0) + c) * 64 + plane];
    }
    SFX_WAVE_SYNC();
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      w[0][c] = a0[c];
      xf[c * 64 + lane] = v[1][c];
      xf[(6 + c) * 64 + lane] = v[2][c];
    }
    pack(0);
    SFX_WAVE_SYNC();
    // l >= 1: b0 = slot 2 (partner of r = 1), b1 = slot 1 (partner of r = 2);
    // l = 0:  b0 = slot 2 (partner of r = 2), b1 = slot 1 (partner of r = 3)
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      b0[c] = xf[(6 + c) * 64 + plane];
      b1[c] = xf[c * 64 + plane];
    }
    SFX_WAVE_SYNC();
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      w[1][c] = l0 ? a1[c] : b0[c];
      w[2][c] = l0 ? b0[c] : b1[c];
      w[3][c] = l0 ? b1[c] : a1[c];
    }
    pack(1);
    pack(2);
    pack(3);
  } else {
    // mode MP - m of slot r is held by lane (L - l) % L in slot 3 - r (l >= 1)
    // or 4 - r (l = 0, r >= 1); layout xf[(6 slot + c) * 64 + lane]
    const int lane = pt * L + l, plane = pt * L + (L - l) % L;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        SFX_CHECK("ybar swap", (6 * r + c) * 64 + lane, 64 * 24);
        xf[(6 * r + c) * 64 + lane] = v[r][c];
      }
    SFX_WAVE_SYNC();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rp = (l != 0) ? 3 - r : ((4 - r) & 3);
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        SFX_CHECK("ybar swap", (6 * rp + c) * 64 + plane, 64 * 24);
        w[r][c] = xf[(6 * rp + c) * 64 + plane];
      }
    }
    SFX_WAVE_SYNC();
#pragma unroll
    for (int r = 0; r < 4; ++r) pack(r);
  }
}

// SPLIT = false: transpose of T_U d = P[(curl U) x W d + (curl d) x U]
// (A = B = U, both parts summed into one output).
// SPLIT = true: transpose, times coef, of the two Lorentz passes
// (curl A) x W dB and (curl dH) x B: the W^T part (pbar x curl A) is the
// adjoint for dB (elo), the R^T part (B x pbar) the adjoint for dH (elo2).
template <int MP, bool SPLIT, int NW>
SFX_DEV void adjoint_body(const AdjArgs& A, int tid, int64_t bid, double2* smem) {
  using S = Shape<MP>;
  using LY = GaussLds<MP, NW>;
  using AL = AdjLds<MP, SPLIT, NW>;
  constexpr int KR = AdjRed<SPLIT>::K;
  (void)KR;   // bounds checks only
  const GaussArgs& a = A.g;
  constexpr int L = S::L, PW = S::PW, E = LY::E, N = S::N, NG = groups_per_wave(NW);
  constexpr int GX = AL::GX;
  constexpr int NT = 64 * NW;
  const int wave = tid >> 6, lane = tid & 63;
  const int pt = lane / L, l = lane % L;
  double* smd = reinterpret_cast<double*>(smem);
  double2* tw = smem + LY::TW;
  double* tab = smd + 2 * LY::TAB;
  double* lmode = smd + 2 * LY::MODE;
  double2* geo = smem + LY::GEO;
  int* nid = reinterpret_cast<int*>(smem + LY::NID);
  const int64_t e0 = bid * E;
  gauss_prologue<MP, NW>(a, tid, e0, smem);
  SFX_SYNC();
  const double2* wN = tw;
  const double2* wMP = tw + 3 * MP;

  // ---- this lane's Gauss point in each of the wave's groups (group wave + 4 gi)
  GaussGeom g[NG];
  bool has[NG], valid[NG];
  int el[NG];
  int64_t p[NG];
#pragma unroll
  for (int gi = 0; gi < NG; ++gi) {
    const int grp = wave + NW * gi;
    has[gi] = grp < GAUSS_GROUPS;
    const int lp = (has[gi] ? grp : 0) * PW + pt;
    el[gi] = lp / 7;
    p[gi] = e0 * 7 + lp;
    valid[gi] = has[gi] && (e0 + el[gi]) < a.me;
    gauss_geom(geo + 6 * el[gi], tab, valid[gi] ? lp % 7 : 0, g[gi]);
  }
  double* big = smd + 2 * LY::BIG;
  double* xf = big + AL::XCH + (size_t)wave * AL::RW;
  double2* lds = reinterpret_cast<double2*>(xf);

  // ---- base fields at the Gauss points: (curl A, B), spectra of both groups
  double2 Z[NG][3][4], B[NG][3][4];
#pragma unroll
  for (int gi = 0; gi < NG; ++gi)
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
      for (int r = 0; r < 4; ++r) { Z[gi][s][r] = make_double2(0, 0); B[gi][s][r] = make_double2(0, 0); }
  {
    NodalFetch<6, MP, NW> fa, fb;
    nodal_fetch<6, MP, NW>(a, a.U, a.nU, e0, 0, tid, nid, fa);
    if (SPLIT) nodal_fetch<6, MP, NW>(a, A.UB, A.nUB, e0, 0, tid, nid, fb);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      // chunk c into buffer c % 2 (chunk c - 2 there was interpolated by
      // every wave before the last barrier), fetch chunk c + 1, one barrier
      double* nodA = big + (c & 1) * (SPLIT ? 2 : 1) * LY::NOD6;
      nodal_store<6, MP, NW>(fa, tid, nodA, LY::ROW6);
      if (SPLIT) nodal_store<6, MP, NW>(fb, tid, nodA + LY::NOD6, LY::ROW6);
      if (c < 3) {
        nodal_fetch<6, MP, NW>(a, a.U, a.nU, e0, c + 1, tid, nid, fa);
        if (SPLIT) nodal_fetch<6, MP, NW>(a, A.UB, A.nUB, e0, c + 1, tid, nid, fb);
      }
      SFX_SYNC();
      const int m = c * L + l;
#pragma unroll
      for (int gi = 0; gi < NG; ++gi) {
        if (!has[gi]) continue;
        double own[12];
        if (m < a.M && valid[gi]) {
          if (SPLIT) pass_fields<GV_PAIRS, MP>(a, nodA, nodA + LY::NOD6, 6 * el[gi], l, lmode[m], g[gi], 1.0, own);
          else pass_fields<GV_CROSS_CURL, MP>(a, nodA, nullptr, 6 * el[gi], l, lmode[m], g[gi], 1.0, own);
        } else {
#pragma unroll
          for (int f = 0; f < 12; ++f) own[f] = 0.0;
        }
        double par[12];
        partner_swap<12, MP>(own, par, pt, l, xf);
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const double o4[4] = {own[4 * s], own[4 * s + 1], own[4 * s + 2], own[4 * s + 3]};
          const double p4[4] = {par[4 * s], par[4 * s + 1], par[4 * s + 2], par[4 * s + 3]};
          pack_chunk<MP>(c, l, o4, p4, Z[gi][s], B[gi][s]);
        }
      }
    }
  }

  // ---- per group: pbar and the base fields on the grid, the two products, Psi^T
  double oa[NG][3][2][4], ob[NG][3][2][4];
#pragma unroll
  for (int gi = 0; gi < NG; ++gi) {
    if (!has[gi]) continue;
    // pbar = Phi^T ybar on the grid: 2 signals x 3 branches
    double pb[3][4][3];
    if constexpr (AL::PARK) {
      static_assert(NG == 1, "parking holds one group per wave");
      // the last PK of the 48 base-spectrum doubles (Z then B,
      // [signal][slot][re, im]) wait in LDS while pbar is formed; the
      // staging area is free once every wave has interpolated from it
      SFX_SYNC();
#pragma unroll
      for (int k = 48 - AL::PK; k < 48; ++k) {
        const double2 z = k < 24 ? Z[gi][k / 8][(k % 8) / 2] : B[gi][(k - 24) / 8][(k % 8) / 2];
        park_slot<AL>(big, k - (48 - AL::PK), tid) = (k & 1) ? z.y : z.x;
      }
    }
    {
      double2 Zy[2][4], By[2][4];
      load_ybar<MP, AL::PARK>(A, p[gi], valid[gi], pt, l, xf, Zy, By);
      double2 y[6][4];
#pragma unroll
      for (int s = 0; s < 2; ++s) prestep<MP>(Zy[s], By[s], l, y[3 * s], y[3 * s + 1], y[3 * s + 2], wN);
      inverse_fft<MP, 6, GX>(y, pt, l, lds, wMP);
#pragma unroll
      for (int j1 = 0; j1 < 3; ++j1)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          pb[j1][r][0] = y[j1][r].x;
          pb[j1][r][1] = y[j1][r].y;
          pb[j1][r][2] = y[3 + j1][r].x;
        }
    }
    if constexpr (AL::PARK) {
      // base spectra back from LDS, the last PK of pbar's 36 doubles
      // ([branch][slot][component]) out.  The reloads go through an index
      // the compiler cannot tie to the stores; it would forward the stored
      // values and keep them in registers otherwise.
      int tz = tid;
      SFX_OPAQUE(tz);
#pragma unroll
      for (int k = 48 - AL::PK; k < 48; ++k) {
        const double v = park_slot<AL>(big, k - (48 - AL::PK), tz);
        double2& z = k < 24 ? Z[gi][k / 8][(k % 8) / 2] : B[gi][(k - 24) / 8][(k % 8) / 2];
        if (k & 1) z.y = v;
        else z.x = v;
      }
#pragma unroll
      for (int j1 = 0; j1 < 3; ++j1)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const int f = (j1 * 4 + r) * 3 + c;
            if (f >= 36 - AL::PK) park_slot<AL>(big, f - (36 - AL::PK), tid) = pb[j1][r][c];
          }
    }
    double2 q[9][4];
    {
      double2 y[9][4];
#pragma unroll
      for (int s = 0; s < 3; ++s) prestep<MP>(Z[gi][s], B[gi][s], l, y[3 * s], y[3 * s + 1], y[3 * s + 2], wN);
      inverse_fft<MP, 9, GX>(y, pt, l, lds, wMP);
      int tp = tid;
      SFX_OPAQUE(tp);
      // fields (rA_r, rA_th), (rA_z, wB_r), (wB_th, wB_z);
      // outputs wbar = pbar x rA, rbar = wB x pbar packed as (wb_r, wb_th), (wb_z, rb_r), (rb_th, rb_z)
#pragma unroll
      for (int j1 = 0; j1 < 3; ++j1)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double r0 = y[j1][r].x, r1 = y[j1][r].y, r2 = y[3 + j1][r].x;
          const double w0 = y[3 + j1][r].y, w1 = y[6 + j1][r].x, w2 = y[6 + j1][r].y;
          double pv[3];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const int f = (j1 * 4 + r) * 3 + c;
            // (an if, not a conditional expression: a conditional over two
            // lvalues selects between their addresses and keeps pb in scratch)
            if (AL::PARK && f >= 36 - AL::PK) pv[c] = park_slot<AL>(big, f - (36 - AL::PK), tp);
            else pv[c] = pb[j1][r][c];
          }
          const double p0 = A.coef * pv[0], p1 = A.coef * pv[1], p2 = A.coef * pv[2];
          const double wb0 = p1 * r2 - p2 * r1, wb1 = p2 * r0 - p0 * r2, wb2 = p0 * r1 - p1 * r0;
          const double rb0 = w1 * p2 - w2 * p1, rb1 = w2 * p0 - w0 * p2, rb2 = w0 * p1 - w1 * p0;
          q[j1][r] = make_double2(wb0, wb1);
          q[3 + j1][r] = make_double2(wb2, rb0);
          q[6 + j1][r] = make_double2(rb1, rb2);
        }
    }
    forward_fft<MP, 9, GX>(q, pt, l, lds, wMP);
    double2 lo[3][4], hi[3][4];
#pragma unroll
    for (int t = 0; t < 3; ++t) poststep<MP>(q[3 * t], q[3 * t + 1], q[3 * t + 2], l, lo[t], hi[t], wN);
    pair_modes<MP, 3, GX>(lo, hi, pt, l, lds, oa[gi], ob[gi]);
  }

  // ---- Psi^T scaling, curl transpose, reference-coordinate adjoints,
  //      then the sum over the element's Gauss points through LDS
  double* red = big;
  // the geometry again (24 FMAs), once for the four slots, rather than kept
  // live through the transforms
  GaussGeom gr[NG];
#pragma unroll
  for (int gi = 0; gi < NG; ++gi) {
    const int lpg = (wave + NW * gi) * PW + pt;
    int elg = lpg / 7;
    SFX_OPAQUE(elg);
    gauss_geom(geo + 6 * elg, tab, valid[gi] ? lpg % 7 : 0, gr[gi]);
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    SFX_SYNC();   // the reduction area aliases the staging and wave regions
    const int m = l + L * r;
    const double sa = (m == 0) ? (double)N : 0.5 * N;
    const double sb = (m == 0) ? 0.0 : 0.5 * N;
#pragma unroll
    for (int gi = 0; gi < NG; ++gi) {
      if (!has[gi]) continue;
      const bool ok = valid[gi] && m < a.M;
      double Wb[6], Rb[6];
      // real fields: t0 = (wb_r, wb_th), t1 = (wb_z, rb_r), t2 = (rb_th, rb_z)
      Wb[0] = sa * oa[gi][0][0][r]; Wb[1] = sb * ob[gi][0][0][r];
      Wb[2] = sa * oa[gi][0][1][r]; Wb[3] = sb * ob[gi][0][1][r];
      Wb[4] = sa * oa[gi][1][0][r]; Wb[5] = sb * ob[gi][1][0][r];
      Rb[0] = sa * oa[gi][1][1][r]; Rb[1] = sb * ob[gi][1][1][r];
      Rb[2] = sa * oa[gi][2][0][r]; Rb[3] = sb * ob[gi][2][0][r];
      Rb[4] = sa * oa[gi][2][1][r]; Rb[5] = sb * ob[gi][2][1][r];
      const GaussGeom& gg = gr[gi];
      const double mr = lmode[m < MP ? m : 0] * gg.iray, ir = gg.iray;
      double Dr[6] = {0, 0, 0, 0, 0, 0}, Dz[6] = {0, 0, 0, 0, 0, 0};
      double Ww[6];   // value weights of the W^T part alone (split form)
#pragma unroll
      for (int k = 0; k < 6; ++k) Ww[k] = Wb[k];
      if (SPLIT) {
#pragma unroll
        for (int k = 0; k < 6; ++k) Wb[k] = 0.0;
      }
      // transpose of curl6 (subroutine_ns_with_velocity.F90:594-608)
      Wb[5] += mr * Rb[0]; Dz[2] -= Rb[0];
      Wb[4] -= mr * Rb[1]; Dz[3] -= Rb[1];
      Dz[0] += Rb[2];      Dr[4] -= Rb[2];
      Dz[1] += Rb[3];      Dr[5] -= Rb[3];
      Wb[2] += ir * Rb[4]; Dr[2] += Rb[4]; Wb[1] -= mr * Rb[4];
      Wb[3] += ir * Rb[5]; Dr[3] += Rb[5]; Wb[0] += mr * Rb[5];
      const int lp = (wave + NW * gi) * PW + pt;
      double* dst = red + (lp * L + l);
      constexpr int RP = AL::SK;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const double xb = fma(gg.a11, Dr[k], gg.a21 * Dz[k]);
        const double yb = fma(gg.a12, Dr[k], gg.a22 * Dz[k]);
        SFX_CHECK("red write", (KR - 1) * RP + (lp * L + l), AL::RED);
        if (SPLIT) {
          dst[(4 * k) * RP] = ok ? Ww[k] : 0.0;
          dst[(4 * k + 1) * RP] = ok ? Wb[k] : 0.0;
          dst[(4 * k + 2) * RP] = ok ? xb : 0.0;
          dst[(4 * k + 3) * RP] = ok ? yb : 0.0;
        } else {
          dst[(3 * k) * RP] = ok ? Wb[k] : 0.0;
          dst[(3 * k + 1) * RP] = ok ? xb : 0.0;
          dst[(3 * k + 2) * RP] = ok ? yb : 0.0;
        }
      }
    }
    SFX_SYNC();
    // Work item (element, component k, mode offset mi) for the six local
    // nodes of the element: the item reads its weights of component k from
    // LDS once for all six nodes, and the shape tables (uniform across the
    // wave) come through the scalar cache into scalar registers, not from
    // LDS.  Each sum runs over the 7 Gauss points in order, and mode is
    // fastest: consecutive lanes store consecutive modes of one entry (elo
    // layout: entry, component, mode).
    for (int it = tid; it < E * 6 * L; it += NT) {
      const int mi = it % L;
      const int k = (it / L) % 6;
      const int e = it / (6 * L);
      const int mm = mi + L * r;
      const int64_t eg = e0 + e;
      if (eg >= a.me || mm >= a.M) continue;
      double acc[6] = {0, 0, 0, 0, 0, 0}, acc2[6] = {0, 0, 0, 0, 0, 0};
      constexpr int RP = AL::SK;
      constexpr int KW = SPLIT ? 4 : 3;
      // the shape tables (global memory, read-only here) through a constant
      // address-space pointer: uniform indices make them scalar loads; the
      // pointer is redefined here so the 126 values are loaded in this loop
      // rather than hoisted out of it (they would not fit in scalar registers)
      const SFX_KARG double* tb = (const SFX_KARG double*)a.tab;
      SFX_OPAQUE_S(tb);
#pragma unroll
      for (int lg2 = 0; lg2 < 7; ++lg2) {
        const double* src = red + ((7 * e + lg2) * L + mi);
        double v[KW];
#pragma unroll
        for (int j = 0; j < KW; ++j) {
          SFX_CHECK("red read", (KW * k + j) * RP + ((7 * e + lg2) * L + mi), AL::RED);
          v[j] = src[(KW * k + j) * RP];
        }
#pragma unroll
        for (int n = 0; n < 6; ++n) {
          const double w = tb[n + 6 * lg2], dx = tb[42 + 2 * (n + 6 * lg2)];
          const double dy = tb[42 + 2 * (n + 6 * lg2) + 1];
          if constexpr (SPLIT) {
            acc[n] = fma(w, v[0], acc[n]);
            acc2[n] = fma(w, v[1], fma(dx, v[2], fma(dy, v[3], acc2[n])));
          } else {
            acc[n] = fma(w, v[0], fma(dx, v[1], fma(dy, v[2], acc[n])));
          }
        }
      }
#pragma unroll
      for (int n = 0; n < 6; ++n) {
        const int64_t o = ((6 * eg + n) * 6 + k) * (int64_t)a.M + mm;
        SFX_CHECK("elo store", o, A.nelo);
        A.elo[o] = acc[n];
        if (SPLIT) A.elo2[o] = acc2[n];
      }
    }
  }
}

// dbar(n, k, m) = sum over the (element, local node) pairs of node n, in
// element order, of elo(6e+i, k, m).  One thread per (node, mode).
struct GatherArgs {
  const int64_t* adj_ptr;
  const int32_t* adj_ent;
  const double* elo;        // index ((6e+i)*6 + k)*M + m
  double* out;              // (np, 6, M)
  int64_t me, np, n_adj, nelo, nout;
  int M;
};

// Work-group bid of the gather (sfx_gather.h): dbar(n, k, m).
SFX_DEV void adjoint_gather_block(const GatherArgs& a, int64_t bid, int tid, double* tile) {
  gather_body<6>(a.adj_ptr, a.adj_ent, a.n_adj, a.elo, a.nelo, a.np, a.M, bid, tid, tile,
                 [&](int k, int m, int64_t n, double v) {
                   const int64_t o = n + a.np * (k + 6 * (int64_t)m);
                   SFX_CHECK("adjoint out", o, a.nout);
                   a.out[o] = v;
                 });
}

}  // namespace sfx
