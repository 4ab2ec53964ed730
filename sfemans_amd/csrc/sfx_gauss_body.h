// Fused Gauss-point nonlinear terms: P2 interpolation (+curl, gradient,
// divergence) of nodal Fourier fields, de-aliased azimuthal product, Gauss-
// point output.  Body compiled for gfx950 and for the CPU emulator.
//
// Variants (reference routines replaced, SOURCES/SIMULATION/...):
//   GV_CROSS_CURL  smb_cross_prod_gauss_sft_par (subroutine_ns_with_velocity.F90:514-630):
//                  V_out = (curl u + 2 Omega) x u
//   GV_CURLH_B     smb_CurlH_cross_B_gauss_sft_par (:632-727): V_out = (curl H) x B
//   GV_UGRADC      smb_ugradc_gauss_fft_par (subroutine_temperature_with_T.F90:247-350):
//                  c_out = (hc grad c).u + (div u)(hc c)
//   GV_CURL_CROSS  one pass coef * (curl pc) x pv, stored or added to V_out
//   GV_PAIRS       sum over passes of coef * (curl pc) x pv: (curl H) x B
//                  (smb_CurlH_cross_B_gauss_sft_par, :632-727); the MHD term
//                  rotv_v = (curl u) x u - c_L (curl H) x B formed at
//                  subroutine_ns_with_velocity.F90:227, 246-248, in one launch; and
//                  the tangents (curl U) x d + (curl d) x U (+ Lorentz terms).
//
// Work decomposition (point tiles aside, see gauss_point_tiles): a work-group
// owns E = 256/MP consecutive elements, i.e.
// 7E consecutive Gauss points (index (m-1)*7 + l), cut into 7 groups of
// PW = 256/MP points (L = MP/4 lanes each, see sfx_engine.h).  The work-group
// has NW waves (gauss_waves): with NW = 4 wave w handles groups w and w + 4
// in up to 512 registers (one wave per SIMD); with NW = 7 each wave handles
// one group in 256 registers (two waves per SIMD).  The nodal data of the E
// elements is staged in LDS one chunk of L modes at a time.
// Lane l interpolates mode cL + l of its Gauss point; the partner part of the
// packed spectrum (mode MP - m) is taken from the partner lane through LDS.
// Geometry is recomputed per Gauss point from the six node coordinates (the
// isoparametric Jacobian, exact for curved P2 elements), radius from the six
// nodes as the reference.
#pragma once
#include "sfx_engine.h"
#include "sfx_fftprod_body.h"
#include "sfx_geom.h"
#include "sfx_kernels.h"

namespace sfx {

template <int V> struct GVT;
// NSIG: complex signals per pass; NOUT: real outputs; MAXP: passes (runtime
// count a.npass for GV_PAIRS: each pass adds coef * (curl pc) x pv).
template <> struct GVT<GV_CROSS_CURL> { static constexpr int NSIG = 3, NOUT = 3, MAXP = 1; };
template <> struct GVT<GV_PAIRS>      { static constexpr int NSIG = 3, NOUT = 3, MAXP = 4; };
template <> struct GVT<GV_CURL_CROSS> { static constexpr int NSIG = 3, NOUT = 3, MAXP = 1; };
template <> struct GVT<GV_UGRADC>     { static constexpr int NSIG = 4, NOUT = 1, MAXP = 1; };

constexpr int GAUSS_GROUPS = 7;   // point groups per work-group
constexpr int GAUSS_WAVES = 4;    // default wave count (batched launches)
// Waves per work-group of the fused kernel for variant V at size MP.  With 7
// waves a wave holds one group and a CU runs two waves per SIMD (256
// registers each), which hides more latency; with 4 waves a wave holds two
// groups in 512 registers at one wave per SIMD.  7 where measured faster:
// u.grad c; the cross-curl up to MP = 128 (at M = 128, 200,704 triangles,
// the fused right-hand side 19.1 -> 17.8 ms and the Gauss output 16.4 ->
// 16.2 ms against 4 waves, despite 152 bytes per lane of spill); the
// single-pass (curl a) x b up to MP = 64.  4 elsewhere.
template <int MP, int V>
constexpr int gauss_waves() {
  return (V == GV_UGRADC || (V == GV_CROSS_CURL && MP <= 128) || (V == GV_CURL_CROSS && MP <= 64)) ? 7 : 4;
}
constexpr int groups_per_wave(int nw) { return (GAUSS_GROUPS + nw - 1) / nw; }

// Point tiles (the plain Gauss-point output of the cross-curl at 16 <= MP
// <= 64): a work-group of GAUSS_PT_WAVES waves takes GAUSS_PT_WAVES * PW
// consecutive Gauss points instead of E whole elements.  With 16 points at
// MP = 64 each plane of the output gets one whole 128-byte line per
// work-group (28 points of whole elements straddle lines, and the partial
// lines cost 1.7x the output bytes of HBM writes), and the 76-77 KB of LDS
// let two work-groups share a CU (256 registers per lane, two waves per
// SIMD).  The outputs leave through LDS, each store instruction writing 4
// planes x 16 points (a lane's own stores touch 16 planes x 4 points, 16
// cache lines).  Measured at 200,704 triangles: 5.72 -> 5.56 ms at M = 64,
// 2.87 -> 2.64 ms at M = 32, 1.06 -> 1.05 ms at M = 16; staged stores 5.56
// -> 5.28 and 2.66 -> 2.52 ms; output bit for bit the same.
constexpr int GAUSS_PT_WAVES = 4;
template <int MP, int V>
constexpr bool gauss_point_tiles() { return V == GV_CROSS_CURL && MP >= 16 && MP <= 64; }
// Work-groups of the launch over me elements.
template <int MP, int V, int NW>
constexpr int64_t gauss_blocks(int64_t me) {
  return gauss_point_tiles<MP, V>() ? (7 * me + NW * (64 / (MP / 4)) - 1) / (NW * (64 / (MP / 4)))
                                    : (me + 256 / MP - 1) / (256 / MP);
}
constexpr int waves_per_simd(int nw) { return nw > 4 ? 2 : 1; }

// LDS layout (double2 units) of the fused kernel with NW waves.  The nodal
// staging holds NCH chunks (1: one chunk at a time; 4: all modes at once)
// of a 6-column field and a second field of NCB columns (6, 2 or 0).
template <int MP, int NW = GAUSS_WAVES, int NCH = 1, int NCB = 6>
struct GaussLds {
  static constexpr int E = 256 / MP;                    // elements per work-group
  static constexpr int L = MP / 4;
  static constexpr int TW = 0;                          // twiddles, 4 MP
  static constexpr int TAB = 4 * MP;                    // ww(6,7), dd(2,6,7): 126 doubles
  static constexpr int MODE = TAB + 64;                 // list_mode as doubles, MP
  static constexpr int GEO = MODE + MP;                 // per element 6 nodes (r,z)
  static constexpr int NID = GEO + E * 6;               // per element 6 node indices (int)
  static constexpr int BIG = NID + (E * 6 + 3) / 4;
  static constexpr int ROW6 = L * 6 + 2;                // doubles per node slot, 6 columns
  static constexpr int ROW2 = L * 2 + 2;                // doubles per node slot, 2 columns
  static constexpr int NOD6 = E * 6 * ROW6;             // doubles of one staged 6-column field
  static constexpr int NOD2 = E * 6 * ROW2;
  // per-wave region in doubles: 2 exchanged signals (2 x 256 complex) or the
  // 64 lanes x 16 doubles of the partner swap, whichever is larger
  static constexpr int RW = (4 * Shape<MP>::WAVE_CPLX > 64 * 16) ? 4 * Shape<MP>::WAVE_CPLX : 64 * 16;
  // nodal staging and the wave regions do not overlap
  static constexpr int NODC = NOD6 + (NCB == 6 ? NOD6 : NCB == 2 ? NOD2 : 0);   // doubles per chunk
  static constexpr int NODAL = NCH * NODC;              // doubles
  static constexpr int XCH = NODAL;                     // doubles from BIG: first wave region
  static constexpr int TOTAL = BIG + (NODAL + NW * RW + 1) / 2;
};

// LDS per CU (bytes) the work-group of the fused kernel may use.
constexpr int GAUSS_LDS_BYTES = 160 * 1024;

// Layout of the fused kernel for variant V.  By default the nodal data is
// staged one chunk at a time into two alternating buffers, with the next
// chunk prefetched into registers, so its loads overlap the interpolation
// of the current one and one barrier per chunk suffices.  With
// STAGE_ALL (used by the batched launch, measured faster there) all four
// chunks are staged at once where that fits in LDS and the variant makes a
// single pass over its fields: one barrier, every load in flight together.
template <int MP, int V, int NW, bool STAGE_ALL = false>
struct GaussCfg {
  static constexpr int NCB = (V == GV_UGRADC) ? 2 : ((V == GV_PAIRS || V == GV_CURL_CROSS) ? 6 : 0);
  static constexpr bool ALL =
      STAGE_ALL && V != GV_PAIRS && GaussLds<MP, NW, 4, NCB>::TOTAL * 16 <= GAUSS_LDS_BYTES;
  // chunk buffers when not staging all at once: two where they fit
  static constexpr int NBUF = GaussLds<MP, NW, 2, NCB>::TOTAL * 16 <= GAUSS_LDS_BYTES ? 2 : 1;
  using LY = GaussLds<MP, NW, ALL ? 4 : NBUF, NCB>;
};

struct GaussGeom {
  const double* w;       // shape values at this Gauss point (LDS table)
  const double* d;       // reference derivatives (dx, dy) pairs (LDS table)
  double a11, a12, a21, a22;   // d/dr = a11 d/dx + a12 d/dy, d/dz = a21 d/dx + a22 d/dy
  double iray;
};

// Interpolation at one Gauss point and one mode from a staged node row set:
// values and reference derivatives, then the affine map to (r, z) derivatives
// (dw = J^{-T} dd, gauss_points_2d.F90:108-121).
// nod: staged block; s0: first node slot of the element; row: doubles per slot.
template <int NC, int MP>
SFX_DEV void interp_mode(const double* nod, int s0, int row, int mi, const GaussGeom& g,
                         double (&W)[NC], double (&Dr)[NC], double (&Dz)[NC]) {
  double Dx[NC], Dy[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) { W[k] = 0; Dx[k] = 0; Dy[k] = 0; }
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    // column-major slot: the lanes of a point read consecutive modes
    const double* v = nod + (s0 + n) * row + mi;
    const double wn = g.w[n], dx = g.d[2 * n], dy = g.d[2 * n + 1];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const double x = v[k * (MP / 4)];
      W[k] = fma(wn, x, W[k]);
      Dx[k] = fma(dx, x, Dx[k]);
      Dy[k] = fma(dy, x, Dy[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    Dr[k] = fma(g.a11, Dx[k], g.a12 * Dy[k]);
    Dz[k] = fma(g.a21, Dx[k], g.a22 * Dy[k]);
  }
}

// Curl in (cos, sin) form (subroutine_ns_with_velocity.F90:594-608).
SFX_DEV void curl6(const double (&W)[6], const double (&Dr)[6], const double (&Dz)[6], double mr,
                   double ir, double (&R)[6]) {
  R[0] = mr * W[5] - Dz[2];
  R[3] = Dz[1] - Dr[5];
  R[4] = ir * W[2] + Dr[2] - mr * W[1];
  R[1] = -mr * W[4] - Dz[3];
  R[2] = Dz[0] - Dr[4];
  R[5] = ir * W[3] + Dr[3] + mr * W[0];
}

// Geometry of Gauss point l of the element with node coordinates geo[0..5]:
// the Jacobian of the isoparametric map from all six nodes at this point,
// dr(k,k1) = SUM(rr(k,jj(:,m))*dd(k1,:,l)), inverted as dw = J^{-T} dd
// (gauss_points_2d.F90:108-121), so curved P2 elements (midpoints moved onto
// a curved boundary, refine_mesh.F90:342-353) are exact; and the radius
// SUM(rr(1,jj(:,m))*ww(:,l)) (subroutine_ns_with_velocity.F90:574).
SFX_DEV void gauss_geom(const double2* geo, const double* tab, int l, GaussGeom& g) {
  g.w = tab + 6 * l;
  g.d = tab + 42 + 12 * l;
  double rn[6], zn[6], ray = 0;
#pragma unroll
  for (int n = 0; n < 6; ++n) {
    rn[n] = geo[n].x;
    zn[n] = geo[n].y;
    ray = fma(rn[n], g.w[n], ray);
  }
  p2_inverse_jacobian(rn, zn, g.d, g.a11, g.a12, g.a21, g.a22);
  g.iray = 1.0 / ray;
}

// g[i] for a run-time i, field by field through selects (indexing the array
// with a run-time value would place it in scratch memory).
template <int N>
SFX_DEV GaussGeom pick_geom(const GaussGeom (&g)[N], int i) {
  GaussGeom r = g[0];
#pragma unroll
  for (int k = 1; k < N; ++k) {
    const bool on = (i == k);
    r.w = on ? g[k].w : r.w;
    r.d = on ? g[k].d : r.d;
    r.a11 = on ? g[k].a11 : r.a11;
    r.a12 = on ? g[k].a12 : r.a12;
    r.a21 = on ? g[k].a21 : r.a21;
    r.a22 = on ? g[k].a22 : r.a22;
    r.iray = on ? g[k].iray : r.iray;
  }
  return r;
}

template <typename T, int N>
SFX_DEV T pick(const T (&v)[N], int i) {
  T r = v[0];
#pragma unroll
  for (int k = 1; k < N; ++k) r = (i == k) ? v[k] : r;
  return r;
}

// Stage the nodal values of modes [c*L, c*L+L) of a (np, NC, M) array for
// the E elements of this work-group: nod[(slot)*row + mi*NC + k].
template <int NC, int MP, int NW>
struct NodalFetch {
  static constexpr int TOT = GaussLds<MP>::E * 6 * NC * (MP / 4);
  static constexpr int NT = 64 * NW;
  static constexpr int IT = (TOT + NT - 1) / NT;
  double v[IT];
};

// Issue the loads of chunk c (modes [c*L, c*L+L)) of a (np, NC, M) nodal array
// for the E elements of this work-group into registers; unconditional loads at
// clamped indices, so they all issue back to back and overlap other work.
template <int NC, int MP, int NW>
SFX_DEV void nodal_fetch(const GaussArgs& a, const double* X, int64_t nX, int64_t e0, int c, int tid,
                         const int* nid, NodalFetch<NC, MP, NW>& f) {
  using NF_ = NodalFetch<NC, MP, NW>;
  constexpr int L = MP / 4, S = GaussLds<MP>::E * 6;
#pragma unroll
  for (int it = 0; it < NF_::IT; ++it) {
    const int idx = tid + it * NF_::NT;
    const int id = idx < NF_::TOT ? idx : 0;
    const int s = id % S;
    const int rest = id / S;
    const int k = rest % NC;
    const int mi = rest / NC;
    const int64_t e = e0 + s / 6;
    const int m = c * L + mi;
    const bool ok = idx < NF_::TOT && e < a.me && m < a.M;
    const int64_t gi = (int64_t)nid[s] + a.np * (k + (int64_t)NC * (m < a.M ? m : 0));
    SFX_CHECK("nodal load", gi, nX);
    const double x = X[gi];
    f.v[it] = ok ? x : 0.0;
  }
}

// Store a fetched chunk to LDS: nod[slot*row + k*L + mi] (columns of L
// modes, so that the lanes of a point read consecutive doubles).
template <int NC, int MP, int NW>
SFX_DEV void nodal_store(const NodalFetch<NC, MP, NW>& f, int tid, double* nod, int row) {
  using NF_ = NodalFetch<NC, MP, NW>;
  constexpr int S = GaussLds<MP>::E * 6;
#pragma unroll
  for (int it = 0; it < NF_::IT; ++it) {
    const int idx = tid + it * NF_::NT;
    if (idx < NF_::TOT) {
      const int s = idx % S;
      const int rest = idx / S;
      const int li = s * row + (rest % NC) * (MP / 4) + rest / NC;
      SFX_CHECK("nodal stage", li, S * row);
      nod[li] = f.v[it];
    }
  }
}

// Exchange NF doubles per lane with the partner lane (mode MP - m <-> m).
// Layout xf[f * 64 + lane]: consecutive lanes, consecutive doubles.
template <int NF, int MP>
SFX_DEV void partner_swap(const double (&own)[NF], double (&par)[NF], int pt, int l, double* xf) {
  constexpr int L = MP / 4;
  const int me_ = pt * L + l;
  const int pa = pt * L + ((L - l) % L);
#pragma unroll
  for (int f = 0; f < NF; ++f) { SFX_CHECK("field swap", f * 64 + me_, 64 * NF); xf[f * 64 + me_] = own[f]; }
  SFX_WAVE_SYNC();
#pragma unroll
  for (int f = 0; f < NF; ++f) { SFX_CHECK("field swap", f * 64 + pa, 64 * NF); par[f] = xf[f * 64 + pa]; }
  SFX_WAVE_SYNC();
}

// Accumulate the packed spectra of complex signal (f1, f2) from the own and
// partner (cos, sin) values of chunk c.  own/par: [a1, b1, a2, b2].
template <int MP>
SFX_DEV void pack_chunk(int c, int l, const double (&o)[4], const double (&p)[4], double2 (&Z)[4],
                        double2 (&B)[4]) {
  const bool m0 = (c == 0 && l == 0);
  const double2 zv = make_double2(m0 ? o[0] : 0.5 * (o[0] + o[3]), m0 ? o[2] : 0.5 * (o[2] - o[1]));
  const double2 bv = make_double2(0.5 * (p[0] - p[3]), 0.5 * (p[1] + p[2]));
  // lanes l >= 1: partner of slot 3-c; lane 0: partner of slot 4-c (c >= 1).
  // Selects over all four slots keep the arrays in registers even where c is
  // not a compile-time constant.
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    Z[r] = sel2(r == c, zv, Z[r]);
    const bool tb = (l != 0) ? (r == 3 - c) : (c >= 1 && r == 4 - c);
    B[r] = sel2(tb, bv, B[r]);
  }
}

// Real fields (cos, sin pairs) of one pass at own mode; NF = 4*NSIG doubles
// ordered signal by signal as [a1, b1, a2, b2].  nodA holds the field whose
// curl is taken (or u for u.grad c), nodB the second field (nullptr: same).
template <int V, int MP>
SFX_DEV void pass_fields(const GaussArgs& a, const double* nodA, const double* nodB, int s0, int mi,
                         double mval, const GaussGeom& g, double hc, double (&F)[4 * GVT<V>::NSIG]) {
  using LY = GaussLds<MP>;
  const double mr = mval * g.iray;
  if constexpr (V == GV_CROSS_CURL || V == GV_PAIRS || V == GV_CURL_CROSS) {
    double W[6], Dr[6], Dz[6], R[6];
    interp_mode<6, MP>(nodA, s0, LY::ROW6, mi, g, W, Dr, Dz);
    curl6(W, Dr, Dz, mr, g.iray, R);
    if constexpr (V == GV_CROSS_CURL) {
      // precession: RotV += 2 Omega (mode 1: r_c and theta_s, mode 0: z_c)
      if (mval == 1.0) { R[0] += a.om[0]; R[3] += a.om[1]; }
      if (mval == 0.0) { R[4] += a.om[2]; }
    }
    if ((V == GV_PAIRS || V == GV_CURL_CROSS) && nodB != nullptr) {
      double Drb[6], Dzb[6];
      interp_mode<6, MP>(nodB, s0, LY::ROW6, mi, g, W, Drb, Dzb);
    }
    // signals (R_r, R_theta), (R_z, W_r), (W_theta, W_z)
    F[0] = R[0]; F[1] = R[1]; F[2] = R[2]; F[3] = R[3];
    F[4] = R[4]; F[5] = R[5]; F[6] = W[0]; F[7] = W[1];
    F[8] = W[2]; F[9] = W[3]; F[10] = W[4]; F[11] = W[5];
  } else {  // GV_UGRADC
    double W[6], Dr[6], Dz[6], C[2], Cr[2], Cz[2];
    interp_mode<6, MP>(nodA, s0, LY::ROW6, mi, g, W, Dr, Dz);
    interp_mode<2, MP>(nodB, s0, LY::ROW2, mi, g, C, Cr, Cz);
    const double d0 = Dr[0] + W[0] * g.iray + mr * W[3] + Dz[4];
    const double d1 = Dr[1] + W[1] * g.iray - mr * W[2] + Dz[5];
    // signals (G_r, G_theta), (G_z, W_r), (W_theta, W_z), (Div, C)
    F[0] = hc * Cr[0]; F[1] = hc * Cr[1]; F[2] = hc * (mr * C[1]); F[3] = hc * (-mr * C[0]);
    F[4] = hc * Cz[0]; F[5] = hc * Cz[1]; F[6] = W[0]; F[7] = W[1];
    F[8] = W[2]; F[9] = W[3]; F[10] = W[4]; F[11] = W[5];
    F[12] = d0; F[13] = d1; F[14] = hc * C[0]; F[15] = hc * C[1];
  }
}

// Per work-group tables in LDS: twiddles, reference-element tables, mode
// numbers, node coordinates and node numbers of the E elements.  All loads
// issue together (clamped indices), then the LDS stores.
template <int MP, int NW>
SFX_DEV void gauss_prologue(const GaussArgs& a, int tid, int64_t e0, double2* smem) {
  using LY = GaussLds<MP, NW>;
  constexpr int E = LY::E, NT = 64 * NW;
  double* smd = reinterpret_cast<double*>(smem);
  double2* tw = smem + LY::TW;
  double* tab = smd + 2 * LY::TAB;
  double* lmode = smd + 2 * LY::MODE;
  double2* geo = smem + LY::GEO;
  int* nid = reinterpret_cast<int*>(smem + LY::NID);
  {
    constexpr int ITW = (4 * MP + NT - 1) / NT;
    constexpr int ITG = (E * 6 + NT - 1) / NT;
    double2 twv[ITW], rz[ITG];
    int node[ITG];
#pragma unroll
    for (int it = 0; it < ITW; ++it) {
      const int i = tid + it * NT;
      const int ic = i < 4 * MP ? i : 0;
      twv[it] = (ic < 3 * MP) ? a.wN[ic] : a.wMP[ic - 3 * MP];
    }
    const double tv = a.tab[tid < 126 ? tid : 0];
    const int lmv = a.list_mode[tid < a.M ? tid : 0];
#pragma unroll
    for (int it = 0; it < ITG; ++it) {
      const int i = tid + it * NT;
      const int ig = i < E * 6 ? i : 0;
      const int64_t eg = e0 + ig / 6;
      const int64_t egc = eg < a.me ? eg : 0;
      node[it] = a.jj[6 * egc + ig % 6] - 1;
      SFX_CHECK("rr load", 2 * (int64_t)node[it] + 1, 2 * a.np);
      rz[it] = make_double2(a.rr[2 * (int64_t)node[it]], a.rr[2 * (int64_t)node[it] + 1]);
    }
#pragma unroll
    for (int it = 0; it < ITW; ++it) {
      const int i = tid + it * NT;
      if (i < 4 * MP) tw[i] = twv[it];
    }
    if (tid < 126) tab[tid] = tv;
    if (tid < MP) lmode[tid] = (tid < a.M) ? (double)lmv : 0.0;
#pragma unroll
    for (int it = 0; it < ITG; ++it) {
      const int i = tid + it * NT;
      if (i < E * 6) {
        const bool ok = e0 + i / 6 < a.me;
        geo[i] = sel2(ok, rz[it], make_double2(1.0, 0.0));
        nid[i] = ok ? node[it] : 0;
      }
    }
  }
}

// RHS: instead of storing V_out at the Gauss points, form the element sums
// of the assembly in the work-group, through LDS one mode slot at a time,
// in the order of the element kernels; a gather then sums the nodes'
// entries.  Vector products: rhs_3x3 (assembling_rhs.F90:73-95, as
// rhs_elem_body), elo(6e+i, k, m) = SUM_l (ww(i,l) wgt(e,l)) V_out(7e+l, k, m),
// gathered into vb_145/vb_236 (rhs_gather_block).  u.grad c: qs_00_gauss
// of the Gauss source alone (fem_rhs_axi.F90:958-975, as qs00_elem_body),
// elo(6e+i, k + 2m) = SUM_l ww(i,l) (c_out(7e+l, k, m) wgt(e,l)), gathered
// into vect(np, 2, m) (qs00_gather_block).
// PT: point tiles.  The work-group takes NW * PW consecutive Gauss points
// (one group per wave) instead of E whole elements: the points of at most
// E elements, the first one e0 = p0 / 7, all of them staged.  For the
// Gauss-point output only (the element sums of RHS need whole elements).
template <int MP, int V, int NW, bool STAGE_ALL = false, bool RHS = false, bool PT = false>
SFX_DEV void gauss_body(const GaussArgs& a, int tid, int64_t bid, double2* smem) {
  using S = Shape<MP>;
  using LY = typename GaussCfg<MP, V, NW, STAGE_ALL>::LY;
  constexpr bool ALL = GaussCfg<MP, V, NW, STAGE_ALL>::ALL;
  constexpr int L = S::L, PW = S::PW, E = LY::E, NG = PT ? 1 : groups_per_wave(NW);
  static_assert(!PT || (!RHS && !STAGE_ALL && (NW * PW + 5) / 7 + 1 <= E), "point tiles: Gauss output only");
  constexpr int NSIG = GVT<V>::NSIG, NO = GVT<V>::NOUT;
  // Outputs staged in LDS and stored in whole runs of the work-group's points
  // per plane (one group per wave): the point tiles, and the single-pass
  // (curl a) x b of the pass lists (tangent at 50,000 triangles, M = 32:
  // 4.22 -> 4.06 ms; neutral at M = 64; u.grad c gains nothing at M = 64).
  constexpr bool SOUT = !RHS && !STAGE_ALL && NG == 1 && (PT || V == GV_CURL_CROSS);
  constexpr int NF = 4 * NSIG;
  constexpr int NOC = (NO + 1) / 2;
  constexpr int GX = 2;
  constexpr int NT = 64 * NW;
  const int wave = tid >> 6, lane = tid & 63;
  const int pt = lane / L, l = lane % L;
  double* smd = reinterpret_cast<double*>(smem);

  // ---- per work-group tables
  double2* tw = smem + LY::TW;
  double* tab = smd + 2 * LY::TAB;
  double* lmode = smd + 2 * LY::MODE;
  double2* geo = smem + LY::GEO;
  int* nid = reinterpret_cast<int*>(smem + LY::NID);
  const int64_t p0 = PT ? bid * (int64_t)(NW * PW) : bid * (int64_t)(7 * E);   // first Gauss point
  const int64_t e0 = p0 / 7;
  gauss_prologue<MP, NW>(a, tid, e0, smem);
  static_assert(126 <= NT && MP <= NT, "prologue assumes one table entry per thread");
  SFX_SYNC();
  const double2* wN = tw;
  const double2* wMP = tw + 3 * MP;

  // ---- this lane's Gauss point in each of the wave's groups (group wave + NW*gi)
  GaussGeom g[NG];
  bool has[NG], valid[NG];
  int el[NG];
  int64_t p[NG];
  double hc[NG];
#pragma unroll
  for (int gi = 0; gi < NG; ++gi) {
    const int grp = wave + NW * gi;
    has[gi] = PT || grp < GAUSS_GROUPS;
    const int lp = (has[gi] ? grp : 0) * PW + pt;    // local point
    const int64_t pg = p0 + lp;
    el[gi] = (int)(pg / 7 - e0);
    const int lg = (int)(pg % 7);
    p[gi] = pg;
    valid[gi] = has[gi] && pg < 7 * a.me;
    gauss_geom(geo + 6 * el[gi], tab, valid[gi] ? lg : 0, g[gi]);
    hc[gi] = (V == GV_UGRADC && valid[gi]) ? a.heat_capa[e0 + el[gi]] : 1.0;
  }

  double* big = smd + 2 * LY::BIG;
  double* xf = big + LY::XCH + (size_t)wave * LY::RW;
  double2* lds = reinterpret_cast<double2*>(xf);

  // Groups are handled together (NGS = NG) or, for the pass list, one per
  // sweep over the passes (NGS = 1), which halves the live state at the cost
  // of staging the nodal chunks once per group.
  constexpr bool GROUP_OUTER = (V == GV_PAIRS);
  constexpr int NGS = GROUP_OUTER ? 1 : NG;
  const int npass = (V == GV_PAIRS) ? a.npass : 1;
  static_assert(!RHS || ((V == GV_CROSS_CURL || V == GV_UGRADC) && !STAGE_ALL), "RHS: cross-curl or u.grad c");
  // RHS: the coefficients of every group (NKR columns), kept until the element sums
  constexpr int NKR = 2 * NO;
  double ro[RHS ? NG : 1][RHS ? NKR : 1][4];
  for (int go = 0; go < (GROUP_OUTER ? NG : 1); ++go) {
    double acc[NGS][3][4][3];   // physical product accumulated over the passes
#pragma unroll
    for (int gs = 0; gs < NGS; ++gs)
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[gs][j][r][0] = acc[gs][j][r][1] = acc[gs][j][r][2] = 0.0;
    double2 q[NGS][3 * NOC][4];
    for (int pass = 0; pass < npass; ++pass) {
      double2 Z[NGS][NSIG][4], B[NGS][NSIG][4];
#pragma unroll
      for (int gs = 0; gs < NGS; ++gs)
#pragma unroll
        for (int s = 0; s < NSIG; ++s)
#pragma unroll
          for (int r = 0; r < 4; ++r) { Z[gs][s][r] = make_double2(0, 0); B[gs][s][r] = make_double2(0, 0); }
      // nodal arrays of this pass: A (6 columns) and B (6 or 2 columns, or none)
      constexpr int NCB = (V == GV_UGRADC) ? 2 : 6;
      // (the pass entries are picked with selects: indexing the argument
      // arrays with the loop counter would copy them to scratch memory)
      const double* XA = a.U;
      const double* XB = a.U2;
      int64_t nXA = a.nU, nXB = a.nU2;
      double coef = 1.0;
      if (V == GV_CURL_CROSS) {
        XA = a.pc[0]; XB = a.pv[0]; nXA = a.npc[0]; nXB = a.npv[0]; coef = a.coef[0];
      }
      if (V == GV_PAIRS) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const bool on = (pass == k);
          XA = on ? a.pc[k] : XA;
          XB = on ? a.pv[k] : XB;
          nXA = on ? a.npc[k] : nXA;
          nXB = on ? a.npv[k] : nXB;
          coef = on ? a.coef[k] : coef;
        }
      }
      const bool useB = (V == GV_UGRADC) || ((V == GV_PAIRS || V == GV_CURL_CROSS) && XB != XA);
      NodalFetch<6, MP, NW> fa[ALL ? 4 : 1];
      NodalFetch<NCB, MP, NW> fb[ALL ? 4 : 1];
      if constexpr (ALL) {
        // every chunk at once: all loads issue back to back, one barrier
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          nodal_fetch<6, MP, NW>(a, XA, nXA, e0, c, tid, nid, fa[c]);
          if (useB) nodal_fetch<NCB, MP, NW>(a, XB, nXB, e0, c, tid, nid, fb[c]);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          nodal_store<6, MP, NW>(fa[c], tid, big + c * LY::NODC, LY::ROW6);
          if (useB) nodal_store<NCB, MP, NW>(fb[c], tid, big + c * LY::NODC + LY::NOD6, NCB == 6 ? LY::ROW6 : LY::ROW2);
        }
        SFX_SYNC();
      } else {
        nodal_fetch<6, MP, NW>(a, XA, nXA, e0, 0, tid, nid, fa[0]);
        if (useB) nodal_fetch<NCB, MP, NW>(a, XB, nXB, e0, 0, tid, nid, fb[0]);
      }
      constexpr int NBUF = GaussCfg<MP, V, NW, STAGE_ALL>::NBUF;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        double* nodA = big + (ALL ? c : (NBUF == 2 ? (c & 1) : 0)) * LY::NODC;
        double* nodB = nodA + LY::NOD6;
        if constexpr (!ALL) {
          // two buffers: chunk c into buffer c % 2 (its previous chunk, c - 2,
          // was interpolated by every wave before the last barrier), start
          // fetching chunk c + 1, and one barrier makes chunk c visible.  One
          // buffer: wait until every wave is done with chunk c - 1 first.
          if (NBUF == 1) SFX_SYNC();
          nodal_store<6, MP, NW>(fa[0], tid, nodA, LY::ROW6);
          if (useB) nodal_store<NCB, MP, NW>(fb[0], tid, nodB, NCB == 6 ? LY::ROW6 : LY::ROW2);
          if (NBUF == 1) SFX_SYNC();
          if (c < 3) {
            nodal_fetch<6, MP, NW>(a, XA, nXA, e0, c + 1, tid, nid, fa[0]);
            if (useB) nodal_fetch<NCB, MP, NW>(a, XB, nXB, e0, c + 1, tid, nid, fb[0]);
          }
          if (NBUF == 2) SFX_SYNC();
        }
        const int m = c * L + l;
#pragma unroll
        for (int gs = 0; gs < NGS; ++gs) {
          const int gi = GROUP_OUTER ? go : gs;
          if (!pick(has, gi)) continue;   // uniform over the wave
          double own[NF];
          if (m < a.M && pick(valid, gi)) {
            pass_fields<V, MP>(a, nodA, useB ? nodB : nullptr, 6 * pick(el, gi), l, lmode[m], pick_geom(g, gi),
                               pick(hc, gi), own);
          } else {
#pragma unroll
            for (int f = 0; f < NF; ++f) own[f] = 0.0;
          }
          double par[NF];
          partner_swap<NF, MP>(own, par, pt, l, xf);
#pragma unroll
          for (int s = 0; s < NSIG; ++s) {
            const double o4[4] = {own[4 * s], own[4 * s + 1], own[4 * s + 2], own[4 * s + 3]};
            const double p4[4] = {par[4 * s], par[4 * s + 1], par[4 * s + 2], par[4 * s + 3]};
            pack_chunk<MP>(c, l, o4, p4, Z[gs][s], B[gs][s]);
          }
        }
      }
      // ---- inverse transforms and pointwise op of this pass, group by group
#pragma unroll
      for (int gs = 0; gs < NGS; ++gs) {
        const int gi = GROUP_OUTER ? go : gs;
        if (!pick(has, gi)) continue;
        double2 y[3 * NSIG][4];
#pragma unroll
        for (int s = 0; s < NSIG; ++s) prestep<MP>(Z[gs][s], B[gs][s], l, y[3 * s], y[3 * s + 1], y[3 * s + 2], wN);
        inverse_fft<MP, 3 * NSIG, GX>(y, pt, l, lds, wMP);
#pragma unroll
        for (int j1 = 0; j1 < 3; ++j1)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            double u[8];
#pragma unroll
            for (int s = 0; s < NSIG; ++s) { u[2 * s] = y[3 * s + j1][r].x; u[2 * s + 1] = y[3 * s + j1][r].y; }
            if constexpr (V == GV_UGRADC) {
              const double o0 = u[0] * u[3] + u[1] * u[4] + u[2] * u[5] + u[6] * u[7];
              q[gs][j1][r] = make_double2(o0, 0.0);
            } else {
              const double cx = u[1] * u[5] - u[2] * u[4];
              const double cy = u[2] * u[3] - u[0] * u[5];
              const double cz = u[0] * u[4] - u[1] * u[3];
              acc[gs][j1][r][0] = fma(coef, cx, acc[gs][j1][r][0]);
              acc[gs][j1][r][1] = fma(coef, cy, acc[gs][j1][r][1]);
              acc[gs][j1][r][2] = fma(coef, cz, acc[gs][j1][r][2]);
            }
          }
      }
    }
    // ---- forward transforms, pairing and stores, group by group
#pragma unroll
    for (int gs = 0; gs < NGS; ++gs) {
      const int gi = GROUP_OUTER ? go : gs;
      if (!pick(has, gi)) continue;
      if constexpr (V != GV_UGRADC) {
#pragma unroll
        for (int j1 = 0; j1 < 3; ++j1)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            q[gs][j1][r] = make_double2(acc[gs][j1][r][0], acc[gs][j1][r][1]);
            q[gs][3 + j1][r] = make_double2(acc[gs][j1][r][2], 0.0);
          }
      }
      forward_fft<MP, 3 * NOC, GX>(q[gs], pt, l, lds, wMP);
      double2 lo[NOC][4], hi[NOC][4];
#pragma unroll
      for (int t = 0; t < NOC; ++t) poststep<MP>(q[gs][3 * t], q[gs][3 * t + 1], q[gs][3 * t + 2], l, lo[t], hi[t], wN);
      double oa[NOC][2][4], ob[NOC][2][4];
      pair_modes<MP, NOC>(lo, hi, pt, l, lds, oa, ob);
      if constexpr (RHS) {
        // column 2o + h' of component o: cos (h' = 0) and sin (h' = 1)
#pragma unroll
        for (int o = 0; o < NO; ++o)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            ro[gi][2 * o][r] = oa[o / 2][o % 2][r];
            ro[gi][2 * o + 1][r] = ob[o / 2][o % 2][r];
          }
        continue;
      }
      if constexpr (SOUT) {
        // the work-group's outputs go through LDS (st[P * SR + lp] for
        // output plane P = 2o + h + 2 NO m), so that the stores below write
        // whole runs of the tile's points per plane
        constexpr int TP = PT ? NW * PW : 7 * E, SR = TP + 1;
        static_assert(2 * NO * MP * SR <= LY::NODAL + NW * LY::RW, "output staging");
        double* st = big;
        const int lp = wave * PW + pt;
        SFX_SYNC();   // every wave is done with the staging and wave regions
#pragma unroll
        for (int t = 0; t < NOC; ++t)
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int o = 2 * t + h;
            if (o >= NO) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int P = 2 * o + 2 * NO * (l + L * r);
              SFX_CHECK("out stage", (P + 1) * SR + lp, 2 * NO * MP * SR);
              st[P * SR + lp] = oa[t][h][r];
              st[(P + 1) * SR + lp] = ob[t][h][r];
            }
          }
        continue;
      }
      if (!pick(valid, gi)) continue;
      const int64_t pg = pick(p, gi);
#pragma unroll
      for (int t = 0; t < NOC; ++t)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int o = 2 * t + h;
          if (o >= NO) continue;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int m = l + L * r;
            if (m >= a.M) continue;
            const int64_t base = pg + a.G * (int64_t)(2 * o + 2 * NO * m);
            SFX_CHECK("gauss out", base + a.G, a.nout);
            if (V == GV_CURL_CROSS && a.accum) {
              a.Vout[base] += oa[t][h][r];
              a.Vout[base + a.G] += ob[t][h][r];
            } else {
              a.Vout[base] = oa[t][h][r];
              a.Vout[base + a.G] = ob[t][h][r];
            }
          }
        }
    }
  }
  if constexpr (SOUT) {
    constexpr int TP = PT ? NW * PW : 7 * E, SR = TP + 1;
    const double* st = big;
    SFX_SYNC();
    const int items = 2 * NO * a.M * TP;
    for (int it = tid; it < items; it += NT) {
      const int lp = it % TP, P = it / TP;
      const int64_t pg = p0 + lp;
      if (pg >= 7 * a.me) continue;
      SFX_CHECK("out stage read", P * SR + lp, 2 * NO * MP * SR);
      SFX_CHECK("gauss out", pg + a.G * (int64_t)P, a.nout);
      if (V == GV_CURL_CROSS && a.accum) a.Vout[pg + a.G * (int64_t)P] += st[P * SR + lp];
      else a.Vout[pg + a.G * (int64_t)P] = st[P * SR + lp];
    }
  }
  if constexpr (RHS) {
    // red[(lp*L + mi)*NKR + k]: column k of local point lp at mode offset mi
    // of the current slot; wl[lp] = wgt(e, l) of the point.  Both alias the
    // staging and wave regions, which every wave has finished with at the
    // first barrier.
    constexpr int NPT = 7 * E;
    double* red = big;
    double* wl = big + NPT * L * NKR;
    static_assert(NPT * L * NKR + NPT <= LY::NODAL + NW * LY::RW, "RHS reduction area");
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      SFX_SYNC();
#pragma unroll
      for (int gi = 0; gi < NG; ++gi) {
        if (!has[gi]) continue;
        const int lp = (wave + NW * gi) * PW + pt;
        const bool ok = valid[gi] && l + L * r < a.M;
#pragma unroll
        for (int k = 0; k < NKR; ++k) {
          SFX_CHECK("rhs red write", (lp * L + l) * NKR + k, NPT * L * NKR);
          red[(lp * L + l) * NKR + k] = ok ? ro[gi][k][r] : 0.0;
        }
        if (r == 0 && l == 0) {
          const int64_t wi = valid[gi] ? 7 * e0 + lp : 0;   // points past the last element read entry 0
          SFX_CHECK("rhs wgt", wi, 7 * a.me);
          const double wv = a.wgt[wi];
          wl[lp] = valid[gi] ? wv : 0.0;
        }
      }
      SFX_SYNC();
      // item (element e, local node n, mode offset mi): 6 components
      for (int it = tid; it < E * 6 * L; it += NT) {
        const int mi = it % L;
        const int n = (it / L) % 6;
        const int e = it / (6 * L);
        const int mm = mi + L * r;
        const int64_t eg = e0 + e;
        if (eg >= a.me || mm >= a.M) continue;
        double acc[NKR];
#pragma unroll
        for (int k = 0; k < NKR; ++k) acc[k] = 0.0;
        // qs_00_gauss's nodal source at the element's nodes, and hc(e)
        double ffl[V == GV_UGRADC ? NKR : 1][6];
        double hce = 1.0;
        if constexpr (V == GV_UGRADC) {
          hce = a.heat_capa[eg];
#pragma unroll
          for (int k = 0; k < NKR; ++k)
#pragma unroll
            for (int j = 0; j < 6; ++j) {
              const int64_t fi = (int64_t)nid[6 * e + j] + a.np * (int64_t)(k + 2 * mm);
              SFX_CHECK("rhs ff load", fi, a.ffn ? a.nffn : 1 << 30);
              ffl[k][j] = a.ffn ? a.ffn[fi] : 0.0;
            }
        }
#pragma unroll
        for (int lg = 0; lg < 7; ++lg) {
          const double* src = red + ((7 * e + lg) * L + mi) * NKR;
          if constexpr (V == GV_UGRADC) {
#pragma unroll
            for (int k = 0; k < NKR; ++k) {
              double sf = 0;
#pragma unroll
              for (int j = 0; j < 6; ++j) sf = fma(ffl[k][j], tab[j + 6 * lg], sf);
              acc[k] = fma(tab[n + 6 * lg], (hce * sf + src[k]) * wl[7 * e + lg], acc[k]);
            }
          } else {
            const double c = tab[n + 6 * lg] * wl[7 * e + lg];
#pragma unroll
            for (int k = 0; k < NKR; ++k) acc[k] = fma(c, src[k], acc[k]);
          }
        }
#pragma unroll
        for (int k = 0; k < NKR; ++k) {
          const int64_t o = (V == GV_UGRADC) ? (6 * eg + n) * (2 * (int64_t)a.M) + k + 2 * mm
                                             : ((6 * eg + n) * 6 + k) * (int64_t)a.M + mm;
          SFX_CHECK("rhs elo store", o, a.nelo);
          a.elo[o] = acc[k];
        }
      }
    }
  }
}

}  // namespace sfx
