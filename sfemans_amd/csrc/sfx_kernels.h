// Internal kernel interfaces shared by the HIP translation units and the C-ABI.
#pragma once
#include "sfx_portable.h"

namespace sfx {

enum { OP_CROSS = 0, OP_DOT = 1, OP_PROD = 2, OP_SVECT1 = 3, OP_TENSOR = 4 };
enum { GV_CROSS_CURL = 0, GV_PAIRS = 1, GV_UGRADC = 2, GV_CURL_CROSS = 3 };

// Twiddles for one power-of-two MP: wN[j] = e^{2 pi i j/(3 MP)}, wMP[j] = e^{2 pi i j/MP}.
struct Twiddles {
  const double2* wN;
  const double2* wMP;
};

struct FftProdArgs {
  const double* V1;   // (np, nb1, M)
  const double* V2;   // (np, nb2, M)
  double* Vout;       // (np, 2*NO, M), or (3, np, 6, M) for the tensor product
  int64_t np;
  int M, MP, nb1, nb2;
  const double2* wN;
  const double2* wMP;
  int64_t n1, n2, nout;   // element counts of V1, V2, Vout (bounds checks on the CPU build)
};

// Fused Gauss-point nonlinear terms (see sfx_gauss_body.h).
struct GaussArgs {
  const int32_t* jj;        // (6, me), 1-based node numbers
  const double* rr;         // (2, np) node coordinates (r, z)
  int64_t me, np, G;        // elements, nodes, Gauss points (7 me)
  const double* U;          // first nodal field  (np, 6, M): u or H
  const double* U2;         // second nodal field (np, 6 or 2, M): B, c or H
  const double* U3;         // third nodal field  (np, 6, M): B (MHD)
  int64_t nU, nU2, nU3;     // element counts (bounds checks on the CPU build)
  const double* heat_capa;  // (me) for u.grad c
  const int32_t* list_mode; // (M) mode numbers
  const double* tab;        // ww(6,7) then dd(2,6,7) of the reference element
  double* Vout;             // (G, 2*NO, M)
  int64_t nout;
  int M, MP;
  double c_lorentz;
  int npass;                // GV_PAIRS: passes, each adds coef * (curl pc) x pv
  int accum;                // GV_CURL_CROSS: 1 adds the result to Vout, 0 stores it
  const double* pc[4];
  const double* pv[4];
  int64_t npc[4], npv[4];
  double coef[4];
  double om[3];             // precession: 2*Omega added to RotV (mode 1: r_c, theta_s; mode 0: z_c)
  const double2* wN;
  const double2* wMP;
  // element sums of rhs_3x3 formed in the kernel (the fused nodal -> RHS form):
  // elo(6e+i, k, m) = SUM_l ww(i,l) wgt(e,l) V_out(7e+l, k, m), no Gauss array
  const double* wgt;        // (7, me): rj * ray
  double* elo;              // index ((6e+i)*6 + k)*M + m (u.grad c: (6e+i)*2M + k + 2m)
  int64_t nelo;
  const double* ffn;        // u.grad c fused to qs_00_gauss: its nodal source (np, 2, M), or null
  int64_t nffn;
};

struct RhsArgs;
struct InnerArgs;
struct PcgArgs;
struct BdyArgs;
struct AdjArgs;
struct GatherArgs;
#ifndef SFX_EMULATE
hipError_t launch_adjoint(const AdjArgs& a, const GatherArgs& ga, const GatherArgs* ga2, hipStream_t st);
hipError_t launch_rhs3x3(const RhsArgs& a, hipStream_t st);
hipError_t launch_rhs_gather(const RhsArgs& a, hipStream_t st);
hipError_t launch_gauss_rhs(int variant, const GaussArgs& a, hipStream_t st);
hipError_t launch_qs00_gather(const RhsArgs& a, hipStream_t st);
hipError_t launch_rhs_ns(const RhsArgs& a, hipStream_t st);
hipError_t launch_qs00(const RhsArgs& a, hipStream_t st);
hipError_t launch_rhs_bdy(const BdyArgs& a, hipStream_t st);
hipError_t launch_rhs_ns_gauss(const RhsArgs& a, hipStream_t st);
hipError_t launch_gauss_values(const InnerArgs& a, hipStream_t st);
int64_t dotsf_blocks(int64_t me, int ncols);
hipError_t launch_dotsf(const InnerArgs& a, double* result, hipStream_t st);
hipError_t launch_coldot(const double* A, const double* B, int64_t np, int ncols, double* out, hipStream_t st);
hipError_t launch_pcg(int which, const PcgArgs& a, hipStream_t st);
hipError_t launch_scale_cols(const InnerArgs& a, hipStream_t st);
hipError_t launch_gauss(int variant, const GaussArgs& a, hipStream_t st);
hipError_t launch_fftprod(int op, const FftProdArgs& a, hipStream_t st);
// Exact-grid products (sfx_exact_body.h): angle pass, then r2c, or the maximum into max_out.
struct ExactArgs;
hipError_t launch_exact(int op, const ExactArgs& a, double* max_out, hipStream_t st);
hipError_t launch_exact_velstat(const ExactArgs& a, double* part, double* out, hipStream_t st);
int64_t exact_velstat_blocks(int nin, int64_t np, int K);
hipError_t launch_exact_checkif(const double* flag, int32_t* ok, hipStream_t st);
int64_t exact_grid_blocks(int op, int nin, int64_t np, int K);
// FFT_COMPUTE_ENTROPY_VISC_MOM: element maxima (pmax: exact_emom_blocks doubles),
// the maximum of |u| into dmax, then the real-space outputs.
hipError_t launch_exact_emom(const ExactArgs& a, double* dmax, hipStream_t st);
int64_t exact_emom_blocks(int64_t np, int lg, int K);
#endif

// Pass lists (GV_PAIRS) run as one single-pass GV_CURL_CROSS launch per
// pass, each adding to the output, where that kernel uses seven waves of 256
// registers (MP <= 64, measured faster than one launch of the pass list,
// whose accumulator needs 512 registers per wave); as one launch otherwise.
inline bool pairs_as_passes(int MP) { return MP <= 64; }

// Smallest supported power of two >= max(M, 4); 0 if M > 256.
inline int mp_for(int M) {
  int mp = 4;
  while (mp < M) mp *= 2;
  return mp <= 256 ? mp : 0;
}

}  // namespace sfx
