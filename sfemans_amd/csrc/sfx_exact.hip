// Kernels of the exact-grid products (bodies in sfx_exact_body.h).
#include "sfx_engine.h"
#include "sfx_exact_body.h"
#include "sfx_exact_setup.h"
#include "sfx_kernels.h"

namespace sfx {

constexpr int EXACT_THREADS = 256;

// Logical block L = xcd_block(...) -> (point block L / nch, angle block L % nch):
// the angle blocks of one point block run on one XCD, so the inputs they all
// read come from that XCD's L2 after the first.
template <int OP, int NI>
__global__ void __launch_bounds__(EXACT_THREADS) exact_grid_kernel(ExactArgs a, int nch) {
  const int64_t L = xcd_block(blockIdx.x, gridDim.x);
  const int64_t p = (L / nch) * EXACT_THREADS + threadIdx.x;
  double vmax = 0.0;
  exact_grid_body<OP, NI>(a, p, (int)(L % nch), vmax);
  if constexpr (exact_is_reduction(OP)) {
    __shared__ double red[EXACT_THREADS];
    red[threadIdx.x] = vmax;
    __syncthreads();
    for (int s = EXACT_THREADS / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
      __syncthreads();
    }
    if (threadIdx.x == 0) a.pmax[blockIdx.x] = red[0];
  }
}

// FFT_PAR_REAL: the values of one field at 256 points and 16 angle pairs,
// written through an LDS tile so that 16 consecutive lanes store 16
// consecutive angles of one point (128-byte runs of its row), first the
// angles j0..j0+15, then their mirrors N-j0-15..N-j0.
constexpr int EXACT_REAL_JC = 16;
__global__ void __launch_bounds__(EXACT_THREADS) exact_real_kernel(ExactArgs a, int nch) {
  __shared__ double tile[EXACT_THREADS][EXACT_REAL_JC + 1];
  const int64_t L = xcd_block(blockIdx.x, gridDim.x);
  const int64_t p0 = (L / nch) * EXACT_THREADS;
  const int j0 = (int)(L % nch) * EXACT_REAL_JC;
  const int t = threadIdx.x;
  const int64_t p = p0 + t;
  const int64_t N = 2 * a.K + 1;
  double E[1][EXACT_REAL_JC], O[1][EXACT_REAL_JC];
  if (p < a.np) {
    exact_c2r<1, EXACT_REAL_JC>(a, p, j0, E, O);
  } else {
#pragma unroll
    for (int q = 0; q < EXACT_REAL_JC; ++q) E[0][q] = O[0][q] = 0.0;
  }
  for (int half = 0; half < 2; ++half) {
#pragma unroll
    for (int q = 0; q < EXACT_REAL_JC; ++q) tile[t][q] = half == 0 ? E[0][q] + O[0][q] : E[0][q] - O[0][q];
    __syncthreads();
    for (int i = t; i < EXACT_THREADS * EXACT_REAL_JC; i += EXACT_THREADS) {
      const int pt = i / EXACT_REAL_JC, q = i % EXACT_REAL_JC;
      const int64_t pp = p0 + pt;
      const int jj = j0 + q;
      if (pp >= a.np || jj > a.K || (half == 1 && jj == 0)) continue;
      const int64_t ang = half == 0 ? jj : N - jj;
      a.rout1[pp * a.rstride + ang] = tile[pt][q];
    }
    __syncthreads();
  }
}

// Logical block -> (point block, output field and mode block).
__global__ void __launch_bounds__(EXACT_THREADS) exact_r2c_kernel(ExactArgs a, int nmc) {
  const int64_t L = xcd_block(blockIdx.x, gridDim.x);
  const int64_t per = (int64_t)a.nout * nmc;
  const int r = (int)(L % per);
  exact_r2c_body(a, (L / per) * EXACT_THREADS + threadIdx.x, r / nmc, r % nmc);
}

// Per element and angle, the maximum of |v| over the element's lg points
// (:2572-2577): a work-group holds EXACT_ENORM_E consecutive elements and one
// block of 16 angle pairs; every thread computes |v| of its point at its
// angles into LDS, then the element maxima are taken from there.
constexpr int EXACT_ENORM_E = 32;
__global__ void __launch_bounds__(512) exact_enorm_kernel(ExactArgs a, int nch, int64_t me) {
  extern __shared__ double red[];   // [thread][2 * EXACT_NORM_JC + 1]
  constexpr int RW = 2 * EXACT_NORM_JC + 1;
  const int64_t L = xcd_block(blockIdx.x, gridDim.x);
  const int64_t eb = L / nch;
  const int jc = (int)(L % nch), j0 = jc * EXACT_NORM_JC;
  const int lg = a.lg, t = threadIdx.x;
  const int64_t p = eb * EXACT_ENORM_E * lg + t;
  double nrm[2 * EXACT_NORM_JC];
  if (p < a.np) {
    exact_norm_body(a, p, j0, nrm);
  } else {
#pragma unroll
    for (int v = 0; v < 2 * EXACT_NORM_JC; ++v) nrm[v] = 0.0;
  }
#pragma unroll
  for (int v = 0; v < 2 * EXACT_NORM_JC; ++v) red[t * RW + v] = nrm[v];
  __syncthreads();
  const int N = 2 * a.K + 1;
  for (int item = t; item < EXACT_ENORM_E * 2 * EXACT_NORM_JC; item += blockDim.x) {
    const int le = item / (2 * EXACT_NORM_JC), v = item % (2 * EXACT_NORM_JC);
    const int64_t e = eb * EXACT_ENORM_E + le;
    const int q = v % EXACT_NORM_JC, j = j0 + q;
    const bool mirror = v >= EXACT_NORM_JC;
    if (e >= me || j > a.K || (mirror && j == 0)) continue;
    double x = red[(le * lg) * RW + v];
    for (int l = 1; l < lg; ++l) x = fmax(x, red[(le * lg + l) * RW + v]);
    a.nint[e * N + (mirror ? N - j : j)] = x;
  }
}

// FFT_COMPUTE_ENTROPY_VISC_MOM, first pass: per element and angle the maxima
// of |u|, |m| and the raw viscosity over the element's points (the angle
// maxima are taken when they are read), and each work-group's maximum of |u|
// into pmax.  Work-group: EXACT_ENORM_E elements x EXACT_EMOM_JC angle pairs.
__global__ void __launch_bounds__(512) exact_emom_kernel(ExactArgs a, int nch, int64_t me) {
  extern __shared__ double red[];   // [thread][3 * 2 * EXACT_EMOM_JC + 1]
  constexpr int NV = 2 * EXACT_EMOM_JC, RW = 3 * NV + 1;
  const int64_t L = xcd_block(blockIdx.x, gridDim.x);
  const int64_t eb = L / nch;
  const int j0 = (int)(L % nch) * EXACT_EMOM_JC;
  const int lg = a.lg, t = threadIdx.x;
  const int64_t p = eb * EXACT_ENORM_E * lg + t;
  double nv[NV], nm[NV], vr[NV];
  if (p < a.np) {
    exact_emom_body(a, p, j0, nv, nm, vr);
  } else {
#pragma unroll
    for (int v = 0; v < NV; ++v) nv[v] = nm[v] = vr[v] = 0.0;
  }
  double bmax = 0.0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    red[t * RW + v] = nv[v];
    red[t * RW + NV + v] = nm[v];
    red[t * RW + 2 * NV + v] = vr[v];
    bmax = fmax(bmax, nv[v]);
  }
  __syncthreads();
  const int N = 2 * a.K + 1;
  for (int item = t; item < EXACT_ENORM_E * 3 * NV; item += blockDim.x) {
    const int le = item / (3 * NV), w = item % (3 * NV), k = w / NV, v = w % NV;
    const int64_t e = eb * EXACT_ENORM_E + le;
    const int j = j0 + v % EXACT_EMOM_JC;
    const bool mirror = v >= EXACT_EMOM_JC;
    if (e >= me || j > a.K || (mirror && j == 0)) continue;
    double x = red[(le * lg) * RW + w];
    for (int l = 1; l < lg; ++l) x = fmax(x, red[(le * lg + l) * RW + w]);
    double* dst = k == 0 ? a.nint : k == 1 ? a.nint2 : a.nint3;
    dst[e * N + (mirror ? N - j : j)] = x;
  }
  __syncthreads();
  // the work-group's maximum of |u| (reuses the first column of red)
  red[t * RW] = bmax;
  __syncthreads();
  if (t == 0) {
    double m = 0.0;
    for (int i = 0; i < (int)blockDim.x; ++i) m = fmax(m, red[i * RW]);
    a.pmax[blockIdx.x] = m;
  }
}

// Second pass: one thread per (point, angle), angles fastest.
__global__ void __launch_bounds__(256) exact_emom_out_kernel(ExactArgs a) {
  const int N = 2 * a.K + 1;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= a.np * N) return;
  exact_emom_out_body(a, i / N, (int)(i % N));
}

// Maximum of n values into out[0] (one work-group).
__global__ void __launch_bounds__(EXACT_THREADS) exact_max_kernel(const double* v, int64_t n, double* out,
                                                                  double init) {
  __shared__ double red[EXACT_THREADS];
  double m = init;
  for (int64_t i = threadIdx.x; i < n; i += EXACT_THREADS) m = fmax(m, v[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = EXACT_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// FFT_MAX_MIN_VEL_DCL in one pass: three maxima per work-group into
// part[3*block + k], then exact_velstat_final_kernel.
template <int NI>
__global__ void __launch_bounds__(EXACT_THREADS) exact_velstat_kernel(ExactArgs a, int nch, double* part) {
  const int64_t L = xcd_block(blockIdx.x, gridDim.x);
  const int64_t p = (L / nch) * EXACT_THREADS + threadIdx.x;
  double r[3] = {0.0, -INFINITY, 0.0};
  exact_velstat_body<NI>(a, p, (int)(L % nch), r);
  __shared__ double red[3][EXACT_THREADS];
#pragma unroll
  for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = r[k];
  __syncthreads();
  for (int s = EXACT_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = fmax(red[k][threadIdx.x], red[k][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x < 3) part[3 * (int64_t)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// out[0] = max |v|; out[1] = min of the positive |v| if |v| > 0 somewhere on
// the last angle, else 0 (the reference's loops, fft_parallel.F90:4563-4576).
__global__ void __launch_bounds__(EXACT_THREADS) exact_velstat_final_kernel(const double* part, int64_t n,
                                                                            double* out) {
  __shared__ double red[3][EXACT_THREADS];
  double r[3] = {0.0, -INFINITY, 0.0};
  for (int64_t i = threadIdx.x; i < n; i += EXACT_THREADS)
#pragma unroll
    for (int k = 0; k < 3; ++k) r[k] = fmax(r[k], part[3 * i + k]);
#pragma unroll
  for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = r[k];
  __syncthreads();
  for (int s = EXACT_THREADS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
#pragma unroll
      for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = fmax(red[k][threadIdx.x], red[k][threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = red[0][0];
    out[1] = red[2][0] >= 1.0 ? -red[1][0] : 0.0;
  }
}

int64_t exact_velstat_blocks(int nin, int64_t np, int K) { return exact_grid_blocks(EX_MAXVEL, nin, np, K); }

hipError_t launch_exact_velstat(const ExactArgs& a, double* part, double* out, hipStream_t st) {
  const int JC = exact_jc(EX_MAXVEL, a.nin);
  const int nch = (a.K + 1 + JC - 1) / JC;
  const int64_t nb = exact_grid_blocks(EX_MAXVEL, a.nin, a.np, a.K);
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  if (nb > 0) {
    if (a.nin == 3) hipLaunchKernelGGL(exact_velstat_kernel<3>, dim3((unsigned)nb), dim3(EXACT_THREADS), 0, st, a,
                                       nch, part);
    else hipLaunchKernelGGL(exact_velstat_kernel<1>, dim3((unsigned)nb), dim3(EXACT_THREADS), 0, st, a, nch, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(exact_velstat_final_kernel, dim3(1), dim3(EXACT_THREADS), 0, st, (const double*)part, nb, out);
  return hipGetLastError();
}

// FFT_CHECK_INTERFACE: interface_ok = 0 if a violation was found (flag null:
// nothing was checked).
__global__ void exact_checkif_kernel(const double* flag, int32_t* ok) {
  if (threadIdx.x == 0) ok[0] = (flag != nullptr && flag[0] >= 1.0) ? 0 : 1;
}

hipError_t launch_exact_checkif(const double* flag, int32_t* ok, hipStream_t st) {
  hipLaunchKernelGGL(exact_checkif_kernel, dim3(1), dim3(64), 0, st, flag, ok);
  return hipGetLastError();
}

// Blocks of the angle pass (also the size of a reduction's partial array).
int64_t exact_grid_blocks(int op, int nin, int64_t np, int K) {
  const int JC = exact_jc(op, nin);
  return ((np + EXACT_THREADS - 1) / EXACT_THREADS) * ((K + 1 + JC - 1) / JC);
}

hipError_t launch_exact_emom(const ExactArgs& a, double* dmax, hipStream_t st) {
  const int64_t me = a.np / a.lg;
  const int nch = (a.K + 1 + EXACT_EMOM_JC - 1) / EXACT_EMOM_JC;
  const int64_t nb = ((me + EXACT_ENORM_E - 1) / EXACT_ENORM_E) * nch;
  const int threads = EXACT_ENORM_E * a.lg;
  if (a.lg < 1 || a.lg > 16 || me * a.lg != a.np || nb > 0x7fffffffLL) return hipErrorInvalidValue;
  if (nb == 0) return hipSuccess;
  const size_t shmem = (size_t)threads * (3 * 2 * EXACT_EMOM_JC + 1) * sizeof(double);
  hipLaunchKernelGGL(exact_emom_kernel, dim3((unsigned)nb), dim3(threads), shmem, st, a, nch, me);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(exact_max_kernel, dim3(1), dim3(EXACT_THREADS), 0, st, (const double*)a.pmax, nb, dmax, 0.0);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const int64_t n = a.np * (2 * (int64_t)a.K + 1);
  const int64_t nb2 = (n + 255) / 256;
  if (nb2 > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(exact_emom_out_kernel, dim3((unsigned)nb2), dim3(256), 0, st, a);
  return hipGetLastError();
}

int64_t exact_emom_blocks(int64_t np, int lg, int K) {
  return ((np / lg + EXACT_ENORM_E - 1) / EXACT_ENORM_E) * ((K + 1 + EXACT_EMOM_JC - 1) / EXACT_EMOM_JC);
}

template <int OP, int NI>
static hipError_t grid_launch(const ExactArgs& a, hipStream_t st, int64_t* nblk_out) {
  constexpr int JC = exact_jc(OP, NI);
  const int nch = (a.K + 1 + JC - 1) / JC;
  const int64_t nblk = exact_grid_blocks(OP, NI, a.np, a.K);
  *nblk_out = nblk;
  if (nblk == 0) return hipSuccess;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((exact_grid_kernel<OP, NI>), dim3((unsigned)nblk), dim3(EXACT_THREADS), 0, st, a, nch);
  return hipGetLastError();
}

hipError_t launch_exact(int op, const ExactArgs& a, double* max_out, hipStream_t st) {
  hipError_t e = hipErrorInvalidValue;
  int64_t nblk = 0;
  switch (op) {
    case EX_DIV: e = grid_launch<EX_DIV, 2>(a, st, &nblk); break;
    case EX_SV2: e = grid_launch<EX_SV2, 4>(a, st, &nblk); break;
    case EX_SQRT: e = grid_launch<EX_SQRT, 1>(a, st, &nblk); break;
    case EX_TENSION: e = grid_launch<EX_TENSION, 6>(a, st, &nblk); break;
    case EX_SVCLAMP: e = grid_launch<EX_SVCLAMP, 4>(a, st, &nblk); break;
    case EX_CLAMP01: e = grid_launch<EX_CLAMP01, 1>(a, st, &nblk); break;
    case EX_ENTRO: e = grid_launch<EX_ENTRO, 4>(a, st, &nblk); break;
    case EX_GRADMOM: e = grid_launch<EX_GRADMOM, 9>(a, st, &nblk); break;
    case EX_COMPR: e = grid_launch<EX_COMPR, 7>(a, st, &nblk); break;
    case EX_CLS: e = grid_launch<EX_CLS, 4>(a, st, &nblk); break;
    case EX_REAL: {
      const int nch = (a.K + 1 + EXACT_REAL_JC - 1) / EXACT_REAL_JC;
      const int64_t nb = ((a.np + EXACT_THREADS - 1) / EXACT_THREADS) * nch;
      if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
      if (nb > 0) hipLaunchKernelGGL(exact_real_kernel, dim3((unsigned)nb), dim3(EXACT_THREADS), 0, st, a, nch);
      return hipGetLastError();
    }
    case EX_ETA: e = grid_launch<EX_ETA, 1>(a, st, &nblk); break;
    case EX_ETA3: e = grid_launch<EX_ETA3, 3>(a, st, &nblk); break;
    case EX_FROMREAL: e = grid_launch<EX_FROMREAL, 1>(a, st, &nblk); break;
    case EX_EVISC: {
      // element maxima of |v| first, then the products
      const int64_t me = a.np / a.lg;
      const int nch = (a.K + 1 + EXACT_NORM_JC - 1) / EXACT_NORM_JC;
      const int64_t nb = ((me + EXACT_ENORM_E - 1) / EXACT_ENORM_E) * nch;
      const int threads = EXACT_ENORM_E * a.lg;
      if (a.lg < 1 || a.lg > 16 || me * a.lg != a.np || nb > 0x7fffffffLL) return hipErrorInvalidValue;
      if (nb > 0) {
        const size_t shmem = (size_t)threads * (2 * EXACT_NORM_JC + 1) * sizeof(double);
        hipLaunchKernelGGL(exact_enorm_kernel, dim3((unsigned)nb), dim3(threads), shmem, st, a, nch, me);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
      }
      // the viscosity on the grid, then the products (fewer input re-reads
      // than one fifteen-input pass)
      ExactArgs v, w;
      exact_evisc_split(a, a.rout2, v, w);
      if ((e = grid_launch<EX_EVISCV, 6>(v, st, &nblk)) != hipSuccess) return e;
      if ((e = grid_launch<EX_GRADMOM, 9>(w, st, &nblk)) != hipSuccess) return e;
      const int nmc = (w.M + EXACT_MC - 1) / EXACT_MC;
      const int64_t nb2 = ((w.np + EXACT_THREADS - 1) / EXACT_THREADS) * (int64_t)w.nout * nmc;
      if (nb2 > 0x7fffffffLL) return hipErrorInvalidValue;
      if (nb2 > 0) hipLaunchKernelGGL(exact_r2c_kernel, dim3((unsigned)nb2), dim3(EXACT_THREADS), 0, st, w, nmc);
      return hipGetLastError();
    }
    case EX_HEAVI:
      switch (a.nin) {
        case 1: e = grid_launch<EX_HEAVI, 1>(a, st, &nblk); break;
        case 2: e = grid_launch<EX_HEAVI, 2>(a, st, &nblk); break;
        case 3: e = grid_launch<EX_HEAVI, 3>(a, st, &nblk); break;
        case 4: e = grid_launch<EX_HEAVI, 4>(a, st, &nblk); break;
      }
      break;
    case EX_MAXVEL:
      if (a.nin == 3) e = grid_launch<EX_MAXVEL, 3>(a, st, &nblk);
      else if (a.nin == 1) e = grid_launch<EX_MAXVEL, 1>(a, st, &nblk);
      break;
    case EX_CHECKIF:
      switch (a.nin) {
        case 2: e = grid_launch<EX_CHECKIF, 2>(a, st, &nblk); break;
        case 3: e = grid_launch<EX_CHECKIF, 3>(a, st, &nblk); break;
        case 4: e = grid_launch<EX_CHECKIF, 4>(a, st, &nblk); break;
      }
      break;
  }
  if (e != hipSuccess) return e;
  if (exact_is_reduction(op)) {
    hipLaunchKernelGGL(exact_max_kernel, dim3(1), dim3(EXACT_THREADS), 0, st, (const double*)a.pmax, nblk, max_out,
                       0.0);
    return hipGetLastError();
  }
  const int nmc = (a.M + EXACT_MC - 1) / EXACT_MC;
  const int64_t nb2 = ((a.np + EXACT_THREADS - 1) / EXACT_THREADS) * (int64_t)a.nout * nmc;
  if (nb2 == 0) return hipSuccess;
  if (nb2 > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(exact_r2c_kernel, dim3((unsigned)nb2), dim3(EXACT_THREADS), 0, st, a, nmc);
  return hipGetLastError();
}

}  // namespace sfx
