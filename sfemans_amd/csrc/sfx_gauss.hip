// Launchers of the fused Gauss-point kernels (body in sfx_gauss_body.h).
#include "sfx_gauss_body.h"

namespace sfx {

// NW waves per work-group, held to the register budget of waves_per_simd(NW)
// waves per SIMD.
template <int MP, int V, int NW = gauss_waves<MP, V>()>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(waves_per_simd(NW), waves_per_simd(NW))))
gauss_kernel(GaussArgs a) {
  extern __shared__ double2 smem[];
  gauss_body<MP, V, NW>(a, threadIdx.x, xcd_block(blockIdx.x, gridDim.x), smem);
}

// Point tiles (see gauss_point_tiles): GAUSS_PT_WAVES waves, two per SIMD.
template <int MP, int V>
__global__ void __launch_bounds__(64 * GAUSS_PT_WAVES) __attribute__((amdgpu_waves_per_eu(2, 2)))
gauss_pt_kernel(GaussArgs a) {
  extern __shared__ double2 smem[];
  gauss_body<MP, V, GAUSS_PT_WAVES, false, false, true>(a, threadIdx.x, xcd_block(blockIdx.x, gridDim.x), smem);
}

template <int MP, int V>
static hipError_t launch_mp(const GaussArgs& a, hipStream_t st) {
  if constexpr (gauss_point_tiles<MP, V>()) {
    constexpr int NW = GAUSS_PT_WAVES;
    const int64_t nblk = gauss_blocks<MP, V, NW>(a.me);
    if (nblk == 0) return hipSuccess;
    const size_t shmem = (size_t)GaussCfg<MP, V, NW>::LY::TOTAL * sizeof(double2);
    hipLaunchKernelGGL((gauss_pt_kernel<MP, V>), dim3((unsigned)nblk), dim3(64 * NW), shmem, st, a);
  } else {
    constexpr int NW = gauss_waves<MP, V>();
    const int64_t nblk = gauss_blocks<MP, V, NW>(a.me);
    if (nblk == 0) return hipSuccess;
    const size_t shmem = (size_t)GaussCfg<MP, V, NW>::LY::TOTAL * sizeof(double2);
    hipLaunchKernelGGL((gauss_kernel<MP, V>), dim3((unsigned)nblk), dim3(64 * NW), shmem, st, a);
  }
  return hipGetLastError();
}

template <int V>
static hipError_t launch_v(const GaussArgs& a, hipStream_t st) {
  switch (a.MP) {
    case 4: return launch_mp<4, V>(a, st);
    case 8: return launch_mp<8, V>(a, st);
    case 16: return launch_mp<16, V>(a, st);
    case 32: return launch_mp<32, V>(a, st);
    case 64: return launch_mp<64, V>(a, st);
    case 128: return launch_mp<128, V>(a, st);
    case 256: return launch_mp<256, V>(a, st);
  }
  return hipErrorInvalidValue;
}

// The fused nodal -> assembly element sums (RHS form of the body):
// cross-curl -> rhs_3x3, u.grad c -> qs_00_gauss.
template <int MP, int V, int NW = gauss_waves<MP, V>()>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(waves_per_simd(NW), waves_per_simd(NW))))
gauss_rhs_kernel(GaussArgs a) {
  extern __shared__ double2 smem[];
  gauss_body<MP, V, NW, false, true>(a, threadIdx.x, xcd_block(blockIdx.x, gridDim.x), smem);
}

template <int MP, int V>
static hipError_t launch_rhs_mp(const GaussArgs& a, hipStream_t st) {
  constexpr int NW = gauss_waves<MP, V>();
  constexpr int E = GaussLds<MP, NW>::E;
  const int64_t nblk = (a.me + E - 1) / E;
  if (nblk == 0) return hipSuccess;
  const size_t shmem = (size_t)GaussCfg<MP, V, NW>::LY::TOTAL * sizeof(double2);
  hipLaunchKernelGGL((gauss_rhs_kernel<MP, V>), dim3((unsigned)nblk), dim3(64 * NW), shmem, st, a);
  return hipGetLastError();
}

template <int V>
static hipError_t launch_rhs_v(const GaussArgs& a, hipStream_t st) {
  switch (a.MP) {
    case 4: return launch_rhs_mp<4, V>(a, st);
    case 8: return launch_rhs_mp<8, V>(a, st);
    case 16: return launch_rhs_mp<16, V>(a, st);
    case 32: return launch_rhs_mp<32, V>(a, st);
    case 64: return launch_rhs_mp<64, V>(a, st);
    case 128: return launch_rhs_mp<128, V>(a, st);
    case 256: return launch_rhs_mp<256, V>(a, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_gauss_rhs(int variant, const GaussArgs& a, hipStream_t st) {
  if (variant == GV_CROSS_CURL) return launch_rhs_v<GV_CROSS_CURL>(a, st);
  if (variant == GV_UGRADC) return launch_rhs_v<GV_UGRADC>(a, st);
  return hipErrorInvalidValue;
}

hipError_t launch_gauss(int variant, const GaussArgs& a, hipStream_t st) {
  switch (variant) {
    case GV_CROSS_CURL: return launch_v<GV_CROSS_CURL>(a, st);
    case GV_PAIRS: return launch_v<GV_PAIRS>(a, st);
    case GV_UGRADC: return launch_v<GV_UGRADC>(a, st);
    case GV_CURL_CROSS: return launch_v<GV_CURL_CROSS>(a, st);
  }
  return hipErrorInvalidValue;
}

}  // namespace sfx
