// Launchers of the adjoint kernels (bodies in sfx_adjoint_body.h).
#include "sfx_adjoint_body.h"

namespace sfx {

template <int MP, bool SPLIT, int NW = adjoint_waves<MP, SPLIT>()>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(waves_per_simd(NW), waves_per_simd(NW))))
adjoint_kernel(AdjArgs a) {
  extern __shared__ double2 smem[];
  adjoint_body<MP, SPLIT, NW>(a, threadIdx.x, xcd_block(blockIdx.x, gridDim.x), smem);
}

__global__ void __launch_bounds__(GATHER_THREADS) adjoint_gather_kernel(GatherArgs a) {
  __shared__ double tile[gather_tile_doubles<6>()];
  adjoint_gather_block(a, xcd_block(blockIdx.x, gridDim.x), threadIdx.x, tile);
}

template <int MP, bool SPLIT>
static hipError_t launch_mp(const AdjArgs& a, hipStream_t st) {
  constexpr int NW = adjoint_waves<MP, SPLIT>();
  constexpr int E = GaussLds<MP, NW>::E;
  const int64_t nblk = (a.g.me + E - 1) / E;
  if (nblk == 0) return hipSuccess;
  const size_t shmem = (size_t)AdjLds<MP, SPLIT, NW>::TOTAL * sizeof(double2);
  hipLaunchKernelGGL((adjoint_kernel<MP, SPLIT>), dim3((unsigned)nblk), dim3(64 * NW), shmem, st, a);
  return hipGetLastError();
}

template <bool SPLIT>
static hipError_t launch_split(const AdjArgs& a, hipStream_t st) {
  switch (a.g.MP) {
    case 4: return launch_mp<4, SPLIT>(a, st);
    case 8: return launch_mp<8, SPLIT>(a, st);
    case 16: return launch_mp<16, SPLIT>(a, st);
    case 32: return launch_mp<32, SPLIT>(a, st);
    case 64: return launch_mp<64, SPLIT>(a, st);
    case 128: return launch_mp<128, SPLIT>(a, st);
    case 256: return launch_mp<256, SPLIT>(a, st);
  }
  return hipErrorInvalidValue;
}

static hipError_t launch_gather(const GatherArgs& ga, hipStream_t st) {
  if (ga.np == 0 || ga.M == 0) return hipSuccess;
  const int64_t nb = (ga.np + GATHER_NODES - 1) / GATHER_NODES;
  hipLaunchKernelGGL(adjoint_gather_kernel, dim3((unsigned)nb), dim3(GATHER_THREADS), 0, st, ga);
  return hipGetLastError();
}

// Element kernel, then one gather per output (two in the split form).
hipError_t launch_adjoint(const AdjArgs& a, const GatherArgs& ga, const GatherArgs* ga2, hipStream_t st) {
  hipError_t e = ga2 ? launch_split<true>(a, st) : launch_split<false>(a, st);
  if (e != hipSuccess) return e;
  if ((e = launch_gather(ga, st)) != hipSuccess) return e;
  return ga2 ? launch_gather(*ga2, st) : hipSuccess;
}

}  // namespace sfx
