// Persistent batched launch of the fused cross-curl kernel (body in sfx_batch_body.h).
#include "sfx_batch_body.h"

namespace sfx {

__global__ void __launch_bounds__(64 * BATCH_WAVES)
__attribute__((amdgpu_waves_per_eu(waves_per_simd(BATCH_WAVES), waves_per_simd(BATCH_WAVES)))) batch_kernel(BatchArgs b) {
  extern __shared__ double2 smem[];
  batch_body(b, threadIdx.x, blockIdx.x, smem);
}

// BY_SIZE schedule: one ordinary grid per grid size, tile t0 + block.
template <int MP, int NW = gauss_waves<MP, GV_CROSS_CURL>()>
__global__ void __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(waves_per_simd(NW), waves_per_simd(NW))))
batch_class_kernel(BatchArgs b, int32_t t0) {
  extern __shared__ double2 smem[];
  batch_class_body<MP>(b, t0, threadIdx.x, xcd_block(blockIdx.x, gridDim.x), smem);
}

template <int MP>
static hipError_t launch_class(const BatchArgs& b, int32_t t0, int32_t nt, hipStream_t st) {
  constexpr int NW = gauss_waves<MP, GV_CROSS_CURL>();
  const size_t shmem = (size_t)GaussCfg<MP, GV_CROSS_CURL, NW>::LY::TOTAL * sizeof(double2);
  hipLaunchKernelGGL((batch_class_kernel<MP>), dim3((unsigned)nt), dim3(64 * NW), shmem, st, b, t0);
  return hipGetLastError();
}

// wg_ptr[c] .. wg_ptr[c + 1]: the tiles of grid size class_mp[c].
hipError_t launch_batch_classes(const BatchArgs& b, const int32_t* wg_ptr_host, const int* class_mp,
                                hipStream_t st) {
  for (int c = 0; c < b.n_wg; ++c) {
    const int32_t t0 = wg_ptr_host[c], nt = wg_ptr_host[c + 1] - t0;
    if (nt <= 0) continue;
    hipError_t e = hipErrorInvalidValue;
    switch (class_mp[c]) {
      case 4: e = launch_class<4>(b, t0, nt, st); break;
      case 8: e = launch_class<8>(b, t0, nt, st); break;
      case 16: e = launch_class<16>(b, t0, nt, st); break;
      case 32: e = launch_class<32>(b, t0, nt, st); break;
      case 64: e = launch_class<64>(b, t0, nt, st); break;
      case 128: e = launch_class<128>(b, t0, nt, st); break;
      case 256: e = launch_class<256>(b, t0, nt, st); break;
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// smem_cplx: LDS of the largest fused-kernel layout among the problems.
hipError_t launch_batch(const BatchArgs& b, size_t smem_cplx, hipStream_t st) {
  if (b.n_wg == 0) return hipSuccess;
  const size_t shmem = smem_cplx * sizeof(double2);
  if (shmem == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(batch_kernel, dim3((unsigned)b.n_wg), dim3(64 * BATCH_WAVES), shmem, st, b);
  return hipGetLastError();
}

}  // namespace sfx
