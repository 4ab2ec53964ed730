// Launchers of the right-hand-side assembly kernels (body in sfx_rhs_body.h).
#include "sfx_engine.h"
#include "sfx_kernels.h"
#include "sfx_rhs_body.h"

namespace sfx {

// Consecutive blocks go to the same XCD (xcd_block): the partial-line
// stores of neighbouring nodes or elements then merge in one L2.
__global__ void __launch_bounds__(256) rhs_elem_kernel(RhsArgs a) {
  rhs_elem_body(a, xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x);
}

template <bool TENSOR>
__global__ void __launch_bounds__(256) rhs_ns_elem_kernel(RhsArgs a) {
  rhs_ns_elem_body<TENSOR>(a, xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x);
}

// The Gauss values of the NS right-hand side before the penalty product.
__global__ void __launch_bounds__(256) rhs_ns_gauss_kernel(RhsArgs a) {
  rhs_ns_elem_body<false, true>(a, xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x);
}

__global__ void __launch_bounds__(GATHER_THREADS) rhs_gather_kernel(RhsArgs a) {
  __shared__ double tile[gather_tile_doubles<6>()];
  rhs_gather_block(a, xcd_block(blockIdx.x, gridDim.x), threadIdx.x, tile);
}

__global__ void __launch_bounds__(256) qs00_elem_kernel(RhsArgs a) {
  qs00_elem_body(a, xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x);
}

__global__ void __launch_bounds__(GATHER_THREADS) qs00_gather_kernel(RhsArgs a) {
  __shared__ double tile[gather_tile_doubles<1>()];
  qs00_gather_block(a, xcd_block(blockIdx.x, gridDim.x), threadIdx.x, tile);
}

// rhs_3x3: element contributions, then the node-centred gather.
hipError_t launch_rhs3x3(const RhsArgs& a, hipStream_t st) {
  const int64_t ne = a.me > 0 && a.M > 0 ? elem_items(a.me, a.M) : 0, nn = a.np * a.M;
  if (ne > 0) hipLaunchKernelGGL(rhs_elem_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nn == 0) return e;
  const int64_t nb = (a.np + GATHER_NODES - 1) / GATHER_NODES;
  hipLaunchKernelGGL(rhs_gather_kernel, dim3((unsigned)nb), dim3(GATHER_THREADS), 0, st, a);
  return hipGetLastError();
}

// The gather alone: element contributions already in a.elo.
hipError_t launch_rhs_gather(const RhsArgs& a, hipStream_t st) {
  if (a.np * a.M == 0) return hipSuccess;
  const int64_t nb = (a.np + GATHER_NODES - 1) / GATHER_NODES;
  hipLaunchKernelGGL(rhs_gather_kernel, dim3((unsigned)nb), dim3(GATHER_THREADS), 0, st, a);
  return hipGetLastError();
}

// The qs_00_gauss gather alone: element contributions already in a.elo.
hipError_t launch_qs00_gather(const RhsArgs& a, hipStream_t st) {
  if (a.np * a.M == 0) return hipSuccess;
  const int64_t nb = (a.np + GATHER_NODES - 1) / GATHER_NODES;
  hipLaunchKernelGGL(qs00_gather_kernel, dim3((unsigned)nb), dim3(GATHER_THREADS), 0, st, a);
  return hipGetLastError();
}

// Full NS right-hand side fused with its element share, then the gather.
hipError_t launch_rhs_ns(const RhsArgs& a, hipStream_t st) {
  const int64_t ne = a.me > 0 && a.M > 0 ? elem_items(a.me, a.M) : 0, nn = a.np * a.M;
  if (ne > 0) {
    if (a.tens)
      hipLaunchKernelGGL(rhs_ns_elem_kernel<true>, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, a);
    else
      hipLaunchKernelGGL(rhs_ns_elem_kernel<false>, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, a);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nn == 0) return e;
  const int64_t nb = (a.np + GATHER_NODES - 1) / GATHER_NODES;
  hipLaunchKernelGGL(rhs_gather_kernel, dim3((unsigned)nb), dim3(GATHER_THREADS), 0, st, a);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) rhs_bdy_kernel(BdyArgs a) {
  rhs_bdy_body(a, (int64_t)blockIdx.x * 256 + threadIdx.x);
}

// Boundary term of rhs_3x3, added to vb after the volume gather.
hipError_t launch_rhs_bdy(const BdyArgs& a, hipStream_t st) {
  const int64_t n = a.nb * (int64_t)a.M;
  if (n > 0) hipLaunchKernelGGL(rhs_bdy_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_rhs_ns_gauss(const RhsArgs& a, hipStream_t st) {
  const int64_t ne = a.me > 0 && a.M > 0 ? elem_items(a.me, a.M) : 0;
  if (ne > 0) hipLaunchKernelGGL(rhs_ns_gauss_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_qs00(const RhsArgs& a, hipStream_t st) {
  const int64_t ne = a.me > 0 && a.M > 0 ? elem_items(a.me, a.M) : 0, nn = a.np * a.M;
  if (ne > 0) hipLaunchKernelGGL(qs00_elem_kernel, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, st, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nn == 0) return e;
  const int64_t nb = (a.np + GATHER_NODES - 1) / GATHER_NODES;
  hipLaunchKernelGGL(qs00_gather_kernel, dim3((unsigned)nb), dim3(GATHER_THREADS), 0, st, a);
  return hipGetLastError();
}

}  // namespace sfx
