// Launchers of the inner-product, mass-matrix and column kernels (body in
// sfx_inner_body.h), and the allocation-free form of dot_product_SF.
#include <mutex>

#include "../../include/sfx.h"
#include "sfx_api_internal.h"
#include "sfx_engine.h"
#include "sfx_inner_body.h"
#include "sfx_kernels.h"

namespace sfx {

__global__ void __launch_bounds__(256) gauss_values_kernel(InnerArgs a) {
  gauss_values_body(a, xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x);
}

__global__ void __launch_bounds__(INNER_THREADS) dotsf_kernel(InnerArgs a, int64_t nitems) {
  __shared__ double sh[INNER_THREADS];
  dotsf_block(a, blockIdx.x, gridDim.x, nitems, threadIdx.x, sh);
}

__global__ void __launch_bounds__(INNER_THREADS) sum_partials_kernel(const double* partial, int64_t n, double scale,
                                                                     double* out) {
  __shared__ double sh[INNER_THREADS];
  sum_partials_block(partial, n, scale, out, threadIdx.x, sh);
}

__global__ void __launch_bounds__(INNER_THREADS) coldot_kernel(const double* A, const double* B, int64_t np,
                                                               double* out) {
  __shared__ double sh[INNER_THREADS];
  coldot_block(A, B, np, out, blockIdx.x, threadIdx.x, sh);
}

__global__ void __launch_bounds__(256) pcg_init_kernel(PcgArgs a) {
  pcg_init_item(a, (int64_t)blockIdx.x * 256 + threadIdx.x);
}
__global__ void __launch_bounds__(256) pcg_step_kernel(PcgArgs a) {
  pcg_step_item(a, (int64_t)blockIdx.x * 256 + threadIdx.x);
}
__global__ void __launch_bounds__(256) pcg_dir_kernel(PcgArgs a) {
  pcg_dir_item(a, (int64_t)blockIdx.x * 256 + threadIdx.x);
}
__global__ void __launch_bounds__(256) scale_cols_kernel(InnerArgs a) {
  scale_cols_item(a, (int64_t)blockIdx.x * 256 + threadIdx.x);
}

static unsigned blocks_of(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_gauss_values(const InnerArgs& a, hipStream_t st) {
  const int64_t n = a.me > 0 && a.ncols > 0 ? elem_items(a.me, a.ncols) : 0;
  if (n > 0) hipLaunchKernelGGL(gauss_values_kernel, dim3(blocks_of(n, 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

int64_t dotsf_blocks(int64_t me, int ncols) {
  const int64_t nb = me > 0 && ncols > 0 ? (elem_items(me, ncols) + INNER_THREADS - 1) / INNER_THREADS : 0;
  return nb < DOTSF_MAX_BLOCKS ? nb : DOTSF_MAX_BLOCKS;
}

hipError_t launch_dotsf(const InnerArgs& a, double* result, hipStream_t st) {
  const int64_t nb = dotsf_blocks(a.me, a.ncols);
  const int64_t nitems = a.me > 0 && a.ncols > 0 ? elem_items(a.me, a.ncols) : 0;
  if (nb > 0) hipLaunchKernelGGL(dotsf_kernel, dim3((unsigned)nb), dim3(INNER_THREADS), 0, st, a, nitems);
  hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(INNER_THREADS), 0, st, (const double*)a.partial, nb,
                     2.0 * 3.14159265358979323846, result);
  return hipGetLastError();
}

hipError_t launch_coldot(const double* A, const double* B, int64_t np, int ncols, double* out, hipStream_t st) {
  if (ncols > 0) hipLaunchKernelGGL(coldot_kernel, dim3((unsigned)ncols), dim3(INNER_THREADS), 0, st, A, B, np, out);
  return hipGetLastError();
}

hipError_t launch_pcg(int which, const PcgArgs& a, hipStream_t st) {
  const int64_t n = a.np * (int64_t)a.ncols;
  if (n <= 0) return hipSuccess;
  if (which == 0) hipLaunchKernelGGL(pcg_init_kernel, dim3(blocks_of(n, 256)), dim3(256), 0, st, a);
  if (which == 1) hipLaunchKernelGGL(pcg_step_kernel, dim3(blocks_of(n, 256)), dim3(256), 0, st, a);
  if (which == 2) hipLaunchKernelGGL(pcg_dir_kernel, dim3(blocks_of(n, 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_scale_cols(const InnerArgs& a, hipStream_t st) {
  const int64_t n = a.np * (int64_t)a.ncols;
  if (n > 0) hipLaunchKernelGGL(scale_cols_kernel, dim3(blocks_of(n, 256)), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace sfx

// ---- dot_product_SF without allocation: the partial sums of the work-groups
// in a fixed device array (DOTSF_MAX_BLOCKS doubles, 16 KiB, held by every
// device the library runs on until the process ends).  The host side never
// uses the array's name as an address: the device address comes from
// hipGetSymbolAddress, once per device.
static __device__ __align__(256) double scratch[sfx::DOTSF_MAX_BLOCKS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mutex;
double* scratch_ptr[MAX_DEVICES] = {};
}  // namespace

extern "C" int sfx_dot_product_sf_noalloc(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c,
                                          int32_t nb_field, const double* v, const double* w, double* result_dev,
                                          void* stream) {
  using namespace sfx;
  clear_err();
  if (!mesh) return fail(SFX_ERR_ARG, "dot_product_SF (no allocation): null mesh");
  if (m_max_c < 0 || nb_field <= 0) return fail(SFX_ERR_ARG, "dot_product_SF (no allocation): bad sizes");
  const int64_t nb = dotsf_blocks(mesh->me, nb_field * m_max_c);
  if (nb > DOTSF_MAX_BLOCKS)
    return fail(SFX_ERR_ARG, "dot_product_SF (no allocation): %lld partial sums exceed the %d of the fixed scratch",
                (long long)nb, DOTSF_MAX_BLOCKS);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "dot_product_SF (no allocation): hipGetDevice");
  if (dev < 0 || dev >= MAX_DEVICES) return fail(SFX_ERR_UNSUPPORTED, "dot_product_SF (no allocation): device %d", dev);
  double* ptr = nullptr;
  {
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
      // this device's copy of the array, looked up with the device current
      void* p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess) return hip_fail(e, "dot_product_SF (no allocation): hipGetSymbolAddress");
      if (!p) return fail(SFX_ERR_HIP, "dot_product_SF (no allocation): hipGetSymbolAddress gave a null address");
      scratch_ptr[dev] = static_cast<double*>(p);
    }
    ptr = scratch_ptr[dev];
  }
  return sfx_dot_product_sf_dev(mesh, list_mode_dev, m_max_c, nb_field, v, w, result_dev, ptr, stream);
}
