// The inner product of SFEMaNS fields, the r-weighted mass matrix, and the
// L2-projected tangent with its adjoint in that inner product (C-ABI).
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/sfx.h"
#include "sfx_api_internal.h"
#include "sfx_inner_body.h"
#include "sfx_kernels.h"

using namespace sfx;

namespace {

InnerArgs inner_args(sfx_mesh_t mesh, const int32_t* lmode, int nc, int ncols) {
  InnerArgs a;
  std::memset(&a, 0, sizeof(a));
  a.jj = mesh->jj; a.tab = mesh->tab; a.wgt = mesh->wgt; a.lmode = lmode;
  a.nc = nc; a.ncols = ncols; a.me = mesh->me; a.np = mesh->np; a.G = 7 * mesh->me;
  a.w0 = 1.0; a.wm = 1.0;
  return a;
}

int64_t solve_work(sfx_mesh_t mesh, int64_t ncols) {
  return 4 * mesh->np * ncols + 4 * ncols + 6 * mesh->me * ncols;
}

// the mode weights of dot_product_SF: 1 for mode 0, 1/2 otherwise (fem_tn_ns_mhd.F90:115-118)
constexpr double W_MODE0 = 1.0, W_MODE = 0.5;

}  // namespace

extern "C" {

int64_t sfx_dot_product_sf_work(sfx_mesh_t mesh, int32_t m_max_c, int32_t nb_field) {
  if (!mesh || m_max_c < 0 || nb_field < 0) return -1;
  return dotsf_blocks(mesh->me, nb_field * m_max_c) + 1;
}

int sfx_dot_product_sf_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, int32_t nb_field,
                           const double* v, const double* w, double* result_dev, double* work, void* stream) {
  clear_err();
  if (!mesh) return fail(SFX_ERR_ARG, "dot_product_SF: null mesh");
  if (m_max_c < 0 || nb_field <= 0) return fail(SFX_ERR_ARG, "dot_product_SF: bad sizes");
  if (!result_dev || !work) return fail(SFX_ERR_ARG, "dot_product_SF: null result or work");
  if (m_max_c > 0 && (!list_mode_dev || !v || !w)) return fail(SFX_ERR_ARG, "dot_product_SF: null array");
  InnerArgs a = inner_args(mesh, list_mode_dev, nb_field, nb_field * m_max_c);
  a.w0 = W_MODE0; a.wm = W_MODE;
  a.v = v; a.w = w; a.partial = work;
  a.n_z = mesh->np * (int64_t)a.ncols; a.n_part = dotsf_blocks(mesh->me, a.ncols);
  hipError_t e = launch_dotsf(a, result_dev, (hipStream_t)stream);
  return e == hipSuccess ? SFX_OK : hip_fail(e, "dot_product_SF");
}

int sfx_dot_product_sf(sfx_mesh_t mesh, const int32_t* list_mode, int32_t m_max_c, int32_t nb_field, const double* v,
                       const double* w, double* result) {
  clear_err();
  if (!mesh || !result) return fail(SFX_ERR_ARG, "dot_product_SF: null mesh or result");
  if (m_max_c < 0 || nb_field <= 0) return fail(SFX_ERR_ARG, "dot_product_SF: bad sizes");
  const size_t n = (size_t)mesh->np * nb_field * m_max_c;
  const int64_t nw = sfx_dot_product_sf_work(mesh, m_max_c, nb_field);
  void *dv, *dw, *dl, *dr, *dk;
  int rc;
  if ((rc = staging(0, n * 8 + 8, &dv)) || (rc = staging(1, n * 8 + 8, &dw)) ||
      (rc = staging(2, (size_t)m_max_c * 4 + 4, &dl)) || (rc = staging(3, 8, &dr)) ||
      (rc = staging(4, (size_t)nw * 8, &dk)))
    return rc;
  hipError_t e;
  if ((n && ((e = hipMemcpy(dv, v, n * 8, hipMemcpyHostToDevice)) != hipSuccess ||
             (e = hipMemcpy(dw, w, n * 8, hipMemcpyHostToDevice)) != hipSuccess)) ||
      (m_max_c && (e = hipMemcpy(dl, list_mode, (size_t)m_max_c * 4, hipMemcpyHostToDevice)) != hipSuccess))
    return hip_fail(e, "H2D");
  if ((rc = sfx_dot_product_sf_dev(mesh, (const int32_t*)dl, m_max_c, nb_field, (const double*)dv, (const double*)dw,
                                   (double*)dr, (double*)dk, nullptr)))
    return rc;
  if ((e = hipMemcpy(result, dr, 8, hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(e, "D2H");
  return SFX_OK;
}

int sfx_mass_apply_dev(sfx_mesh_t mesh, int32_t ncols, const double* x, double* y, double* work_elem, void* stream) {
  clear_err();
  if (!mesh) return fail(SFX_ERR_ARG, "mass_apply: null mesh");
  if (!x) return fail(SFX_ERR_ARG, "mass_apply: null x");
  // qs_00_gauss with heat capacity 1 and no Gauss source: SUM_l ww(i,l) rj ray SUM(x ww)
  return sfx_qs_00_gauss_dev(mesh, ncols, nullptr, x, nullptr, work_elem, y, stream);
}

int64_t sfx_mass_solve_work(sfx_mesh_t mesh, int32_t ncols) {
  if (!mesh || ncols < 0) return -1;
  return solve_work(mesh, ncols);
}

int sfx_mass_solve_dev(sfx_mesh_t mesh, int32_t ncols, const double* b, double* x, double rtol, int32_t max_iter,
                       double* work, int32_t* iters, void* stream) {
  clear_err();
  if (!mesh) return fail(SFX_ERR_ARG, "mass_solve: null mesh");
  if (ncols < 0) return fail(SFX_ERR_ARG, "mass_solve: ncols < 0");
  if (ncols == 0 || mesh->np == 0) {
    if (iters) *iters = 0;
    return SFX_OK;
  }
  if (!b || !x || !work) return fail(SFX_ERR_ARG, "mass_solve: null array");
  if (rtol <= 0.0) rtol = 1e-14;
  if (max_iter <= 0) max_iter = 1000;
  hipStream_t st = (hipStream_t)stream;
  const int64_t nn = mesh->np * (int64_t)ncols;
  double* r = work;
  double* z = r + nn;
  double* p = z + nn;
  double* q = p + nn;
  double* pq = q + nn;
  double* rzA = pq + ncols;
  double* rzB = rzA + ncols;
  double* rz0 = rzB + ncols;
  double* welem = rz0 + ncols;
  PcgArgs a;
  std::memset(&a, 0, sizeof(a));
  a.np = mesh->np; a.ncols = ncols; a.dinv = mesh->mdinv;
  a.x = x; a.r = r; a.z = z; a.p = p; a.q = q; a.b = b; a.pq = pq;
  hipError_t e;
  if ((e = launch_pcg(0, a, st)) != hipSuccess || (e = launch_coldot(r, z, mesh->np, ncols, rzA, st)) != hipSuccess ||
      (e = hipMemcpyAsync(rz0, rzA, ncols * sizeof(double), hipMemcpyDeviceToDevice, st)) != hipSuccess)
    return hip_fail(e, "mass_solve: start");
  std::vector<double> h_rz(ncols), h_rz0(ncols);
  if ((e = hipMemcpyAsync(h_rz0.data(), rz0, ncols * sizeof(double), hipMemcpyDeviceToHost, st)) != hipSuccess ||
      (e = hipStreamSynchronize(st)) != hipSuccess)
    return hip_fail(e, "mass_solve: start");
  const double tol2 = rtol * rtol;
  int it = 0;
  bool done = true;
  for (int c = 0; c < ncols; ++c) done = done && h_rz0[c] == 0.0;
  while (!done && it < max_iter) {
    int rc = sfx_mass_apply_dev(mesh, ncols, p, q, welem, stream);
    if (rc) return rc;
    a.rz = rzA; a.rz_new = rzB;
    if ((e = launch_coldot(p, q, mesh->np, ncols, pq, st)) != hipSuccess || (e = launch_pcg(1, a, st)) != hipSuccess ||
        (e = launch_coldot(r, z, mesh->np, ncols, rzB, st)) != hipSuccess || (e = launch_pcg(2, a, st)) != hipSuccess)
      return hip_fail(e, "mass_solve: iteration");
    std::swap(rzA, rzB);
    ++it;
    if (it % 8 == 0 || it == max_iter) {
      if ((e = hipMemcpyAsync(h_rz.data(), rzA, ncols * sizeof(double), hipMemcpyDeviceToHost, st)) != hipSuccess ||
          (e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(e, "mass_solve: convergence check");
      done = true;
      for (int c = 0; c < ncols; ++c) done = done && (h_rz[c] <= tol2 * h_rz0[c]);
    }
  }
  if (iters) *iters = it;
  if (!done) return fail(SFX_ERR_ARG, "mass_solve: no convergence to rtol=%g in %d iterations", rtol, max_iter);
  return SFX_OK;
}

int64_t sfx_nl_proj_work(sfx_mesh_t mesh, int32_t m_max_c) {
  if (!mesh || m_max_c < 0) return -1;
  const int64_t nc = 6 * (int64_t)m_max_c;
  return 7 * mesh->me * nc + mesh->np * nc + solve_work(mesh, nc);
}

int sfx_nl_tangent_proj_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* U,
                            const double* dU, double* out, double rtol, double* work, void* stream) {
  clear_err();
  if (!mesh) return fail(SFX_ERR_ARG, "nl_tangent_proj: null mesh");
  if (!work) return fail(SFX_ERR_ARG, "nl_tangent_proj: null work");
  const int64_t nc = 6 * (int64_t)m_max_c;
  double* g = work;
  double* rhs = g + 7 * mesh->me * nc;
  double* sw = rhs + mesh->np * nc;
  // T_U dU at the Gauss points, its Galerkin right-hand side R (qs_00_gauss of
  // the Gauss values alone), then the mass solve
  int rc = sfx_nl_tangent_dev(mesh, list_mode_dev, m_max_c, U, dU, g, stream);
  if (rc) return rc;
  if ((rc = sfx_qs_00_gauss_dev(mesh, (int32_t)nc, nullptr, nullptr, g, sw + 4 * mesh->np * nc + 4 * nc, rhs, stream)))
    return rc;
  return sfx_mass_solve_dev(mesh, (int32_t)nc, rhs, out, rtol, 0, sw, nullptr, stream);
}

int sfx_nl_adjoint_sf_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* U,
                          const double* ybar, double* out, double rtol, double* work, void* stream) {
  clear_err();
  if (!mesh) return fail(SFX_ERR_ARG, "nl_adjoint_sf: null mesh");
  if (!work) return fail(SFX_ERR_ARG, "nl_adjoint_sf: null work");
  if (m_max_c <= 0) return fail(SFX_ERR_ARG, "nl_adjoint_sf: SIZE(list_mode) == 0");
  hipStream_t st = (hipStream_t)stream;
  const int nc = 6 * m_max_c;
  double* g = work;
  double* t = g + 7 * mesh->me * (int64_t)nc;
  double* sw = t + mesh->np * (int64_t)nc;
  double* welem = sw + 4 * mesh->np * (int64_t)nc + 4 * nc;   // 36 me m doubles, shared with the solve
  // R^T (w ybar): Gauss values times rj ray, mode weights 1 and 1/2
  InnerArgs a = inner_args(mesh, list_mode_dev, 6, nc);
  a.w0 = W_MODE0; a.wm = W_MODE; a.weighted = 1;
  a.z = ybar; a.out = g; a.n_z = mesh->np * (int64_t)nc; a.n_out = 7 * mesh->me * (int64_t)nc;
  hipError_t e = launch_gauss_values(a, st);
  if (e != hipSuccess) return hip_fail(e, "nl_adjoint_sf");
  // the Euclidean transpose of the tangent
  int rc = sfx_nl_adjoint_dev(mesh, list_mode_dev, m_max_c, U, g, welem, t, stream);
  if (rc) return rc;
  // w^{-1}, then the mass solve
  InnerArgs s = inner_args(mesh, list_mode_dev, 6, nc);
  s.w0 = 1.0 / W_MODE0; s.wm = 1.0 / W_MODE;
  s.z = t; s.out = t;
  if ((e = launch_scale_cols(s, st)) != hipSuccess) return hip_fail(e, "nl_adjoint_sf");
  return sfx_mass_solve_dev(mesh, nc, t, out, rtol, 0, sw, nullptr, stream);
}

}  // extern "C"
