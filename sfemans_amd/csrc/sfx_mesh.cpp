// Mesh registration and the Gauss-point entry points of the C-ABI.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/sfx.h"
#include "sfx_api_internal.h"
#include "sfx_kernels.h"
#include "sfx_bdy_setup.h"
#include "sfx_p2tab.h"
#include "sfx_rhs_body.h"
#include "sfx_adjoint_body.h"


using namespace sfx;

static int mesh_upload(const int32_t* jj, const double* rr, int64_t me, int64_t np, sfx_mesh_t* out) {
  sfx_mesh_s* m = new sfx_mesh_s();
  (void)hipGetDevice(&m->device);
  m->me = me;
  m->np = np;
  double tab[126];
  p2_reference_tables(tab);
  // node -> (element, local node) adjacency in element order (deterministic sums)
  std::vector<int64_t> ptr(np + 1, 0);
  for (int64_t e = 0; e < me; ++e)
    for (int n = 0; n < 6; ++n) ptr[jj[6 * e + n]] += 1;   // jj is 1-based: counts land at node+1
  for (int64_t i = 0; i < np; ++i) ptr[i + 1] += ptr[i];
  std::vector<int32_t> ent(ptr[np]);
  std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
  for (int64_t e = 0; e < me; ++e)
    for (int n = 0; n < 6; ++n) ent[fill[jj[6 * e + n] - 1]++] = (int32_t)(6 * e + n);
  std::vector<double> wgt(7 * me);
  p2_assembly_weights(jj, rr, me, tab, wgt.data());
  // diagonal of the r-weighted mass matrix, SUM ww(i,l)^2 rj ray (the Jacobi
  // preconditioner of the mass solves)
  std::vector<double> mdinv(np, 0.0);
  for (int64_t e = 0; e < me; ++e)
    for (int n = 0; n < 6; ++n)
      for (int l = 0; l < 7; ++l) mdinv[jj[6 * e + n] - 1] += tab[n + 6 * l] * tab[n + 6 * l] * wgt[7 * e + l];
  for (int64_t i = 0; i < np; ++i) mdinv[i] = mdinv[i] > 0.0 ? 1.0 / mdinv[i] : 0.0;
  m->n_adj = (int64_t)ent.size();
  hipError_t e;
  if ((e = hipMalloc(&m->wgt, 7 * me * sizeof(double) + 8)) != hipSuccess ||
      (e = hipMalloc(&m->jj, 6 * me * sizeof(int32_t) + 4)) != hipSuccess ||
      (e = hipMalloc(&m->rr, 2 * np * sizeof(double) + 8)) != hipSuccess ||
      (e = hipMalloc(&m->tab, 126 * sizeof(double))) != hipSuccess ||
      (e = hipMalloc(&m->adj_ptr, (np + 1) * sizeof(int64_t))) != hipSuccess ||
      (e = hipMalloc(&m->adj_ent, ent.size() * sizeof(int32_t) + 4)) != hipSuccess ||
      (e = hipMalloc(&m->mdinv, np * sizeof(double) + 8)) != hipSuccess) {
    sfx_mesh_destroy(m);
    return fail(SFX_ERR_NOMEM, "sfx_mesh_create: hipMalloc: %s", hipGetErrorString(e));
  }
  if ((e = hipMemcpy(m->jj, jj, 6 * me * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->rr, rr, 2 * np * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(m->tab, tab, sizeof(tab), hipMemcpyHostToDevice)) != hipSuccess ||
      (me && (e = hipMemcpy(m->wgt, wgt.data(), 7 * me * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) ||
      (e = hipMemcpy(m->adj_ptr, ptr.data(), (np + 1) * sizeof(int64_t), hipMemcpyHostToDevice)) != hipSuccess ||
      (ent.size() && (e = hipMemcpy(m->adj_ent, ent.data(), ent.size() * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess) ||
      (np && (e = hipMemcpy(m->mdinv, mdinv.data(), np * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)) {
    sfx_mesh_destroy(m);
    return hip_fail(e, "sfx_mesh_create: upload");
  }
  *out = m;
  return SFX_OK;
}

extern "C" {

int sfx_mesh_create(const int32_t* jj, const double* rr, int32_t n_w, int32_t l_G, int64_t me, int64_t np,
                    sfx_mesh_t* out) {
  clear_err();
  if (!out) return fail(SFX_ERR_ARG, "sfx_mesh_create: null handle pointer");
  *out = nullptr;
  if (n_w != 6 || l_G != 7)
    return fail(SFX_ERR_UNSUPPORTED, "sfx_mesh_create: only P2 triangles (n_w=6, l_G=7) are supported");
  if (me < 0 || np < 0 || (me > 0 && (!jj || !rr))) return fail(SFX_ERR_ARG, "sfx_mesh_create: bad sizes");
  double tab[126];
  p2_reference_tables(tab);
  for (int64_t e = 0; e < me; ++e) {
    for (int n = 0; n < 6; ++n) {
      const int32_t v = jj[6 * e + n];
      if (v < 1 || v > np) return fail(SFX_ERR_ARG, "sfx_mesh_create: jj(%d,%lld)=%d out of 1..np", n + 1, (long long)e + 1, v);
    }
    // curved elements are allowed (the Jacobian is taken per Gauss point);
    // the reference divides by det J at every Gauss point (gauss_points_2d.F90:115-120)
    for (int l = 0; l < 7; ++l) {
      const double det = p2_jacobian_det(jj + 6 * e, rr, tab, l);
      if (!(std::fabs(det) > 0.0) || !std::isfinite(det))
        return fail(SFX_ERR_ARG, "sfx_mesh_create: element %lld is degenerate (det J = %g at Gauss point %d)",
                    (long long)e + 1, det, l + 1);
    }
  }
  return mesh_upload(jj, rr, me, np, out);
}

int sfx_mesh_destroy(sfx_mesh_t m) {
  if (!m) return SFX_OK;
  (void)hipFree(m->jj);
  (void)hipFree(m->rr);
  (void)hipFree(m->tab);
  (void)hipFree(m->adj_ptr);
  (void)hipFree(m->adj_ent);
  (void)hipFree(m->wgt);
  (void)hipFree(m->mdinv);
  (void)hipFree(m->bptr);
  (void)hipFree(m->bent);
  (void)hipFree(m->bnode);
  (void)hipFree(m->bw);
  delete m;
  return SFX_OK;
}

int sfx_mesh_set_boundary(sfx_mesh_t m, const int32_t* jjs, int64_t mes) {
  clear_err();
  if (!m) return fail(SFX_ERR_ARG, "sfx_mesh_set_boundary: null mesh");
  if (mes < 0 || (mes > 0 && !jjs)) return fail(SFX_ERR_ARG, "sfx_mesh_set_boundary: bad sizes");
  for (int64_t s = 0; s < 3 * mes; ++s)
    if (jjs[s] < 1 || jjs[s] > m->np)
      return fail(SFX_ERR_ARG, "sfx_mesh_set_boundary: jjs(%d,%lld)=%d out of 1..np", (int)(s % 3) + 1,
                  (long long)(s / 3) + 1, jjs[s]);
  std::vector<double> rr(2 * m->np);
  hipError_t e;
  if (m->np && (e = hipMemcpy(rr.data(), m->rr, 2 * m->np * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
    return hip_fail(e, "sfx_mesh_set_boundary: coordinates");
  BdyTables t;
  build_boundary_tables(jjs, mes, rr.data(), m->np, t);
  (void)hipFree(m->bptr); (void)hipFree(m->bent); (void)hipFree(m->bnode); (void)hipFree(m->bw);
  m->bptr = nullptr; m->bent = nullptr; m->bnode = nullptr; m->bw = nullptr; m->mes = -1;
  const int64_t nb = (int64_t)t.node.size();
  if ((e = hipMalloc(&m->bptr, (nb + 1) * sizeof(int64_t))) != hipSuccess ||
      (e = hipMalloc(&m->bent, t.ent.size() * sizeof(int32_t) + 4)) != hipSuccess ||
      (e = hipMalloc(&m->bnode, nb * sizeof(int64_t) + 8)) != hipSuccess ||
      (e = hipMalloc(&m->bw, t.bw.size() * sizeof(double) + 8)) != hipSuccess)
    return fail(SFX_ERR_NOMEM, "sfx_mesh_set_boundary: hipMalloc: %s", hipGetErrorString(e));
  if ((e = hipMemcpy(m->bptr, t.ptr.data(), (nb + 1) * sizeof(int64_t), hipMemcpyHostToDevice)) != hipSuccess ||
      (nb && (e = hipMemcpy(m->bent, t.ent.data(), t.ent.size() * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess) ||
      (nb && (e = hipMemcpy(m->bnode, t.node.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice)) != hipSuccess) ||
      (mes && (e = hipMemcpy(m->bw, t.bw.data(), t.bw.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess))
    return hip_fail(e, "sfx_mesh_set_boundary: upload");
  m->mes = mes;
  m->nb = nb;
  m->n_bent = (int64_t)t.ent.size();
  return SFX_OK;
}

int sfx_mesh_info(sfx_mesh_t m, int64_t* me, int64_t* np) {
  clear_err();
  if (!m) return fail(SFX_ERR_ARG, "sfx_mesh_info: null mesh");
  if (me) *me = m->me;
  if (np) *np = m->np;
  return SFX_OK;
}

}  // extern "C"

// ---- Gauss-point entry points
namespace sfx {

// Curl-type passes of GV_PAIRS: out += coef * (curl pc) x pv.
struct Passes {
  int n = 0;
  const double* pc[4] = {nullptr, nullptr, nullptr, nullptr};
  const double* pv[4] = {nullptr, nullptr, nullptr, nullptr};
  double coef[4] = {0, 0, 0, 0};
  void add(const double* c, const double* v, double k) { pc[n] = c; pv[n] = v; coef[n] = k; ++n; }
};

// Node-centred assembly arguments of a mesh (element sums -> gather).
static RhsArgs rhs_args(sfx_mesh_t mesh, int32_t M, double scale) {
  RhsArgs a;
  std::memset(&a, 0, sizeof(a));
  a.adj_ptr = mesh->adj_ptr; a.adj_ent = mesh->adj_ent; a.wgt = mesh->wgt; a.tab = mesh->tab;
  a.jj = mesh->jj; a.me = mesh->me; a.np = mesh->np; a.G = 7 * mesh->me; a.M = M; a.scale = scale;
  a.n_adj = mesh->n_adj;
  return a;
}

// Mesh, modes and twiddles of a fused Gauss-point launch.
static int gauss_setup(const char* nm, sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t M, GaussArgs& a) {
  if (!mesh) return fail(SFX_ERR_ARG, "%s: null mesh", nm);
  if (M <= 0) return fail(SFX_ERR_ARG, "%s: SIZE(list_mode) == 0", nm);
  if (M > 256) return fail(SFX_ERR_UNSUPPORTED, "%s: m_max_c=%d > 256", nm, M);
  std::memset(&a, 0, sizeof(a));
  a.jj = mesh->jj; a.rr = mesh->rr; a.me = mesh->me; a.np = mesh->np; a.G = 7 * mesh->me;
  a.list_mode = list_mode_dev; a.tab = mesh->tab;
  a.M = M; a.MP = mp_for(M);
  Twiddles tw;
  int rc = get_twiddles(a.MP, &tw);
  if (rc) return rc;
  a.wN = tw.wN; a.wMP = tw.wMP;
  return SFX_OK;
}

int run_gauss_dev(int variant, const char* nm, sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t M,
                  const double* U, int ncU, const double* U2, int ncU2, const Passes* ps,
                  const double* heat_capa_dev, double* out, int ncout, const double om[3], hipStream_t st) {
  GaussArgs a;
  int rc = gauss_setup(nm, mesh, list_mode_dev, M, a);
  if (rc) return rc;
  a.U = U; a.U2 = U2;
  a.nU = mesh->np * ncU * (int64_t)M; a.nU2 = mesh->np * ncU2 * (int64_t)M;
  if (ps) {
    a.npass = ps->n;
    for (int i = 0; i < ps->n; ++i) {
      a.pc[i] = ps->pc[i]; a.pv[i] = ps->pv[i]; a.coef[i] = ps->coef[i];
      a.npc[i] = a.npv[i] = mesh->np * 6 * (int64_t)M;
    }
  }
  a.heat_capa = heat_capa_dev;
  a.Vout = out; a.nout = a.G * ncout * (int64_t)M;
  for (int i = 0; i < 3; ++i) a.om[i] = om ? om[i] : 0.0;
  if (mesh->me == 0) return SFX_OK;
  if (variant == GV_PAIRS && pairs_as_passes(a.MP)) {
    // one launch per pass, the first storing, the others adding to the output
    const GaussArgs all = a;
    for (int i = 0; i < all.npass; ++i) {
      a.npass = 1;
      a.pc[0] = all.pc[i]; a.pv[0] = all.pv[i]; a.coef[0] = all.coef[i];
      a.npc[0] = all.npc[i]; a.npv[0] = all.npv[i];
      a.accum = i > 0;
      hipError_t e = launch_gauss(GV_CURL_CROSS, a, st);
      if (e != hipSuccess) return hip_fail(e, nm);
    }
    return SFX_OK;
  }
  hipError_t e = launch_gauss(variant, a, st);
  if (e != hipSuccess) return hip_fail(e, nm);
  return SFX_OK;
}

// Host path: copy the distinct nodal arrays of the call to the device (slots
// 4..), run, copy the Gauss output back.
struct HostIn {
  const double* h;
  int nc;
};

static int run_gauss_host(int variant, const char* nm, sfx_mesh_t mesh, const int32_t* list_mode, int32_t M,
                          const HostIn* in, int nin, const int* pass_c, const int* pass_v, const double* coef,
                          int npass, const double* heat_capa, double* out, int ncout, const double om[3]) {
  if (!mesh) return fail(SFX_ERR_ARG, "%s: null mesh", nm);
  if (M <= 0) return fail(SFX_ERR_ARG, "%s: SIZE(list_mode) == 0", nm);
  const size_t nO = (size_t)7 * mesh->me * ncout * M;
  void *dO, *dL, *dH = nullptr;
  void* din[8] = {nullptr};
  int rc;
  if ((rc = staging(0, nO * 8 + 8, &dO)) || (rc = staging(1, M * 4 + 64, &dL))) return rc;
  if (heat_capa && (rc = staging(2, mesh->me * 8 + 8, &dH))) return rc;
  hipError_t e;
  for (int i = 0; i < nin; ++i) {
    const size_t n = (size_t)mesh->np * in[i].nc * M;
    if ((rc = staging(4 + i, n * 8 + 8, &din[i]))) return rc;
    if ((e = hipMemcpy(din[i], in[i].h, n * 8, hipMemcpyHostToDevice)) != hipSuccess) return hip_fail(e, "H2D");
  }
  if (heat_capa && (e = hipMemcpy(dH, heat_capa, mesh->me * 8, hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(e, "H2D");
  if ((e = hipMemcpy(dL, list_mode, M * 4, hipMemcpyHostToDevice)) != hipSuccess) return hip_fail(e, "H2D");
  Passes ps;
  for (int i = 0; i < npass; ++i) ps.add((const double*)din[pass_c[i]], (const double*)din[pass_v[i]], coef[i]);
  rc = run_gauss_dev(variant, nm, mesh, (const int32_t*)dL, M, (const double*)din[0], in[0].nc,
                     nin > 1 ? (const double*)din[1] : nullptr, nin > 1 ? in[1].nc : 0, npass ? &ps : nullptr,
                     (const double*)dH, (double*)dO, ncout, om, nullptr);
  if (rc) return rc;
  if ((e = hipMemcpy(out, dO, nO * 8, hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(e, "D2H");
  return SFX_OK;
}

static void precession_om(int32_t precession, double taux, double angle_s_pi, double om[3]) {
  om[0] = om[1] = om[2] = 0.0;
  if (!precession) return;
  const double PI = std::acos(-1.0);
  // Omega at nodes (subroutine_ns_with_velocity.F90:540-548); RotV += 2 Omega (:621)
  om[0] = 2.0 * (taux * std::sin(angle_s_pi * PI));
  om[1] = 2.0 * (-taux * std::sin(angle_s_pi * PI));
  om[2] = 2.0 * (taux * std::cos(angle_s_pi * PI));
}

}  // namespace sfx

extern "C" {

int sfx_smb_cross_prod_gauss_sft_par(sfx_mesh_t mesh, const int32_t* list_mode, int32_t m_max_c,
                                     const double* V_in, double* V_out, int32_t precession_in,
                                     double taux_precession, double angle_s_pi) {
  clear_err();
  double om[3];
  precession_om(precession_in, taux_precession, angle_s_pi, om);
  const HostIn in[1] = {{V_in, 6}};
  return run_gauss_host(GV_CROSS_CURL, "smb_cross_prod_gauss_sft_par", mesh, list_mode, m_max_c, in, 1, nullptr,
                        nullptr, nullptr, 0, nullptr, V_out, 6, om);
}

int sfx_smb_cross_prod_gauss_sft_par_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c,
                                         const double* V_in, double* V_out, int32_t precession_in,
                                         double taux_precession, double angle_s_pi, void* stream) {
  clear_err();
  double om[3];
  precession_om(precession_in, taux_precession, angle_s_pi, om);
  return run_gauss_dev(GV_CROSS_CURL, "smb_cross_prod_gauss_sft_par", mesh, list_mode_dev, m_max_c, V_in, 6,
                       nullptr, 0, nullptr, nullptr, V_out, 6, om, (hipStream_t)stream);
}

int sfx_nl_rhs_3x3_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* V_in,
                       double* vb_145, double* vb_236, double scale, int32_t precession_in, double taux_precession,
                       double angle_s_pi, double* work_elem, void* stream) {
  clear_err();
  const char* nm = "nl_rhs_3x3";
  GaussArgs a;
  int rc = gauss_setup(nm, mesh, list_mode_dev, m_max_c, a);
  if (rc) return rc;
  if (!V_in || !vb_145 || !vb_236) return fail(SFX_ERR_ARG, "%s: null array", nm);
  if (!work_elem && mesh->me > 0) return fail(SFX_ERR_ARG, "%s: null work_elem", nm);
  precession_om(precession_in, taux_precession, angle_s_pi, a.om);
  a.U = V_in; a.nU = mesh->np * 6 * (int64_t)m_max_c;
  a.wgt = mesh->wgt; a.elo = work_elem; a.nelo = 36 * mesh->me * (int64_t)m_max_c;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = launch_gauss_rhs(GV_CROSS_CURL, a, st);
  if (e != hipSuccess) return hip_fail(e, nm);
  RhsArgs r = rhs_args(mesh, m_max_c, scale);
  r.v145 = vb_145; r.v236 = vb_236; r.elo = work_elem;
  r.n_out = mesh->np * 3 * (int64_t)m_max_c; r.n_elo = a.nelo;
  e = launch_rhs_gather(r, st);
  return e == hipSuccess ? SFX_OK : hip_fail(e, nm);
}

int sfx_nl_qs_00_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* heat_capa_dev,
                     const double* V_in, const double* c_in, const double* ff, double* vect, double scale,
                     double* work_elem, void* stream) {
  clear_err();
  const char* nm = "nl_qs_00";
  GaussArgs a;
  int rc = gauss_setup(nm, mesh, list_mode_dev, m_max_c, a);
  if (rc) return rc;
  if (!heat_capa_dev) return fail(SFX_ERR_ARG, "%s: heat_capa is required", nm);
  if (!V_in || !c_in || !vect) return fail(SFX_ERR_ARG, "%s: null array", nm);
  if (!work_elem && mesh->me > 0) return fail(SFX_ERR_ARG, "%s: null work_elem", nm);
  a.U = V_in; a.nU = mesh->np * 6 * (int64_t)m_max_c;
  a.U2 = c_in; a.nU2 = mesh->np * 2 * (int64_t)m_max_c;
  a.heat_capa = heat_capa_dev;
  a.ffn = ff; a.nffn = mesh->np * 2 * (int64_t)m_max_c;
  a.wgt = mesh->wgt; a.elo = work_elem; a.nelo = 6 * mesh->me * 2 * (int64_t)m_max_c;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = launch_gauss_rhs(GV_UGRADC, a, st);
  if (e != hipSuccess) return hip_fail(e, nm);
  RhsArgs r = rhs_args(mesh, 2 * m_max_c, scale);
  r.vect = vect; r.elo = work_elem;
  r.n_out = mesh->np * 2 * (int64_t)m_max_c; r.n_elo = a.nelo;
  e = launch_qs00_gather(r, st);
  return e == hipSuccess ? SFX_OK : hip_fail(e, nm);
}

int sfx_smb_curlh_cross_b_gauss_sft_par(sfx_mesh_t mesh, const int32_t* list_mode, int32_t m_max_c,
                                        const double* V_in, const double* W_in, double* V_out) {
  clear_err();
  const HostIn in[2] = {{V_in, 6}, {W_in, 6}};
  const int pc[1] = {0}, pv[1] = {1};
  const double k[1] = {1.0};
  return run_gauss_host(GV_PAIRS, "smb_CurlH_cross_B_gauss_sft_par", mesh, list_mode, m_max_c, in, 2, pc, pv, k, 1,
                        nullptr, V_out, 6, nullptr);
}

int sfx_smb_curlh_cross_b_gauss_sft_par_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c,
                                            const double* V_in, const double* W_in, double* V_out, void* stream) {
  clear_err();
  Passes ps;
  ps.add(V_in, W_in, 1.0);
  return run_gauss_dev(GV_PAIRS, "smb_CurlH_cross_B_gauss_sft_par", mesh, list_mode_dev, m_max_c, V_in, 6, W_in,
                       6, &ps, nullptr, V_out, 6, nullptr, (hipStream_t)stream);
}

int sfx_smb_ugradc_gauss_fft_par(sfx_mesh_t mesh, const int32_t* list_mode, int32_t m_max_c,
                                 const double* heat_capa_in, const double* V_in, const double* c_in,
                                 double* c_out) {
  clear_err();
  if (!heat_capa_in) return fail(SFX_ERR_ARG, "smb_ugradc_gauss_fft_par: heat_capa_in is required");
  const HostIn in[2] = {{V_in, 6}, {c_in, 2}};
  return run_gauss_host(GV_UGRADC, "smb_ugradc_gauss_fft_par", mesh, list_mode, m_max_c, in, 2, nullptr, nullptr,
                        nullptr, 0, heat_capa_in, c_out, 2, nullptr);
}

int sfx_smb_ugradc_gauss_fft_par_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c,
                                     const double* heat_capa_dev, const double* V_in, const double* c_in,
                                     double* c_out, void* stream) {
  clear_err();
  if (!heat_capa_dev) return fail(SFX_ERR_ARG, "smb_ugradc_gauss_fft_par: heat_capa_in is required");
  return run_gauss_dev(GV_UGRADC, "smb_ugradc_gauss_fft_par", mesh, list_mode_dev, m_max_c, V_in, 6, c_in, 2,
                       nullptr, heat_capa_dev, c_out, 2, nullptr, (hipStream_t)stream);
}

int sfx_nl_rotv_mhd(sfx_mesh_t mesh, const int32_t* list_mode, int32_t m_max_c, const double* u,
                    const double* H, const double* B, double coeff_lorentz, double* rotv_v) {
  clear_err();
  const HostIn in[3] = {{u, 6}, {H, 6}, {B, 6}};
  const int pc[2] = {1, 0}, pv[2] = {2, 0};
  const double k[2] = {-coeff_lorentz, 1.0};
  return run_gauss_host(GV_PAIRS, "nl_rotv_mhd", mesh, list_mode, m_max_c, in, 3, pc, pv, k, 2, nullptr, rotv_v,
                        6, nullptr);
}

int sfx_nl_rotv_mhd_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* u,
                        const double* H, const double* B, double coeff_lorentz, double* rotv_v, void* stream) {
  clear_err();
  Passes ps;
  ps.add(H, B, -coeff_lorentz);
  ps.add(u, u, 1.0);
  return run_gauss_dev(GV_PAIRS, "nl_rotv_mhd", mesh, list_mode_dev, m_max_c, u, 6, nullptr, 0, &ps, nullptr,
                       rotv_v, 6, nullptr, (hipStream_t)stream);
}

int sfx_nl_tangent(sfx_mesh_t mesh, const int32_t* list_mode, int32_t m_max_c, const double* U,
                   const double* dU, double* out) {
  clear_err();
  const HostIn in[2] = {{U, 6}, {dU, 6}};
  const int pc[2] = {0, 1}, pv[2] = {1, 0};
  const double k[2] = {1.0, 1.0};
  return run_gauss_host(GV_PAIRS, "nl_tangent", mesh, list_mode, m_max_c, in, 2, pc, pv, k, 2, nullptr, out, 6,
                        nullptr);
}

int sfx_nl_tangent_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* U,
                       const double* dU, double* out, void* stream) {
  clear_err();
  Passes ps;
  ps.add(U, dU, 1.0);
  ps.add(dU, U, 1.0);
  return run_gauss_dev(GV_PAIRS, "nl_tangent", mesh, list_mode_dev, m_max_c, U, 6, nullptr, 0, &ps, nullptr, out,
                       6, nullptr, (hipStream_t)stream);
}

}  // extern "C"

namespace sfx {

// Transpose of the linearized pass pair (curl A) x W d + (curl d) x B, times
// coef, applied to ybar.  B == nullptr: A = B and both parts go to out_w
// (the convective adjoint).  Otherwise the W^T part goes to out_w and the
// R^T part to out_r (the Lorentz adjoint for dB and dH).  work_elem holds
// 36*me*M doubles per output.
static int adjoint_dev(const char* nm, sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c,
                       const double* Afield, const double* Bfield, double coef, const double* ybar,
                       double* work_elem, double* out_w, double* out_r, hipStream_t st) {
  if (!mesh) return fail(SFX_ERR_ARG, "%s: null mesh", nm);
  if (m_max_c <= 0) return fail(SFX_ERR_ARG, "%s: SIZE(list_mode) == 0", nm);
  if (m_max_c > 256) return fail(SFX_ERR_UNSUPPORTED, "%s: m_max_c=%d > 256", nm, m_max_c);
  if (!work_elem) return fail(SFX_ERR_ARG, "%s: null work_elem", nm);
  const bool split = Bfield != nullptr;
  AdjArgs A;
  std::memset(&A, 0, sizeof(A));
  GaussArgs& a = A.g;
  a.jj = mesh->jj; a.rr = mesh->rr; a.me = mesh->me; a.np = mesh->np; a.G = 7 * mesh->me;
  a.U = Afield; a.nU = mesh->np * 6 * (int64_t)m_max_c;
  a.list_mode = list_mode_dev; a.tab = mesh->tab;
  a.M = m_max_c; a.MP = mp_for(m_max_c);
  Twiddles tw;
  int rc = get_twiddles(a.MP, &tw);
  if (rc) return rc;
  a.wN = tw.wN; a.wMP = tw.wMP;
  A.UB = Bfield; A.nUB = split ? a.nU : 0;
  A.coef = coef;
  A.ybar = ybar; A.nybar = a.G * 6 * (int64_t)m_max_c;
  A.nelo = 6 * mesh->me * 6 * (int64_t)m_max_c;
  A.elo = work_elem;
  A.elo2 = split ? work_elem + A.nelo : nullptr;
  GatherArgs ga;
  ga.adj_ptr = mesh->adj_ptr; ga.adj_ent = mesh->adj_ent; ga.elo = A.elo; ga.out = out_w;
  ga.me = mesh->me; ga.np = mesh->np; ga.n_adj = mesh->n_adj; ga.nelo = A.nelo;
  ga.nout = mesh->np * 6 * (int64_t)m_max_c; ga.M = m_max_c;
  GatherArgs ga2 = ga;
  ga2.elo = A.elo2; ga2.out = out_r;
  if (mesh->me == 0) return SFX_OK;
  hipError_t e = launch_adjoint(A, ga, split ? &ga2 : nullptr, st);
  return e == hipSuccess ? SFX_OK : hip_fail(e, nm);
}

}  // namespace sfx

extern "C" {

int sfx_nl_adjoint_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* U,
                       const double* ybar, double* work_elem, double* dbar, void* stream) {
  clear_err();
  return adjoint_dev("nl_adjoint", mesh, list_mode_dev, m_max_c, U, nullptr, 1.0, ybar, work_elem, dbar, nullptr,
                     (hipStream_t)stream);
}

int sfx_nl_adjoint_lorentz_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* H0,
                               const double* B0, double coeff_lorentz, const double* ybar, double* work_elem,
                               double* dHbar, double* dBbar, void* stream) {
  clear_err();
  if (!H0 || !B0) return fail(SFX_ERR_ARG, "nl_adjoint_lorentz: null base field");
  return adjoint_dev("nl_adjoint_lorentz", mesh, list_mode_dev, m_max_c, H0, B0, -coeff_lorentz, ybar, work_elem,
                     dBbar, dHbar, (hipStream_t)stream);
}

int sfx_nl_adjoint_mhd_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* U,
                           const double* H0, const double* B0, double coeff_lorentz, const double* ybar,
                           double* work_elem, double* dUbar, double* dHbar, double* dBbar, void* stream) {
  clear_err();
  int rc = sfx_nl_adjoint_dev(mesh, list_mode_dev, m_max_c, U, ybar, work_elem, dUbar, stream);
  if (rc) return rc;
  return sfx_nl_adjoint_lorentz_dev(mesh, list_mode_dev, m_max_c, H0, B0, coeff_lorentz, ybar, work_elem, dHbar,
                                    dBbar, stream);
}

int sfx_nl_adjoint(sfx_mesh_t mesh, const int32_t* list_mode, int32_t m_max_c, const double* U, const double* ybar,
                   double* dbar) {
  clear_err();
  if (!mesh) return fail(SFX_ERR_ARG, "nl_adjoint: null mesh");
  if (m_max_c <= 0) return fail(SFX_ERR_ARG, "nl_adjoint: SIZE(list_mode) == 0");
  const size_t nU = (size_t)mesh->np * 6 * m_max_c, nY = (size_t)7 * mesh->me * 6 * m_max_c;
  const size_t nE = (size_t)6 * mesh->me * 6 * m_max_c;
  void *dU, *dY, *dE, *dO, *dL;
  int rc;
  if ((rc = staging(0, nU * 8 + 8, &dU)) || (rc = staging(1, nY * 8 + 8, &dY)) || (rc = staging(2, nE * 8 + 8, &dE)) ||
      (rc = staging(3, nU * 8 + 8, &dO)) || (rc = staging(4, m_max_c * 4 + 64, &dL)))
    return rc;
  hipError_t e;
  if ((e = hipMemcpy(dU, U, nU * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(dY, ybar, nY * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(dL, list_mode, m_max_c * 4, hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(e, "H2D");
  if ((rc = sfx_nl_adjoint_dev(mesh, (const int32_t*)dL, m_max_c, (const double*)dU, (const double*)dY, (double*)dE,
                               (double*)dO, nullptr)))
    return rc;
  if ((e = hipMemcpy(dbar, dO, nU * 8, hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(e, "D2H");
  return SFX_OK;
}

int sfx_nl_tangent_mhd_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* U,
                           const double* H0, const double* B0, const double* dU, const double* dH,
                           const double* dB, double coeff_lorentz, double* out, void* stream) {
  clear_err();
  Passes ps;
  ps.add(U, dU, 1.0);
  ps.add(dU, U, 1.0);
  ps.add(H0, dB, -coeff_lorentz);
  ps.add(dH, B0, -coeff_lorentz);
  return run_gauss_dev(GV_PAIRS, "nl_tangent_mhd", mesh, list_mode_dev, m_max_c, U, 6, nullptr, 0, &ps, nullptr,
                       out, 6, nullptr, (hipStream_t)stream);
}

}  // extern "C"

// ---- right-hand-side assembly entry points
namespace sfx {

int rhs3x3_dev(sfx_mesh_t mesh, int32_t M, const double* rhs, double* v145, double* v236, double scale,
               double* work_elem, hipStream_t st) {
  if (!work_elem && mesh->me > 0 && M > 0) return fail(SFX_ERR_ARG, "rhs_3x3: null work_elem");
  RhsArgs a = rhs_args(mesh, M, scale);
  a.rhs = rhs; a.v145 = v145; a.v236 = v236; a.elo = work_elem;
  a.n_rhs = a.G * 6 * (int64_t)M; a.n_out = a.np * 3 * (int64_t)M; a.n_elo = 36 * a.me * (int64_t)M;
  hipError_t e = launch_rhs3x3(a, st);
  return e == hipSuccess ? SFX_OK : hip_fail(e, "rhs_3x3");
}

int rhs_ns_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t M, const double* rotv_v, const double* V1m,
               const double* ff, const double* P, const double* P_inc, const double* tensor, double* work_elem,
               double* v145, double* v236, hipStream_t st, double rsign = -1.0) {
  if (!work_elem && mesh->me > 0 && M > 0) return fail(SFX_ERR_ARG, "rhs_ns_3x3: null work_elem");
  if (!list_mode_dev && M > 0) return fail(SFX_ERR_ARG, "rhs_ns_3x3: null list_mode");
  RhsArgs a = rhs_args(mesh, M, 1.0);
  a.rr = mesh->rr; a.lmode = list_mode_dev;
  a.rhs = rotv_v; a.v1m = V1m; a.fsrc = ff; a.pres = P; a.pinc = P_inc; a.tens = tensor; a.rsign = rsign;
  a.v145 = v145; a.v236 = v236; a.elo = work_elem;
  a.n_rhs = a.G * 6 * (int64_t)M; a.n_v1m = a.n_fsrc = a.np * 6 * (int64_t)M; a.n_tens = 3 * a.n_rhs;
  a.n_pres = a.n_pinc = a.np * 2 * (int64_t)M;
  a.n_out = a.np * 3 * (int64_t)M; a.n_elo = 36 * a.me * (int64_t)M;
  hipError_t e = launch_rhs_ns(a, st);
  return e == hipSuccess ? SFX_OK : hip_fail(e, "rhs_ns_3x3");
}

}  // namespace sfx

extern "C" {

int sfx_rhs_ns_3x3_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* rotv_v,
                       const double* V1m, const double* ff, const double* P, const double* P_inc,
                       const double* opt_tensor, double* work_elem, double* vb_145, double* vb_236, void* stream) {
  clear_err();
  if (!mesh) return fail(SFX_ERR_ARG, "rhs_ns_3x3: null mesh");
  if (m_max_c < 0) return fail(SFX_ERR_ARG, "rhs_ns_3x3: m_max_c < 0");
  return rhs_ns_dev(mesh, list_mode_dev, m_max_c, rotv_v, V1m, ff, P, P_inc, opt_tensor, work_elem, vb_145,
                    vb_236, (hipStream_t)stream);
}

int sfx_rhs_ns_3x3_penal_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, int32_t mode0_slot,
                             const double* rotv_v, const double* V1m, const double* ff, const double* P,
                             const double* P_inc, const double* penal_grid, const double* imposed_vel,
                             const double* opt_tensor, double* work_gauss, double* work_elem, double* vb_145,
                             double* vb_236, void* stream) {
  clear_err();
  const char* nm = "rhs_ns_3x3 (penalty)";
  if (!mesh) return fail(SFX_ERR_ARG, "%s: null mesh", nm);
  if (m_max_c <= 0) return fail(SFX_ERR_ARG, "%s: SIZE(list_mode) == 0", nm);
  if (!penal_grid) return fail(SFX_ERR_ARG, "%s: null penal_grid", nm);
  if (!work_gauss && mesh->me > 0) return fail(SFX_ERR_ARG, "%s: null work_gauss", nm);
  if (mode0_slot < -1 || mode0_slot >= m_max_c) return fail(SFX_ERR_ARG, "%s: mode0_slot out of range", nm);
  if (!work_elem && mesh->me > 0) return fail(SFX_ERR_ARG, "%s: null work_elem", nm);
  hipStream_t st = (hipStream_t)stream;
  const int64_t G = 7 * mesh->me, nG = G * 6 * (int64_t)m_max_c;
  double* g0 = work_gauss;
  double* g1 = work_gauss + nG;
  // (1) ft + fp_inc + fs - rotv_v at the Gauss points, minus the imposed
  //     velocity on the mode-0 slot (computing_rhs_gauss.F90:54-121)
  RhsArgs a = rhs_args(mesh, m_max_c, 1.0);
  a.rr = mesh->rr; a.lmode = list_mode_dev;
  a.rhs = rotv_v; a.v1m = V1m; a.fsrc = ff; a.pinc = P_inc; a.rsign = -1.0;
  a.imp = imposed_vel; a.imp_slot = imposed_vel ? mode0_slot : -1; a.isign = -1.0; a.gout = g0;
  a.n_rhs = nG; a.n_v1m = a.n_fsrc = a.np * 6 * (int64_t)m_max_c; a.n_pinc = a.np * 2 * (int64_t)m_max_c;
  a.n_imp = a.np * 6; a.n_gout = nG;
  hipError_t e = launch_rhs_ns_gauss(a, st);
  if (e != hipSuccess) return hip_fail(e, nm);
  // (2) the product with the penalty function on the grid, m_max_pad = 3*m/2
  //     (FFT_PAR_VAR_ETA_PROD_GAUSS_DCL, :124-129)
  int rc = sfx_fft_par_var_eta_prod_dcl_dev(penal_grid, g0, g1, G, 6, m_max_c, (3 * m_max_c) / 2, stream);
  if (rc) return rc;
  // (3) plus the imposed velocity on the mode-0 slot, plus fp, assembled by
  //     rhs_3x3 with its optional tensor (:132-145)
  RhsArgs b = rhs_args(mesh, m_max_c, 1.0);
  b.rr = mesh->rr; b.lmode = list_mode_dev;
  b.rhs = g1; b.pres = P; b.tens = opt_tensor; b.rsign = 1.0;
  b.imp = imposed_vel; b.imp_slot = imposed_vel ? mode0_slot : -1; b.isign = 1.0;
  b.v145 = vb_145; b.v236 = vb_236; b.elo = work_elem;
  b.n_rhs = nG; b.n_pres = b.np * 2 * (int64_t)m_max_c; b.n_tens = 3 * nG; b.n_imp = b.np * 6;
  b.n_out = b.np * 3 * (int64_t)m_max_c; b.n_elo = 36 * b.me * (int64_t)m_max_c;
  e = launch_rhs_ns(b, st);
  return e == hipSuccess ? SFX_OK : hip_fail(e, nm);
}

int sfx_rhs_3x3_dev(sfx_mesh_t mesh, int32_t m_max_c, const double* rhs_gauss, double* vb_145, double* vb_236,
                    double scale, double* work_elem, void* stream) {
  clear_err();
  if (!mesh) return fail(SFX_ERR_ARG, "rhs_3x3: null mesh");
  if (m_max_c < 0) return fail(SFX_ERR_ARG, "rhs_3x3: m_max_c < 0");
  return rhs3x3_dev(mesh, m_max_c, rhs_gauss, vb_145, vb_236, scale, work_elem, (hipStream_t)stream);
}

int sfx_rhs_3x3_bdy_dev(sfx_mesh_t mesh, int32_t m_max_c, const double* tensor_bdy, double* vb_145,
                        double* vb_236, double scale, void* stream) {
  clear_err();
  if (!mesh) return fail(SFX_ERR_ARG, "rhs_3x3 boundary term: null mesh");
  if (mesh->mes < 0)
    return fail(SFX_ERR_ARG, "rhs_3x3 boundary term: no boundary sides registered (sfx_mesh_set_boundary)");
  if (m_max_c < 0) return fail(SFX_ERR_ARG, "rhs_3x3 boundary term: m_max_c < 0");
  if (mesh->nb == 0 || m_max_c == 0) return SFX_OK;
  if (!tensor_bdy || !vb_145 || !vb_236) return fail(SFX_ERR_ARG, "rhs_3x3 boundary term: null array");
  BdyArgs a;
  std::memset(&a, 0, sizeof(a));
  double d[9], p[3];
  p2_edge_tables(a.wws, d, p);
  a.bptr = mesh->bptr; a.bent = mesh->bent; a.bnode = mesh->bnode; a.bw = mesh->bw;
  a.nb = mesh->nb; a.mes = mesh->mes; a.np = mesh->np; a.M = m_max_c; a.scale = scale;
  a.tb = tensor_bdy; a.v145 = vb_145; a.v236 = vb_236;
  a.n_tb = 3 * mesh->mes * 6 * (int64_t)m_max_c; a.n_out = mesh->np * 3 * (int64_t)m_max_c; a.n_ent = mesh->n_bent;
  hipError_t e = launch_rhs_bdy(a, (hipStream_t)stream);
  return e == hipSuccess ? SFX_OK : hip_fail(e, "rhs_3x3 boundary term");
}

int sfx_rhs_3x3_mode(sfx_mesh_t mesh, int32_t mode, const double* rhs_in, const double* opt_tensor,
                     const double* opt_tensor_scaln_bdy, double* vb_145, double* vb_236) {
  clear_err();
  if (!mesh) return fail(SFX_ERR_ARG, "rhs_3x3: null mesh");
  if (opt_tensor_scaln_bdy && mesh->mes < 0)
    return fail(SFX_ERR_ARG, "rhs_3x3: opt_tensor_scaln_bdy needs the boundary sides (sfx_mesh_set_boundary)");
  const size_t nG = (size_t)7 * mesh->me * 6, nN = (size_t)mesh->np * 3;
  const size_t nB = opt_tensor_scaln_bdy ? (size_t)3 * mesh->mes * 6 : 0;
  void *dr, *d1, *d2, *dw, *dt, *db, *dm;
  int rc;
  if ((rc = staging(0, nG * 8 + 8, &dr)) || (rc = staging(1, nN * 8 + 8, &d1)) || (rc = staging(2, nN * 8 + 8, &d2)) ||
      (rc = staging(3, (size_t)36 * mesh->me * 8 + 8, &dw)) || (rc = staging(4, (opt_tensor ? 3 * nG : 0) * 8 + 8, &dt)) ||
      (rc = staging(5, nB * 8 + 8, &db)) || (rc = staging(6, 8, &dm)))
    return rc;
  hipError_t e;
  if ((e = hipMemcpy(dr, rhs_in, nG * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (opt_tensor && (e = hipMemcpy(dt, opt_tensor, 3 * nG * 8, hipMemcpyHostToDevice)) != hipSuccess) ||
      (nB && (e = hipMemcpy(db, opt_tensor_scaln_bdy, nB * 8, hipMemcpyHostToDevice)) != hipSuccess) ||
      (e = hipMemcpy(dm, &mode, sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(e, "H2D");
  // with opt_tensor: the NS right-hand-side element kernel with rhs_in as its
  // Gauss field (sign +1) and every other field absent, plus the tensor terms
  if (opt_tensor) {
    if ((rc = rhs_ns_dev(mesh, (const int32_t*)dm, 1, (const double*)dr, nullptr, nullptr, nullptr, nullptr,
                         (const double*)dt, (double*)dw, (double*)d1, (double*)d2, nullptr, 1.0)))
      return rc;
  } else if ((rc = rhs3x3_dev(mesh, 1, (const double*)dr, (double*)d1, (double*)d2, 1.0, (double*)dw, nullptr))) {
    return rc;
  }
  if (nB && (rc = sfx_rhs_3x3_bdy_dev(mesh, 1, (const double*)db, (double*)d1, (double*)d2, 1.0, nullptr)))
    return rc;
  if ((e = hipMemcpy(vb_145, d1, nN * 8, hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(e, "D2H");
  if ((e = hipMemcpy(vb_236, d2, nN * 8, hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(e, "D2H");
  return SFX_OK;
}

int sfx_rhs_3x3(sfx_mesh_t mesh, int32_t mode, const double* rhs_in, double* vb_145, double* vb_236) {
  clear_err();
  (void)mode;   // only used by the opt_tensor terms of the reference
  if (!mesh) return fail(SFX_ERR_ARG, "rhs_3x3: null mesh");
  const size_t nG = (size_t)7 * mesh->me * 6, nN = (size_t)mesh->np * 3;
  void *dr, *d1, *d2, *dw;
  int rc;
  if ((rc = staging(0, nG * 8 + 8, &dr)) || (rc = staging(1, nN * 8 + 8, &d1)) || (rc = staging(2, nN * 8 + 8, &d2)) ||
      (rc = staging(3, (size_t)36 * mesh->me * 8 + 8, &dw)))
    return rc;
  hipError_t e;
  if ((e = hipMemcpy(dr, rhs_in, nG * 8, hipMemcpyHostToDevice)) != hipSuccess) return hip_fail(e, "H2D");
  if ((rc = rhs3x3_dev(mesh, 1, (const double*)dr, (double*)d1, (double*)d2, 1.0, (double*)dw, nullptr))) return rc;
  if ((e = hipMemcpy(vb_145, d1, nN * 8, hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(e, "D2H");
  if ((e = hipMemcpy(vb_236, d2, nN * 8, hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(e, "D2H");
  return SFX_OK;
}

int sfx_qs_00_gauss_dev(sfx_mesh_t mesh, int32_t ncols, const double* heat_capa, const double* ff,
                        const double* ff_gauss, double* work_elem, double* vect, void* stream) {
  clear_err();
  if (!mesh) return fail(SFX_ERR_ARG, "qs_00_gauss: null mesh");
  if (ncols < 0) return fail(SFX_ERR_ARG, "qs_00_gauss: ncols < 0");
  if (!work_elem && mesh->me > 0 && ncols > 0) return fail(SFX_ERR_ARG, "qs_00_gauss: null work_elem");
  RhsArgs a = rhs_args(mesh, ncols, 1.0);
  a.heat_capa = heat_capa; a.ff = ff; a.ffg = ff_gauss; a.vect = vect; a.elo = work_elem;
  a.n_ff = a.np * (int64_t)ncols; a.n_ffg = a.G * (int64_t)ncols; a.n_out = a.np * (int64_t)ncols;
  a.n_elo = 6 * a.me * (int64_t)ncols;
  hipError_t e = launch_qs00(a, (hipStream_t)stream);
  return e == hipSuccess ? SFX_OK : hip_fail(e, "qs_00_gauss");
}

int sfx_qs_00_gauss(sfx_mesh_t mesh, const double* heat_capa, const double* ff, const double* ff_gauss,
                    double* vect) {
  clear_err();
  if (!mesh) return fail(SFX_ERR_ARG, "qs_00_gauss: null mesh");
  const size_t nG = (size_t)7 * mesh->me, nN = (size_t)mesh->np;
  void *dh, *df, *dg, *dv, *dw;
  int rc;
  if ((rc = staging(0, mesh->me * 8 + 8, &dh)) || (rc = staging(1, nN * 8 + 8, &df)) ||
      (rc = staging(2, nG * 8 + 8, &dg)) || (rc = staging(3, nN * 8 + 8, &dv)) ||
      (rc = staging(4, (size_t)6 * mesh->me * 8 + 8, &dw)))
    return rc;
  hipError_t e;
  if ((e = hipMemcpy(dh, heat_capa, mesh->me * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(df, ff, nN * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(dg, ff_gauss, nG * 8, hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(e, "H2D");
  if ((rc = sfx_qs_00_gauss_dev(mesh, 1, (const double*)dh, (const double*)df, (const double*)dg, (double*)dw,
                                (double*)dv, nullptr)))
    return rc;
  if ((e = hipMemcpy(vect, dv, nN * 8, hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(e, "D2H");
  return SFX_OK;
}

int sfx_nl_rhs_mhd_dev(sfx_mesh_t mesh, const int32_t* list_mode_dev, int32_t m_max_c, const double* u,
                       const double* H, const double* B, double coeff_lorentz, double* work_gauss,
                       double* work_elem, double* vb_145, double* vb_236, void* stream) {
  clear_err();
  int rc = sfx_nl_rotv_mhd_dev(mesh, list_mode_dev, m_max_c, u, H, B, coeff_lorentz, work_gauss, stream);
  if (rc) return rc;
  // the nonlinear share of rhs_gauss is -rotv_v (computing_rhs_gauss.F90:100)
  return rhs3x3_dev(mesh, m_max_c, work_gauss, vb_145, vb_236, -1.0, work_elem, (hipStream_t)stream);
}

}  // extern "C"
