// C-ABI of the exact-grid products: DFT tables, scratch, argument checks.
//
// The device-pointer paths keep their scratch per (device, stream)
// (stream_scratch): calls on different streams never share it.  The
// host-pointer paths stage their arrays in the per-device staging slots.
#include <math.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/sfx.h"
#include "sfx_api_internal.h"
#include "sfx_exact_body.h"
#include "sfx_exact_setup.h"

namespace sfx {

// ---- c2r and r2c tables per (device, M, K), built on the host in long double
static std::mutex g_exact_mu;
static std::map<std::tuple<int, int, int>, double*> g_exact_tabs;

static int exact_tables(int M, int K, const double** ctab, const double** rtab) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  const size_t nc = exact_ctab_size(M, K), nr = exact_rtab_size(M, K);
  std::lock_guard<std::mutex> lk(g_exact_mu);
  auto key = std::make_tuple(dev, M, K);
  auto it = g_exact_tabs.find(key);
  if (it == g_exact_tabs.end()) {
    std::vector<double> h;
    exact_host_tables(M, K, h);
    double* d = nullptr;
    e = hipMalloc(&d, (nc + nr) * sizeof(double));
    if (e != hipSuccess) return fail(SFX_ERR_NOMEM, "hipMalloc(exact tables): %s", hipGetErrorString(e));
    e = hipMemcpy(d, h.data(), (nc + nr) * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(exact tables)");
    it = g_exact_tabs.emplace(key, d).first;
  }
  *ctab = it->second;
  *rtab = it->second + nc;
  return SFX_OK;
}

// Staging slots of the exact-grid products (the FFT products use 0-2).
enum { SLOT_EX_IN1 = 3, SLOT_EX_IN2 = 4, SLOT_EX_OUT = 5, SLOT_EX_W = 6, SLOT_EX_MAX = 7, SLOT_EX_IN3 = 8,
       SLOT_EX_OUT2 = 9, SLOT_EX_IN4 = 10, SLOT_EX_NINT = 11, SLOT_EX_IN5 = 12, SLOT_EX_OUT3 = 14,
       SLOT_EX_AUX = 15, SLOT_EX_STAT = 16, SLOT_EX_FLAG = 17 };

// Common set-up: sizes, tables, scratch.  `a.in`, `a.out`, nin, nout filled by the caller.
static int exact_prepare(const char* nm, ExactArgs& a, int op, int64_t np, int M, int m_max_pad,
                         hipStream_t st) {
  if (M <= 0) return fail(SFX_ERR_ARG, "BUG in %s: m_max_c==0", nm);
  if (np < 0) return fail(SFX_ERR_ARG, "%s: np < 0", nm);
  if (M > 256) return fail(SFX_ERR_UNSUPPORTED, "%s: m_max_c=%d > 256", nm, M);
  if (m_max_pad < M) return fail(SFX_ERR_ARG, "%s: m_max_pad=%d < m_max=%d", nm, m_max_pad, M);
  if (m_max_pad > 4096) return fail(SFX_ERR_UNSUPPORTED, "%s: m_max_pad=%d > 4096", nm, m_max_pad);
  exact_setup_sizes(a, np, M, m_max_pad);
  int rc = exact_tables(M, a.K, &a.ctab, &a.rtab);
  if (rc) return rc;
  void* w = nullptr;
  if (exact_is_reduction(op)) {
    const int64_t nb = exact_grid_blocks(op, a.nin, np, a.K);
    if ((rc = stream_scratch(SLOT_EX_W, (size_t)(nb > 0 ? nb : 1) * 8, st, &w))) return rc;
    a.pmax = (double*)w;
  } else {
    a.n_W = 2 * (int64_t)a.nout * (a.K + 1) * np;
    if ((rc = stream_scratch(SLOT_EX_W, (size_t)(a.n_W > 0 ? a.n_W : 1) * 8, st, &w))) return rc;
    a.W = (double*)w;
  }
  return SFX_OK;
}

static int exact_run(const char* nm, int op, const ExactArgs& a, double* max_out, hipStream_t st) {
  if (a.np == 0 && !exact_is_reduction(op)) return SFX_OK;
  hipError_t e = launch_exact(op, a, max_out, st);
  return e == hipSuccess ? SFX_OK : hip_fail(e, nm);
}

// ---- argument sets of each product on device arrays

// c1, c2, c_out: (np, 2, M)
static int div_dev(const double* c1, const double* c2, double* c, int64_t np, int M, int mp, hipStream_t st) {
  ExactArgs a;
  exact_setup_div(a, c1, c2, c, np, M);
  int rc = exact_prepare("FFT_PAR_DIV_DCL", a, EX_DIV, np, M, mp, st);
  return rc ? rc : exact_run("FFT_PAR_DIV_DCL", EX_DIV, a, nullptr, st);
}

// V (np, 6, M) / s (np, 2, M) -> (np, 6, M)
static int sv2_dev(const double* V, const double* s, double* out, int64_t np, int M, int mp, hipStream_t st) {
  ExactArgs a;
  exact_setup_sv2(a, V, s, out, np, M);
  int rc = exact_prepare("FFT_SCALAR_VECT_DCL", a, EX_SV2, np, M, mp, st);
  return rc ? rc : exact_run("FFT_SCALAR_VECT_DCL", EX_SV2, a, nullptr, st);
}

// c (np, nb, M) -> first scalar of out (np, 2, M)
static int sqrt_dev(const double* c, double* out, int64_t np, int nb, int M, int mp, hipStream_t st) {
  ExactArgs a;
  exact_setup_sqrt(a, c, out, np, nb, M);
  int rc = exact_prepare("FFT_SQUAREROOT_SCALAR", a, EX_SQRT, np, M, mp, st);
  return rc ? rc : exact_run("FFT_SQUAREROOT_SCALAR", EX_SQRT, a, nullptr, st);
}

// V (np, 6, M), s (np, 2, M) clamped to [min, max] of bounds -> (np, 6, M)
static int svclamp_dev(const double* bounds, int nbd, const double* V, const double* s, double* out, int64_t np,
                       int M, int pb, int mp, hipStream_t st) {
  if (!bounds || nbd < 1) return fail(SFX_ERR_ARG, "FFT_SCALAR_VECT_NO_OVERSHOOT: empty scalar_bounds");
  double lo = bounds[0], hi = bounds[0];
  for (int i = 1; i < nbd; ++i) {
    lo = fmin(lo, bounds[i]);
    hi = fmax(hi, bounds[i]);
  }
  ExactArgs a;
  exact_setup_svclamp(a, V, s, out, np, M, pb, lo, hi);
  int rc = exact_prepare("FFT_SCALAR_VECT_NO_OVERSHOOT", a, EX_SVCLAMP, np, M, mp, st);
  return rc ? rc : exact_run("FFT_SCALAR_VECT_NO_OVERSHOOT", EX_SVCLAMP, a, nullptr, st);
}

// first scalar of c (np, nb, M) clamped to [0, 1] on the grid -> first scalar of out
static int clamp01_dev(const double* c, double* out, int64_t np, int nb, int M, int mp, hipStream_t st) {
  ExactArgs a;
  exact_setup_sqrt(a, c, out, np, nb, M);
  int rc = exact_prepare("FFT_NO_OVERSHOOT_LEVEL_SET", a, EX_CLAMP01, np, M, mp, st);
  return rc ? rc : exact_run("FFT_NO_OVERSHOOT_LEVEL_SET", EX_CLAMP01, a, nullptr, st);
}

static int entro_dev(const double* V, const double* c, const double* creal, double coeff, double* Vout,
                     double* cout, int64_t np, int M, int mp, hipStream_t st) {
  if (!creal && np > 0) return fail(SFX_ERR_ARG, "FFT_PAR_ENTRO_VISC_DCL: null c_in_real");
  ExactArgs a;
  exact_setup_entro(a, V, c, creal, coeff, Vout, cout, np, M, mp);
  int rc = exact_prepare("FFT_PAR_ENTRO_VISC_DCL", a, EX_ENTRO, np, M, mp, st);
  return rc ? rc : exact_run("FFT_PAR_ENTRO_VISC_DCL", EX_ENTRO, a, nullptr, st);
}

static int compr_dev(const double* V1, const double* V2, const double* c, const double* creal, double coeff,
                     double comp, double* Vout, double* cout, int64_t np, int M, int mp, hipStream_t st) {
  if (!creal && np > 0) return fail(SFX_ERR_ARG, "FFT_PAR_COMPR_ENTRO_VISC_DCL: null c_in_real");
  ExactArgs a;
  exact_setup_compr(a, V1, V2, c, creal, coeff, comp, Vout, cout, np, M, mp);
  int rc = exact_prepare("FFT_PAR_COMPR_ENTRO_VISC_DCL", a, EX_COMPR, np, M, mp, st);
  return rc ? rc : exact_run("FFT_PAR_COMPR_ENTRO_VISC_DCL", EX_COMPR, a, nullptr, st);
}

static int evisc_dev(const double* const (&V)[5], const double* hloc, double* const (&O)[3], int64_t np, int M,
                     int mp, int lg, double c1, double c2, double c4, double res, hipStream_t st) {
  const char* nm = "FFT_COMPUTE_ENTROPY_VISC";
  if (lg < 1 || lg > 16) return fail(SFX_ERR_UNSUPPORTED, "%s: l_G=%d (1..16)", nm, lg);
  if (np % lg != 0) return fail(SFX_ERR_ARG, "%s: np=%lld is not a multiple of l_G=%d", nm, (long long)np, lg);
  if (!hloc && np > 0) return fail(SFX_ERR_ARG, "%s: null hloc_gauss", nm);
  if (mp < 1 || mp > 4096) return fail(SFX_ERR_ARG, "%s: m_max_pad=%d", nm, mp);
  ExactArgs a;
  exact_setup_evisc(a, V, hloc, O, np, M, lg, c1, c2, c4, res);
  a.n_nint = (np / lg) * (2 * (int64_t)mp - 1);
  void* w = nullptr;
  int rc = stream_scratch(SLOT_EX_NINT, (size_t)(a.n_nint > 0 ? a.n_nint : 1) * 8, st, &w);
  if (rc) return rc;
  a.nint = (double*)w;
  // the viscosity on the grid between the two product passes (N, np)
  void* vr = nullptr;
  if ((rc = stream_scratch(SLOT_EX_AUX, (size_t)(2 * (int64_t)mp - 1) * (np > 0 ? np : 1) * 8, st, &vr))) return rc;
  a.rout2 = (double*)vr;
  if ((rc = exact_prepare(nm, a, EX_EVISC, np, M, mp, st))) return rc;
  return exact_run(nm, EX_EVISC, a, nullptr, st);
}

static int emom_dev(const double* V1, const double* V2, const double* V3, const double* c, const double* hloc,
                    double* out1, double* out2, int64_t np, int M, int mp, int lg, double c2, double c4,
                    hipStream_t st) {
  const char* nm = "FFT_COMPUTE_ENTROPY_VISC_MOM";
  if (M <= 0) return fail(SFX_ERR_ARG, "BUG in %s: m_max_c==0", nm);
  if (M > 256) return fail(SFX_ERR_UNSUPPORTED, "%s: m_max_c=%d > 256", nm, M);
  if (mp < M || mp > 4096) return fail(SFX_ERR_ARG, "%s: m_max_pad=%d", nm, mp);
  if (lg < 1 || lg > 16) return fail(SFX_ERR_UNSUPPORTED, "%s: l_G=%d (1..16)", nm, lg);
  if (np < 0 || np % lg != 0) return fail(SFX_ERR_ARG, "%s: np=%lld is not a multiple of l_G=%d", nm, (long long)np, lg);
  if (np == 0) return SFX_OK;
  if (!hloc || !out1) return fail(SFX_ERR_ARG, "%s: null hloc_gauss or c1_real_out", nm);
  ExactArgs a;
  exact_setup_emom(a, V1, V2, V3, c, hloc, out1, out2, np, M, lg, c2, c4);
  exact_setup_sizes(a, np, M, mp);
  int rc = exact_tables(M, a.K, &a.ctab, &a.rtab);
  if (rc) return rc;
  const int64_t ne = (np / lg) * (2 * (int64_t)mp - 1);
  void *w = nullptr, *pm = nullptr, *dm = nullptr;
  if ((rc = stream_scratch(SLOT_EX_NINT, (size_t)3 * ne * 8, st, &w)) ||
      (rc = stream_scratch(SLOT_EX_AUX, (size_t)exact_emom_blocks(np, lg, a.K) * 8, st, &pm)) ||
      (rc = stream_scratch(SLOT_EX_MAX, 8, st, &dm)))
    return rc;
  a.nint = (double*)w;
  a.nint2 = a.nint + ne;
  a.nint3 = a.nint2 + ne;
  a.n_nint = ne;
  a.pmax = (double*)pm;
  a.dmax = (const double*)dm;
  hipError_t e = launch_exact_emom(a, (double*)dm, st);
  return e == hipSuccess ? SFX_OK : hip_fail(e, nm);
}

static int gradmom_dev(const double* V1, const double* V2, const double* V3, const double* creal, double coeff,
                       double* out, int64_t np, int M, int mp, hipStream_t st) {
  if (!creal && np > 0) return fail(SFX_ERR_ARG, "FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM: null c_in_real");
  ExactArgs a;
  exact_setup_gradmom(a, V1, V2, V3, creal, coeff, out, np, M, mp);
  int rc = exact_prepare("FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM", a, EX_GRADMOM, np, M, mp, st);
  return rc ? rc : exact_run("FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM", EX_GRADMOM, a, nullptr, st);
}

// V1 (n_int, np, 2, M), values (n_int + 1) -> (np, 2, M)
static int heavi_dev(const double* V1, const double* values, int n_values, double* out, int64_t np, int M,
                     int mp, double eps, hipStream_t st) {
  const int n_int = n_values - 1;
  if (n_int < 1) return fail(SFX_ERR_ARG, "FFT_HEAVISIDE_DCL: SIZE(values) < 2");
  if (n_int > EXACT_MAX_INT)
    return fail(SFX_ERR_UNSUPPORTED, "FFT_HEAVISIDE_DCL: %d interfaces > %d", n_int, EXACT_MAX_INT);
  ExactArgs a;
  exact_setup_heavi(a, V1, values, n_values, out, np, M, eps);
  int rc = exact_prepare("FFT_HEAVISIDE_DCL", a, EX_HEAVI, np, M, mp, st);
  return rc ? rc : exact_run("FFT_HEAVISIDE_DCL", EX_HEAVI, a, nullptr, st);
}

// V1, V2 (np, 6, M) -> V_out (3, np, 6, M)
static int tension_dev(const double* V1, const double* V2, double* out, int64_t np, int M, int mp, hipStream_t st) {
  ExactArgs a;
  exact_setup_tension(a, V1, V2, out, np, M);
  int rc = exact_prepare("FFT_TENSOR_DCL", a, EX_TENSION, np, M, mp, st);
  return rc ? rc : exact_run("FFT_TENSOR_DCL", EX_TENSION, a, nullptr, st);
}

// V1 (np, nb, M), nb = 6 (|v|) or 2 (|c|) -> max_out[0] (device)
static int maxvel_dev(const double* V1, double* max_out, int64_t np, int nb, int M, int mp, hipStream_t st) {
  ExactArgs a;
  exact_setup_maxvel(a, V1, np, nb, M);
  int rc = exact_prepare("FFT_MAX_VEL_DCL", a, EX_MAXVEL, np, M, mp, st);
  return rc ? rc : exact_run("FFT_MAX_VEL_DCL", EX_MAXVEL, a, max_out, st);
}

static int cls_dev(const double* V1, const double* V2, const double* c, double* cout, int64_t np, int M, int mp,
                   double comp, double c4, double hmin, hipStream_t st) {
  const char* nm = "FFT_COMPRESSION_LEVEL_SET_DCL";
  // max |velocity| over points and angles (the element and angle maxima the
  // reference takes first do not change it)
  void* dm = nullptr;
  int rc = stream_scratch(SLOT_EX_MAX, 8, st, &dm);
  if (rc) return rc;
  if ((rc = maxvel_dev(V2, (double*)dm, np, 6, M, mp, st))) return rc;
  ExactArgs a;
  exact_setup_cls(a, V1, c, cout, (const double*)dm, np, M, comp, c4, hmin);
  if ((rc = exact_prepare(nm, a, EX_CLS, np, M, mp, st))) return rc;
  return exact_run(nm, EX_CLS, a, nullptr, st);
}

// V (np, nb, M) -> real planes out(N, nb/2, np), N = 2*mp - 1: one pass per field
static int real_dev(const double* V, double* out, int64_t np, int nb, int M, int mp, hipStream_t st) {
  const char* nm = "FFT_PAR_REAL";
  if (nb % 2 != 0 || M == 0) return fail(SFX_ERR_ARG, "Bug in FFT_PAR_REAL: MOD(nb_field,2)/=0 .OR. m_max_c==0");
  const int N = 2 * mp - 1;
  for (int f = 0; f < nb / 2; ++f) {
    ExactArgs a;
    exact_setup_real(a, V, f, nb, out, np, M, N);
    int rc = exact_prepare(nm, a, EX_REAL, np, M, mp, st);
    if (rc || (rc = exact_run(nm, EX_REAL, a, nullptr, st))) return rc;
  }
  return SFX_OK;
}

// V1 (np, nb, M) -> out[0] = max |v|, out[1] = the reference's minimum (device)
static int maxminvel_dev(const double* V1, double* out, int64_t np, int nb, int M, int mp, hipStream_t st) {
  const char* nm = "FFT_MAX_MIN_VEL_DCL";
  ExactArgs a;
  exact_setup_maxvel(a, V1, np, nb, M);
  int rc = exact_prepare(nm, a, EX_MAXVEL, np, M, mp, st);
  if (rc) return rc;
  // one pass: three maxima per work-group, then one reduction block
  void* part = nullptr;
  const int64_t nblk = exact_velstat_blocks(a.nin, np, a.K);
  if ((rc = stream_scratch(SLOT_EX_STAT, (size_t)(nblk > 0 ? nblk : 1) * 3 * 8, st, &part))) return rc;
  hipError_t e = launch_exact_velstat(a, (double*)part, out, st);
  return e == hipSuccess ? SFX_OK : hip_fail(e, nm);
}

// V1 (nb_fluids-1, np, 2, M) -> ok[0] (device int)
static int checkif_dev(const double* V1, int nb_fluids, int32_t* ok, int64_t np, int M, int mp, hipStream_t st) {
  const char* nm = "FFT_CHECK_INTERFACE";
  const int n_int = nb_fluids - 1;
  if (n_int < 1) return fail(SFX_ERR_ARG, "%s: nb_fluids=%d < 2", nm, nb_fluids);
  if (n_int > EXACT_MAX_INT) return fail(SFX_ERR_UNSUPPORTED, "%s: %d level sets > %d", nm, n_int, EXACT_MAX_INT);
  void* f = nullptr;
  int rc = stream_scratch(SLOT_EX_FLAG, 8, st, &f);
  if (rc) return rc;
  if (n_int == 1) {
    f = nullptr;   // two fluids: nothing to compare (:2390-2392), interface_ok = 1
  } else {
    ExactArgs a;
    exact_setup_checkif(a, V1, n_int, np, M);
    if ((rc = exact_prepare(nm, a, EX_CHECKIF, np, M, mp, st)) || (rc = exact_run(nm, EX_CHECKIF, a, (double*)f, st)))
      return rc;
  }
  hipError_t e = launch_exact_checkif((const double*)f, ok, st);
  return e == hipSuccess ? SFX_OK : hip_fail(e, nm);
}

// eta (N, np), c (np, nb, M) -> out (np, nb, M), nb = 2 or 6
static int eta_dev(const double* eta, const double* c, double* out, int64_t np, int nb, int M, int mp,
                   hipStream_t st) {
  const char* nm = "FFT_PAR_VAR_ETA_PROD_DCL";
  if (M == 0) return fail(SFX_ERR_ARG, " BUG in FFT_PAR_VAR_ETA_PROD_T_DCL ");
  if (nb != 2 && nb != 6)
    return fail(SFX_ERR_ARG, "error in nb_field of c_in while calling FFT_PAR_VAR_ETA_PROD_T_DCL");
  if (!eta && np > 0) return fail(SFX_ERR_ARG, "%s: null eta", nm);
  const int op = nb == 2 ? EX_ETA : EX_ETA3;
  ExactArgs a;
  exact_setup_eta(a, eta, c, out, np, nb, M, mp);
  int rc = exact_prepare(nm, a, op, np, M, mp, st);
  return rc ? rc : exact_run(nm, op, a, nullptr, st);
}

// real (N, np) values on the grid -> c (np, 2, M)
static int fromreal_dev(const double* real, double* c, int64_t np, int M, int mp, hipStream_t st) {
  const char* nm = "FFT_PAR_FROM_REAL";
  if (!real && np > 0) return fail(SFX_ERR_ARG, "%s: null real_in", nm);
  ExactArgs a;
  exact_setup_fromreal(a, real, c, np, M, mp);
  int rc = exact_prepare(nm, a, EX_FROMREAL, np, M, mp, st);
  return rc ? rc : exact_run(nm, EX_FROMREAL, a, nullptr, st);
}

// ---- host-pointer paths: stage, run, copy back, synchronise
struct HostRun {
  double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
};

static int to_dev(int slot, const double* h, size_t n, void** d) {
  int rc = staging(slot, n * 8, d);
  if (rc) return rc;
  hipError_t e = n ? hipMemcpy(*d, h, n * 8, hipMemcpyHostToDevice) : hipSuccess;
  return e == hipSuccess ? SFX_OK : hip_fail(e, "H2D");
}

static int from_dev(double* h, const void* d, size_t n) {
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "exact-grid product");
  e = n ? hipMemcpy(h, d, n * 8, hipMemcpyDeviceToHost) : hipSuccess;
  return e == hipSuccess ? SFX_OK : hip_fail(e, "D2H");
}

static void set_temps(double* temps, const HostRun& t) {
  if (temps) {
    temps[0] = (t.t1 - t.t0) + (t.t3 - t.t2);
    temps[1] = t.t2 - t.t1;
    temps[2] = 0.0;
  }
}

// In-place scalar products (square root, clamp to [0, 1]): the result overwrites
// the first scalar of its own input, so the kernels read a copy.
static int inplace_dev(int op, const char* nm, double* c1_inout, int64_t np, int32_t nb_field, int32_t m_max_c,
                       int32_t m_max_pad, hipStream_t st) {
  if (nb_field % 2 != 0 || nb_field < 2 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in %s ", nm);
  const size_t n = (size_t)np * nb_field * (m_max_c > 0 ? m_max_c : 0);
  void* d = nullptr;
  int rc = stream_scratch(SLOT_EX_IN1, n * 8, st, &d);
  if (rc) return rc;
  hipError_t e = n ? hipMemcpyAsync(d, c1_inout, n * 8, hipMemcpyDeviceToDevice, st) : hipSuccess;
  if (e != hipSuccess) return hip_fail(e, nm);
  return op == EX_SQRT ? sqrt_dev((const double*)d, c1_inout, np, nb_field, m_max_c, m_max_pad, st)
                       : clamp01_dev((const double*)d, c1_inout, np, nb_field, m_max_c, m_max_pad, st);
}

static int inplace_host(int op, const char* nm, double* c1_inout, int64_t np, int32_t nb_field, int32_t m_max_c,
                        int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad, double* temps) {
  if (temps) temps[0] = temps[1] = temps[2] = 0.0;
  if (nb_field % 2 != 0 || nb_field < 2 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in %s ", nm);
  int rc = check_procs(nm, np, nb_procs, bloc_size);
  if (rc) return rc;
  const size_t n = (size_t)np * nb_field * m_max_c;
  void *d1, *d3;
  HostRun t;
  t.t0 = now_s();
  if ((rc = to_dev(SLOT_EX_IN1, c1_inout, n, &d1)) || (rc = to_dev(SLOT_EX_OUT, c1_inout, n, &d3))) return rc;
  t.t1 = now_s();
  rc = op == EX_SQRT ? sqrt_dev((const double*)d1, (double*)d3, np, nb_field, m_max_c, m_max_pad, nullptr)
                     : clamp01_dev((const double*)d1, (double*)d3, np, nb_field, m_max_c, m_max_pad, nullptr);
  if (rc) return rc;
  if ((rc = from_dev(nullptr, d3, 0))) return rc;
  t.t2 = now_s();
  if ((rc = from_dev(c1_inout, d3, n))) return rc;
  t.t3 = now_s();
  set_temps(temps, t);
  return SFX_OK;
}

}  // namespace sfx

using namespace sfx;

extern "C" {

int sfx_fft_par_div_dcl_dev(const double* c1_in, const double* c2_in, double* c_out, int64_t np, int32_t m_max_c,
                            int32_t m_max_pad, void* stream) {
  clear_err();
  return div_dev(c1_in, c2_in, c_out, np, m_max_c, m_max_pad, (hipStream_t)stream);
}

int sfx_fft_par_div_dcl(const double* c1_in, const double* c2_in, double* c_out, int64_t np, int32_t m_max_c,
                        int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad, double* temps) {
  clear_err();
  if (temps) temps[0] = temps[1] = temps[2] = 0.0;
  int rc = check_procs("FFT_PAR_DIV_DCL", np, nb_procs, bloc_size);
  if (rc) return rc;
  if (m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_PAR_DIV_DCL ");
  const size_t n = (size_t)np * 2 * (m_max_c > 0 ? m_max_c : 0);
  void *d1, *d2, *d3;
  HostRun t;
  t.t0 = now_s();
  if ((rc = to_dev(SLOT_EX_IN1, c1_in, n, &d1)) || (rc = to_dev(SLOT_EX_IN2, c2_in, n, &d2))) return rc;
  if ((rc = staging(SLOT_EX_OUT, n * 8, &d3))) return rc;
  t.t1 = now_s();
  if ((rc = div_dev((const double*)d1, (const double*)d2, (double*)d3, np, m_max_c, m_max_pad, nullptr))) return rc;
  if ((rc = from_dev(nullptr, d3, 0))) return rc;
  t.t2 = now_s();
  if ((rc = from_dev(c_out, d3, n))) return rc;
  t.t3 = now_s();
  set_temps(temps, t);
  return SFX_OK;
}

int sfx_fft_heaviside_dcl_dev(const double* V1_in, const double* values, double* V_out, int32_t n_values,
                              int64_t np, int32_t m_max_c, int32_t m_max_pad, double coeff_tanh, void* stream) {
  clear_err();
  if (!values) return fail(SFX_ERR_ARG, "FFT_HEAVISIDE_DCL: null values");
  return heavi_dev(V1_in, values, n_values, V_out, np, m_max_c, m_max_pad, coeff_tanh, (hipStream_t)stream);
}

int sfx_fft_heaviside_dcl(const double* V1_in, const double* values, double* V_out, int32_t n_values, int64_t np,
                          int32_t nb_field, int32_t m_max_c, int32_t nb_procs, int64_t bloc_size,
                          int32_t m_max_pad, double coeff_tanh, double* temps) {
  clear_err();
  if (temps) temps[0] = temps[1] = temps[2] = 0.0;
  if (nb_field % 2 != 0 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_HEAVISIDE_DCL ");
  if (nb_field != 2) return fail(SFX_ERR_ARG, "FFT_HEAVISIDE_DCL: the level sets are scalars (nb_field == 2)");
  if (!values || n_values < 2) return fail(SFX_ERR_ARG, "FFT_HEAVISIDE_DCL: SIZE(values) < 2");
  int rc = check_procs("FFT_HEAVISIDE_DCL", np, nb_procs, bloc_size);
  if (rc) return rc;
  const size_t nin = (size_t)(n_values - 1) * np * 2 * (m_max_c > 0 ? m_max_c : 0);
  const size_t nout = (size_t)np * 2 * (m_max_c > 0 ? m_max_c : 0);
  void *d1, *d3;
  HostRun t;
  t.t0 = now_s();
  if ((rc = to_dev(SLOT_EX_IN1, V1_in, nin, &d1)) || (rc = staging(SLOT_EX_OUT, nout * 8, &d3))) return rc;
  t.t1 = now_s();
  if ((rc = heavi_dev((const double*)d1, values, n_values, (double*)d3, np, m_max_c, m_max_pad, coeff_tanh,
                      nullptr)))
    return rc;
  if ((rc = from_dev(nullptr, d3, 0))) return rc;
  t.t2 = now_s();
  if ((rc = from_dev(V_out, d3, nout))) return rc;
  t.t3 = now_s();
  set_temps(temps, t);
  return SFX_OK;
}

int sfx_fft_squareroot_scalar_dev(double* c1_inout, int64_t np, int32_t nb_field, int32_t m_max_c,
                                  int32_t m_max_pad, void* stream) {
  clear_err();
  return inplace_dev(EX_SQRT, "FFT_SQUAREROOT_SCALAR", c1_inout, np, nb_field, m_max_c, m_max_pad,
                     (hipStream_t)stream);
}

int sfx_fft_squareroot_scalar(double* c1_inout, int64_t np, int32_t nb_field, int32_t m_max_c, int32_t nb_procs,
                              int64_t bloc_size, int32_t m_max_pad, double* temps) {
  clear_err();
  return inplace_host(EX_SQRT, "FFT_SQUAREROOT_SCALAR", c1_inout, np, nb_field, m_max_c, nb_procs, bloc_size,
                      m_max_pad, temps);
}

int sfx_fft_no_overshoot_level_set_dev(double* c1_inout, int64_t np, int32_t nb_field, int32_t m_max_c,
                                       int32_t m_max_pad, void* stream) {
  clear_err();
  return inplace_dev(EX_CLAMP01, "FFT_NO_OVERSHOOT_LEVEL_SET", c1_inout, np, nb_field, m_max_c, m_max_pad,
                     (hipStream_t)stream);
}

int sfx_fft_no_overshoot_level_set(double* c1_inout, int64_t np, int32_t nb_field, int32_t m_max_c,
                                   int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad, double* temps) {
  clear_err();
  return inplace_host(EX_CLAMP01, "FFT_NO_OVERSHOOT_LEVEL_SET", c1_inout, np, nb_field, m_max_c, nb_procs,
                      bloc_size, m_max_pad, temps);
}

int sfx_fft_scalar_vect_no_overshoot_dev(const double* scalar_bounds, int32_t n_bounds, const double* V1_in,
                                         const double* V2_in, double* V_out, int64_t np, int32_t m_max_c,
                                         int32_t pb, int32_t m_max_pad, void* stream) {
  clear_err();
  if (pb != 1 && pb != 2) return fail(SFX_ERR_ARG, "error in problem type while calling FFT_SCALAR_VECT_NO_OVERSHOOT ");
  return svclamp_dev(scalar_bounds, n_bounds, V1_in, V2_in, V_out, np, m_max_c, pb, m_max_pad, (hipStream_t)stream);
}

int sfx_fft_scalar_vect_no_overshoot(const double* scalar_bounds, int32_t n_bounds, const double* V1_in,
                                     const double* V2_in, double* V_out, int64_t np, int32_t nb_field1,
                                     int32_t nb_field2, int32_t m_max_c, int32_t pb, int32_t nb_procs,
                                     int64_t bloc_size, int32_t m_max_pad, double* temps) {
  clear_err();
  if (temps) temps[0] = temps[1] = temps[2] = 0.0;
  if (nb_field1 % 2 != 0 || nb_field2 % 2 != 0 || m_max_c == 0)
    return fail(SFX_ERR_ARG, " BUG in FFT_SCALAR_VECT_NO_OVERSHOOT ");
  if (nb_field1 != 6 || nb_field2 != 2)
    return fail(SFX_ERR_ARG, "error in dimension of INPUT variables while calling FFT_SCALAR_VECT_NO_OVERSHOOT ");
  if (pb != 1 && pb != 2) return fail(SFX_ERR_ARG, "error in problem type while calling FFT_SCALAR_VECT_NO_OVERSHOOT ");
  int rc = check_procs("FFT_SCALAR_VECT_NO_OVERSHOOT", np, nb_procs, bloc_size);
  if (rc) return rc;
  const size_t n1 = (size_t)np * 6 * m_max_c, n2 = (size_t)np * 2 * m_max_c;
  void *d1, *d2, *d3;
  HostRun t;
  t.t0 = now_s();
  if ((rc = to_dev(SLOT_EX_IN1, V1_in, n1, &d1)) || (rc = to_dev(SLOT_EX_IN2, V2_in, n2, &d2)) ||
      (rc = staging(SLOT_EX_OUT, n1 * 8, &d3)))
    return rc;
  t.t1 = now_s();
  if ((rc = svclamp_dev(scalar_bounds, n_bounds, (const double*)d1, (const double*)d2, (double*)d3, np, m_max_c,
                        pb, m_max_pad, nullptr)))
    return rc;
  if ((rc = from_dev(nullptr, d3, 0))) return rc;
  t.t2 = now_s();
  if ((rc = from_dev(V_out, d3, n1))) return rc;
  t.t3 = now_s();
  set_temps(temps, t);
  return SFX_OK;
}

int sfx_fft_par_entro_visc_dcl_dev(const double* V1_in, const double* c1_in, const double* c_in_real,
                                   double coeff1_in_level, double* V_out, double* c_out, int64_t np,
                                   int32_t m_max_c, int32_t m_max_pad, void* stream) {
  clear_err();
  return entro_dev(V1_in, c1_in, c_in_real, coeff1_in_level, V_out, c_out, np, m_max_c, m_max_pad,
                   (hipStream_t)stream);
}

int sfx_fft_par_entro_visc_dcl(const double* V1_in, const double* c1_in, const double* c_in_real,
                               const double* hloc_gauss, double coeff1_in_level, double* V_out, double* c_out,
                               int64_t np, int32_t nb_field1, int32_t nb_field2, int32_t m_max_c, int32_t nb_procs,
                               int64_t bloc_size, int32_t m_max_pad, double* temps) {
  clear_err();
  (void)hloc_gauss;   // read by the reference but not used by this product
  if (temps) temps[0] = temps[1] = temps[2] = 0.0;
  if (nb_field1 % 2 != 0 || nb_field2 % 2 != 0 || m_max_c == 0)
    return fail(SFX_ERR_ARG, " BUG in FFT_PAR_ENTRO_VISC_DCL ");
  if (nb_field1 != 6 || nb_field2 != 2)
    return fail(SFX_ERR_ARG, "error in problem type while calling FFT_PAR_ENTRO_VISC_DCL");
  int rc = check_procs("FFT_PAR_ENTRO_VISC_DCL", np, nb_procs, bloc_size);
  if (rc) return rc;
  if (m_max_pad < 1 || m_max_pad > 4096) return fail(SFX_ERR_ARG, "FFT_PAR_ENTRO_VISC_DCL: m_max_pad=%d", m_max_pad);
  const size_t n1 = (size_t)np * 6 * m_max_c, n2 = (size_t)np * 2 * m_max_c, nr = (size_t)np * (2 * m_max_pad - 1);
  void *d1, *d2, *dr, *d3, *d4;
  HostRun t;
  t.t0 = now_s();
  if ((rc = to_dev(SLOT_EX_IN1, V1_in, n1, &d1)) || (rc = to_dev(SLOT_EX_IN2, c1_in, n2, &d2)) ||
      (rc = to_dev(SLOT_EX_IN3, c_in_real, nr, &dr)) || (rc = staging(SLOT_EX_OUT, n1 * 8, &d3)) ||
      (rc = staging(SLOT_EX_OUT2, n2 * 8, &d4)))
    return rc;
  t.t1 = now_s();
  if ((rc = entro_dev((const double*)d1, (const double*)d2, (const double*)dr, coeff1_in_level, (double*)d3,
                      (double*)d4, np, m_max_c, m_max_pad, nullptr)))
    return rc;
  if ((rc = from_dev(nullptr, d3, 0))) return rc;
  t.t2 = now_s();
  if ((rc = from_dev(V_out, d3, n1)) || (rc = from_dev(c_out, d4, n2))) return rc;
  t.t3 = now_s();
  set_temps(temps, t);
  return SFX_OK;
}

int sfx_fft_par_compr_entro_visc_dcl_dev(const double* V1_in, const double* V2_in, const double* c1_in,
                                         const double* c_in_real, double coeff1_in_level,
                                         double level_set_comp_factor, double* V_out, double* c_out, int64_t np,
                                         int32_t m_max_c, int32_t m_max_pad, void* stream) {
  clear_err();
  return compr_dev(V1_in, V2_in, c1_in, c_in_real, coeff1_in_level, level_set_comp_factor, V_out, c_out, np,
                   m_max_c, m_max_pad, (hipStream_t)stream);
}

int sfx_fft_par_compr_entro_visc_dcl(const double* V1_in, const double* V2_in, const double* c1_in,
                                     const double* c_in_real, const double* hloc_gauss, double coeff1_in_level,
                                     double level_set_comp_factor, double* V_out, double* c_out, int64_t np,
                                     int32_t nb_field1, int32_t nb_field2, int32_t nb_field3, int32_t m_max_c,
                                     int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad, double* temps) {
  clear_err();
  (void)hloc_gauss;   // read by the reference but not used by this product
  if (temps) temps[0] = temps[1] = temps[2] = 0.0;
  if (m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_PAR_COMPR_ENTRO_VISC_DCL ");
  if (nb_field1 != 6 || nb_field2 != 6 || nb_field3 != 2)
    return fail(SFX_ERR_ARG, "error in problem type while calling FFT_PAR_COMPR_ENTRO_VISC_DCL");
  int rc = check_procs("FFT_PAR_COMPR_ENTRO_VISC_DCL", np, nb_procs, bloc_size);
  if (rc) return rc;
  if (m_max_pad < 1 || m_max_pad > 4096)
    return fail(SFX_ERR_ARG, "FFT_PAR_COMPR_ENTRO_VISC_DCL: m_max_pad=%d", m_max_pad);
  const size_t n1 = (size_t)np * 6 * m_max_c, n2 = (size_t)np * 2 * m_max_c, nr = (size_t)np * (2 * m_max_pad - 1);
  void *d1, *d2, *dc, *dr, *d3, *d4;
  HostRun t;
  t.t0 = now_s();
  if ((rc = to_dev(SLOT_EX_IN1, V1_in, n1, &d1)) || (rc = to_dev(SLOT_EX_IN2, V2_in, n1, &d2)) ||
      (rc = to_dev(SLOT_EX_IN4, c1_in, n2, &dc)) || (rc = to_dev(SLOT_EX_IN3, c_in_real, nr, &dr)) ||
      (rc = staging(SLOT_EX_OUT, n1 * 8, &d3)) || (rc = staging(SLOT_EX_OUT2, n2 * 8, &d4)))
    return rc;
  t.t1 = now_s();
  if ((rc = compr_dev((const double*)d1, (const double*)d2, (const double*)dc, (const double*)dr, coeff1_in_level,
                      level_set_comp_factor, (double*)d3, (double*)d4, np, m_max_c, m_max_pad, nullptr)))
    return rc;
  if ((rc = from_dev(nullptr, d3, 0))) return rc;
  t.t2 = now_s();
  if ((rc = from_dev(V_out, d3, n1)) || (rc = from_dev(c_out, d4, n2))) return rc;
  t.t3 = now_s();
  set_temps(temps, t);
  return SFX_OK;
}

int sfx_fft_compute_entropy_visc_dev(const double* V1_in, const double* V2_in, const double* V3_in,
                                    const double* V4_in, const double* V5_in, const double* hloc_gauss,
                                    double* V1_out, double* V2_out, double* V3_out, int64_t np, int32_t m_max_c,
                                    int32_t m_max_pad, double residual_normalization, int32_t l_G,
                                    double les_coeff1, double les_coeff2, double les_coeff4, void* stream) {
  clear_err();
  const double* V[5] = {V1_in, V2_in, V3_in, V4_in, V5_in};
  double* O[3] = {V1_out, V2_out, V3_out};
  return evisc_dev(V, hloc_gauss, O, np, m_max_c, m_max_pad, l_G, les_coeff1, les_coeff2, les_coeff4,
                   residual_normalization, (hipStream_t)stream);
}

int sfx_fft_compute_entropy_visc(const double* V1_in, const double* V2_in, const double* V3_in, const double* V4_in,
                                 const double* V5_in, const double* hloc_gauss, double* V1_out, double* V2_out,
                                 double* V3_out, int64_t np, int32_t nb_field, int32_t m_max_c, int32_t nb_procs,
                                 int64_t bloc_size, int32_t m_max_pad, double residual_normalization, int32_t l_G,
                                 double les_coeff1, double les_coeff2, double les_coeff4, double* temps) {
  clear_err();
  if (temps) temps[0] = temps[1] = temps[2] = 0.0;
  if (nb_field != 6 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_COMPUTE_ENTROPY_VISC ");
  int rc = check_procs("FFT_COMPUTE_ENTROPY_VISC", np, nb_procs, bloc_size);
  if (rc) return rc;
  const size_t n1 = (size_t)np * 6 * m_max_c;
  void* d[5];
  void *dh, *o1, *o2, *o3;
  HostRun t;
  t.t0 = now_s();
  const double* src[5] = {V1_in, V2_in, V3_in, V4_in, V5_in};
  const int slots[5] = {SLOT_EX_IN1, SLOT_EX_IN2, SLOT_EX_IN3, SLOT_EX_IN4, SLOT_EX_IN5};
  for (int i = 0; i < 5; ++i)
    if ((rc = to_dev(slots[i], src[i], n1, &d[i]))) return rc;
  if ((rc = to_dev(13, hloc_gauss, (size_t)np, &dh)) || (rc = staging(SLOT_EX_OUT, n1 * 8, &o1)) ||
      (rc = staging(SLOT_EX_OUT2, n1 * 8, &o2)) || (rc = staging(SLOT_EX_OUT3, n1 * 8, &o3)))
    return rc;
  t.t1 = now_s();
  const double* V[5] = {(const double*)d[0], (const double*)d[1], (const double*)d[2], (const double*)d[3],
                        (const double*)d[4]};
  double* O[3] = {(double*)o1, (double*)o2, (double*)o3};
  if ((rc = evisc_dev(V, (const double*)dh, O, np, m_max_c, m_max_pad, l_G, les_coeff1, les_coeff2, les_coeff4,
                      residual_normalization, nullptr)))
    return rc;
  if ((rc = from_dev(nullptr, o1, 0))) return rc;
  t.t2 = now_s();
  if ((rc = from_dev(V1_out, o1, n1)) || (rc = from_dev(V2_out, o2, n1)) || (rc = from_dev(V3_out, o3, n1))) return rc;
  t.t3 = now_s();
  set_temps(temps, t);
  return SFX_OK;
}

int sfx_fft_compute_entropy_visc_mom_dev(const double* V1_in, const double* V2_in, const double* V3_in,
                                        const double* c1_in, const double* hloc_gauss, double* c1_real_out,
                                        double* opt_c2_real_out, int64_t np, int32_t m_max_c, int32_t m_max_pad,
                                        int32_t l_G, double les_coeff2, double les_coeff4, void* stream) {
  clear_err();
  return emom_dev(V1_in, V2_in, V3_in, c1_in, hloc_gauss, c1_real_out, opt_c2_real_out, np, m_max_c, m_max_pad,
                  l_G, les_coeff2, les_coeff4, (hipStream_t)stream);
}

int sfx_fft_compute_entropy_visc_mom(const double* V1_in, const double* V2_in, const double* V3_in,
                                     const double* c1_in, const double* hloc_gauss, double* c1_real_out,
                                     double* opt_c2_real_out, int64_t np, int32_t nb_field, int32_t m_max_c,
                                     int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad, int32_t l_G,
                                     double les_coeff2, double les_coeff4, double* temps) {
  clear_err();
  if (temps) temps[0] = temps[1] = temps[2] = 0.0;
  if (nb_field != 6 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_COMPUTE_ENTROPY_VISC_MOM ");
  int rc = check_procs("FFT_COMPUTE_ENTROPY_VISC_MOM", np, nb_procs, bloc_size);
  if (rc) return rc;
  if (m_max_pad < 1 || m_max_pad > 4096) return fail(SFX_ERR_ARG, "FFT_COMPUTE_ENTROPY_VISC_MOM: m_max_pad");
  const size_t n1 = (size_t)np * 6 * m_max_c, n2 = (size_t)np * 2 * m_max_c, nr = (size_t)np * (2 * m_max_pad - 1);
  void *d1, *d2, *d3, *dc, *dh, *o1, *o2 = nullptr;
  HostRun t;
  t.t0 = now_s();
  if ((rc = to_dev(SLOT_EX_IN1, V1_in, n1, &d1)) || (rc = to_dev(SLOT_EX_IN2, V2_in, n1, &d2)) ||
      (rc = to_dev(SLOT_EX_IN3, V3_in, n1, &d3)) || (rc = to_dev(SLOT_EX_IN4, c1_in, n2, &dc)) ||
      (rc = to_dev(13, hloc_gauss, (size_t)np, &dh)) || (rc = staging(SLOT_EX_OUT, nr * 8, &o1)) ||
      (opt_c2_real_out && (rc = staging(SLOT_EX_OUT2, nr * 8, &o2))))
    return rc;
  t.t1 = now_s();
  if ((rc = emom_dev((const double*)d1, (const double*)d2, (const double*)d3, (const double*)dc, (const double*)dh,
                     (double*)o1, (double*)o2, np, m_max_c, m_max_pad, l_G, les_coeff2, les_coeff4, nullptr)))
    return rc;
  if ((rc = from_dev(nullptr, o1, 0))) return rc;
  t.t2 = now_s();
  if ((rc = from_dev(c1_real_out, o1, nr))) return rc;
  if (opt_c2_real_out && (rc = from_dev(opt_c2_real_out, o2, nr))) return rc;
  t.t3 = now_s();
  set_temps(temps, t);
  return SFX_OK;
}

int sfx_fft_compression_level_set_dcl_dev(const double* V1_in, const double* V2_in, const double* c_in,
                                          double* c_out, int64_t np, int32_t m_max_c, int32_t m_max_pad,
                                          double level_set_comp_factor, double les_coeff4, double h_min_distance,
                                          void* stream) {
  clear_err();
  return cls_dev(V1_in, V2_in, c_in, c_out, np, m_max_c, m_max_pad, level_set_comp_factor, les_coeff4,
                 h_min_distance, (hipStream_t)stream);
}

int sfx_fft_compression_level_set_dcl(const double* V1_in, const double* V2_in, const double* c_in, double* c_out,
                                      const double* hloc_gauss, int32_t l_G, int64_t np, int32_t nb_field,
                                      int32_t m_max_c, int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad,
                                      double level_set_comp_factor, double les_coeff4, double h_min_distance,
                                      double* temps) {
  clear_err();
  (void)hloc_gauss;   // read by the reference but not used by this product
  (void)l_G;          // the element maxima it takes do not change the maximum used
  if (temps) temps[0] = temps[1] = temps[2] = 0.0;
  if (nb_field != 6 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_COMPRESSION_LEVEL_SET_DCL ");
  int rc = check_procs("FFT_COMPRESSION_LEVEL_SET_DCL", np, nb_procs, bloc_size);
  if (rc) return rc;
  const size_t n1 = (size_t)np * 6 * m_max_c, n2 = (size_t)np * 2 * m_max_c;
  void *d1, *d2, *dc, *o;
  HostRun t;
  t.t0 = now_s();
  if ((rc = to_dev(SLOT_EX_IN1, V1_in, n1, &d1)) || (rc = to_dev(SLOT_EX_IN2, V2_in, n1, &d2)) ||
      (rc = to_dev(SLOT_EX_IN4, c_in, n2, &dc)) || (rc = staging(SLOT_EX_OUT, n2 * 8, &o)))
    return rc;
  t.t1 = now_s();
  if ((rc = cls_dev((const double*)d1, (const double*)d2, (const double*)dc, (double*)o, np, m_max_c, m_max_pad,
                    level_set_comp_factor, les_coeff4, h_min_distance, nullptr)))
    return rc;
  if ((rc = from_dev(nullptr, o, 0))) return rc;
  t.t2 = now_s();
  if ((rc = from_dev(c_out, o, n2))) return rc;
  t.t3 = now_s();
  set_temps(temps, t);
  return SFX_OK;
}

int sfx_fft_compute_entropy_visc_grad_mom_dev(const double* V1_in, const double* V2_in, const double* V3_in,
                                              const double* c_in_real, double les_coeff1_mom, double* V_out,
                                              int64_t np, int32_t m_max_c, int32_t m_max_pad, void* stream) {
  clear_err();
  return gradmom_dev(V1_in, V2_in, V3_in, c_in_real, les_coeff1_mom, V_out, np, m_max_c, m_max_pad,
                     (hipStream_t)stream);
}

int sfx_fft_compute_entropy_visc_grad_mom(const double* V1_in, const double* V2_in, const double* V3_in,
                                          const double* c_in_real, double les_coeff1_mom, double* V_out, int64_t np,
                                          int32_t nb_field, int32_t m_max_c, int32_t nb_procs, int64_t bloc_size,
                                          int32_t m_max_pad, double* temps) {
  clear_err();
  if (temps) temps[0] = temps[1] = temps[2] = 0.0;
  if (nb_field != 6 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM ");
  int rc = check_procs("FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM", np, nb_procs, bloc_size);
  if (rc) return rc;
  if (m_max_pad < 1 || m_max_pad > 4096)
    return fail(SFX_ERR_ARG, "FFT_COMPUTE_ENTROPY_VISC_GRAD_MOM: m_max_pad=%d", m_max_pad);
  const size_t n1 = (size_t)np * 6 * m_max_c, nr = (size_t)np * (2 * m_max_pad - 1), no = 3 * n1;
  void *d1, *d2, *d3, *dr, *dout;
  HostRun t;
  t.t0 = now_s();
  if ((rc = to_dev(SLOT_EX_IN1, V1_in, n1, &d1)) || (rc = to_dev(SLOT_EX_IN2, V2_in, n1, &d2)) ||
      (rc = to_dev(SLOT_EX_IN4, V3_in, n1, &d3)) || (rc = to_dev(SLOT_EX_IN3, c_in_real, nr, &dr)) ||
      (rc = staging(SLOT_EX_OUT, no * 8, &dout)))
    return rc;
  t.t1 = now_s();
  if ((rc = gradmom_dev((const double*)d1, (const double*)d2, (const double*)d3, (const double*)dr,
                        les_coeff1_mom, (double*)dout, np, m_max_c, m_max_pad, nullptr)))
    return rc;
  if ((rc = from_dev(nullptr, dout, 0))) return rc;
  t.t2 = now_s();
  if ((rc = from_dev(V_out, dout, no))) return rc;
  t.t3 = now_s();
  set_temps(temps, t);
  return SFX_OK;
}

int sfx_fft_max_vel_dcl_dev(const double* V1_in, double* V_out, int64_t np, int32_t nb_field, int32_t m_max_c,
                            int32_t m_max_pad, void* stream) {
  clear_err();
  if (nb_field % 2 != 0 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_MAX_VEL_DCL ");
  if (nb_field != 6 && nb_field != 2)
    return fail(SFX_ERR_UNSUPPORTED, "FFT_MAX_VEL_DCL: nb_field=%d (6 or 2)", nb_field);
  return maxvel_dev(V1_in, V_out, np, nb_field, m_max_c, m_max_pad, (hipStream_t)stream);
}

int sfx_fft_max_vel_dcl(const double* V1_in, double* V_out, int64_t np, int32_t nb_field, int32_t m_max_c,
                        int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad) {
  clear_err();
  if (!V_out) return fail(SFX_ERR_ARG, "FFT_MAX_VEL_DCL: null V_out");
  if (nb_field % 2 != 0 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_MAX_VEL_DCL ");
  if (nb_field != 6 && nb_field != 2)
    return fail(SFX_ERR_UNSUPPORTED, "FFT_MAX_VEL_DCL: nb_field=%d (6 or 2)", nb_field);
  int rc = check_procs("FFT_MAX_VEL_DCL", np, nb_procs, bloc_size);
  if (rc) return rc;
  void *d1, *dm;
  if ((rc = to_dev(SLOT_EX_IN1, V1_in, (size_t)np * nb_field * m_max_c, &d1)) ||
      (rc = staging(SLOT_EX_MAX, 8, &dm)))
    return rc;
  if ((rc = maxvel_dev((const double*)d1, (double*)dm, np, nb_field, m_max_c, m_max_pad, nullptr))) return rc;
  return from_dev(V_out, dm, 1);
}

int sfx_fft_max_min_vel_dcl_dev(const double* V1_in, double* V_out, int64_t np, int32_t nb_field, int32_t m_max_c,
                                int32_t m_max_pad, void* stream) {
  clear_err();
  if (nb_field % 2 != 0 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_MAX_MIN_VEL_DCL ");
  return maxminvel_dev(V1_in, V_out, np, nb_field, m_max_c, m_max_pad, (hipStream_t)stream);
}

int sfx_fft_max_min_vel_dcl(const double* V1_in, double* V_out, int64_t np, int32_t nb_field, int32_t m_max_c,
                            int32_t nb_procs, int64_t bloc_size, int32_t m_max_pad) {
  clear_err();
  if (!V_out) return fail(SFX_ERR_ARG, "FFT_MAX_MIN_VEL_DCL: null V_out");
  if (nb_field % 2 != 0 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_MAX_MIN_VEL_DCL ");
  int rc = check_procs("FFT_MAX_MIN_VEL_DCL", np, nb_procs, bloc_size);
  if (rc) return rc;
  void *d1, *dm;
  if ((rc = to_dev(SLOT_EX_IN1, V1_in, (size_t)np * nb_field * m_max_c, &d1)) ||
      (rc = staging(SLOT_EX_MAX, 16, &dm)))
    return rc;
  if ((rc = maxminvel_dev((const double*)d1, (double*)dm, np, nb_field, m_max_c, m_max_pad, nullptr))) return rc;
  return from_dev(V_out, dm, 2);
}

int sfx_fft_check_interface_dev(const double* V1_in, int32_t nb_fluids, int32_t* interface_ok, int64_t np,
                                int32_t nb_field, int32_t m_max_c, int32_t m_max_pad, void* stream) {
  clear_err();
  if (nb_field % 2 != 0 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_CHECK_INTERFACE ");
  if (nb_field != 2) return fail(SFX_ERR_ARG, "BUG in FFT_CHECK_INTERFACE: pb with V1_in dimensions");
  return checkif_dev(V1_in, nb_fluids, interface_ok, np, m_max_c, m_max_pad, (hipStream_t)stream);
}

int sfx_fft_check_interface(const double* V1_in, int32_t nb_fluids, int32_t* interface_ok, int64_t np,
                            int32_t nb_field, int32_t m_max_c, int32_t nb_procs, int64_t bloc_size,
                            int32_t m_max_pad, double* temps) {
  clear_err();
  if (temps) temps[0] = temps[1] = temps[2] = 0.0;
  if (!interface_ok) return fail(SFX_ERR_ARG, "FFT_CHECK_INTERFACE: null interface_ok");
  if (nb_field % 2 != 0 || m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_CHECK_INTERFACE ");
  if (nb_field != 2) return fail(SFX_ERR_ARG, "BUG in FFT_CHECK_INTERFACE: pb with V1_in dimensions");
  int rc = check_procs("FFT_CHECK_INTERFACE", np, nb_procs, bloc_size);
  if (rc) return rc;
  if (nb_fluids < 2) return fail(SFX_ERR_ARG, "FFT_CHECK_INTERFACE: nb_fluids=%d < 2", nb_fluids);
  const size_t n1 = (size_t)(nb_fluids - 1) * np * 2 * m_max_c;
  void *d1, *dk;
  HostRun t;
  t.t0 = now_s();
  if ((rc = to_dev(SLOT_EX_IN1, V1_in, n1, &d1)) || (rc = staging(SLOT_EX_MAX, 8, &dk))) return rc;
  t.t1 = now_s();
  if ((rc = checkif_dev((const double*)d1, nb_fluids, (int32_t*)dk, np, m_max_c, m_max_pad, nullptr))) return rc;
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "FFT_CHECK_INTERFACE");
  t.t2 = now_s();
  e = hipMemcpy(interface_ok, dk, sizeof(int32_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "D2H");
  t.t3 = now_s();
  set_temps(temps, t);
  return SFX_OK;
}

int sfx_fft_par_var_eta_prod_dcl_dev(const double* eta_grid, const double* c_in, double* c_out, int64_t np,
                                     int32_t nb_field, int32_t m_max_c, int32_t m_max_pad, void* stream) {
  clear_err();
  return eta_dev(eta_grid, c_in, c_out, np, nb_field, m_max_c, m_max_pad, (hipStream_t)stream);
}

int sfx_fft_par_var_eta_prod_dcl(const double* eta_grid, const double* c_in, double* c_out, int64_t np,
                                 int32_t nb_field, int32_t m_max_c, int32_t nb_procs, int64_t bloc_size,
                                 int32_t m_max_pad, double* temps) {
  clear_err();
  if (temps) temps[0] = temps[1] = temps[2] = 0.0;
  int rc = check_procs("FFT_PAR_VAR_ETA_PROD_DCL", np, nb_procs, bloc_size);
  if (rc) return rc;
  if (m_max_c == 0) return fail(SFX_ERR_ARG, " BUG in FFT_PAR_VAR_ETA_PROD_T_DCL ");
  if (nb_field != 2 && nb_field != 6)
    return fail(SFX_ERR_ARG, "error in nb_field of c_in while calling FFT_PAR_VAR_ETA_PROD_T_DCL");
  const size_t nc = (size_t)np * nb_field * m_max_c, ne = (size_t)(2 * (int64_t)m_max_pad - 1) * np;
  void *de, *dc, *o;
  HostRun t;
  t.t0 = now_s();
  if ((rc = to_dev(SLOT_EX_IN1, eta_grid, ne, &de)) || (rc = to_dev(SLOT_EX_IN2, c_in, nc, &dc)) ||
      (rc = staging(SLOT_EX_OUT, nc * 8, &o)))
    return rc;
  t.t1 = now_s();
  if ((rc = eta_dev((const double*)de, (const double*)dc, (double*)o, np, nb_field, m_max_c, m_max_pad, nullptr)))
    return rc;
  if ((rc = from_dev(nullptr, o, 0))) return rc;
  t.t2 = now_s();
  if ((rc = from_dev(c_out, o, nc))) return rc;
  t.t3 = now_s();
  set_temps(temps, t);
  return SFX_OK;
}

int sfx_fft_par_from_real_dev(const double* real_in, double* c_out, int64_t np, int32_t m_max_c, int32_t m_max_pad,
                              void* stream) {
  clear_err();
  return fromreal_dev(real_in, c_out, np, m_max_c, m_max_pad, (hipStream_t)stream);
}

int sfx_fft_par_from_real(const double* real_in, double* c_out, int64_t np, int32_t m_max_c, int32_t m_max_pad) {
  clear_err();
  if (m_max_c <= 0) return fail(SFX_ERR_ARG, "FFT_PAR_FROM_REAL: m_max_c == 0");
  if (m_max_pad < m_max_c) return fail(SFX_ERR_ARG, "FFT_PAR_FROM_REAL: m_max_pad < m_max_c");
  const size_t nr = (size_t)(2 * (int64_t)m_max_pad - 1) * np, nc = (size_t)np * 2 * m_max_c;
  void *dr, *o;
  int rc;
  if ((rc = to_dev(SLOT_EX_IN1, real_in, nr, &dr)) || (rc = staging(SLOT_EX_OUT, nc * 8, &o))) return rc;
  if ((rc = fromreal_dev((const double*)dr, (double*)o, np, m_max_c, m_max_pad, nullptr))) return rc;
  return from_dev(c_out, o, nc);
}

int32_t sfx_fft_par_real_n(int32_t m_max_c, int32_t opt_nb_plane) {
  // fft_parallel.F90:76-84 with nb_procs = 1 (m_max = m_max_c)
  const int32_t mp = (opt_nb_plane > 2 * m_max_c - 1) ? (opt_nb_plane + 1) / 2 : m_max_c;
  return 2 * mp - 1;
}

int sfx_fft_par_real_dev(const double* V1_in, double* V_out, int64_t np, int32_t nb_field, int32_t m_max_c,
                         int32_t opt_nb_plane, void* stream) {
  clear_err();
  const int N = sfx_fft_par_real_n(m_max_c, opt_nb_plane);
  return real_dev(V1_in, V_out, np, nb_field, m_max_c, (N + 1) / 2, (hipStream_t)stream);
}

int sfx_fft_par_real(const double* V1_in, double* V_out, int64_t np, int32_t nb_field, int32_t m_max_c,
                     int32_t nb_procs, int32_t opt_nb_plane) {
  clear_err();
  if (nb_field % 2 != 0 || m_max_c == 0)
    return fail(SFX_ERR_ARG, "Bug in FFT_PAR_REAL: MOD(nb_field,2)/=0 .OR. m_max_c==0");
  if (nb_procs != 1)
    return fail(SFX_ERR_UNSUPPORTED, "FFT_PAR_REAL: nb_procs=%d (one rank: no MPI_ALLTOALL)", nb_procs);
  if (!V_out) return fail(SFX_ERR_ARG, "FFT_PAR_REAL: null V_out");
  const int N = sfx_fft_par_real_n(m_max_c, opt_nb_plane);
  const size_t nin = (size_t)np * nb_field * m_max_c, nout = (size_t)N * (nb_field / 2) * np;
  void *d1, *o;
  int rc;
  if ((rc = to_dev(SLOT_EX_IN1, V1_in, nin, &d1)) || (rc = staging(SLOT_EX_OUT, nout * 8 + 8, &o))) return rc;
  if ((rc = real_dev((const double*)d1, (double*)o, np, nb_field, m_max_c, (N + 1) / 2, nullptr))) return rc;
  return from_dev(V_out, o, nout);
}

}  // extern "C"

// The pb = 2 scalar-vector product and the tension tensor, reached from the
// FFT_SCALAR_VECT_DCL / FFT_TENSOR_DCL entry points in sfx_api.cpp.
namespace sfx {

int exact_sv2_dev(const double* V1, const double* V2, double* out, int64_t np, int M, int mp, hipStream_t st) {
  return sv2_dev(V1, V2, out, np, M, mp, st);
}

int exact_tension_dev(const double* V1, const double* V2, double* out, int64_t np, int M, int mp, hipStream_t st) {
  return tension_dev(V1, V2, out, np, M, mp, st);
}

int exact_binary_host(int op, const double* V1, const double* V2, double* out, int64_t np, int M, int mp,
                      double* temps) {
  if (temps) temps[0] = temps[1] = temps[2] = 0.0;
  const size_t n1 = (size_t)np * 6 * M, n2 = (size_t)np * (op == EX_SV2 ? 2 : 6) * M;
  const size_t no = (size_t)np * (op == EX_SV2 ? 6 : 18) * M;
  void *d1, *d2, *d3;
  int rc;
  HostRun t;
  t.t0 = now_s();
  if ((rc = to_dev(SLOT_EX_IN1, V1, n1, &d1)) || (rc = to_dev(SLOT_EX_IN2, V2, n2, &d2)) ||
      (rc = staging(SLOT_EX_OUT, no * 8, &d3)))
    return rc;
  t.t1 = now_s();
  rc = op == EX_SV2 ? sv2_dev((const double*)d1, (const double*)d2, (double*)d3, np, M, mp, nullptr)
                    : tension_dev((const double*)d1, (const double*)d2, (double*)d3, np, M, mp, nullptr);
  if (rc) return rc;
  if ((rc = from_dev(nullptr, d3, 0))) return rc;
  t.t2 = now_s();
  if ((rc = from_dev(out, d3, no))) return rc;
  t.t3 = now_s();
  set_temps(temps, t);
  return SFX_OK;
}

}  // namespace sfx
