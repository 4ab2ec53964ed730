// Batched persistent launch: scheduler, autotuner and hipGraph capture.
//
// A batch is a fixed set of independent smb_cross_prod_gauss_sft_par problems
// (meshes of different sizes, different mode counts).  The work is cut into
// tiles (one element block of one problem, see sfx_batch_body.h) and dealt to
// a fixed number of persistent work-groups by one of three policies:
//
//   SFX_BATCH_LPT         longest-processing-time first: tiles by decreasing
//                         estimated cost, each to the least-loaded list;
//   SFX_BATCH_CONTIGUOUS  tiles sorted by grid size then problem, cut into
//                         equal-cost contiguous ranges (neighbouring elements
//                         of one problem stay on one work-group);
//   SFX_BATCH_ROUND_ROBIN tile i to list i mod n_wg, in problem order;
//   SFX_BATCH_BY_SIZE     one ordinary grid per grid size, launched one after
//                         the other (each with the fused kernel's own layout
//                         for that size; no persistent loop).
//
// The autotuner times every (policy, work-group count) candidate on the
// caller's stream with HIP events and keeps the fastest; results are cached
// per batch shape for the life of the process.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "../../include/sfx.h"
#include "sfx_api_internal.h"
#include "sfx_batch_body.h"

struct sfx_batch_s {
  int device = 0;
  int n_cu = 0;
  std::vector<sfx::GaussArgs> probs;
  std::vector<sfx::BatchTile> tiles_natural;   // problem order
  std::vector<double> tile_cost;               // per natural tile
  size_t smem_cplx = 0;   // largest fused-kernel LDS layout among the problems
  sfx::GaussArgs* d_probs = nullptr;
  sfx::BatchTile* d_tiles = nullptr;
  int32_t* d_wg_ptr = nullptr;
  int32_t policy = -1, n_wg = 0;
  std::vector<int32_t> wg_ptr_host;   // the lists' bounds (BY_SIZE: per grid size)
  std::vector<int> class_mp;          // BY_SIZE: grid size of each list
  // BY_SIZE: the grids of the grid sizes run as parallel branches, forked from
  // the caller's stream onto side streams (default priority) and joined back
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> ev;
  double best_ms = -1.0;
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace sfx {
hipError_t launch_batch(const BatchArgs& b, size_t smem_cplx, hipStream_t st);
hipError_t launch_batch_classes(const BatchArgs& b, const int32_t* wg_ptr_host, const int* class_mp,
                                hipStream_t st);
}

using namespace sfx;

namespace {

std::mutex g_tune_mu;
std::map<std::vector<int64_t>, std::pair<int32_t, int32_t>> g_tuned;   // shape -> (policy, n_wg)

std::vector<int64_t> shape_key(const sfx_batch_s* b) {
  std::vector<int64_t> k;
  for (const auto& a : b->probs) { k.push_back(a.me); k.push_back(a.M); }
  return k;
}

void clear_graph(sfx_batch_s* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// Build the per-work-group lists for (policy, n_wg) and upload them.
int schedule(sfx_batch_s* b, int32_t policy, int32_t n_wg) {
  if (n_wg <= 0) return fail(SFX_ERR_ARG, "sfx_batch_schedule: n_wg=%d must be positive", n_wg);
  std::vector<int> mp;
  for (const auto& a : b->probs) mp.push_back(a.MP);
  std::vector<BatchTile> tiles;
  std::vector<int32_t> ptr;
  const int32_t used = batch_schedule(b->tiles_natural, b->tile_cost, mp, policy, n_wg, tiles, ptr);
  if (used < 0) return fail(SFX_ERR_ARG, "sfx_batch_schedule: unknown policy %d", policy);
  const int64_t nt = (int64_t)tiles.size();
  // the lists may be in use by an earlier launch on any stream
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "sfx_batch_schedule");
  if (nt && (e = hipMemcpy(b->d_tiles, tiles.data(), nt * sizeof(BatchTile), hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(e, "sfx_batch_schedule");
  if ((e = hipMemcpy(b->d_wg_ptr, ptr.data(), ptr.size() * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(e, "sfx_batch_schedule");
  clear_graph(b);
  b->policy = policy;
  b->n_wg = used;
  b->wg_ptr_host = ptr;
  b->class_mp.clear();
  for (int32_t c = 0; c < used; ++c)
    b->class_mp.push_back(ptr[c + 1] > ptr[c] ? b->probs[tiles[ptr[c]].prob].MP : 0);
  return SFX_OK;
}

BatchArgs batch_args(const sfx_batch_s* b) {
  BatchArgs a;
  a.probs = b->d_probs;
  a.tiles = b->d_tiles;
  a.wg_ptr = b->d_wg_ptr;
  a.n_prob = (int32_t)b->probs.size();
  a.n_wg = b->n_wg;
  a.n_tiles = (int64_t)b->tiles_natural.size();
  return a;
}

// BY_SIZE: grid size c's grid on branch c (branch 0: the caller's stream).
hipError_t launch_by_size(sfx_batch_s* b, hipStream_t st) {
  const BatchArgs a = batch_args(b);
  const int nc = b->n_wg;
  hipError_t e;
  while ((int)b->side.size() < nc) {
    hipStream_t s;
    if ((e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) return e;
    b->side.push_back(s);
  }
  while ((int)b->ev.size() < nc + 1) {
    hipEvent_t v;
    if ((e = hipEventCreateWithFlags(&v, hipEventDisableTiming)) != hipSuccess) return e;
    b->ev.push_back(v);
  }
  if ((e = hipEventRecord(b->ev[nc], st)) != hipSuccess) return e;   // fork point
  for (int c = 0; c < nc; ++c) {
    hipStream_t sc = (c == 0) ? st : b->side[c];
    if (c > 0 && (e = hipStreamWaitEvent(sc, b->ev[nc], 0)) != hipSuccess) return e;
    BatchArgs one = a;
    one.n_wg = 1;
    const int32_t ptr1[2] = {b->wg_ptr_host[c], b->wg_ptr_host[c + 1]};
    if ((e = launch_batch_classes(one, ptr1, &b->class_mp[c], sc)) != hipSuccess) return e;
    if (c > 0 && (e = hipEventRecord(b->ev[c], sc)) != hipSuccess) return e;
  }
  for (int c = 1; c < nc; ++c)
    if ((e = hipStreamWaitEvent(st, b->ev[c], 0)) != hipSuccess) return e;   // join
  return hipSuccess;
}

hipError_t launch_any(sfx_batch_s* b, hipStream_t st) {
  if (b->policy == SFX_BATCH_BY_SIZE) return launch_by_size(b, st);
  return launch_batch(batch_args(b), b->smem_cplx, st);
}

int launch(sfx_batch_s* b, hipStream_t st) {
  hipError_t e = launch_any(b, st);
  if (e != hipSuccess) return hip_fail(e, "sfx_batch_run");
  return SFX_OK;
}

void destroy(sfx_batch_s* b) {
  clear_graph(b);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  for (hipStream_t s : b->side) (void)hipStreamDestroy(s);
  for (hipEvent_t v : b->ev) (void)hipEventDestroy(v);
  (void)hipFree(b->d_probs);
  (void)hipFree(b->d_tiles);
  (void)hipFree(b->d_wg_ptr);
  delete b;
}

}  // namespace

extern "C" {

int sfx_batch_create(const sfx_batch_problem_t* probs, int32_t n_prob, sfx_batch_t* out) {
  clear_err();
  if (!out) return fail(SFX_ERR_ARG, "sfx_batch_create: null output handle");
  *out = nullptr;
  if (n_prob <= 0 || !probs) return fail(SFX_ERR_ARG, "sfx_batch_create: n_prob=%d", n_prob);
  sfx_batch_s* b = new sfx_batch_s();
  (void)hipGetDevice(&b->device);
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, b->device);
  if (e != hipSuccess) { destroy(b); return hip_fail(e, "sfx_batch_create"); }
  b->n_cu = prop.multiProcessorCount;
  for (int32_t i = 0; i < n_prob; ++i) {
    const sfx_batch_problem_t& p = probs[i];
    if (!p.mesh) { destroy(b); return fail(SFX_ERR_ARG, "sfx_batch_create: problem %d has a null mesh", i); }
    if (p.m_max_c <= 0 || p.m_max_c > 256) {
      destroy(b);
      return fail(p.m_max_c <= 0 ? SFX_ERR_ARG : SFX_ERR_UNSUPPORTED,
                  "sfx_batch_create: problem %d: m_max_c=%d not in [1,256]", i, p.m_max_c);
    }
    if (!p.list_mode_dev || !p.V_in || !p.V_out) {
      destroy(b);
      return fail(SFX_ERR_ARG, "sfx_batch_create: problem %d has a null array", i);
    }
    GaussArgs a;
    std::memset(&a, 0, sizeof(a));
    const sfx_mesh_s* m = p.mesh;
    a.jj = m->jj; a.rr = m->rr; a.me = m->me; a.np = m->np; a.G = 7 * m->me;
    a.U = p.V_in; a.nU = m->np * 6 * (int64_t)p.m_max_c;
    a.list_mode = p.list_mode_dev; a.tab = m->tab;
    a.Vout = p.V_out; a.nout = a.G * 6 * (int64_t)p.m_max_c;
    a.M = p.m_max_c; a.MP = mp_for(p.m_max_c);
    Twiddles tw;
    int rc = get_twiddles(a.MP, &tw);
    if (rc) { destroy(b); return rc; }
    a.wN = tw.wN; a.wMP = tw.wMP;
    b->probs.push_back(a);
    b->smem_cplx = std::max(b->smem_cplx, batch_smem_cplx(a.MP));
  }
  {
    std::vector<int64_t> me;
    std::vector<int> mp;
    for (const auto& a : b->probs) { me.push_back(a.me); mp.push_back(a.MP); }
    batch_tiles(me, mp, b->tiles_natural, b->tile_cost);
  }
  const int64_t nt = (int64_t)b->tiles_natural.size();
  if ((e = hipMalloc(&b->d_probs, n_prob * sizeof(GaussArgs))) != hipSuccess ||
      (e = hipMalloc(&b->d_tiles, (nt + 1) * sizeof(BatchTile))) != hipSuccess ||
      (e = hipMalloc(&b->d_wg_ptr, (nt + 2) * sizeof(int32_t))) != hipSuccess ||
      (e = hipMemcpy(b->d_probs, b->probs.data(), n_prob * sizeof(GaussArgs), hipMemcpyHostToDevice)) != hipSuccess) {
    destroy(b);
    return hip_fail(e, "sfx_batch_create");
  }
  int rc = schedule(b, SFX_BATCH_LPT, std::max(1, b->n_cu));
  if (rc) { destroy(b); return rc; }
  *out = b;
  return SFX_OK;
}

int sfx_batch_schedule(sfx_batch_t b, int32_t policy, int32_t n_wg) {
  clear_err();
  if (!b) return fail(SFX_ERR_ARG, "sfx_batch_schedule: null batch");
  return schedule(b, policy, n_wg);
}

int sfx_batch_autotune(sfx_batch_t b, int32_t reps, void* stream) {
  clear_err();
  if (!b) return fail(SFX_ERR_ARG, "sfx_batch_autotune: null batch");
  if (reps <= 0) return fail(SFX_ERR_ARG, "sfx_batch_autotune: reps=%d", reps);
  hipStream_t st = (hipStream_t)stream;
  const auto key = shape_key(b);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) return schedule(b, it->second.first, it->second.second);
  }
  const int64_t nt = (int64_t)b->tiles_natural.size();
  std::vector<int32_t> wgs;
  for (int k : {1, 2, 4}) wgs.push_back((int32_t)std::min<int64_t>(nt, (int64_t)k * b->n_cu));
  wgs.push_back((int32_t)nt);   // one tile per work-group: the non-persistent launch
  std::sort(wgs.begin(), wgs.end());
  wgs.erase(std::unique(wgs.begin(), wgs.end()), wgs.end());
  hipEvent_t e0, e1;
  hipError_t e;
  if ((e = hipEventCreate(&e0)) != hipSuccess) return hip_fail(e, "sfx_batch_autotune");
  if ((e = hipEventCreate(&e1)) != hipSuccess) { (void)hipEventDestroy(e0); return hip_fail(e, "sfx_batch_autotune"); }
  int32_t best_p = SFX_BATCH_LPT, best_w = wgs.front();
  double best = 1e300;
  int rc = SFX_OK;
  for (int32_t pol : {SFX_BATCH_LPT, SFX_BATCH_CONTIGUOUS, SFX_BATCH_ROUND_ROBIN, SFX_BATCH_BY_SIZE}) {
    for (int32_t w : wgs) {
      if (pol == SFX_BATCH_BY_SIZE && w != wgs.front()) break;   // n_wg does not apply
      if ((rc = schedule(b, pol, w))) break;
      if ((rc = launch(b, st))) break;   // warm-up
      if ((e = hipEventRecord(e0, st)) != hipSuccess) { rc = hip_fail(e, "sfx_batch_autotune"); break; }
      for (int r = 0; r < reps && !rc; ++r) rc = launch(b, st);
      if (rc) break;
      if ((e = hipEventRecord(e1, st)) != hipSuccess || (e = hipEventSynchronize(e1)) != hipSuccess) {
        rc = hip_fail(e, "sfx_batch_autotune");
        break;
      }
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      const double per = ms / reps;
      if (per < best) { best = per; best_p = pol; best_w = w; }
    }
    if (rc) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  rc = schedule(b, best_p, best_w);
  if (rc) return rc;
  b->best_ms = best;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[key] = {best_p, best_w};
  return SFX_OK;
}

int sfx_batch_capture(sfx_batch_t b) {
  clear_err();
  if (!b) return fail(SFX_ERR_ARG, "sfx_batch_capture: null batch");
  clear_graph(b);
  hipError_t e;
  if (!b->cap_stream && (e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking)) != hipSuccess)
    return hip_fail(e, "sfx_batch_capture");
  if ((e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
    return hip_fail(e, "sfx_batch_capture");
  const hipError_t el = launch_any(b, b->cap_stream);
  e = hipStreamEndCapture(b->cap_stream, &b->graph);
  if (el != hipSuccess) { clear_graph(b); return hip_fail(el, "sfx_batch_capture (launch)"); }
  if (e != hipSuccess) { clear_graph(b); return hip_fail(e, "sfx_batch_capture"); }
  if ((e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0)) != hipSuccess) {
    clear_graph(b);
    return hip_fail(e, "sfx_batch_capture (instantiate)");
  }
  return SFX_OK;
}

int sfx_batch_run(sfx_batch_t b, void* stream) {
  clear_err();
  if (!b) return fail(SFX_ERR_ARG, "sfx_batch_run: null batch");
  if (b->exec) {
    hipError_t e = hipGraphLaunch(b->exec, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "sfx_batch_run (graph)");
    return SFX_OK;
  }
  return launch(b, (hipStream_t)stream);
}

int sfx_batch_info(sfx_batch_t b, int32_t* policy, int32_t* n_wg, int64_t* n_tiles, double* tuned_ms,
                   int32_t* captured) {
  clear_err();
  if (!b) return fail(SFX_ERR_ARG, "sfx_batch_info: null batch");
  if (policy) *policy = b->policy;
  if (n_wg) *n_wg = b->n_wg;
  if (n_tiles) *n_tiles = (int64_t)b->tiles_natural.size();
  if (tuned_ms) *tuned_ms = b->best_ms;
  if (captured) *captured = b->exec ? 1 : 0;
  return SFX_OK;
}

int sfx_batch_destroy(sfx_batch_t b) {
  clear_err();
  if (b) destroy(b);
  return SFX_OK;
}

}  // extern "C"
