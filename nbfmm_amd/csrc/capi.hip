// C ABI of nbfmm_amd (include/nbfmm_amd.h): solver context, stage dispatch,
// buffer access in the reference layouts, fused step and graph replay.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/nbfmm_amd.h"
#include "common.hpp"
#include "launch.hpp"

using namespace nbfmm_amd;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(NBFMM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define NB_HIP(call)                                  \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return hip_fail(e_, #call); \
  } while (0)

int cdiv(int a, int b) { return (a + b - 1) / b; }

// A sorted state kept for a trajectory's adjoint (solver or batch arrays).
struct SortArrays {
  float4* part;
  uint32_t* key;
  int *head, *occ, *nocc;
  float4* mpole;
};

}  // namespace

int nbfmm_amd::report_error(int code, const std::string& msg) { return fail(code, msg); }

// A stream of the library's own on which graph replays asked for on the
// null stream run (a graph cannot be captured on the null stream), ordered
// after the work queued on the null stream before the call and before the
// work queued after it.  Made on first use.
struct SideStream {
  hipStream_t s = nullptr;
  hipEvent_t in = nullptr, out = nullptr;
  hipError_t begin() {
    hipError_t e;
    if (!s && (e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) return e;
    if (!in && (e = hipEventCreateWithFlags(&in, hipEventDisableTiming)) != hipSuccess) return e;
    if (!out && (e = hipEventCreateWithFlags(&out, hipEventDisableTiming)) != hipSuccess) return e;
    if ((e = hipEventRecord(in, nullptr)) != hipSuccess) return e;
    return hipStreamWaitEvent(s, in, 0);
  }
  hipError_t end() {
    hipError_t e;
    if ((e = hipEventRecord(out, s)) != hipSuccess) return e;
    return hipStreamWaitEvent(nullptr, out, 0);
  }
  void release() {
    if (s) hipStreamDestroy(s);
    if (in) hipEventDestroy(in);
    if (out) hipEventDestroy(out);
    s = nullptr;
    in = out = nullptr;
  }
};

// The trajectory backward's second stream: it recomputes the segment before
// while the caller's stream runs the adjoint steps of the current one.
// `start` orders it after the work queued on the caller's stream before the
// call; `done[p]` / `kept[p]` mark the end of the adjoint steps / of the
// recompute that last used (filled) kept-state set p.  Made on first use.
struct RecomputeStream {
  hipStream_t s = nullptr;
  hipEvent_t start = nullptr, done[2] = {nullptr, nullptr}, kept[2] = {nullptr, nullptr};
  hipError_t make() {
    hipError_t e;
    if (!s && (e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking)) != hipSuccess) return e;
    for (hipEvent_t* ev : {&start, &done[0], &done[1], &kept[0], &kept[1]})
      if (!*ev && (e = hipEventCreateWithFlags(ev, hipEventDisableTiming)) != hipSuccess) return e;
    return hipSuccess;
  }
  void release() {
    if (s) hipStreamDestroy(s);
    for (hipEvent_t* ev : {&start, &done[0], &done[1], &kept[0], &kept[1]}) {
      if (*ev) hipEventDestroy(*ev);
      *ev = nullptr;
    }
    s = nullptr;
  }
};

// Stages of the fused pipeline timed in profile mode: sort, upward pass
// (P2M + M2M), M2L, P2P (with L2L, L2P, un-permute and the Verlet epilogue).
constexpr int kProfStages = 4;
constexpr int kProfEvents = kProfStages + 1;

// Steps captured in the multi-step graph of a run: a replayed graph starts
// about 8 us after the previous one ends, while the kernels inside one follow
// each other without a gap.
constexpr int kRunGraphSteps = 8;

// Captures `reps` consecutive calls of `body` on st into a graph and
// instantiates it.
template <class F>
int capture_graph(hipStream_t st, int reps, F&& body, hipGraph_t* graph, hipGraphExec_t* exec) {
  NB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = NBFMM_OK;
  for (int r = 0; r < reps && rc == NBFMM_OK; ++r) rc = body();
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  *graph = g;
  NB_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return NBFMM_OK;
}

// Replays `steps` steps: `steps - 1` of kind [0] (the last kind when not
// chained), the multi-step graph wherever it fits, then one of kind [1].
int replay_run(hipGraphExec_t* exec, bool chain, int steps, hipStream_t st) {
  int inner = chain ? steps - 1 : steps;
  const int mid = chain ? 0 : 1;
  while (inner >= kRunGraphSteps) {
    NB_HIP(hipGraphLaunch(exec[2], st));
    inner -= kRunGraphSteps;
  }
  for (; inner > 0; --inner) NB_HIP(hipGraphLaunch(exec[mid], st));
  if (chain) NB_HIP(hipGraphLaunch(exec[1], st));
  return NBFMM_OK;
}

struct nbfmm_solver {
  int L = 0, B = 0, max_n = 0;
  float limits[4] = {0, 0, 0, 0};
  float cw = 0, ch = 0, eps = 0;
  int device = 0;
  int cus = 256;            // compute units of the device
  int p2p_grid = 0;         // P2P blocks (each takes every p2p_grid-th leaf of the size-ordered list)
  int adj_grid = 0;         // adjoint P2P blocks
  int p2p_variant = kDefaultP2PVariant;     // P2P kernel variant (staging, unroll, waves per block)
  int p2p_bpc = kDefaultP2PBlocksPerCU;     // P2P blocks per compute unit
  SysView v{};  // arrays and sizes; v.n, v.pos0, v.w0 are set per call
  std::vector<void*> allocs;
  // profiling of the fused pipeline: kProfEvents events per solve/step
  // (before the sort, after it, after the upward pass, after M2L, after P2P)
  bool profile = false;
  std::vector<hipEvent_t> ev;
  size_t ev_used = 0;
  double stage_ms[kProfStages] = {0.0, 0.0, 0.0, 0.0};
  int prof_steps = 0;
  unsigned long long* d_pairs = nullptr;
  int* d_kept = nullptr;  // star count after a prune
  // sorted states a trajectory backward keeps for the steps it recomputes
  // (one per recomputed step of a segment; made by the first backward that
  // needs them, held until the solver is destroyed)
  using SortSlot = SortArrays;
  std::vector<SortSlot> slots;
  int ncells_tree = 0;
  // warm-started sort: key/part/head hold a stable sort of warm_n particles
  // (-1: none) that the next sort of the same n starts from
  bool warm_enabled = true;
  int warm_n = -1;
  // small trees: the sort, P2M and M2M of solve/step/run in one block
  bool front_enabled = true;
  // graphs of a run: [0] a step followed by another (its P2P makes the next
  // sort's first pass), [1] the last step, [2] kRunGraphSteps steps of kind [0]
  hipGraphExec_t graph_exec[3] = {nullptr, nullptr, nullptr};
  hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
  SideStream side;  // replays asked for on the null stream run here
  RecomputeStream recompute;  // the trajectory backward's recompute
  struct {
    int n = -1;
    float tick = 0;
    const void *xcur = nullptr, *xpre = nullptr, *w = nullptr, *acc = nullptr;
    hipStream_t stream = nullptr;
    bool warm = false;  // captured with the warm-started sort
  } gkey;
};

namespace {

template <class T>
int dev_alloc(nbfmm_solver* s, T** p, size_t count) {
  void* q = nullptr;
  const size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 16;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return fail(NBFMM_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  s->allocs.push_back(q);
  e = hipMemset(q, 0, bytes);
  if (e != hipSuccess) return hip_fail(e, "hipMemset");
  *p = (T*)q;
  return NBFMM_OK;
}

void free_all(nbfmm_solver* s) {
  for (int k = 0; k < 3; ++k) {
    if (s->graph_exec[k]) hipGraphExecDestroy(s->graph_exec[k]);
    if (s->graph[k]) hipGraphDestroy(s->graph[k]);
  }
  for (auto& e : s->ev) hipEventDestroy(e);
  s->side.release();
  s->recompute.release();
  for (void* p : s->allocs) hipFree(p);
  s->allocs.clear();
}

SysView view(nbfmm_solver* s, int n) {
  SysView v = s->v;
  v.n = n;
  return v;
}

CallIO call_io(const float* pos0, const float* w0, float* acc0 = nullptr, float* xcur = nullptr,
               float* xpre = nullptr) {
  CallIO io;
  io.pos0 = (const float2*)pos0;
  io.w0 = w0;
  io.acc0 = (float2*)acc0;
  io.xcur = (float2*)xcur;
  io.xpre = (float2*)xpre;
  return io;
}

int check_n(nbfmm_solver* s, int n) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  if (n < 0) return fail(NBFMM_ERR_INVALID_ARGUMENT, "num_particle < 0");
  if (n > s->max_n) return fail(NBFMM_ERR_CAPACITY, "num_particle exceeds max_num_particle");
  return NBFMM_OK;
}

hipStream_t as_stream(void* p) { return (hipStream_t)p; }

int variant_waves(int variant) {
  int stage = 0, unroll = 0, waves = 0;
  p2p_variant_info(variant, &stage, &unroll, &waves);
  return waves;
}

// P2P grid of `variant` for `bpc` blocks per CU, never more blocks than
// occupied leaves can use.
// P2P grid: at most bpc blocks per CU, and no more than there can be runs
// (one per 64 sorted particles, plus a partial one per row), so that on a
// small grid every run (a split one with it, helped by its block's waves)
// can start on a block of its own.
int p2p_grid_for(const nbfmm_solver* s, int variant, int bpc) {
  const int want = std::max(cdiv(std::min(s->B * s->B, s->max_n), variant_waves(variant)),
                            cdiv(s->max_n, kWave) + s->B);
  int g = want < s->cus * bpc ? want : s->cus * bpc;
  return g < 1 ? 1 : g;
}

int adj_grid_for(const nbfmm_solver* s) {
  const int C = s->B * s->B;
  const int max_occ = C < s->max_n ? C : s->max_n;
  const int want = cdiv(max_occ, kAdjP2PWaves);
  const int g = want < s->cus * kAdjP2PBlocksPerCU ? want : s->cus * kAdjP2PBlocksPerCU;
  return g < 1 ? 1 : g;
}

// A captured step graph holds the launch configuration it was made with.
void drop_graph(nbfmm_solver* s) {
  for (int k = 0; k < 3; ++k) {
    if (s->graph_exec[k]) { hipGraphExecDestroy(s->graph_exec[k]); s->graph_exec[k] = nullptr; }
    if (s->graph[k]) { hipGraphDestroy(s->graph[k]); s->graph[k] = nullptr; }
  }
  s->gkey.n = -1;
}

// The next group of kProfEvents profiling events (made 64 groups at a time).
int prof_events(nbfmm_solver* s, hipEvent_t** out) {
  if (s->ev_used + kProfEvents > s->ev.size()) {
    for (int k = 0; k < 64 * kProfEvents; ++k) {
      hipEvent_t e;
      NB_HIP(hipEventCreate(&e));
      s->ev.push_back(e);
    }
  }
  *out = s->ev.data() + s->ev_used;
  s->ev_used += kProfEvents;
  return NBFMM_OK;
}

// Sort stage of the fused paths: warm-started from the previous sort when
// that was of the same n, else the cold sort.  Either way key/part/head then
// hold the sort of these n particles.  gathered: the previous step's P2P
// (kP2PStepSort, inside a run) made the warm sort's first pass.
int sort_stage(nbfmm_solver* s, const SysView& v, const CallIO& io, hipStream_t st, bool gathered = false) {
  if (s->warm_enabled && v.n > 0 && s->warm_n == v.n)
    NB_HIP(launch_predo_warm(v, io, 1, st, gathered));
  else
    NB_HIP(launch_predo(v, io, 1, st));
  s->warm_n = v.n > 0 ? v.n : -1;
  return NBFMM_OK;
}

// Whether solve/step/run of n particles take the one-block front (sort,
// P2M, M2M) of small trees; its sort is cold, so their steps do not chain
// the next sort's first pass.
bool use_front(const nbfmm_solver* s, int n) { return s->front_enabled && small_front_fits(s->L, n); }

// The fused pipeline shared by solve and step.
int run_pipeline(nbfmm_solver* s, int n, float tick, int mode, const CallIO& io, hipStream_t st,
                 bool gathered = false) {
  SysView v = view(s, n);
  v.tick = tick;
  hipEvent_t* E = nullptr;
  if (s->profile && n > 0) {
    if (int rc = prof_events(s, &E)) return rc;
    NB_HIP(hipEventRecord(E[0], st));
  }
  const bool front = use_front(s, n);
  if (front) {
    NB_HIP(launch_small_front(v, io, st));
    s->warm_n = n;
    if (E) NB_HIP(hipEventRecord(E[1], st));
  } else {
    int rc = sort_stage(s, v, io, st, gathered);
    if (rc) return rc;
    if (n == 0) return NBFMM_OK;
    if (E) NB_HIP(hipEventRecord(E[1], st));
    NB_HIP(launch_upward(v, st, false));  // P2M + M2M of the tiles
  }
  if (E) NB_HIP(hipEventRecord(E[2], st));
  NB_HIP(launch_m2l_fused(v, st));  // the M2M levels above the tiles + M2L
  if (E) NB_HIP(hipEventRecord(E[3], st));
  // L2L and L2P are folded into the P2P kernel's epilogue
  NB_HIP(launch_p2p(v, mode, io, s->p2p_grid, s->p2p_variant, st));
  if (E) NB_HIP(hipEventRecord(E[4], st));
  return NBFMM_OK;
}

}  // namespace

extern "C" {

int nbfmm_version(void) { return 10200; }

const char* nbfmm_last_error(void) { return g_last_error.c_str(); }

int nbfmm_solver_create(int num_level, int max_num_particle, const float position_limits[4], float epsilon,
                        nbfmm_solver** out) {
  if (!out) return fail(NBFMM_ERR_INVALID_ARGUMENT, "out is null");
  *out = nullptr;
  if (num_level < 0 || num_level > kMaxLevel) return fail(NBFMM_ERR_INVALID_ARGUMENT, "num_level must be in [0, 8]");
  if (max_num_particle <= 0) return fail(NBFMM_ERR_INVALID_ARGUMENT, "max_num_particle must be > 0");
  if (!position_limits) return fail(NBFMM_ERR_INVALID_ARGUMENT, "position_limits is null");
  if (!(position_limits[0] < position_limits[2] && position_limits[1] < position_limits[3]))
    return fail(NBFMM_ERR_INVALID_ARGUMENT, "position_limits must satisfy x_min < x_max and y_min < y_max");
  if (!(epsilon > 0.0f) || !std::isfinite(epsilon)) return fail(NBFMM_ERR_INVALID_ARGUMENT, "epsilon must be > 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NBFMM_ERR_NO_DEVICE, "no HIP device");

  nbfmm_solver* s = new nbfmm_solver();
  s->L = num_level;
  s->B = 1 << (num_level + 1);
  s->max_n = max_num_particle;
  std::memcpy(s->limits, position_limits, sizeof(s->limits));
  // predo.cu:112-113: cell size computed in float on the host
  s->cw = (position_limits[2] - position_limits[0]) / (float)s->B;
  s->ch = (position_limits[3] - position_limits[1]) / (float)s->B;
  s->eps = epsilon;
  hipGetDevice(&s->device);
  hipDeviceProp_t prop;
  int cus = 256;
  if (hipGetDeviceProperties(&prop, s->device) == hipSuccess && prop.multiProcessorCount > 0)
    cus = prop.multiProcessorCount;

  SysView& v = s->v;
  v.L = s->L;
  v.B = s->B;
  v.lgB = s->L + 1;
  v.xmin = position_limits[0];
  v.ymin = position_limits[1];
  v.cw = s->cw;
  v.ch = s->ch;
  v.eps = epsilon;
  const int C = s->B * s->B;
  const int levels = s->L < 1 ? 1 : s->L;
  int off = 0;
  for (int l = 0; l <= kMaxLevel; ++l) {
    if (v.lvl_off(l) != off) {  // the closed form against the running sum; cannot differ
      delete s;
      return fail(NBFMM_ERR_INVALID_ARGUMENT, "level offsets");
    }
    if (l < levels) off += (s->B >> l) * (s->B >> l);
  }
  const int ncells_tree = off;
  s->ncells_tree = ncells_tree;
  const int ntiles = cdiv(max_num_particle, kSortTile);
  const size_t N = (size_t)max_num_particle;
  int rc = NBFMM_OK;
#define ALLOC(ptr, cnt) \
  if ((rc = dev_alloc(s, &(ptr), (cnt))) != NBFMM_OK) { free_all(s); delete s; return rc; }
  ALLOC(v.tkey, N);
  ALLOC(v.tpart, N);
  ALLOC(v.key, N);
  ALLOC(v.part, N);
  ALLOC(v.eff, N);
  ALLOC(v.head, (size_t)C + 1);
  ALLOC(v.rowstart, (size_t)s->B + 1);
  ALLOC(v.occ, (size_t)C);
  ALLOC(v.nocc, 1);
  ALLOC(v.hist, (size_t)s->B * ntiles);
  ALLOC(v.dtotal, (size_t)s->B);
  ALLOC(v.mpole, (size_t)ncells_tree);
  ALLOC(v.local, (size_t)ncells_tree);
  ALLOC(v.gsort, N);
  ALLOC(v.gup, (size_t)ncells_tree);
  ALLOC(v.gmp, (size_t)ncells_tree);
  ALLOC(v.flags, 1);
  ALLOC(s->d_pairs, 1);
  ALLOC(s->d_kept, 1);
  ALLOC(v.wstate, (size_t)kWarmInts);
  ALLOC(v.wrel, N);
  ALLOC(v.wmgrp, N);
  ALLOC(v.wcin, (size_t)C);
  ALLOC(v.wcout, (size_t)C);
  ALLOC(v.wtouch, (size_t)C);
  ALLOC(v.wtold, (size_t)C);
  ALLOC(v.wtlist, (size_t)C);
  ALLOC(v.wcnew, (size_t)C);
  ALLOC(v.winoff, (size_t)C);
  ALLOC(v.wfill, (size_t)C);
  ALLOC(v.wrowtot, 4 * ((size_t)s->B + 1));
#undef ALLOC
  // P2P: enough waves to fill every CU
  s->cus = cus;
  default_p2p_config(s->max_n, &s->p2p_variant, &s->p2p_bpc);
  s->p2p_grid = p2p_grid_for(s, s->p2p_variant, s->p2p_bpc);
  s->adj_grid = adj_grid_for(s);
  *out = s;
  return NBFMM_OK;
}

int nbfmm_solver_destroy(nbfmm_solver* solver) {
  if (!solver) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  hipDeviceSynchronize();
  free_all(solver);
  delete solver;
  return NBFMM_OK;
}

int nbfmm_solver_info(const nbfmm_solver* s, int* num_level, int* base_dim, int* max_num_particle) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  if (num_level) *num_level = s->L;
  if (base_dim) *base_dim = s->B;
  if (max_num_particle) *max_num_particle = s->max_n;
  return NBFMM_OK;
}

int nbfmm_solve(nbfmm_solver* s, int n, const float* pos0, const float* w0, float* eff0, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  if (n > 0 && (!pos0 || !w0 || !eff0)) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null array");
  return run_pipeline(s, n, 0.0f, kP2PFinal, call_io(pos0, w0, eff0), as_stream(stream));
}

int nbfmm_predo(nbfmm_solver* s, int n, const float* pos0, const float* w0, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  if (n > 0 && (!pos0 || !w0)) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null array");
  NB_HIP(launch_predo(view(s, n), call_io(pos0, w0), 0, as_stream(stream)));
  s->warm_n = n > 0 ? n : -1;
  return NBFMM_OK;
}

int nbfmm_p2p(nbfmm_solver* s, int n, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  NB_HIP(launch_p2p(view(s, n), kP2POnly, call_io(nullptr, nullptr), s->p2p_grid, s->p2p_variant,
                    as_stream(stream)));
  return NBFMM_OK;
}

int nbfmm_p2m(nbfmm_solver* s, int n, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  NB_HIP(launch_p2m(view(s, n), as_stream(stream)));
  return NBFMM_OK;
}

int nbfmm_m2m(nbfmm_solver* s, void* stream) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  NB_HIP(launch_m2m(view(s, 0), as_stream(stream)));
  return NBFMM_OK;
}

int nbfmm_m2l(nbfmm_solver* s, void* stream) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  NB_HIP(launch_m2l(view(s, 0), as_stream(stream)));
  return NBFMM_OK;
}

int nbfmm_l2l(nbfmm_solver* s, void* stream) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  NB_HIP(launch_l2l(view(s, 0), as_stream(stream)));
  return NBFMM_OK;
}

int nbfmm_l2p(nbfmm_solver* s, int n, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  NB_HIP(launch_l2p(view(s, n), as_stream(stream)));
  return NBFMM_OK;
}

int nbfmm_postdo(nbfmm_solver* s, int n, float* eff0, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  if (n > 0 && !eff0) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null array");
  NB_HIP(launch_postdo(view(s, n), (float2*)eff0, as_stream(stream)));
  return NBFMM_OK;
}

int nbfmm_buffer_bytes(const nbfmm_solver* s, int id, int count, size_t* bytes) {
  if (!s || !bytes) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null argument");
  const size_t C = (size_t)s->B * s->B;
  const size_t cube = C * (size_t)(s->L < 1 ? 1 : s->L);
  switch (id) {
    case NBFMM_BUF_POSITION:
    case NBFMM_BUF_EFFECT:
    case NBFMM_BUF_INDEX: *bytes = (size_t)count * 8; break;
    case NBFMM_BUF_WEIGHT:
    case NBFMM_BUF_PERM: *bytes = (size_t)count * 4; break;
    case NBFMM_BUF_HEAD: *bytes = (C + 1) * 4; break;
    case NBFMM_BUF_CELL_POSITION:
    case NBFMM_BUF_CELL_EFFECT: *bytes = cube * 8; break;
    case NBFMM_BUF_CELL_WEIGHT: *bytes = cube * 4; break;
    default: return fail(NBFMM_ERR_INVALID_ARGUMENT, "unknown buffer id");
  }
  return NBFMM_OK;
}

static int buffer_io(nbfmm_solver* s, int id, int count, void* ptr, bool write, hipStream_t st) {
  if (!s || !ptr) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null argument");
  const SysView v = view(s, count);
  switch (id) {
    case NBFMM_BUF_POSITION:
    case NBFMM_BUF_EFFECT:
    case NBFMM_BUF_WEIGHT:
    case NBFMM_BUF_INDEX:
    case NBFMM_BUF_PERM:
      if (count < 0 || count > s->max_n) return fail(NBFMM_ERR_CAPACITY, "count exceeds max_num_particle");
      NB_HIP(write ? launch_part_write(v, id, count, ptr, st) : launch_part_read(v, id, count, ptr, st));
      return NBFMM_OK;
    case NBFMM_BUF_HEAD: {
      const size_t bytes = ((size_t)s->B * s->B + 1) * 4;
      if (write) {
        NB_HIP(hipMemcpyAsync(s->v.head, ptr, bytes, hipMemcpyDeviceToDevice, st));
        NB_HIP(launch_occ_from_head(v, st));
      } else {
        NB_HIP(hipMemcpyAsync(ptr, s->v.head, bytes, hipMemcpyDeviceToDevice, st));
      }
      return NBFMM_OK;
    }
    case NBFMM_BUF_CELL_POSITION:
    case NBFMM_BUF_CELL_WEIGHT:
    case NBFMM_BUF_CELL_EFFECT: {
      const int which = id == NBFMM_BUF_CELL_POSITION ? 0 : (id == NBFMM_BUF_CELL_WEIGHT ? 1 : 2);
      NB_HIP(write ? launch_cube_write(v, which, ptr, st) : launch_cube_read(v, which, ptr, st));
      return NBFMM_OK;
    }
    default: return fail(NBFMM_ERR_INVALID_ARGUMENT, "unknown buffer id");
  }
}

int nbfmm_buffer_read(nbfmm_solver* s, int id, int count, void* dst, void* stream) {
  return buffer_io(s, id, count, dst, false, as_stream(stream));
}

int nbfmm_buffer_write(nbfmm_solver* s, int id, int count, const void* src, void* stream) {
  if (s) {
    // the stored sort may no longer be one: the next sort starts cold, and a
    // step graph captured with the warm start is dropped
    s->warm_n = -1;
    drop_graph(s);
  }
  return buffer_io(s, id, count, const_cast<void*>(src), true, as_stream(stream));
}

int nbfmm_prune(nbfmm_solver* s, int n, float* xcur, float* xpre, float* w, int* num_kept, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  if (!num_kept) return fail(NBFMM_ERR_INVALID_ARGUMENT, "num_kept is null");
  if (n > 0 && (!xcur || !xpre || !w)) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null array");
  *num_kept = 0;
  if (n == 0) return NBFMM_OK;
  hipStream_t st = as_stream(stream);
  const SysView& v = s->v;
  const float4 lim = make_float4(s->limits[0], s->limits[1], s->limits[2], s->limits[3]);
  NB_HIP(launch_prune(n, lim, (float2*)xcur, (float2*)xpre, w, v.hist, v.tpart, (float*)v.gsort, s->d_kept, st));
  int kept = 0;
  NB_HIP(hipMemcpyAsync(&kept, s->d_kept, sizeof(int), hipMemcpyDeviceToHost, st));
  NB_HIP(hipStreamSynchronize(st));
  *num_kept = kept;
  return NBFMM_OK;
}

int nbfmm_set_warm_start(nbfmm_solver* s, int enable) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  s->warm_enabled = enable != 0;
  drop_graph(s);
  return NBFMM_OK;
}

int nbfmm_set_small_front(nbfmm_solver* s, int enable) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  s->front_enabled = enable != 0;
  drop_graph(s);
  return NBFMM_OK;
}

int nbfmm_solver_flags(nbfmm_solver* s, unsigned* flags, int clear) {
  if (!s || !flags) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null argument");
  NB_HIP(hipDeviceSynchronize());
  NB_HIP(hipMemcpy(flags, s->v.flags, 4, hipMemcpyDeviceToHost));
  if (clear) NB_HIP(hipMemset(s->v.flags, 0, 4));
  return NBFMM_OK;
}

int nbfmm_update(int n, float tick, float* xcur, float* xpre, const float* acc, void* stream) {
  if (n < 0) return fail(NBFMM_ERR_INVALID_ARGUMENT, "num_star < 0");
  if (n > 0 && (!xcur || !xpre || !acc)) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null array");
  NB_HIP(launch_update(n, tick, (float2*)xcur, (float2*)xpre, (const float2*)acc, as_stream(stream)));
  return NBFMM_OK;
}

int nbfmm_step(nbfmm_solver* s, int n, float tick, float* xcur, float* xpre, const float* w, float* acc,
               void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  if (n > 0 && (!xcur || !xpre || !w || !acc)) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null array");
  return run_pipeline(s, n, tick, kP2PStep, call_io(xcur, w, acc, xcur, xpre), as_stream(stream));
}

int nbfmm_run(nbfmm_solver* s, int n, float tick, int steps, float* xcur, float* xpre, const float* w, float* acc,
              int use_graph, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  if (steps < 0) return fail(NBFMM_ERR_INVALID_ARGUMENT, "num_steps < 0");
  hipStream_t st = as_stream(stream);
  if (n > 0 && (!xcur || !xpre || !w || !acc)) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null array");
  if (!use_graph || s->profile || n == 0) {
    // eager: the launches of the replayed steps, one by one (with the warm
    // start on, every step but the last makes the next sort's first pass)
    CallIO io = call_io(xcur, w, acc, xcur, xpre);
    const bool chain = s->warm_enabled && n > 0 && !use_front(s, n);
    for (int k = 0; k < steps; ++k) {
      const int mode = chain && k + 1 < steps ? kP2PStepSort : kP2PStep;
      if ((rc = run_pipeline(s, n, tick, mode, io, st, chain && k > 0))) return rc;
    }
    return NBFMM_OK;
  }
  if (steps == 0) return NBFMM_OK;
  if (!st) {
    // graphs are captured and replayed on the library's side stream, ordered
    // with the null stream on both sides
    NB_HIP(s->side.begin());
    rc = nbfmm_run(s, n, tick, steps, xcur, xpre, w, acc, use_graph, (void*)s->side.s);
    const hipError_t e = s->side.end();
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "side stream");
    return NBFMM_OK;
  }
  // The first step is eager: it sorts from the caller's arrays (warm from an
  // earlier sort of n, else cold).  The other steps replay captured steps.
  // With the warm start on, the P2P of every step but the last also makes
  // the first pass of the next step's sort (binning each new position in the
  // old sorted order, with the mass from the current sort: w is constant
  // during the call), so those sorts start at their second pass; the last
  // step leaves nothing pending, and the next call sorts from the caller's
  // arrays as usual.
  CallIO io = call_io(xcur, w, acc, xcur, xpre);
  const bool chain = s->warm_enabled && !use_front(s, n);  // after the first step's sort, warm_n == n
  if ((rc = run_pipeline(s, n, tick, chain && steps > 1 ? kP2PStepSort : kP2PStep, io, st))) return rc;
  if (--steps == 0) return NBFMM_OK;
  auto& g = s->gkey;
  const bool same = s->graph_exec[1] && g.n == n && g.tick == tick && g.xcur == xcur && g.xpre == xpre &&
                    g.w == w && g.acc == acc && g.stream == st && g.warm == chain;
  if (!same) {
    drop_graph(s);
    // without the warm start every step is the last kind
    const int mid = chain ? kP2PStepSort : kP2PStep;
    auto one = [&](int mode) { return [&, mode]() { return run_pipeline(s, n, tick, mode, io, st, chain); }; };
    if (chain && (rc = capture_graph(st, 1, one(kP2PStepSort), &s->graph[0], &s->graph_exec[0]))) return rc;
    if ((rc = capture_graph(st, 1, one(kP2PStep), &s->graph[1], &s->graph_exec[1]))) return rc;
    if ((rc = capture_graph(st, kRunGraphSteps, one(mid), &s->graph[2], &s->graph_exec[2]))) return rc;
    g.n = n; g.tick = tick; g.xcur = xcur; g.xpre = xpre; g.w = w; g.acc = acc; g.stream = st; g.warm = chain;
  }
  return replay_run(s->graph_exec, chain, steps, st);
}

// The forward state the adjoint needs (sorted particles, head, multipoles
// of every level) at the given positions.
static int adjoint_prepare(nbfmm_solver* s, int n, const CallIO& io, hipStream_t st) {
  const SysView v = view(s, n);
  const int rc = sort_stage(s, v, io, st);
  if (rc) return rc;
  if (n > 0) NB_HIP(launch_upward(v, st, true));
  return NBFMM_OK;
}

int nbfmm_solve_adjoint(nbfmm_solver* s, int n, const float* pos0, const float* w0, const float* grad_eff,
                        float* grad_pos, float* grad_w, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  if (n > 0 && (!pos0 || !w0 || !grad_eff || !grad_pos || !grad_w))
    return fail(NBFMM_ERR_INVALID_ARGUMENT, "null array");
  hipStream_t st = as_stream(stream);
  CallIO io = call_io(pos0, w0);
  io.gin = (const float2*)grad_eff;
  io.gpos = (float2*)grad_pos;
  io.gw = grad_w;
  io.gscale = 1.0f;
  if ((rc = adjoint_prepare(s, n, io, st))) return rc;
  NB_HIP(launch_adjoint(view(s, n), io, 0, s->adj_grid, st));
  return NBFMM_OK;
}

int nbfmm_step_adjoint(nbfmm_solver* s, int n, float tick, const float* x, const float* w, float* lam, float* pi,
                       float* grad_w, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  if (n > 0 && (!x || !w || !lam || !pi || !grad_w)) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null array");
  hipStream_t st = as_stream(stream);
  CallIO io = call_io(x, w);
  io.gin = (const float2*)lam;
  io.gpos = (float2*)lam;
  io.gpre = (float2*)pi;
  io.gw = grad_w;
  io.gscale = tick * tick;
  if ((rc = adjoint_prepare(s, n, io, st))) return rc;
  NB_HIP(launch_adjoint(view(s, n), io, 1, s->adj_grid, st));
  return NBFMM_OK;
}

int nbfmm_trajectory_workspace(int n, int num_steps, int checkpoint_every, size_t* checkpoint_bytes,
                               size_t* work_bytes) {
  if (n < 0 || num_steps < 0 || checkpoint_every <= 0) return fail(NBFMM_ERR_INVALID_ARGUMENT, "bad sizes");
  const size_t state = (size_t)n * 2 * sizeof(float);
  const size_t nck = (size_t)(num_steps / checkpoint_every) + 1;
  if (checkpoint_bytes) *checkpoint_bytes = nck * 2 * state;
  // recompute state: x, x_pre and acceleration scratch
  if (work_bytes) *work_bytes = 3 * state;
  return NBFMM_OK;
}

int nbfmm_trajectory_forward(nbfmm_solver* s, int n, float tick, int num_steps, int K, float* xcur, float* xpre,
                             const float* w, float* acc, float* ckpt, int n_obs, const int* obs_steps, float* obs_pos,
                             void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  if (num_steps < 0 || K <= 0 || n_obs < 0 || (n_obs && (!obs_steps || !obs_pos)) || !ckpt)
    return fail(NBFMM_ERR_INVALID_ARGUMENT, "bad trajectory arguments");
  hipStream_t st = as_stream(stream);
  const size_t sb = (size_t)n * 2 * sizeof(float);
  auto observe = [&](int k) -> int {
    for (int o = 0; o < n_obs; ++o)
      if (obs_steps[o] == k) NB_HIP(hipMemcpyAsync(obs_pos + (size_t)o * n * 2, xcur, sb, hipMemcpyDeviceToDevice, st));
    return NBFMM_OK;
  };
  // the steps between two events (a checkpoint or an observation) as one
  // eager run: the same bits as that many nbfmm_step calls, with the sorts
  // chained (no graph: the caller's arrays, which key a run's graphs, are
  // new on every call)
  for (int k = 0;;) {
    if (k % K == 0) {
      float* c = ckpt + (size_t)(k / K) * 2 * n * 2;
      NB_HIP(hipMemcpyAsync(c, xcur, sb, hipMemcpyDeviceToDevice, st));
      NB_HIP(hipMemcpyAsync(c + (size_t)n * 2, xpre, sb, hipMemcpyDeviceToDevice, st));
    }
    if ((rc = observe(k))) return rc;
    if (k == num_steps) break;
    int next = std::min(num_steps, (k / K + 1) * K);
    for (int o = 0; o < n_obs; ++o)
      if (obs_steps[o] > k && obs_steps[o] < next) next = obs_steps[o];
    if ((rc = nbfmm_run(s, n, tick, next - k, xcur, xpre, w, acc, 0, stream))) return rc;
    k = next;
  }
  return NBFMM_OK;
}

namespace {

// The copies that keep a sorted state (particles, keys, head, occupied
// leaves and their count, multipoles: n, heads, occ, nocc and tree entries
// of the arrays that start at v's pointers) in a slot, for one launch.
CopyRegions sort_copy(const SortArrays& sl, const SysView& v, size_t n, size_t heads, size_t occ, size_t nocc,
                      size_t tree) {
  CopyRegions c{};
  auto put = [&](int r, const void* src, void* dst, size_t bytes) {
    c.src[r] = static_cast<const uint32_t*>(src);
    c.dst[r] = static_cast<uint32_t*>(dst);
    c.words[r] = bytes / sizeof(uint32_t);
  };
  put(0, v.part, sl.part, n * sizeof(float4));
  put(1, v.key, sl.key, n * sizeof(uint32_t));
  put(2, v.head, sl.head, heads * sizeof(int));
  put(3, v.occ, sl.occ, occ * sizeof(int));
  put(4, v.nocc, sl.nocc, nocc * sizeof(int));
  put(5, v.mpole, sl.mpole, tree * sizeof(float4));
  return c;
}

// At least `count` sorted-state slots, each for max_n particles.
int ensure_slots(nbfmm_solver* s, int count) {
  const size_t N = (size_t)s->max_n, C = (size_t)s->B * s->B;
  while ((int)s->slots.size() < count) {
    nbfmm_solver::SortSlot sl{};
    int rc;
    if ((rc = dev_alloc(s, &sl.part, N)) || (rc = dev_alloc(s, &sl.key, N)) || (rc = dev_alloc(s, &sl.head, C + 1)) ||
        (rc = dev_alloc(s, &sl.occ, C)) || (rc = dev_alloc(s, &sl.nocc, 1)) ||
        (rc = dev_alloc(s, &sl.mpole, (size_t)s->ncells_tree)))
      return rc;
    s->slots.push_back(sl);
  }
  return NBFMM_OK;
}

// The solver's sorted state (of its last sort, of n particles) into a slot.
int save_sort(nbfmm_solver* s, int n, const nbfmm_solver::SortSlot& sl, hipStream_t st) {
  const SysView& v = s->v;
  const size_t C = (size_t)s->B * s->B;
  NB_HIP(launch_copy_regions(sort_copy(sl, v, (size_t)n, C + 1, std::min(C, (size_t)n), 1, (size_t)s->ncells_tree),
                             kCopyRegions, st));
  return NBFMM_OK;
}

}  // namespace

// Reverse pass of a checkpointed trajectory.  In: the checkpoints written by
// nbfmm_trajectory_forward, the loss gradient at each observation step
// (obs_grad, n_obs x n float2).  Out: gradients wrt the initial x, x_pre and
// the masses.  Each segment of K steps is recomputed from its checkpoint
// (same kernels, so the same bits as the forward) and then stepped back.
int nbfmm_trajectory_backward(nbfmm_solver* s, int n, float tick, int num_steps, int K, const float* w,
                              const float* ckpt, int n_obs, const int* obs_steps, const float* obs_grad,
                              float* grad_x0, float* grad_xpre0, float* grad_w, float* work, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  if (num_steps < 0 || K <= 0 || n_obs < 0 || (n_obs && (!obs_steps || !obs_grad)) || !ckpt || !work ||
      !grad_x0 || !grad_xpre0 || !grad_w)
    return fail(NBFMM_ERR_INVALID_ARGUMENT, "bad trajectory arguments");
  hipStream_t st = as_stream(stream);
  const size_t sb = (size_t)n * 2 * sizeof(float);
  float* lam = grad_x0;
  float* pi = grad_xpre0;
  float* xc = work;  // recompute state
  float* xp = xc + (size_t)n * 2;
  float* ac = xp + (size_t)n * 2;
  NB_HIP(launch_fill(2 * n, 0.0f, lam, st));
  NB_HIP(launch_fill(2 * n, 0.0f, pi, st));
  NB_HIP(launch_fill(n, 0.0f, grad_w, st));
  auto add_obs = [&](int k) -> int {
    for (int o = 0; o < n_obs; ++o)
      if (obs_steps[o] == k)
        NB_HIP(launch_axpy(2 * n, 1.0f, obs_grad + (size_t)o * n * 2, lam, st));
    return NBFMM_OK;
  };
  if ((rc = add_obs(num_steps))) return rc;
  if (n == 0 || num_steps == 0) {
    for (int k = num_steps - 1; k >= 0; --k)
      if ((rc = add_obs(k))) return rc;
    return NBFMM_OK;
  }
  // Two sets of K kept states: segment sg fills set sg % 2 on the recompute
  // stream R (its steps k0 .. k1-2 again, as nbfmm_run's eager steps, each
  // kept once its sort and tree are built, then the sort and tree of
  // x_{k1-1}, as nbfmm_step_adjoint would build them) while the caller's
  // stream runs the adjoint steps of segment sg + 1 from the other set.  The
  // adjoint steps read only a kept state and their own scratch (the sorted
  // gradient, the subtree sums, the cell gradients), which no forward kernel
  // touches, and R alone uses the solver's sort and tree.
  const int nseg = (num_steps + K - 1) / K;
  const int sets = nseg > 1 ? 2 : 1;  // one segment: nothing to overlap
  if ((rc = ensure_slots(s, sets * K))) return rc;
  RecomputeStream& rs = s->recompute;
  NB_HIP(rs.make());
  hipStream_t R = rs.s;
  NB_HIP(hipEventRecord(rs.start, st));
  NB_HIP(hipEventRecord(rs.done[0], st));
  NB_HIP(hipEventRecord(rs.done[1], st));
  NB_HIP(hipStreamWaitEvent(R, rs.start, 0));
  const bool chain = s->warm_enabled && !use_front(s, n);
  for (int sg = nseg - 1; sg >= 0; --sg) {
    const int p = sg % sets;
    const nbfmm_solver::SortSlot* set = s->slots.data() + (size_t)p * K;
    const int k0 = sg * K;
    const int k1 = std::min(k0 + K, num_steps);
    // R: the set is free once the adjoint steps of segment sg + 2 are done
    NB_HIP(hipStreamWaitEvent(R, rs.done[p], 0));
    const float* c = ckpt + (size_t)sg * 2 * n * 2;
    NB_HIP(hipMemcpyAsync(xc, c, sb, hipMemcpyDeviceToDevice, R));
    NB_HIP(hipMemcpyAsync(xp, c + (size_t)n * 2, sb, hipMemcpyDeviceToDevice, R));
    // with the warm start on, every recomputed step makes the next sort's
    // first pass (the last one's is x_{k1-1}'s): the same bits as separate steps
    CallIO rio = call_io(xc, w, ac, xc, xp);
    for (int k = k0; k + 1 < k1; ++k) {
      if ((rc = run_pipeline(s, n, tick, chain ? kP2PStepSort : kP2PStep, rio, R, chain && k > k0))) return rc;
      if ((rc = save_sort(s, n, set[k - k0], R))) return rc;
    }
    const SysView v = view(s, n);
    if ((rc = sort_stage(s, v, call_io(xc, w), R, chain && k1 - 1 > k0))) return rc;
    NB_HIP(launch_upward(v, R, true));
    if ((rc = save_sort(s, n, set[k1 - 1 - k0], R))) return rc;
    NB_HIP(hipEventRecord(rs.kept[p], R));
    // the caller's stream: steps k1-1 .. k0 back, each on its kept state
    NB_HIP(hipStreamWaitEvent(st, rs.kept[p], 0));
    for (int k = k1 - 1; k >= k0; --k) {
      const nbfmm_solver::SortSlot& sl = set[k - k0];
      SysView va = v;
      va.part = sl.part;
      va.key = sl.key;
      va.head = sl.head;
      va.occ = sl.occ;
      va.nocc = sl.nocc;
      va.mpole = sl.mpole;
      CallIO io = call_io(nullptr, nullptr);  // (the adjoint reads the kept state's particles)
      io.gin = (const float2*)lam;
      io.gpos = (float2*)lam;
      io.gpre = (float2*)pi;
      io.gw = grad_w;
      io.gscale = tick * tick;
      NB_HIP(launch_adjoint(va, io, 1, s->adj_grid, st));
      if ((rc = add_obs(k))) return rc;
    }
    NB_HIP(hipEventRecord(rs.done[p], st));
  }
  return NBFMM_OK;
}

int nbfmm_p2p_variant_count(void) { return p2p_num_variants(); }

int nbfmm_p2p_variant_info(int variant, int* stage, int* unroll, int* waves_per_block) {
  if (variant < 0 || variant >= p2p_num_variants()) return fail(NBFMM_ERR_INVALID_ARGUMENT, "unknown P2P variant");
  int a = 0, b = 0, c = 0;
  p2p_variant_info(variant, &a, &b, &c);
  if (stage) *stage = a;
  if (unroll) *unroll = b;
  if (waves_per_block) *waves_per_block = c;
  return NBFMM_OK;
}

int nbfmm_set_p2p_config(nbfmm_solver* s, int variant, int blocks_per_cu) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  if (variant < 0 || variant >= p2p_num_variants()) return fail(NBFMM_ERR_INVALID_ARGUMENT, "unknown P2P variant");
  if (blocks_per_cu < 1 || blocks_per_cu > 64) return fail(NBFMM_ERR_INVALID_ARGUMENT, "blocks_per_cu in [1, 64]");
  s->p2p_variant = variant;
  s->p2p_bpc = blocks_per_cu;
  s->p2p_grid = p2p_grid_for(s, variant, blocks_per_cu);
  drop_graph(s);
  return NBFMM_OK;
}

int nbfmm_get_p2p_config(const nbfmm_solver* s, int* variant, int* blocks_per_cu) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  if (variant) *variant = s->p2p_variant;
  if (blocks_per_cu) *blocks_per_cu = s->p2p_bpc;
  return NBFMM_OK;
}

int nbfmm_set_p2p_wait_polls(nbfmm_solver* s, int polls) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  if (polls < 1) return fail(NBFMM_ERR_INVALID_ARGUMENT, "polls must be >= 1");
  s->v.wait_polls = polls;
  drop_graph(s);
  return NBFMM_OK;
}

// Times every P2P configuration (variant x blocks per CU) on this input and
// keeps the fastest.  Synchronises.  The configurations give identical
// results, so tuning never changes an answer.
int nbfmm_autotune(nbfmm_solver* s, int n, const float* pos0, const float* w0, int repeats, int* best_variant,
                   int* best_bpc, float* best_ms, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  if (n <= 0 || !pos0 || !w0) return fail(NBFMM_ERR_INVALID_ARGUMENT, "autotune needs particles");
  if (repeats < 1) repeats = 1;
  hipStream_t st = as_stream(stream);
  const SysView v = view(s, n);
  const CallIO io = call_io(pos0, w0);
  if ((rc = sort_stage(s, v, io, st))) return rc;
  NB_HIP(launch_upward(v, st, false));
  NB_HIP(launch_m2l_fused(v, st));
  // Time the step kernel itself (L2L walk, scatter and Verlet epilogue
  // included) with its outputs pointed at scratch that the next predo
  // rewrites: accelerations into eff, x and x_pre into the two halves of the
  // row-pass particle buffer.  The caller's arrays are not touched.
  CallIO tio = io;
  tio.acc0 = v.eff;
  tio.xcur = (float2*)v.tpart;
  tio.xpre = (float2*)v.tpart + s->max_n;
  hipEvent_t e0, e1;
  NB_HIP(hipEventCreate(&e0));
  NB_HIP(hipEventCreate(&e1));
  // grids of 16 to 64 waves per CU (8 per SIMD is the hardware's maximum
  // residency; LDS holds three 8-wave blocks of the 384-source variants, so
  // 24 is one round of them and 32 queues a second round)
  static const int wpcs[] = {16, 24, 32, 48, 64};
  float best = 3.4e38f;
  int bv = s->p2p_variant, bb = s->p2p_bpc;
  hipError_t err = hipSuccess;
  for (int var = 0; var < p2p_num_variants() && err == hipSuccess; ++var) {
    for (int wpc : wpcs) {
      const int bpc = std::max(1, wpc / variant_waves(var));
      const int grid = p2p_grid_for(s, var, bpc);
      err = launch_p2p(v, kP2PStep, tio, grid, var, st);  // warm-up
      if (err == hipSuccess) err = hipEventRecord(e0, st);
      for (int r = 0; r < repeats && err == hipSuccess; ++r) err = launch_p2p(v, kP2PStep, tio, grid, var, st);
      if (err == hipSuccess) err = hipEventRecord(e1, st);
      if (err == hipSuccess) err = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
      if (err != hipSuccess) break;
      ms /= repeats;
      if (ms < best) {
        best = ms;
        bv = var;
        bb = bpc;
      }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (err != hipSuccess) return hip_fail(err, "autotune");
  s->p2p_variant = bv;
  s->p2p_bpc = bb;
  s->p2p_grid = p2p_grid_for(s, bv, bb);
  drop_graph(s);
  if (best_variant) *best_variant = bv;
  if (best_bpc) *best_bpc = bb;
  if (best_ms) *best_ms = best;
  return NBFMM_OK;
}

int nbfmm_profile_enable(nbfmm_solver* s, int enable) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  s->profile = enable != 0;
  return NBFMM_OK;
}

// Folds the recorded event groups into the per-stage sums.
static int prof_collect(nbfmm_solver* s) {
  for (size_t k = 0; k + kProfEvents <= s->ev_used; k += kProfEvents) {
    NB_HIP(hipEventSynchronize(s->ev[k + kProfStages]));
    for (int j = 0; j < kProfStages; ++j) {
      float ms = 0.0f;
      NB_HIP(hipEventElapsedTime(&ms, s->ev[k + j], s->ev[k + j + 1]));
      s->stage_ms[j] += ms;
    }
    s->prof_steps += 1;
  }
  s->ev_used = 0;
  return NBFMM_OK;
}

int nbfmm_p2p_time(nbfmm_solver* s, double* total_ms, int* launches, int reset) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  if (int rc = prof_collect(s)) return rc;
  if (total_ms) *total_ms = s->stage_ms[kProfStages - 1];
  if (launches) *launches = s->prof_steps;
  if (reset) {
    for (double& m : s->stage_ms) m = 0.0;
    s->prof_steps = 0;
  }
  return NBFMM_OK;
}

int nbfmm_stage_times(nbfmm_solver* s, double* ms, int* steps, int reset) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  if (int rc = prof_collect(s)) return rc;
  if (ms)
    for (int j = 0; j < kProfStages; ++j) ms[j] = s->stage_ms[j];
  if (steps) *steps = s->prof_steps;
  if (reset) {
    for (double& m : s->stage_ms) m = 0.0;
    s->prof_steps = 0;
  }
  return NBFMM_OK;
}

int nbfmm_pair_count(nbfmm_solver* s, int n, double* pairs, int* occupied) {
  int rc = check_n(s, n);
  if (rc) return rc;
  NB_HIP(hipDeviceSynchronize());
  NB_HIP(hipMemset(s->d_pairs, 0, 8));
  NB_HIP(launch_pair_count(view(s, n), s->d_pairs, nullptr));
  unsigned long long h = 0;
  int occ = 0;
  NB_HIP(hipMemcpy(&h, s->d_pairs, 8, hipMemcpyDeviceToHost));
  NB_HIP(hipMemcpy(&occ, s->v.nocc, 4, hipMemcpyDeviceToHost));
  if (pairs) *pairs = (double)h;
  if (occupied) *occupied = occ;
  return NBFMM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batch

struct nbfmm_batch {
  int S = 0;
  int total_n = 0;
  size_t n_all = 0, head_all = 0, cell_all = 0, tree_all = 0;  // sizes of the concatenated arrays
  // sorted states kept for the step adjoint (nbfmm_batch_keep_sort), each
  // laid out as the batch's own arrays
  using SortSlot = SortArrays;
  std::vector<SortSlot> slots;
  std::vector<int> part_off;      // S+1
  std::vector<SysView> hv;        // host copies of the views
  std::vector<SysDesc> hdesc;     // per-system sizes and array offsets
  SysDesc* ddesc = nullptr;       // device copy
  BatchLaunch bl{};
  // On by default: measured on 256 systems (7.1M particles) a batched step
  // takes 1.23 ms with the warm-started sort and 1.26 ms with the cold one.
  bool warm_enabled = true;
  bool warm_valid = false;        // key/part/head hold a sort of every system
  int cus = 256;
  int p2p_variant = kDefaultBatchP2PVariant;
  int p2p_bpc = kDefaultBatchP2PBlocksPerCU;
  std::map<std::pair<int, int>, KernelMap> p2p_maps;  // (waves per block, blocks per CU) -> map
  // P2P work of each system, the weight of its share of the P2P grids: n
  // until nbfmm_batch_balance measures pairs + kBalancePerParticle * n
  std::vector<double> work;
  double total_work = 0.0;
  bool balanced = false;
  unsigned long long* d_pairs = nullptr;  // S pair counts
  std::vector<void*> allocs;
  // graphs of a run, as for a solver: [0] a step followed by another, [1] the
  // last step, [2] kRunGraphSteps steps of kind [0]
  hipGraphExec_t graph_exec[3] = {nullptr, nullptr, nullptr};
  hipGraph_t graph[3] = {nullptr, nullptr, nullptr};
  SideStream side;  // replays asked for on the null stream run here
  struct {
    const void *xcur = nullptr, *xpre = nullptr, *w = nullptr, *acc = nullptr;
    hipStream_t stream = nullptr;
    bool valid = false;
    bool warm = false;
  } gkey;
};

namespace {

template <class T>
int batch_alloc(nbfmm_batch* b, T** p, size_t count) {
  void* q = nullptr;
  const size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 16;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return fail(NBFMM_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  b->allocs.push_back(q);
  e = hipMemset(q, 0, bytes);
  if (e != hipSuccess) return hip_fail(e, "hipMemset");
  *p = (T*)q;
  return NBFMM_OK;
}

void batch_free(nbfmm_batch* b) {
  for (int k = 0; k < 3; ++k) {
    if (b->graph_exec[k]) hipGraphExecDestroy(b->graph_exec[k]);
    if (b->graph[k]) hipGraphDestroy(b->graph[k]);
  }
  b->side.release();
  for (void* p : b->allocs) hipFree(p);
  b->allocs.clear();
}

// A captured step graph holds the kernels and launch maps of its capture.
void drop_batch_graph(nbfmm_batch* b) {
  for (int k = 0; k < 3; ++k) {
    if (b->graph_exec[k]) { hipGraphExecDestroy(b->graph_exec[k]); b->graph_exec[k] = nullptr; }
    if (b->graph[k]) { hipGraphDestroy(b->graph[k]); b->graph[k] = nullptr; }
  }
  b->gkey.valid = false;
}

// Builds a launch map from the per-system block counts.
int make_map(nbfmm_batch* b, const std::vector<int>& blocks, KernelMap* km) {
  const int S = (int)blocks.size();
  std::vector<int> start(S + 1, 0);
  for (int s = 0; s < S; ++s) start[s + 1] = start[s] + blocks[s];
  std::vector<int> sys(std::max(start[S], 1), 0);
  for (int s = 0; s < S; ++s)
    for (int k = start[s]; k < start[s + 1]; ++k) sys[k] = s;
  int* d = nullptr;
  int rc = batch_alloc(b, &d, (size_t)(S + 1) + sys.size());
  if (rc) return rc;
  NB_HIP(hipMemcpy(d, start.data(), (S + 1) * sizeof(int), hipMemcpyHostToDevice));
  NB_HIP(hipMemcpy(d + S + 1, sys.data(), sys.size() * sizeof(int), hipMemcpyHostToDevice));
  km->sys.base = b->hv[0];  // system 0's arrays start at every array's base
  km->sys.desc = b->ddesc;
  km->sys.bstart = d;
  km->sys.bsys = d + S + 1;
  km->blocks = start[S];
  return NBFMM_OK;
}

// P2P blocks of `waves` waves, `total_blocks` in all, shared among the
// systems in proportion to their P2P work (b->work) and never more than a
// system's leaves can use.  Each block works only on its own system's runs,
// so a system given fewer blocks than its work needs finishes last.
std::vector<int> p2p_share(const nbfmm_batch* b, int waves, int total_blocks) {
  std::vector<int> blocks(b->S);
  for (int s = 0; s < b->S; ++s) {
    const SysView& v = b->hv[s];
    const int cap = std::max(cdiv(std::min(v.n, v.B * v.B), waves), cdiv(v.n, kWave) + v.B);  // as p2p_grid_for
    const int share = (int)((double)total_blocks * b->work[s] / b->total_work + 0.5);
    int nb = std::max(1, std::min(cap, share));
    // a multiple of 8 from 8 blocks on, so that the system's blocks take
    // their runs in XCD order (xcd_block)
    if (nb >= 8) nb = (nb + 4) / 8 * 8;
    blocks[s] = nb;
  }
  return blocks;
}

int batch_use_p2p(nbfmm_batch* b, int variant, int bpc);

// Pairs a particle's P2P cost is worth beyond its pair evaluations (the
// epilogue, the run set-up): keeps systems with few pairs from being starved.
constexpr double kBalancePerParticle = 32.0;

void drop_batch_graph(nbfmm_batch* b);

// Re-weights the P2P shares by the pairs of the current sort (which must be
// valid for every system), rebuilds the P2P and adjoint P2P launch maps and
// drops a captured step graph.  Synchronises.
int batch_balance(nbfmm_batch* b, hipStream_t st) {
  NB_HIP(hipMemsetAsync(b->d_pairs, 0, sizeof(unsigned long long) * b->S, st));
  for (int s = 0; s < b->S; ++s) NB_HIP(launch_pair_count(b->hv[s], b->d_pairs + s, st));
  std::vector<unsigned long long> pairs(b->S);
  NB_HIP(hipMemcpyAsync(pairs.data(), b->d_pairs, sizeof(unsigned long long) * b->S, hipMemcpyDeviceToHost, st));
  NB_HIP(hipStreamSynchronize(st));
  b->total_work = 0.0;
  for (int s = 0; s < b->S; ++s) {
    b->work[s] = (double)pairs[s] + kBalancePerParticle * b->hv[s].n;
    b->total_work += b->work[s];
  }
  b->p2p_maps.clear();  // their device tables stay in b->allocs until the batch is destroyed
  if (int rc = make_map(b, p2p_share(b, kAdjP2PWaves, b->cus * kAdjP2PBatchBlocksPerCU), &b->bl.adj_p2p)) return rc;
  if (int rc = batch_use_p2p(b, b->p2p_variant, b->p2p_bpc)) return rc;
  drop_batch_graph(b);
  b->balanced = true;
  return NBFMM_OK;
}

// Points bl.p2p at the launch map of (variant, blocks per CU), built on
// first use and kept until the batch is destroyed.
int batch_use_p2p(nbfmm_batch* b, int variant, int bpc) {
  const std::pair<int, int> key(variant_waves(variant), bpc);
  auto it = b->p2p_maps.find(key);
  if (it == b->p2p_maps.end()) {
    KernelMap km{};
    const int rc = make_map(b, p2p_share(b, key.first, b->cus * bpc), &km);
    if (rc) return rc;
    it = b->p2p_maps.emplace(key, km).first;
  }
  b->bl.p2p = it->second;
  b->p2p_variant = variant;
  b->p2p_bpc = bpc;
  return NBFMM_OK;
}

// Sort stage of a batch: warm-started from its previous sort when there is
// one (every system keeps its N), else cold.
int batch_sort(nbfmm_batch* b, const CallIO& io, hipStream_t st, bool gathered = false) {
  if (b->warm_enabled && b->warm_valid)
    NB_HIP(launch_predo_warm_batch(b->bl, io, st, gathered));
  else
    NB_HIP(launch_predo_batch(b->bl, io, st));
  b->warm_valid = true;
  return NBFMM_OK;
}

// The first call of a batch shares its P2P grids by the pairs of its first
// sort (one synchronisation), unless the stream is being captured.
int batch_balance_once(nbfmm_batch* b, hipStream_t st) {
  if (b->balanced) return NBFMM_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return NBFMM_OK;
  return batch_balance(b, st);
}

int batch_pipeline(nbfmm_batch* b, int mode, const CallIO& io, hipStream_t st, bool gathered = false) {
  if (int rc = batch_sort(b, io, st, gathered)) return rc;
  if (int rc = batch_balance_once(b, st)) return rc;
  NB_HIP(launch_tree_batch(b->bl, st, true));
  NB_HIP(launch_p2p_batch(b->bl, mode, io, b->p2p_variant, st));
  return NBFMM_OK;
}

}  // namespace

extern "C" {

int nbfmm_batch_create(int S, const int* num_particle, const int* num_level, const float* limits,
                       const float* epsilon, nbfmm_batch** out) {
  if (!out) return fail(NBFMM_ERR_INVALID_ARGUMENT, "out is null");
  *out = nullptr;
  if (S <= 0 || !num_particle || !num_level || !limits || !epsilon)
    return fail(NBFMM_ERR_INVALID_ARGUMENT, "need num_systems > 0 and non-null arrays");
  long long total = 0;
  for (int s = 0; s < S; ++s) {
    if (num_particle[s] <= 0) return fail(NBFMM_ERR_INVALID_ARGUMENT, "every system needs num_particle > 0");
    if (num_level[s] < 0 || num_level[s] > kMaxLevel)
      return fail(NBFMM_ERR_INVALID_ARGUMENT, "num_level must be in [0, 8]");
    const float* lim = limits + 4 * s;
    if (!(lim[0] < lim[2] && lim[1] < lim[3]))
      return fail(NBFMM_ERR_INVALID_ARGUMENT, "position_limits must satisfy x_min < x_max and y_min < y_max");
    if (!(epsilon[s] > 0.0f) || !std::isfinite(epsilon[s]))
      return fail(NBFMM_ERR_INVALID_ARGUMENT, "epsilon must be > 0");
    total += num_particle[s];
  }
  if (total > 0x7fffffff) return fail(NBFMM_ERR_INVALID_ARGUMENT, "too many particles in total");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NBFMM_ERR_NO_DEVICE, "no HIP device");
  int dev = 0, cus = 256;
  hipGetDevice(&dev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
    cus = prop.multiProcessorCount;

  nbfmm_batch* b = new nbfmm_batch();
  b->cus = cus;
  b->S = S;
  b->total_n = (int)total;
  b->part_off.assign(S + 1, 0);
  b->hv.resize(S);
  b->work.resize(S);
  for (int s = 0; s < S; ++s) b->work[s] = (double)num_particle[s];
  b->total_work = (double)total;
  // sizes of the concatenated scratch arrays
  size_t n_all = 0, head_all = 0, row_all = 0, cell_all = 0, hist_all = 0, tree_all = 0;
  std::vector<size_t> head_off(S), row_off(S), cell_off(S), hist_off(S), tree_off(S);
  for (int s = 0; s < S; ++s) {
    SysView& v = b->hv[s];
    v = SysView{};
    v.split_small = 0;  // the other systems fill the GPU: small systems' runs are folded, not split
    v.n = num_particle[s];
    v.L = num_level[s];
    v.B = 1 << (v.L + 1);
    v.lgB = v.L + 1;
    const float* lim = limits + 4 * s;
    v.xmin = lim[0];
    v.ymin = lim[1];
    v.cw = (lim[2] - lim[0]) / (float)v.B;  // predo.cu:112-113
    v.ch = (lim[3] - lim[1]) / (float)v.B;
    v.eps = epsilon[s];
    v.part_off = (int)n_all;
    b->part_off[s] = (int)n_all;
    const int levels = v.L < 1 ? 1 : v.L;
    int off = 0;
    for (int l = 0; l <= kMaxLevel; ++l) {
      if (v.lvl_off(l) != off) {  // the closed form against the running sum; cannot differ
        batch_free(b);
        delete b;
        return fail(NBFMM_ERR_INVALID_ARGUMENT, "level offsets");
      }
      if (l < levels) off += (v.B >> l) * (v.B >> l);
    }
    const size_t C = (size_t)v.B * v.B;
    head_off[s] = head_all;
    row_off[s] = row_all;
    cell_off[s] = cell_all;
    hist_off[s] = hist_all;
    tree_off[s] = tree_all;
    n_all += v.n;
    head_all += C + 1;
    row_all += v.B + 1;
    cell_all += C;
    hist_all += (size_t)v.B * cdiv(v.n, kSortTile);
    tree_all += off;
  }
  b->part_off[S] = (int)n_all;
  b->n_all = n_all;
  b->head_all = head_all;
  b->cell_all = cell_all;
  b->tree_all = tree_all;
  uint32_t *tkey, *key, *hist, *dtotal, *flags;
  float4 *tpart, *part, *mpole;
  float2 *eff, *local, *gsort, *gup;
  float4* gmp;
  int *head, *rowstart, *occ, *nocc;
  int rc = NBFMM_OK;
#define BALLOC(ptr, cnt) \
  if ((rc = batch_alloc(b, &(ptr), (cnt))) != NBFMM_OK) { batch_free(b); delete b; return rc; }
  BALLOC(tkey, n_all);
  BALLOC(tpart, n_all);
  BALLOC(key, n_all);
  BALLOC(part, n_all);
  BALLOC(eff, n_all);
  BALLOC(head, head_all);
  BALLOC(rowstart, row_all);
  BALLOC(occ, cell_all);
  BALLOC(nocc, (size_t)S);
  BALLOC(hist, hist_all);
  BALLOC(dtotal, row_all);
  BALLOC(mpole, tree_all);
  BALLOC(local, tree_all);
  BALLOC(gsort, n_all);
  BALLOC(gup, tree_all);
  BALLOC(gmp, tree_all);
  BALLOC(flags, (size_t)S);
  BALLOC(b->ddesc, (size_t)S);
  int *wstate, *wrel, *wmgrp, *wcin, *wcout, *wtouch, *wtlist, *wcnew, *winoff, *wfill, *wrowtot;
  int2* wtold;
  BALLOC(wstate, (size_t)S * kWarmInts);
  BALLOC(wrel, n_all);
  BALLOC(wmgrp, n_all);
  BALLOC(wcin, cell_all);
  BALLOC(wcout, cell_all);
  BALLOC(wtouch, cell_all);
  BALLOC(wtlist, cell_all);
  BALLOC(wcnew, cell_all);
  BALLOC(winoff, cell_all);
  BALLOC(wfill, cell_all);
  BALLOC(wtold, cell_all);
  BALLOC(wrowtot, 4 * row_all);
#undef BALLOC
  for (int s = 0; s < S; ++s) {
    SysView& v = b->hv[s];
    const size_t po = (size_t)v.part_off;
    v.tkey = tkey + po;
    v.tpart = tpart + po;
    v.key = key + po;
    v.part = part + po;
    v.eff = eff + po;
    v.head = head + head_off[s];
    v.rowstart = rowstart + row_off[s];
    v.occ = occ + cell_off[s];
    v.nocc = nocc + s;
    v.hist = hist + hist_off[s];
    v.dtotal = dtotal + row_off[s];
    v.mpole = mpole + tree_off[s];
    v.local = local + tree_off[s];
    v.gsort = gsort + po;
    v.gup = gup + tree_off[s];
    v.gmp = gmp + tree_off[s];
    v.flags = flags + s;
    v.wstate = wstate + (size_t)s * kWarmInts;
    v.wrel = wrel + po;
    v.wmgrp = wmgrp + po;
    v.wcin = wcin + cell_off[s];
    v.wcout = wcout + cell_off[s];
    v.wtouch = wtouch + cell_off[s];
    v.wtlist = wtlist + cell_off[s];
    v.wcnew = wcnew + cell_off[s];
    v.winoff = winoff + cell_off[s];
    v.wfill = wfill + cell_off[s];
    v.wtold = wtold + cell_off[s];
    v.wrowtot = wrowtot + 4 * row_off[s];
  }
  b->hdesc.resize(S);
  for (int s = 0; s < S; ++s) {
    const SysView& v = b->hv[s];
    SysDesc& d = b->hdesc[s];
    d.n = v.n;
    d.L = v.L;
    d.B = v.B;
    d.lgB = v.lgB;
    d.xmin = v.xmin;
    d.ymin = v.ymin;
    d.cw = v.cw;
    d.ch = v.ch;
    d.eps = v.eps;
    d.tick = v.tick;
    d.part_off = v.part_off;
    d.head_off = (int)head_off[s];
    d.row_off = (int)row_off[s];
    d.cell_off = (int)cell_off[s];
    d.hist_off = (int)hist_off[s];
    d.tree_off = (int)tree_off[s];
  }
  hipError_t e = hipMemcpy(b->ddesc, b->hdesc.data(), S * sizeof(SysDesc), hipMemcpyHostToDevice);
  if (e != hipSuccess) { batch_free(b); delete b; return hip_fail(e, "hipMemcpy system descriptors"); }
  // launch maps
  std::vector<int> tiles(S), rows(S), up(S), mtop(S), m2l(S), aup(S), am2l(S), asort(S), wmerge(S), one(S, 1);
  for (int s = 0; s < S; ++s) {
    const SysView& v = b->hv[s];
    const int C = v.B * v.B;
    tiles[s] = cdiv(v.n, kSortTile);
    rows[s] = v.B;
    up[s] = up_blocks(v.L, v.B);
    mtop[s] = upward_top_blocks(v.L, v.B);
    m2l[s] = m2l_fused_blocks(v);
    aup[s] = adj_up_blocks(v.L, v.B);
    am2l[s] = adj_m2l_fused_blocks(v);
    asort[s] = cdiv(v.n, 256);
    // one wave per 256 particles (touched leaves are about twice the movers;
    // the waves grid-stride over the list).  Measured on the 256-system batch:
    // 1.079-1.081 ms a step with one wave per 64 particles, 1.071-1.072 with
    // one per 256 (the waves that find no leaf still take a round of the grid)
    wmerge[s] = cdiv(std::min(C, v.n / 256 + 64), 256 / kWave);
    b->bl.max_B = std::max(b->bl.max_B, v.B);
  }
  b->bl.warm_one = true;
  for (int s = 0; s < S; ++s) b->bl.warm_one = b->bl.warm_one && warm_rows_one_fits(b->hv[s].B, b->hv[s].n);
  if ((rc = make_map(b, tiles, &b->bl.tiles)) || (rc = make_map(b, rows, &b->bl.rows)) ||
      (rc = make_map(b, up, &b->bl.up)) || (rc = make_map(b, mtop, &b->bl.m2m_top)) ||
      (rc = make_map(b, m2l, &b->bl.m2l)) || (rc = make_map(b, aup, &b->bl.adj_up)) ||
      (rc = make_map(b, am2l, &b->bl.adj_m2l)) || (rc = make_map(b, asort, &b->bl.adj_sort)) ||
      (rc = make_map(b, wmerge, &b->bl.warm_merge)) || (rc = make_map(b, one, &b->bl.one)) ||
      (rc = batch_alloc(b, &b->d_pairs, (size_t)S)) ||
      (rc = make_map(b, p2p_share(b, kAdjP2PWaves, cus * kAdjP2PBatchBlocksPerCU), &b->bl.adj_p2p)) ||
      (rc = batch_use_p2p(b, b->p2p_variant, b->p2p_bpc))) {
    batch_free(b);
    delete b;
    return rc;
  }
  *out = b;
  return NBFMM_OK;
}

int nbfmm_batch_destroy(nbfmm_batch* b) {
  if (!b) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null batch");
  hipDeviceSynchronize();
  batch_free(b);
  delete b;
  return NBFMM_OK;
}

int nbfmm_batch_info(const nbfmm_batch* b, int* num_systems, int* total_particles) {
  if (!b) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null batch");
  if (num_systems) *num_systems = b->S;
  if (total_particles) *total_particles = b->total_n;
  return NBFMM_OK;
}

int nbfmm_batch_offsets(const nbfmm_batch* b, int* part_off) {
  if (!b || !part_off) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null argument");
  std::memcpy(part_off, b->part_off.data(), (b->S + 1) * sizeof(int));
  return NBFMM_OK;
}

int nbfmm_batch_set_ticks(nbfmm_batch* b, const float* ticks) {
  if (!b || !ticks) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null argument");
  for (int s = 0; s < b->S; ++s) b->hv[s].tick = b->hdesc[s].tick = ticks[s];
  NB_HIP(hipDeviceSynchronize());
  NB_HIP(hipMemcpy(b->ddesc, b->hdesc.data(), b->S * sizeof(SysDesc), hipMemcpyHostToDevice));
  return NBFMM_OK;
}

int nbfmm_batch_solve(nbfmm_batch* b, const float* pos0, const float* w0, float* eff0, void* stream) {
  if (!b || !pos0 || !w0 || !eff0) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null argument");
  return batch_pipeline(b, kP2PFinal, call_io(pos0, w0, eff0), as_stream(stream));
}

int nbfmm_batch_step(nbfmm_batch* b, float* xcur, float* xpre, const float* w, float* acc, void* stream) {
  if (!b || !xcur || !xpre || !w || !acc) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null argument");
  return batch_pipeline(b, kP2PStep, call_io(xcur, w, acc, xcur, xpre), as_stream(stream));
}

int nbfmm_batch_run(nbfmm_batch* b, int steps, float* xcur, float* xpre, const float* w, float* acc, int use_graph,
                    void* stream) {
  if (!b || !xcur || !xpre || !w || !acc) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null argument");
  if (steps < 0) return fail(NBFMM_ERR_INVALID_ARGUMENT, "num_steps < 0");
  hipStream_t st = as_stream(stream);
  int rc = NBFMM_OK;
  if (!use_graph) {
    // eager: the launches of the replayed steps, one by one (with the warm
    // start on, every step but the last makes the next sort's first pass)
    CallIO io = call_io(xcur, w, acc, xcur, xpre);
    const bool chain = b->warm_enabled;
    for (int k = 0; k < steps; ++k) {
      const int mode = chain && k + 1 < steps ? kP2PStepSort : kP2PStep;
      if ((rc = batch_pipeline(b, mode, io, st, chain && k > 0))) return rc;
    }
    return NBFMM_OK;
  }
  if (steps == 0) return NBFMM_OK;
  if (!st) {
    NB_HIP(b->side.begin());
    rc = nbfmm_batch_run(b, steps, xcur, xpre, w, acc, use_graph, (void*)b->side.s);
    const hipError_t e = b->side.end();
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "side stream");
    return NBFMM_OK;
  }
  // As nbfmm_run: the first step is eager; with the warm start on, the P2P
  // of every step but the last makes the next step's first sort pass.
  CallIO io = call_io(xcur, w, acc, xcur, xpre);
  const bool chain = b->warm_enabled;  // after the first step's sort every system has a sort
  if ((rc = batch_pipeline(b, chain && steps > 1 ? kP2PStepSort : kP2PStep, io, st))) return rc;
  if (--steps == 0) return NBFMM_OK;
  auto& g = b->gkey;
  const bool same = g.valid && b->graph_exec[1] && g.xcur == xcur && g.xpre == xpre && g.w == w && g.acc == acc &&
                    g.stream == st && g.warm == chain;
  if (!same) {
    drop_batch_graph(b);
    const int mid = chain ? kP2PStepSort : kP2PStep;
    auto one = [&](int mode) { return [&, mode]() { return batch_pipeline(b, mode, io, st, chain); }; };
    if (chain && (rc = capture_graph(st, 1, one(kP2PStepSort), &b->graph[0], &b->graph_exec[0]))) return rc;
    if ((rc = capture_graph(st, 1, one(kP2PStep), &b->graph[1], &b->graph_exec[1]))) return rc;
    if ((rc = capture_graph(st, kRunGraphSteps, one(mid), &b->graph[2], &b->graph_exec[2]))) return rc;
    g.xcur = xcur; g.xpre = xpre; g.w = w; g.acc = acc; g.stream = st; g.valid = true; g.warm = chain;
  }
  return replay_run(b->graph_exec, chain, steps, st);
}

int nbfmm_batch_set_warm_start(nbfmm_batch* b, int enable) {
  if (!b) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null batch");
  b->warm_enabled = enable != 0;
  drop_batch_graph(b);
  return NBFMM_OK;
}

int nbfmm_batch_solve_adjoint(nbfmm_batch* b, const float* pos0, const float* w0, const float* grad_eff,
                              float* grad_pos, float* grad_w, void* stream) {
  if (!b || !pos0 || !w0 || !grad_eff || !grad_pos || !grad_w) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null argument");
  hipStream_t st = as_stream(stream);
  CallIO io = call_io(pos0, w0);
  io.gin = (const float2*)grad_eff;
  io.gpos = (float2*)grad_pos;
  io.gw = grad_w;
  io.gscale = 1.0f;
  if (int rc = batch_sort(b, io, st)) return rc;
  if (int rc = batch_balance_once(b, st)) return rc;
  NB_HIP(launch_tree_batch(b->bl, st, false));
  NB_HIP(launch_adjoint_batch(b->bl, io, 0, st));
  return NBFMM_OK;
}

int nbfmm_batch_step_adjoint(nbfmm_batch* b, const float* x, const float* w, float* lam, float* pi, float* grad_w,
                             void* stream) {
  if (!b || !x || !w || !lam || !pi || !grad_w) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null argument");
  hipStream_t st = as_stream(stream);
  CallIO io = call_io(x, w);
  io.gin = (const float2*)lam;
  io.gpos = (float2*)lam;
  io.gpre = (float2*)pi;
  io.gw = grad_w;
  io.gscale = -1.0f;  // each system's tick^2
  if (int rc = batch_sort(b, io, st)) return rc;
  if (int rc = batch_balance_once(b, st)) return rc;
  NB_HIP(launch_tree_batch(b->bl, st, false));
  NB_HIP(launch_adjoint_batch(b->bl, io, 1, st));
  return NBFMM_OK;
}

namespace {

constexpr int kMaxKeptSorts = 4096;  // slots 0..4096 of nbfmm_batch_keep_sort

// The batch's current sort and tree into slot `slot` (allocated on first use).
int batch_keep(nbfmm_batch* b, int slot, hipStream_t st) {
  while ((int)b->slots.size() <= slot) {
    nbfmm_batch::SortSlot sl{};
    int rc;
    if ((rc = batch_alloc(b, &sl.part, b->n_all)) || (rc = batch_alloc(b, &sl.key, b->n_all)) ||
        (rc = batch_alloc(b, &sl.head, b->head_all)) || (rc = batch_alloc(b, &sl.occ, b->cell_all)) ||
        (rc = batch_alloc(b, &sl.nocc, (size_t)b->S)) || (rc = batch_alloc(b, &sl.mpole, b->tree_all)))
      return rc;
    b->slots.push_back(sl);
  }
  const nbfmm_batch::SortSlot& sl = b->slots[slot];
  const SysView& v = b->hv[0];  // system 0's view: every array's base
  NB_HIP(launch_copy_regions(sort_copy(sl, v, b->n_all, b->head_all, b->cell_all, (size_t)b->S, b->tree_all),
                             kCopyRegions, st));
  return NBFMM_OK;
}

}  // namespace

int nbfmm_batch_keep_sort(nbfmm_batch* b, int slot, void* stream) {
  if (!b) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null batch");
  if (slot < 0 || slot > kMaxKeptSorts) return fail(NBFMM_ERR_INVALID_ARGUMENT, "slot in [0, 4096]");
  if (!b->warm_valid) return fail(NBFMM_ERR_INVALID_ARGUMENT, "no sort to keep: step or solve first");
  return batch_keep(b, slot, as_stream(stream));
}

int nbfmm_batch_run_keep(nbfmm_batch* b, int steps, int first_slot, float* xcur, float* xpre, const float* w,
                         float* acc, void* stream) {
  if (!b || !xcur || !xpre || !w || !acc) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null argument");
  if (steps < 0 || first_slot < 0 || first_slot + steps > kMaxKeptSorts)
    return fail(NBFMM_ERR_INVALID_ARGUMENT, "kept slots first_slot .. first_slot + num_steps in [0, 4096]");
  hipStream_t st = as_stream(stream);
  // nbfmm_batch_run's eager steps, each kept once its sort and tree are built
  // (the P2P of a step makes only the next sort's first pass, which leaves
  // the kept arrays alone; the last step's is the final positions')
  CallIO io = call_io(xcur, w, acc, xcur, xpre);
  const bool chain = b->warm_enabled;
  for (int k = 0; k < steps; ++k) {
    if (int rc = batch_pipeline(b, chain ? kP2PStepSort : kP2PStep, io, st, chain && k > 0)) return rc;
    if (int rc = batch_keep(b, first_slot + k, st)) return rc;
  }
  // the final positions: their sort and tree, as nbfmm_batch_step_adjoint builds them
  if (int rc = batch_sort(b, io, st, chain && steps > 0)) return rc;
  if (int rc = batch_balance_once(b, st)) return rc;
  NB_HIP(launch_tree_batch(b->bl, st, false));
  return batch_keep(b, first_slot + steps, st);
}

int nbfmm_batch_step_adjoint_kept(nbfmm_batch* b, int slot, float* lam, float* pi, float* grad_w, void* stream) {
  if (!b || !lam || !pi || !grad_w) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null argument");
  if (slot < 0 || slot >= (int)b->slots.size()) return fail(NBFMM_ERR_INVALID_ARGUMENT, "no sort kept in this slot");
  hipStream_t st = as_stream(stream);
  const nbfmm_batch::SortSlot& sl = b->slots[slot];
  // the adjoint's launch maps with every system's sorted arrays in the slot
  BatchLaunch bl = b->bl;
  for (KernelMap* km : {&bl.adj_sort, &bl.adj_up, &bl.adj_m2l, &bl.adj_p2p}) {
    km->sys.base.part = sl.part;
    km->sys.base.key = sl.key;
    km->sys.base.head = sl.head;
    km->sys.base.occ = sl.occ;
    km->sys.base.nocc = sl.nocc;
    km->sys.base.mpole = sl.mpole;
  }
  CallIO io;
  io.gin = (const float2*)lam;
  io.gpos = (float2*)lam;
  io.gpre = (float2*)pi;
  io.gw = grad_w;
  io.gscale = -1.0f;  // each system's tick^2
  NB_HIP(launch_adjoint_batch(bl, io, 1, st));
  return NBFMM_OK;
}

int nbfmm_batch_set_p2p_config(nbfmm_batch* b, int variant, int blocks_per_cu) {
  if (!b) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null batch");
  if (variant < 0 || variant >= p2p_num_variants()) return fail(NBFMM_ERR_INVALID_ARGUMENT, "unknown P2P variant");
  if (blocks_per_cu < 1 || blocks_per_cu > 64) return fail(NBFMM_ERR_INVALID_ARGUMENT, "blocks_per_cu in [1, 64]");
  const int rc = batch_use_p2p(b, variant, blocks_per_cu);
  if (rc) return rc;
  drop_batch_graph(b);
  return NBFMM_OK;
}

int nbfmm_batch_set_p2p_wait_polls(nbfmm_batch* b, int polls) {
  if (!b) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null batch");
  if (polls < 1) return fail(NBFMM_ERR_INVALID_ARGUMENT, "polls must be >= 1");
  // the views every launch map carries (maps made later copy hv[0])
  for (auto& v : b->hv) v.wait_polls = polls;
  for (auto& kv : b->p2p_maps) kv.second.sys.base.wait_polls = polls;
  b->bl.p2p.sys.base.wait_polls = polls;
  drop_batch_graph(b);
  return NBFMM_OK;
}

int nbfmm_batch_balance(nbfmm_batch* b, const float* pos0, const float* w0, void* stream) {
  if (!b || !pos0 || !w0) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null argument");
  hipStream_t st = as_stream(stream);
  if (int rc = batch_sort(b, call_io(pos0, w0), st)) return rc;
  return batch_balance(b, st);
}

int nbfmm_batch_get_p2p_config(const nbfmm_batch* b, int* variant, int* blocks_per_cu) {
  if (!b) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null batch");
  if (variant) *variant = b->p2p_variant;
  if (blocks_per_cu) *blocks_per_cu = b->p2p_bpc;
  return NBFMM_OK;
}

int nbfmm_batch_flags(nbfmm_batch* b, unsigned* flags, int clear) {
  if (!b || !flags) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null argument");
  NB_HIP(hipDeviceSynchronize());
  for (int s = 0; s < b->S; ++s) NB_HIP(hipMemcpy(flags + s, b->hv[s].flags, 4, hipMemcpyDeviceToHost));
  if (clear)
    for (int s = 0; s < b->S; ++s) NB_HIP(hipMemset(b->hv[s].flags, 0, 4));
  return NBFMM_OK;
}

}  // extern "C"
