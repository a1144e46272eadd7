// C ABI of nbfmm_amd (include/nbfmm_amd.h): solver context, stage dispatch,
// buffer access in the reference layouts, fused step and graph replay.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
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

}  // namespace

struct nbfmm_solver {
  int L = 0, B = 0, max_n = 0;
  float limits[4] = {0, 0, 0, 0};
  float cw = 0, ch = 0, eps = 0;
  int device = 0;
  int p2p_grid = 0;
  SysView v{};  // arrays and sizes; v.n, v.pos0, v.w0 are set per call
  std::vector<void*> allocs;
  // profiling of the P2P launch
  bool profile = false;
  std::vector<hipEvent_t> ev;  // pairs (start, stop)
  size_t ev_used = 0;
  double p2p_ms = 0.0;
  int p2p_launches = 0;
  unsigned long long* d_pairs = nullptr;
  // graph of one step
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  struct {
    int n = -1;
    float tick = 0;
    const void *xcur = nullptr, *xpre = nullptr, *w = nullptr, *acc = nullptr;
    hipStream_t stream = nullptr;
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
  if (s->graph_exec) hipGraphExecDestroy(s->graph_exec);
  if (s->graph) hipGraphDestroy(s->graph);
  for (auto& e : s->ev) hipEventDestroy(e);
  for (void* p : s->allocs) hipFree(p);
  s->allocs.clear();
}

SysView view(nbfmm_solver* s, int n, const float* pos0, const float* w0) {
  SysView v = s->v;
  v.n = n;
  v.pos0 = (const float2*)pos0;
  v.w0 = w0;
  return v;
}

int check_n(nbfmm_solver* s, int n) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  if (n < 0) return fail(NBFMM_ERR_INVALID_ARGUMENT, "num_particle < 0");
  if (n > s->max_n) return fail(NBFMM_ERR_CAPACITY, "num_particle exceeds max_num_particle");
  return NBFMM_OK;
}

hipStream_t as_stream(void* p) { return (hipStream_t)p; }

int p2p_timed(nbfmm_solver* s, const SysView& v, int mode, const StepArgs& a, hipStream_t st) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s->profile) {
    if (s->ev_used + 2 > s->ev.size()) {
      for (int k = 0; k < 64; ++k) {
        hipEvent_t e;
        NB_HIP(hipEventCreate(&e));
        s->ev.push_back(e);
      }
    }
    e0 = s->ev[s->ev_used];
    e1 = s->ev[s->ev_used + 1];
    s->ev_used += 2;
    NB_HIP(hipEventRecord(e0, st));
  }
  NB_HIP(launch_p2p(v, mode, a, s->p2p_grid, st));
  if (s->profile) NB_HIP(hipEventRecord(e1, st));
  return NBFMM_OK;
}

// The fused pipeline shared by solve and step.
int run_pipeline(nbfmm_solver* s, int n, const float* pos0, const float* w0, int mode, const StepArgs& a,
                 hipStream_t st) {
  const SysView v = view(s, n, pos0, w0);
  NB_HIP(launch_predo(v, st));
  if (n == 0) return NBFMM_OK;
  NB_HIP(launch_p2m(v, st));
  NB_HIP(launch_m2m(v, st));
  NB_HIP(launch_m2l(v, st));
  NB_HIP(launch_l2l(v, st));
  return p2p_timed(s, v, mode, a, st);
}

}  // namespace

extern "C" {

int nbfmm_version(void) { return 10000; }

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
    v.lvl_off[l] = off;
    if (l < levels) off += (s->B >> l) * (s->B >> l);
  }
  const int ncells_tree = off;
  const int ntiles = cdiv(max_num_particle, kSortTile);
  const size_t N = (size_t)max_num_particle;
  int rc = NBFMM_OK;
#define ALLOC(ptr, cnt) \
  if ((rc = dev_alloc(s, &(ptr), (cnt))) != NBFMM_OK) { free_all(s); delete s; return rc; }
  ALLOC(v.key0, N);
  ALLOC(v.keyA, N);
  ALLOC(v.valA, N);
  ALLOC(v.key, N);
  ALLOC(v.part, N);
  ALLOC(v.eff, N);
  ALLOC(v.count, (size_t)C);
  ALLOC(v.head, (size_t)C + 1);
  ALLOC(v.occ, (size_t)C);
  ALLOC(v.nocc, 1);
  ALLOC(v.part_cnt, (size_t)scan_blocks(s->B));
  ALLOC(v.part_occ, (size_t)scan_blocks(s->B));
  ALLOC(v.hist, (size_t)s->B * ntiles);
  ALLOC(v.dtotal, (size_t)s->B);
  ALLOC(v.mpole, (size_t)ncells_tree);
  ALLOC(v.local, (size_t)ncells_tree);
  ALLOC(v.flags, 1);
  ALLOC(s->d_pairs, 1);
#undef ALLOC
  // P2P: grid-stride over occupied leaves, enough waves to fill every CU
  const int max_occ = C < max_num_particle ? C : max_num_particle;
  const int want = cdiv(max_occ, kP2PThreads / kWave);
  s->p2p_grid = want < cus * 16 ? want : cus * 16;
  if (s->p2p_grid < 1) s->p2p_grid = 1;
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
  StepArgs a{};
  a.acc0 = (float2*)eff0;
  return run_pipeline(s, n, pos0, w0, kP2PFinal, a, as_stream(stream));
}

int nbfmm_predo(nbfmm_solver* s, int n, const float* pos0, const float* w0, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  if (n > 0 && (!pos0 || !w0)) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null array");
  NB_HIP(launch_predo(view(s, n, pos0, w0), as_stream(stream)));
  return NBFMM_OK;
}

int nbfmm_p2p(nbfmm_solver* s, int n, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  StepArgs a{};
  return p2p_timed(s, view(s, n, nullptr, nullptr), kP2POnly, a, as_stream(stream));
}

int nbfmm_p2m(nbfmm_solver* s, int n, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  NB_HIP(launch_p2m(view(s, n, nullptr, nullptr), as_stream(stream)));
  return NBFMM_OK;
}

int nbfmm_m2m(nbfmm_solver* s, void* stream) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  NB_HIP(launch_m2m(view(s, 0, nullptr, nullptr), as_stream(stream)));
  return NBFMM_OK;
}

int nbfmm_m2l(nbfmm_solver* s, void* stream) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  NB_HIP(launch_m2l(view(s, 0, nullptr, nullptr), as_stream(stream)));
  return NBFMM_OK;
}

int nbfmm_l2l(nbfmm_solver* s, void* stream) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  NB_HIP(launch_l2l(view(s, 0, nullptr, nullptr), as_stream(stream)));
  return NBFMM_OK;
}

int nbfmm_l2p(nbfmm_solver* s, int n, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  NB_HIP(launch_l2p(view(s, n, nullptr, nullptr), as_stream(stream)));
  return NBFMM_OK;
}

int nbfmm_postdo(nbfmm_solver* s, int n, float* eff0, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  if (n > 0 && !eff0) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null array");
  NB_HIP(launch_postdo(view(s, n, nullptr, nullptr), (float2*)eff0, as_stream(stream)));
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
  const SysView v = view(s, count, nullptr, nullptr);
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
  return buffer_io(s, id, count, const_cast<void*>(src), true, as_stream(stream));
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
  StepArgs a{};
  a.acc0 = (float2*)acc;
  a.xcur = (float2*)xcur;
  a.xpre = (float2*)xpre;
  a.tick = tick;
  return run_pipeline(s, n, xcur, w, kP2PStep, a, as_stream(stream));
}

int nbfmm_run(nbfmm_solver* s, int n, float tick, int steps, float* xcur, float* xpre, const float* w, float* acc,
              int use_graph, void* stream) {
  int rc = check_n(s, n);
  if (rc) return rc;
  if (steps < 0) return fail(NBFMM_ERR_INVALID_ARGUMENT, "num_steps < 0");
  hipStream_t st = as_stream(stream);
  if (!use_graph || s->profile) {
    for (int k = 0; k < steps; ++k) {
      rc = nbfmm_step(s, n, tick, xcur, xpre, w, acc, stream);
      if (rc) return rc;
    }
    return NBFMM_OK;
  }
  if (!st) return fail(NBFMM_ERR_INVALID_ARGUMENT, "graph replay needs a non-null stream");
  auto& g = s->gkey;
  const bool same = s->graph_exec && g.n == n && g.tick == tick && g.xcur == xcur && g.xpre == xpre &&
                    g.w == w && g.acc == acc && g.stream == st;
  if (!same) {
    if (s->graph_exec) { hipGraphExecDestroy(s->graph_exec); s->graph_exec = nullptr; }
    if (s->graph) { hipGraphDestroy(s->graph); s->graph = nullptr; }
    NB_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    rc = nbfmm_step(s, n, tick, xcur, xpre, w, acc, stream);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) { if (graph) hipGraphDestroy(graph); return rc; }
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    s->graph = graph;
    NB_HIP(hipGraphInstantiate(&s->graph_exec, graph, nullptr, nullptr, 0));
    g.n = n; g.tick = tick; g.xcur = xcur; g.xpre = xpre; g.w = w; g.acc = acc; g.stream = st;
  }
  for (int k = 0; k < steps; ++k) NB_HIP(hipGraphLaunch(s->graph_exec, st));
  return NBFMM_OK;
}

int nbfmm_profile_enable(nbfmm_solver* s, int enable) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  s->profile = enable != 0;
  return NBFMM_OK;
}

int nbfmm_p2p_time(nbfmm_solver* s, double* total_ms, int* launches, int reset) {
  if (!s) return fail(NBFMM_ERR_INVALID_ARGUMENT, "null solver");
  for (size_t k = 0; k + 1 < s->ev_used; k += 2) {
    NB_HIP(hipEventSynchronize(s->ev[k + 1]));
    float ms = 0.0f;
    NB_HIP(hipEventElapsedTime(&ms, s->ev[k], s->ev[k + 1]));
    s->p2p_ms += ms;
    s->p2p_launches += 1;
  }
  s->ev_used = 0;
  if (total_ms) *total_ms = s->p2p_ms;
  if (launches) *launches = s->p2p_launches;
  if (reset) { s->p2p_ms = 0.0; s->p2p_launches = 0; }
  return NBFMM_OK;
}

int nbfmm_pair_count(nbfmm_solver* s, int n, double* pairs, int* occupied) {
  int rc = check_n(s, n);
  if (rc) return rc;
  NB_HIP(hipDeviceSynchronize());
  NB_HIP(hipMemset(s->d_pairs, 0, 8));
  NB_HIP(launch_pair_count(view(s, n, nullptr, nullptr), s->d_pairs, nullptr));
  unsigned long long h = 0;
  int occ = 0;
  NB_HIP(hipMemcpy(&h, s->d_pairs, 8, hipMemcpyDeviceToHost));
  NB_HIP(hipMemcpy(&occ, s->v.nocc, 4, hipMemcpyDeviceToHost));
  if (pairs) *pairs = (double)h;
  if (occupied) *occupied = occ;
  return NBFMM_OK;
}

}  // extern "C"
