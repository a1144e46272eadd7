// Near field (P2P) with the fused L2P, un-permute and Verlet epilogue, plus
// the per-particle stage kernels of the buffer API.
//
// Reference: source/nbfmm/core/solver/p2p.cu:30-176 (p2pNaive / p2pBlock),
// l2p.cu:23-48, postdo.cu:22-45, source/nbfmm/display/stars/update.cu:26-42.
//
// P2P schedule: the sorted particles of every row of leaves are cut into
// runs of 64, one wavefront per run and one target per lane, so a wave's
// lanes are busy whatever the leaf sizes.  The sources of a target are three
// contiguous ranges of the sorted particles, one per neighbour row
// [head[x-1 + yB], head[x+1 + yB + 1]) (the row-range idea of
// p2p.cu:105-118); a run stages the union of its targets' ranges in LDS and
// each lane sums its own sources.  In the final modes the wave also forms
// each target's leaf local (L2L folded in), adds it (L2P) and scatters to the
// original order (postdo), and in the step mode applies the Verlet update.
#include "common.hpp"
#include "launch.hpp"
#include "../../include/nbfmm_amd.h"

namespace nbfmm_amd {

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// Leaf local of cell (cx, cy): the L2L recursion (l2l.cu:21-36) unrolled
// from the top, eff_0 = m2l_0 + (m2l_1 + (... + m2l_{L-1})), which gives
// the bits of the per-level in-place L2L.  Wave-uniform loads.
__device__ __forceinline__ float2 leaf_local(const SysView& v, int cx, int cy) {
#pragma clang fp contract(off)
  int l = v.L - 1;
  float2 acc = v.local[v.lvl_off(l) + (cy >> l) * (v.B >> l) + (cx >> l)];
  for (l = v.L - 2; l >= 0; --l) {
    const float2 e = v.local[v.lvl_off(l) + (cy >> l) * (v.B >> l) + (cx >> l)];
    acc.x = e.x + acc.x;
    acc.y = e.y + acc.y;
  }
  return acc;
}


// Geometry of run j: its targets (one per lane) and, per neighbour row
// q = 0, 1, 2 (rows y-1, y, y+1), the run's union of source ranges
// [us, us + ulen) in the sorted particles and this lane's part [a, b) of it
// (relative to us; empty for idle lanes and rows outside the grid).
struct RunGeom {
  int y, i, xt;
  bool valid;
  float4 pt;
  int us[3], ulen[3];  // wave-uniform
  int a[3], b[3];      // this lane's range in each row's union (relative to us)
  int total;           // sum of ulen: the union's size
};

__device__ __forceinline__ void run_geom(const SysView& v, const int* s_runs, int j, int lane, RunGeom& g) {
  const int B = v.B;
  // everything but the lane's own target is wave-uniform: scalar registers
  g.y = __builtin_amdgcn_readfirstlane(run_row(s_runs, B, j));
  const int T0 = __builtin_amdgcn_readfirstlane(v.head[g.y * B] + kWave * (j - s_runs[g.y]));
  const int nt = __builtin_amdgcn_readfirstlane(min(kWave, v.head[(g.y + 1) * B] - T0));
  g.valid = lane < nt;
  g.i = T0 + (g.valid ? lane : 0);
  g.pt = v.part[g.i];
  g.xt = (int)(v.key[g.i] & (uint32_t)(B - 1));
  const int xf = __builtin_amdgcn_readfirstlane(__shfl(g.xt, 0, kWave));
  const int xl = __builtin_amdgcn_readfirstlane(__shfl(g.xt, nt - 1, kWave));
  g.total = 0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int r = g.y + q - 1;
    if (r < 0 || r >= B) {
      g.us[q] = g.ulen[q] = g.a[q] = g.b[q] = 0;
      continue;
    }
    const int* hr = v.head + r * B;
    g.us[q] = __builtin_amdgcn_readfirstlane(hr[xf > 0 ? xf - 1 : 0]);
    g.ulen[q] = __builtin_amdgcn_readfirstlane(hr[(xl < B - 1 ? xl + 1 : B - 1) + 1]) - g.us[q];
    g.a[q] = g.valid ? hr[g.xt > 0 ? g.xt - 1 : 0] - g.us[q] : 0;
    g.b[q] = g.valid ? hr[(g.xt < B - 1 ? g.xt + 1 : B - 1) + 1] - g.us[q] : 0;
    g.total += g.ulen[q];
  }
}

// Segments of a union of `total` sources: K = 1 up to kP2PSplitMin, else
// min(kP2PMaxSeg, ceil(total / kP2PSegLen)) segments of S sources (a
// multiple of 64), the last one shorter.
__device__ __forceinline__ int p2p_segments(int total) {
  if (total <= kP2PSplitMin) return 1;
  return min(kP2PMaxSeg, (total + kP2PSegLen - 1) / kP2PSegLen);
}
__device__ __forceinline__ int p2p_segment_len(int total, int K) { return (((total + K - 1) / K) + 63) & ~63; }

// Sum of this lane's sources among union positions [sa, sb) (rows in order,
// then sorted order), staged in the wave's LDS buffer STAGE sources at a time.
template <int STAGE, int U>
__device__ __forceinline__ void seg_sum(const SysView& v, float4* src, const RunGeom& g, int lane, int sa, int sb,
                                        float eps, float& ax, float& ay) {
  ax = 0.0f;
  ay = 0.0f;
  const int B = v.B;
  int off = 0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int wb = max(sa - off, 0);
    const int we = min(sb - off, g.ulen[q]);
    const int la = g.a[q], lb = g.b[q];
    for (int w0 = wb; w0 < we; w0 += STAGE) {
      const int cnt = min(STAGE, we - w0);
      // the chunk goes from global memory straight into LDS (each lane's
      // 16 bytes to slot k * 64 + lane), all loads in flight at once and no
      // registers held; indices are clamped to the chunk's last source
      // (cnt >= 1), and the slots from cnt on, which hold copies of it, are
      // never read.  The wait covers every counter: the chunk is in LDS.
      {
        const float4* const gsrc = v.part + g.us[q] + w0;
#pragma unroll
        for (int k = 0; k < STAGE / kWave; ++k)
          __builtin_amdgcn_global_load_lds(
              (const __attribute__((address_space(1))) void*)(gsrc + min(k * kWave + lane, cnt - 1)),
              (__attribute__((address_space(3))) void*)(src + k * kWave), 16, 0, 0);
        __builtin_amdgcn_s_waitcnt(0);
      }
      wave_lds_order();
      const int lo = min(max(la - w0, 0), cnt);
      const int hi = min(max(lb - w0, 0), cnt);
      // loop control on the LDS address alone: one compare and one add
      // per U sources (pfull >= p, so the unsigned address compare never
      // sees a bound below the buffer)
      const float4* p = src + lo;
      const float4* const pend = src + hi;
      const float4* const pfull = src + max(hi - (U - 1), lo);
      for (; p < pfull; p += U) {
        float4 sv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) sv[u] = p[u];
#pragma unroll
        for (int u = 0; u < U; ++u) kernel_accum(g.pt.x, g.pt.y, sv[u].x, sv[u].y, sv[u].z, eps, ax, ay);
      }
      for (; p < pend; ++p) {
        const float4 s1 = *p;
        kernel_accum(g.pt.x, g.pt.y, s1.x, s1.y, s1.z, eps, ax, ay);
      }
      wave_lds_order();
    }
    off += g.ulen[q];
  }
}

// The targets' epilogue.  kP2POnly stores the sorted effect; kP2PFinal adds
// the leaf local (L2L folded in, L2P) and scatters to the original order
// (postdo); kP2PStep also applies the Verlet update.
template <int MODE>
__device__ __forceinline__ void p2p_finish(const SysView& v, const CallIO& io, const RunGeom& g, float ax, float ay) {
#pragma clang fp contract(off)
  if (!g.valid) return;
  if (MODE == kP2POnly) {
    v.eff[g.i] = make_float2(ax, ay);
    return;
  }
  // l2p.cu:34: effect += cell effect; postdo.cu:32: scatter by perm
  const float2 loc = v.L > 0 ? leaf_local(v, g.xt, g.y) : make_float2(0.0f, 0.0f);
  const float ex = ax + loc.x;
  const float ey = ay + loc.y;
  const uint32_t o = __float_as_uint(g.pt.w);
  io.acc0[v.part_off + o] = make_float2(ex, ey);
  if (MODE == kP2PStep) {
    // update.cu:39-41
    float2* xcur = io.xcur + v.part_off;
    float2* xpre = io.xpre + v.part_off;
    const float2 x = xcur[o];
    const float2 xp = xpre[o];
    float2 nx;
    nx.x = 2.0f * x.x - xp.x + ex * v.tick * v.tick;
    nx.y = 2.0f * x.y - xp.y + ey * v.tick * v.tick;
    xcur[o] = nx;
    xpre[o] = x;
    if (io.xs_out) io.xs_out[v.part_off + g.i] = nx;
  }
}

// One wavefront works one run of up to 64 targets of row y, one target per
// lane.  Target t in leaf x_t needs, in each neighbour row r, the sorted
// range [head[r, x_t - 1], head[r, x_t + 2]) (the row-range idea of
// p2p.cu:105-118).  Those ranges lie inside the run's union
// [head[r, x_first - 1], head[r, x_last + 2]), which the wave stages in its
// LDS buffer STAGE sources at a time; each lane then loops over its own part
// of the window, U sources per iteration.  A lane adds its sources in sorted
// order, rows y-1, y, y+1, whatever STAGE and U.
//
// A run whose union exceeds kP2PSplitMin (a cluster of dense leaves) would
// be one long serial chain on one wave, so it is split into K segments of
// the union.  The wave that takes it publishes it in the block's slot; every
// wave of the block that is free takes a segment (a ticket from an LDS
// counter, generation and K in its high bits), sums it, waits until the
// segments before it have been added into the slot's running sum, adds its
// own and passes on; the last one runs the epilogue and frees the slot.
// The wait always ends (tickets go out in order and segment 0 waits for
// nothing; a poll limit raises NBFMM_FLAG_P2P_WAIT instead of hanging).  A
// wave whose heavy run finds the slot taken keeps it pending, helps with the
// published run and publishes its own once the slot is free (after a poll
// limit it sums all K segments itself, in the same order).  All of this
// state is in LDS, and the hardware executes each wave's LDS instructions in
// program order, so the atomics are relaxed and only the compiler is kept
// from reordering (wave_lds_order); no wave reads another's global writes
// here, so no wave waits for its global stores.  A wave that finds
// the run queue empty stays to help while another wave holds a run that may
// still be published (the low bits of the run queue count runs from their
// take until they are known light, published, or summed alone) or a split
// run is published.
// Every variant (stage, unroll, waves) gives the same bits.
template <int MODE, class Sys, int STAGE, int U, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void k_p2p(Sys sys, CallIO io) {
  static_assert(STAGE % kWave == 0, "staging size must be a multiple of 64");
  __shared__ float4 s_src[WAVES][STAGE];
  __shared__ int s_runs[(1 << (kMaxLevel + 1)) + 1];
  __shared__ uint32_t sh[WAVES];
  // the block's run queue: runs taken (bits 31..4) and runs taken but not
  // yet known light, published or summed (bits 3..0, at most one per wave),
  // bumped together so that taking a run costs one LDS atomic
  __shared__ uint32_t s_queue;
  // the block's slot for a split run
  __shared__ float2 s_acc[kWave];  // running sum of the segments added so far
  __shared__ uint32_t s_ticket;    // generation << 22 | K << 16 | next segment
  __shared__ int s_open, s_busy, s_done, s_run;
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int lane = lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  float4* src = s_src[wave];
  const int B = v.B;
  const float eps = v.eps;
  row_runs<WAVES * kWave>(v, s_runs, sh);
  if (threadIdx.x == 0) {
    s_queue = 0;
    s_ticket = 0;
    s_open = 0;
    s_busy = 0;
    s_done = 0;
    s_run = 0;
  }
  __syncthreads();
  const int total = s_runs[B];
  RunGeom g;
  bool exhausted = false;  // this wave found the block's run queue empty
  int idle_polls = 0;
  int pend = -1, pend_k = 1, pend_polls = 0;  // a split run waiting for the slot (wave-uniform)
  // publishes split run jr of kr segments if the slot is free (lane 0's
  // view, made wave-uniform); the run stops counting as active once published
  auto try_publish = [&](int jr, int kr) -> bool {
    int got = 0;
    if (lane == 0) {
      int expected = 0;
      if (__hip_atomic_compare_exchange_strong(&s_busy, &expected, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP)) {
        got = 1;
        s_run = jr;
        s_done = 0;
        const uint32_t gen = ((s_ticket >> 22) + 1u) & 0x3ffu;
        wave_lds_order();
        __hip_atomic_exchange(&s_ticket, (gen << 22) | ((uint32_t)kr << 16), __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_WORKGROUP);
        wave_lds_order();
        __hip_atomic_store(&s_open, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        wave_lds_order();
        __hip_atomic_fetch_add(&s_queue, 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    return __builtin_amdgcn_readfirstlane(got) != 0;
  };
  for (;;) {
    // what this iteration sums: segments [k0, k1) of run j out of K; a
    // ticket (one segment of the published run) has k1 = k0 + 1 <= K
    int j = -1, K = 1, k0 = 0, k1 = 1;
    bool ticket = false;
    // 1. a segment of the published split run, if any is left
    if (__hip_atomic_load(&s_open, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
      uint32_t t = 0;
      if (lane == 0) t = __hip_atomic_fetch_add(&s_ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      wave_lds_order();
      t = __builtin_amdgcn_readfirstlane(t);
      const int k = (int)(t & 0xffffu), Kt = (int)((t >> 16) & 0x3fu);
      if (k < Kt) {
        if (k == Kt - 1 && lane == 0) __hip_atomic_store(&s_open, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        ticket = true;
        j = __builtin_amdgcn_readfirstlane(s_run);
        K = Kt;
        k0 = k;
        k1 = k + 1;
        run_geom(v, s_runs, j, lane, g);
      }
    }
    if (!ticket) {
      if (pend >= 0) {
        // 2. this wave's split run waits for the slot, helping meanwhile;
        // after the poll limit it is summed here alone (same bits)
        if (try_publish(pend, pend_k)) {
          pend = -1;
          continue;
        }
        if (++pend_polls < kP2PWaitPolls) {
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        j = pend;
        K = pend_k;
        k1 = K;
        pend = -1;
        run_geom(v, s_runs, j, lane, g);
      } else {
        // 3. a new run; counted active from before it is taken until it is
        // finished or published, so that a wave that finds the queue empty
        // knows whether a split run may still be published
        if (exhausted) {
          // leave once no run is in progress and no split run is published;
          // new segments can appear only from a run in progress (a wave that
          // stops waiting only stops helping, so the poll limit is no error)
          wave_lds_order();
          if ((__hip_atomic_load(&s_queue, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & 0xfu) == 0 &&
              __hip_atomic_load(&s_busy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
            break;
          if (++idle_polls == kP2PWaitPolls) break;
          __builtin_amdgcn_s_sleep(2);
          continue;
        }
        uint32_t q = 0;
        if (lane == 0) q = __hip_atomic_fetch_add(&s_queue, 17u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        wave_lds_order();
        j = blk + (int)(__builtin_amdgcn_readfirstlane(q) >> 4) * nblk;
        if (j >= total) {
          if (lane == 0) __hip_atomic_fetch_add(&s_queue, 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          exhausted = true;
          continue;
        }
        run_geom(v, s_runs, j, lane, g);
        K = p2p_segments(g.total);
        k1 = K;
        // a light run can publish nothing: it stops counting at once, so
        // that waves without work need not stay for its end
        if (K == 1 && lane == 0)
          __hip_atomic_fetch_add(&s_queue, 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (K > 1) {
          if (!try_publish(j, K)) {
            pend = j;
            pend_k = K;
            pend_polls = 0;
          }
          continue;  // this wave takes segments like any other
        }
      }
    }
    // the segments [k0, k1), added in order
    const int S = p2p_segment_len(g.total, K);
    float ax = 0.0f, ay = 0.0f;
    for (int k = k0; k < k1; ++k) {
      float px, py;
      seg_sum<STAGE, U>(v, src, g, lane, k * S, min(k * S + S, g.total), eps, px, py);
      if (k == k0) {
        ax = px;
        ay = py;
      } else {
        ax = ax + px;
        ay = ay + py;
      }
    }
    if (!ticket) {
      p2p_finish<MODE>(v, io, g, ax, ay);
      // a split run summed here alone (poll limit) counted until now
      if (K > 1 && lane == 0)
        __hip_atomic_fetch_add(&s_queue, 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      continue;
    }
    // a ticket: wait for segments 0 .. k0-1, then add this one
    for (int polls = 0; __hip_atomic_load(&s_done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) != k0;
         ++polls) {
      if (polls == kP2PWaitPolls) {
        if (lane == 0) atomicOr(v.flags, kFlagP2PWait);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    wave_lds_order();
    if (k0 > 0) {
      const float2 o = s_acc[lane];
      ax = o.x + ax;
      ay = o.y + ay;
    }
    if (k0 < K - 1) {
      s_acc[lane] = make_float2(ax, ay);
      wave_lds_order();
      if (lane == 0) __hip_atomic_store(&s_done, k0 + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
      wave_lds_order();
      if (lane == 0) __hip_atomic_store(&s_busy, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      p2p_finish<MODE>(v, io, g, ax, ay);
    }
  }
}

// L2P stage (l2p.cu:23-35): effect[i] += leaf effect of its cell.
__global__ __launch_bounds__(256) void k_l2p(SysView v) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.n) return;
  const float2 c = v.local[v.key[i]];
  float2 e = v.eff[i];
  e.x = e.x + c.x;
  e.y = e.y + c.y;
  v.eff[i] = e;
}

// postdo (postdo.cu:22-33): effect_origin[perm[i]] = effect[i].
__global__ __launch_bounds__(256) void k_postdo(SysView v, float2* eff0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.n) return;
  eff0[__float_as_uint(v.part[i].w)] = v.eff[i];
}

// updateDevice (update.cu:26-42).
__global__ __launch_bounds__(256) void k_update(int n, float tick, float2* xcur, float2* xpre, const float2* acc) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 x = xcur[i];
  const float2 xp = xpre[i];
  const float2 a = acc[i];
  float2 nx;
  nx.x = 2.0f * x.x - xp.x + a.x * tick * tick;
  nx.y = 2.0f * x.y - xp.y + a.y * tick * tick;
  xcur[i] = nx;
  xpre[i] = x;
}

// Particle buffers in the reference layouts.
__global__ __launch_bounds__(256) void k_part_read(SysView v, int which, int count, void* dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  switch (which) {
    case NBFMM_BUF_POSITION: { const float4 p = v.part[i]; ((float2*)dst)[i] = make_float2(p.x, p.y); break; }
    case NBFMM_BUF_WEIGHT: ((float*)dst)[i] = v.part[i].z; break;
    case NBFMM_BUF_PERM: ((int*)dst)[i] = (int)__float_as_uint(v.part[i].w); break;
    case NBFMM_BUF_EFFECT: ((float2*)dst)[i] = v.eff[i]; break;
    case NBFMM_BUF_INDEX: {
      const uint32_t k = v.key[i];
      ((int2*)dst)[i] = make_int2((int)(k & (uint32_t)(v.B - 1)), (int)(k >> v.lgB));
      break;
    }
    default: break;
  }
}

__global__ __launch_bounds__(256) void k_part_write(SysView v, int which, int count, const void* src) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  switch (which) {
    case NBFMM_BUF_POSITION: { const float2 p = ((const float2*)src)[i]; v.part[i].x = p.x; v.part[i].y = p.y; break; }
    case NBFMM_BUF_WEIGHT: v.part[i].z = ((const float*)src)[i]; break;
    case NBFMM_BUF_PERM: v.part[i].w = __uint_as_float((uint32_t)((const int*)src)[i]); break;
    case NBFMM_BUF_EFFECT: v.eff[i] = ((const float2*)src)[i]; break;
    case NBFMM_BUF_INDEX: {
      const int2 x = ((const int2*)src)[i];
      v.key[i] = (uint32_t)(x.y * v.B + x.x);
      break;
    }
    default: break;
  }
}

// Ordered P2P pairs for the current head (self pairs excluded): the work
// figure of the P2P roofline.
__global__ __launch_bounds__(256) void k_pair_count(SysView v, unsigned long long* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long pairs = 0;
  if (c < v.B * v.B) {
    const int B = v.B;
    const int cx = c & (B - 1), cy = c >> v.lgB;
    const int nt = v.head[c + 1] - v.head[c];
    if (nt > 0) {
      const int xl = cx > 0 ? cx - 1 : 0;
      const int xr = cx < B - 1 ? cx + 1 : B - 1;
      long long src = 0;
      for (int y = cy - 1; y <= cy + 1; ++y) {
        if (y < 0 || y >= B) continue;
        src += v.head[xr + y * B + 1] - v.head[xl + y * B];
      }
      pairs = (unsigned long long)nt * (unsigned long long)(src - 1);
    }
  }
  // wave reduction then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) pairs += __shfl_xor(pairs, off, kWave);
  if (lane_id() == 0 && pairs) atomicAdd(out, pairs);
}

// P2P kernel variants (staging size, unroll, waves per block); the
// autotuner picks one.  All give identical results.
struct P2PVariant {
  int stage, unroll, waves;
};
static const P2PVariant kVariants[] = {{384, 4, 4}, {256, 4, 4}, {512, 4, 4}, {384, 8, 4},
                                       {256, 2, 4}, {384, 4, 8}, {320, 4, 4}};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

int p2p_num_variants() { return kNumVariants; }
void p2p_variant_info(int v, int* stage, int* unroll, int* waves) {
  *stage = kVariants[v].stage;
  *unroll = kVariants[v].unroll;
  *waves = kVariants[v].waves;
}

template <int MODE, class Sys>
static void launch_p2p_variant(int variant, dim3 g, const Sys& sys, const CallIO& io, hipStream_t s) {
  const dim3 b(kVariants[variant < 0 || variant >= kNumVariants ? 0 : variant].waves * kWave);
  switch (variant) {
    case 1: hipLaunchKernelGGL((k_p2p<MODE, Sys, 256, 4, 4>), g, b, 0, s, sys, io); break;
    case 2: hipLaunchKernelGGL((k_p2p<MODE, Sys, 512, 4, 4>), g, b, 0, s, sys, io); break;
    case 3: hipLaunchKernelGGL((k_p2p<MODE, Sys, 384, 8, 4>), g, b, 0, s, sys, io); break;
    case 4: hipLaunchKernelGGL((k_p2p<MODE, Sys, 256, 2, 4>), g, b, 0, s, sys, io); break;
    case 5: hipLaunchKernelGGL((k_p2p<MODE, Sys, 384, 4, 8>), g, b, 0, s, sys, io); break;
    case 6: hipLaunchKernelGGL((k_p2p<MODE, Sys, 320, 4, 4>), g, b, 0, s, sys, io); break;
    default: hipLaunchKernelGGL((k_p2p<MODE, Sys, 384, 4, 4>), g, b, 0, s, sys, io); break;
  }
}

hipError_t launch_p2p(const SysView& v, int mode, const CallIO& io, int grid_blocks, int variant, hipStream_t s) {
  if (v.n <= 0) return hipSuccess;
  const dim3 g(grid_blocks);
  const OneSys one{v};
  switch (mode) {
    case kP2POnly: launch_p2p_variant<kP2POnly, OneSys>(variant, g, one, io, s); break;
    case kP2PFinal: launch_p2p_variant<kP2PFinal, OneSys>(variant, g, one, io, s); break;
    default: launch_p2p_variant<kP2PStep, OneSys>(variant, g, one, io, s); break;
  }
  return hipGetLastError();
}

hipError_t launch_p2p_batch(const BatchLaunch& bl, int mode, const CallIO& io, int variant, hipStream_t s) {
  const dim3 g(bl.p2p.blocks);
  if (mode == kP2PFinal)
    launch_p2p_variant<kP2PFinal, ManySys>(variant, g, bl.p2p.sys, io, s);
  else
    launch_p2p_variant<kP2PStep, ManySys>(variant, g, bl.p2p.sys, io, s);
  return hipGetLastError();
}

hipError_t launch_l2p(const SysView& v, hipStream_t s) {
  if (v.L <= 0 || v.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_l2p, dim3(cdiv(v.n, 256)), dim3(256), 0, s, v);
  return hipGetLastError();
}

hipError_t launch_postdo(const SysView& v, float2* eff0, hipStream_t s) {
  if (v.n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_postdo, dim3(cdiv(v.n, 256)), dim3(256), 0, s, v, eff0);
  return hipGetLastError();
}

hipError_t launch_update(int n, float tick, float2* xcur, float2* xpre, const float2* acc, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_update, dim3(cdiv(n, 256)), dim3(256), 0, s, n, tick, xcur, xpre, acc);
  return hipGetLastError();
}

hipError_t launch_part_read(const SysView& v, int which, int count, void* dst, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_part_read, dim3(cdiv(count, 256)), dim3(256), 0, s, v, which, count, dst);
  return hipGetLastError();
}

hipError_t launch_part_write(const SysView& v, int which, int count, const void* src, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_part_write, dim3(cdiv(count, 256)), dim3(256), 0, s, v, which, count, src);
  return hipGetLastError();
}

hipError_t launch_pair_count(const SysView& v, unsigned long long* out, hipStream_t s) {
  hipLaunchKernelGGL(k_pair_count, dim3(cdiv(v.B * v.B, 256)), dim3(256), 0, s, v, out);
  return hipGetLastError();
}

}  // namespace nbfmm_amd
