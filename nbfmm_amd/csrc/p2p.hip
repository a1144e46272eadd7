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
#include "runs.hpp"
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

// L2P, un-permute and the Verlet update of one valid target (step modes):
// writes x_cur, x_pre and, with EFFECT, the effect in the caller's order;
// returns the new position.  A step followed by another step of the same call
// (kP2PStepSort) leaves the effect alone: nothing reads it before the next
// step overwrites it, and only the call's last step's is seen.  (x_cur is
// still written: a sort that switches to the cold passes bins the caller's
// positions.)
template <bool EFFECT>
__device__ __forceinline__ float2 p2p_step_update(const SysView& v, const CallIO& io, const RunGeom& g, float ax,
                                                  float ay) {
#pragma clang fp contract(off)
  // l2p.cu:34: effect += cell effect; postdo.cu:32: scatter by perm
  const float2 loc = v.L > 0 ? leaf_local(v, g.xt, g.y) : make_float2(0.0f, 0.0f);
  const float ex = ax + loc.x;
  const float ey = ay + loc.y;
  const uint32_t o = __float_as_uint(g.pt.w);
  if (EFFECT) io.acc0[v.part_off + o] = make_float2(ex, ey);
  // update.cu:39-41
  float2* xcur = io.xcur + v.part_off;
  float2* xpre = io.xpre + v.part_off;
  // x_cur[o] is the position this step sorted (a step's forces are taken
  // at x_cur, so the sort read it there and carried it in pt): not read again
  const float2 x = make_float2(g.pt.x, g.pt.y);
  const float2 xp = xpre[o];
  float2 nx;
  nx.x = 2.0f * x.x - xp.x + ex * v.tick * v.tick;
  nx.y = 2.0f * x.y - xp.y + ey * v.tick * v.tick;
  xcur[o] = nx;
  xpre[o] = x;
  return nx;
}

// The targets' epilogue.  kP2POnly stores the sorted effect; kP2PFinal adds
// the leaf local (L2L folded in, L2P) and scatters to the original order
// (postdo); kP2PStep also applies the Verlet update; kP2PStepSort then makes
// the first pass of the next step's warm-started sort from the new position
// (what k_warm_gather would do at the start of that step, same bits), and
// leaves the caller's effect array alone.
template <int MODE>
__device__ __forceinline__ void p2p_finish(const SysView& v, const CallIO& io, const RunGeom& g, float ax, float ay) {
#pragma clang fp contract(off)
  if (MODE == kP2PStepSort) {
    // the whole wave takes part in the first pass's mover counts
    float2 nx = make_float2(0.0f, 0.0f);
    if (g.valid) nx = p2p_step_update<false>(v, io, g, ax, ay);
    if (warm_first_pass(v, g.valid, g.i, (uint32_t)(g.y * v.B + g.xt), g.hk, nx, g.pt.z, g.pt.w))
      atomicOr(v.flags, 1u);
    return;
  }
  if (!g.valid) return;
  if (MODE == kP2POnly) {
    v.eff[g.i] = make_float2(ax, ay);
    return;
  }
  if (MODE == kP2PFinal) {
    // l2p.cu:34: effect += cell effect; postdo.cu:32: scatter by perm
    const float2 loc = v.L > 0 ? leaf_local(v, g.xt, g.y) : make_float2(0.0f, 0.0f);
    const uint32_t o = __float_as_uint(g.pt.w);
    io.acc0[v.part_off + o] = make_float2(ax + loc.x, ay + loc.y);
  } else {
    p2p_step_update<true>(v, io, g, ax, ay);
  }
}

// One wavefront works one run of up to 64 targets of row y, one target per
// lane (runs.hpp: the run queue and the split of heavy runs).  Target t in
// leaf x_t needs, in each neighbour row r, the sorted range
// [head[r, x_t - 1], head[r, x_t + 2]) (the row-range idea of
// p2p.cu:105-118).  Those ranges lie inside the run's union, which the wave
// stages in its LDS buffer STAGE sources at a time; each lane then loops over
// its own part of the window, U sources per iteration.  A lane adds its
// sources in sorted order, rows y-1, y, y+1, and the segments of a split run
// in segment order, so every variant (stage, unroll, waves) gives the same
// bits.
template <int MODE, class Sys, int STAGE, int U, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void k_p2p(Sys sys, CallIO io) {
  static_assert(STAGE % kWave == 0, "staging size must be a multiple of 64");
  __shared__ float4 s_src[WAVES][STAGE];
  __shared__ int s_runs[(1 << (kMaxLevel + 1)) + 1];
  __shared__ uint32_t sh[WAVES];
  __shared__ RunSlot<2> slot;
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int lane = lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  float4* src = s_src[wave];
  const float eps = v.eps;
  row_runs<WAVES * kWave>(v, s_runs, sh);
  if (threadIdx.x == 0) run_slot_init(slot);
  __syncthreads();
  run_queue<2, MODE == kP2PStepSort>(
      v, s_runs, slot, blk, nblk, s_runs[v.B], lane,
      [&](const RunGeom& g, int sa, int sb, float* acc) {
        seg_sum<STAGE, U>(v, src, g, lane, sa, sb, eps, acc[0], acc[1]);
      },
      [&](const RunGeom& g, int sa, int sb, float* acc) {
        seg_sum<STAGE, 2>(v, src, g, lane, sa, sb, eps, acc[0], acc[1]);
      },
      [&](const RunGeom& g, const float* acc) { p2p_finish<MODE>(v, io, g, acc[0], acc[1]); });
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
    case kP2PStepSort: launch_p2p_variant<kP2PStepSort, OneSys>(variant, g, one, io, s); break;
    default: launch_p2p_variant<kP2PStep, OneSys>(variant, g, one, io, s); break;
  }
  return hipGetLastError();
}

hipError_t launch_p2p_batch(const BatchLaunch& bl, int mode, const CallIO& io, int variant, hipStream_t s) {
  const dim3 g(bl.p2p.blocks);
  if (mode == kP2PFinal)
    launch_p2p_variant<kP2PFinal, ManySys>(variant, g, bl.p2p.sys, io, s);
  else if (mode == kP2PStepSort)
    launch_p2p_variant<kP2PStepSort, ManySys>(variant, g, bl.p2p.sys, io, s);
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
