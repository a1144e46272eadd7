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


// One wavefront works one run of up to 64 targets of row y, one target per
// lane.  Target t in leaf x_t needs, in each neighbour row r, the sorted
// range [head[r, x_t - 1], head[r, x_t + 2]) (the row-range idea of
// p2p.cu:105-118).  Those ranges lie inside the run's union
// [head[r, x_first - 1], head[r, x_last + 2]), which the wave stages in its
// LDS buffer STAGE sources at a time; each lane then loops over its own part
// of the window, U sources per iteration.  A lane adds its sources in sorted
// order, rows y-1, y, y+1, whatever STAGE and U: every variant gives the
// same bits.  The modes after the sum: kP2POnly stores the sorted effect;
// kP2PFinal adds the leaf local (L2L folded in, L2P) and scatters to the
// original order (postdo); kP2PStep also applies the Verlet update.
template <int MODE, class Sys, int STAGE, int U, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void k_p2p(Sys sys, CallIO io) {
  static_assert(STAGE % kWave == 0, "staging size must be a multiple of 64");
  __shared__ float4 s_src[WAVES][STAGE];
  __shared__ int s_runs[(1 << (kMaxLevel + 1)) + 1];
  __shared__ uint32_t sh[WAVES];
  __shared__ int s_next;
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int lane = lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  float4* src = s_src[wave];
  float2* acc0 = io.acc0 + v.part_off;
  float2* xcur = io.xcur + v.part_off;
  float2* xpre = io.xpre + v.part_off;
  const int B = v.B;
  const float eps = v.eps;
  row_runs<WAVES * kWave>(v, s_runs, sh);
  if (threadIdx.x == 0) s_next = 0;
  __syncthreads();
  const int total = s_runs[B];
  for (int j = blk + block_take(&s_next) * nblk; j < total; j = blk + block_take(&s_next) * nblk) {
    const int y = run_row(s_runs, B, j);
    const int rs = v.head[y * B];
    const int T0 = rs + kWave * (j - s_runs[y]);
    const int nt = min(kWave, v.head[(y + 1) * B] - T0);
    const bool valid = lane < nt;
    const int i = T0 + (valid ? lane : 0);
    const float4 pt = v.part[i];
    const int xt = (int)(v.key[i] & (uint32_t)(B - 1));
    const int xf = __shfl(xt, 0, kWave);
    const int xl = __shfl(xt, nt - 1, kWave);
    float ax = 0.0f, ay = 0.0f;
#pragma unroll
    for (int dr = -1; dr <= 1; ++dr) {
      const int r = y + dr;
      if (r < 0 || r >= B) continue;
      const int* hr = v.head + r * B;
      const int us = hr[xf > 0 ? xf - 1 : 0];
      const int ulen = hr[(xl < B - 1 ? xl + 1 : B - 1) + 1] - us;
      // this lane's range inside the union (empty for idle lanes)
      const int a = valid ? hr[xt > 0 ? xt - 1 : 0] - us : 0;
      const int b = valid ? hr[(xt < B - 1 ? xt + 1 : B - 1) + 1] - us : 0;
      for (int w0 = 0; w0 < ulen; w0 += STAGE) {
        const int cnt = min(STAGE, ulen - w0);
        // all loads of the chunk in flight at once, then the LDS writes, with
        // no branch for the compiler to sink a load into (a load under a
        // branch waits before the next one is issued): indices are clamped
        // to the chunk's last source (cnt >= 1), and the slots from cnt on,
        // which hold copies of it, are never read
        {
          const float4* const gsrc = v.part + us + w0;
          float4 stage[STAGE / kWave];
#pragma unroll
          for (int q = 0; q < STAGE / kWave; ++q) stage[q] = gsrc[min(q * kWave + lane, cnt - 1)];
#pragma unroll
          for (int q = 0; q < STAGE / kWave; ++q) src[q * kWave + lane] = stage[q];
        }
        wave_lds_order();
        const int lo = min(max(a - w0, 0), cnt);
        const int hi = min(max(b - w0, 0), cnt);
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
          for (int u = 0; u < U; ++u) kernel_accum(pt.x, pt.y, sv[u].x, sv[u].y, sv[u].z, eps, ax, ay);
        }
        for (; p < pend; ++p) {
          const float4 s1 = *p;
          kernel_accum(pt.x, pt.y, s1.x, s1.y, s1.z, eps, ax, ay);
        }
        wave_lds_order();
      }
    }
    if (valid) {
      if (MODE == kP2POnly) {
        v.eff[i] = make_float2(ax, ay);
      } else {
#pragma clang fp contract(off)
        // l2p.cu:34: effect += cell effect; postdo.cu:32: scatter by perm
        const float2 loc = v.L > 0 ? leaf_local(v, xt, y) : make_float2(0.0f, 0.0f);
        const float ex = ax + loc.x;
        const float ey = ay + loc.y;
        const uint32_t o = __float_as_uint(pt.w);
        acc0[o] = make_float2(ex, ey);
        if (MODE == kP2PStep) {
          // update.cu:39-41
          const float2 x = xcur[o];
          const float2 xp = xpre[o];
          float2 nx;
          nx.x = 2.0f * x.x - xp.x + ex * v.tick * v.tick;
          nx.y = 2.0f * x.y - xp.y + ey * v.tick * v.tick;
          xcur[o] = nx;
          xpre[o] = x;
          if (io.xs_out) io.xs_out[v.part_off + i] = nx;
        }
      }
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
