// Near field (P2P) with the fused L2P, un-permute and Verlet epilogue, plus
// the per-particle stage kernels of the buffer API.
//
// Reference: source/nbfmm/core/solver/p2p.cu:30-176 (p2pNaive / p2pBlock),
// l2p.cu:23-48, postdo.cu:22-45, source/nbfmm/display/stars/update.cu:26-42.
//
// P2P schedule: one wavefront per occupied leaf cell; the waves take the
// leaves from the per-system work queue, largest size class first.  The
// sources of a leaf are three contiguous ranges of the sorted particles, one
// per neighbour row [head[x-1 + yB], head[x+1 + yB + 1]) (the row-range idea
// of p2p.cu:105-118).  The wave stages them, 256 at a
// time, in its own LDS buffer as float4 {x, y, m, -} and pads the buffer
// with zero-mass sources.  For a chunk of nt <= 64 targets the wave is split
// into K = 64/G copies of G = next_pow2(nt) lanes: lane (t, k) accumulates
// target t against staged sources k, k+K, k+2K, ... four per iteration, and
// the K partial sums are combined by a butterfly over the copies.  In the
// final modes the wave also forms its leaf's local (L2L folded in), adds it
// (L2P) and scatters to the original order (postdo), and in the step mode
// applies the Verlet update.
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
  float2 acc = v.local[v.lvl_off[l] + (cy >> l) * (v.B >> l) + (cx >> l)];
  for (l = v.L - 2; l >= 0; --l) {
    const float2 e = v.local[v.lvl_off[l] + (cy >> l) * (v.B >> l) + (cx >> l)];
    acc.x = e.x + acc.x;
    acc.y = e.y + acc.y;
  }
  return acc;
}


// Lane (t, k) of a chunk: target pt against staged sources k, k+K, k+2K, ...
// (cnt sources, zero-mass padded to a multiple of U*K).  K is a compile-time
// constant so the LDS reads use immediate offsets.  Whatever U and the
// staging size, a lane adds its sources in the same order (k, k+K, ...) and
// the padding zeros only at the end, so every variant gives the same bits.
template <int K, int U>
__device__ __forceinline__ void p2p_lane_sum(const float4* __restrict__ src, int k, int cnt, float4 pt, float eps,
                                             float& ax, float& ay) {
  constexpr int step = U * K;
  const int iters = __builtin_amdgcn_readfirstlane((cnt + step - 1) / step);
  const float4* p = src + k;
  for (int i = 0; i < iters; ++i, p += step) {
    float4 s[U];
#pragma unroll
    for (int u = 0; u < U; ++u) s[u] = p[u * K];
#pragma unroll
    for (int u = 0; u < U; ++u) kernel_accum(pt.x, pt.y, s[u].x, s[u].y, s[u].z, eps, ax, ay);
  }
}

// STAGE sources staged per pass (a multiple of 64*U, so every lane's
// padding falls in the last pass only), U sources per lane per iteration,
// WAVES waves per block.
template <int MODE, class Sys, int STAGE, int U, int WAVES>
__global__ __launch_bounds__(WAVES * kWave) void k_p2p(Sys sys, CallIO io) {
  static_assert(STAGE % (kWave * U) == 0, "staging size must be a multiple of 64*U");
  __shared__ float4 s_src[WAVES][STAGE];
  __shared__ int s_next;
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int lane = lane_id();
  // wave-uniform: keeps the cell bookkeeping and loop counters scalar
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  float4* src = s_src[wave];
  float2* acc0 = io.acc0 + v.part_off;
  float2* xcur = io.xcur + v.part_off;
  float2* xpre = io.xpre + v.part_off;
  const int B = v.B;
  const float eps = v.eps;
  LeafQueue queue;
  queue.load(v);
  if (threadIdx.x == 0) s_next = 0;
  __syncthreads();
  for (int j = blk + block_take(&s_next) * nblk; j < queue.total;) {
    const int c = queue.leaf(v, j);
    j = blk + block_take(&s_next) * nblk;
    const int cx = c & (B - 1), cy = c >> v.lgB;
    const int t0 = v.head[c], t1 = v.head[c + 1];
    const int xl = cx > 0 ? cx - 1 : 0;
    const int xr = cx < B - 1 ? cx + 1 : B - 1;
    // the three neighbour rows as contiguous ranges of the sorted particles
    int rs0 = 0, rs1 = 0, rs2 = 0, n0 = 0, n1 = 0, n2 = 0;
    if (cy > 0) { rs0 = v.head[xl + (cy - 1) * B]; n0 = v.head[xr + (cy - 1) * B + 1] - rs0; }
    rs1 = v.head[xl + cy * B]; n1 = v.head[xr + cy * B + 1] - rs1;
    if (cy < B - 1) { rs2 = v.head[xl + (cy + 1) * B]; n2 = v.head[xr + (cy + 1) * B + 1] - rs2; }
    const int S = n0 + n1 + n2;
    float2 loc = make_float2(0.0f, 0.0f);
    if (MODE != kP2POnly && v.L > 0) loc = leaf_local(v, cx, cy);
    for (int tb = t0; tb < t1; tb += kWave) {
      const int nt = min(kWave, t1 - tb);
      const int lgG = nt <= 1 ? 0 : 32 - __clz(nt - 1);
      const int G = 1 << lgG;
      const int K = kWave >> lgG;
      const int t = lane & (G - 1);
      const int k = lane >> lgG;
      const bool valid = t < nt;
      const float4 pt = v.part[tb + (valid ? t : 0)];
      float ax = 0.0f, ay = 0.0f;
      for (int base = 0; base < S; base += STAGE) {
        const int cnt = min(STAGE, S - base);
        // stage up to STAGE sources; slots past cnt hold zero-mass sources,
        // which add exactly nothing
#pragma unroll
        for (int q = 0; q < STAGE / kWave; ++q) {
          const int slot = q * kWave + lane;
          const int f = base + slot;
          float4 val = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          if (slot < cnt) {
            const int g = f < n0 ? rs0 + f : (f < n0 + n1 ? rs1 + (f - n0) : rs2 + (f - n0 - n1));
            val = v.part[g];
          }
          src[slot] = val;
        }
        wave_lds_order();
        switch (lgG) {
          case 0: p2p_lane_sum<64, U>(src, k, cnt, pt, eps, ax, ay); break;
          case 1: p2p_lane_sum<32, U>(src, k, cnt, pt, eps, ax, ay); break;
          case 2: p2p_lane_sum<16, U>(src, k, cnt, pt, eps, ax, ay); break;
          case 3: p2p_lane_sum<8, U>(src, k, cnt, pt, eps, ax, ay); break;
          case 4: p2p_lane_sum<4, U>(src, k, cnt, pt, eps, ax, ay); break;
          case 5: p2p_lane_sum<2, U>(src, k, cnt, pt, eps, ax, ay); break;
          default: p2p_lane_sum<1, U>(src, k, cnt, pt, eps, ax, ay); break;
        }
        wave_lds_order();
      }
      for (int off = G; off < kWave; off <<= 1) {
        ax += __shfl_xor(ax, off, kWave);
        ay += __shfl_xor(ay, off, kWave);
      }
      if (k == 0 && valid) {
        const int i = tb + t;
        if (MODE == kP2POnly) {
          v.eff[i] = make_float2(ax, ay);
        } else {
#pragma clang fp contract(off)
          // l2p.cu:34: effect += cell effect; postdo.cu:32: scatter by perm
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
static const P2PVariant kVariants[] = {{256, 4, 16}, {256, 4, 8}, {256, 2, 16}, {512, 4, 8},
                                       {512, 8, 8},  {256, 4, 4}, {128, 2, 16}};
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
    case 1: hipLaunchKernelGGL((k_p2p<MODE, Sys, 256, 4, 8>), g, b, 0, s, sys, io); break;
    case 2: hipLaunchKernelGGL((k_p2p<MODE, Sys, 256, 2, 16>), g, b, 0, s, sys, io); break;
    case 3: hipLaunchKernelGGL((k_p2p<MODE, Sys, 512, 4, 8>), g, b, 0, s, sys, io); break;
    case 4: hipLaunchKernelGGL((k_p2p<MODE, Sys, 512, 8, 8>), g, b, 0, s, sys, io); break;
    case 5: hipLaunchKernelGGL((k_p2p<MODE, Sys, 256, 4, 4>), g, b, 0, s, sys, io); break;
    case 6: hipLaunchKernelGGL((k_p2p<MODE, Sys, 128, 2, 16>), g, b, 0, s, sys, io); break;
    default: hipLaunchKernelGGL((k_p2p<MODE, Sys, 256, 4, 16>), g, b, 0, s, sys, io); break;
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
