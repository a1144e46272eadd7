// Removal of out-of-box stars after a step (Stars::prune).
//
// Reference: source/nbfmm/display/stars/prune.cu:28-62.  pruneDevice flags a
// star with x < x_min || y < y_min || x > x_max || y > y_max (a NaN compares
// false and stays); thrust::sort_by_key on the bool flags then moves
// (x_cur, x_pre, weight) so that the kept stars come first, both groups in
// their previous order, and the new star count is the number kept.
//
// Here the same stable partition runs over tiles of 2048 stars in three
// passes: kept count per tile, exclusive scan of the tile counts (one
// block), and a scatter that ranks each tile's stars in index order; the
// result is staged in solver scratch and copied back over the caller's
// arrays.
#include "common.hpp"
#include "launch.hpp"

namespace nbfmm_amd {

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

constexpr int kPruneThreads = 256;
constexpr int kPruneItems = 8;
constexpr int kPruneTile = kPruneThreads * kPruneItems;

__device__ __forceinline__ bool out_of_box(float2 p, float4 lim) {
  return p.x < lim.x || p.y < lim.y || p.x > lim.z || p.y > lim.w;
}

__global__ __launch_bounds__(kPruneThreads) void k_prune_count(int n, float4 lim, const float2* xcur,
                                                               uint32_t* tcount) {
  __shared__ uint32_t sh[kPruneThreads / kWave];
  const int base = blockIdx.x * kPruneTile;
  uint32_t kept = 0;
#pragma unroll
  for (int j = 0; j < kPruneItems; ++j) {
    const int i = base + j * kPruneThreads + (int)threadIdx.x;
    if (i < n && !out_of_box(xcur[i], lim)) ++kept;
  }
  uint32_t tot;
  block_exclusive_scan<kPruneThreads>(kept, sh, &tot);
  if (threadIdx.x == 0) tcount[blockIdx.x] = tot;
}

// In place: tcount[t] becomes the stars kept before tile t; *kept the total.
__global__ __launch_bounds__(kPruneThreads) void k_prune_scan(int ntiles, uint32_t* tcount, int* kept) {
  __shared__ uint32_t sh[kPruneThreads / kWave];
  uint32_t carry = 0;
  for (int t0 = 0; t0 < ntiles; t0 += kPruneThreads) {
    const int t = t0 + (int)threadIdx.x;
    const uint32_t c = t < ntiles ? tcount[t] : 0u;
    uint32_t tot;
    const uint32_t e = block_exclusive_scan<kPruneThreads>(c, sh, &tot);
    if (t < ntiles) tcount[t] = carry + e;
    carry += tot;
  }
  if (threadIdx.x == 0) *kept = (int)carry;
}

// Kept star i goes to (kept before its tile) + (kept before it in the tile);
// a removed one to total_kept + (removed before it).
__global__ __launch_bounds__(kPruneThreads) void k_prune_scatter(int n, float4 lim, const float2* xcur,
                                                                 const float2* xpre, const float* w,
                                                                 const uint32_t* toff, const int* kept,
                                                                 float4* tmp, float* tmpw) {
  __shared__ uint32_t sh[kPruneThreads / kWave];
  const int base = blockIdx.x * kPruneTile;
  const uint32_t kept_before = toff[blockIdx.x];
  const uint32_t gone_before = (uint32_t)base - kept_before;
  const uint32_t total = (uint32_t)*kept;
  uint32_t run_kept = 0;
  for (int j = 0; j < kPruneItems; ++j) {
    const int i = base + j * kPruneThreads + (int)threadIdx.x;
    const bool valid = i < n;
    float2 xc = make_float2(0.0f, 0.0f);
    bool keep = false;
    if (valid) {
      xc = xcur[i];
      keep = !out_of_box(xc, lim);
    }
    uint32_t tot;
    const uint32_t rk = block_exclusive_scan<kPruneThreads>(keep ? 1u : 0u, sh, &tot);
    if (valid) {
      // stars before i in this tile: j * kPruneThreads + threadIdx.x
      const uint32_t before = (uint32_t)(j * kPruneThreads) + threadIdx.x;
      const uint32_t k = run_kept + rk;
      const uint32_t dst = keep ? kept_before + k : total + gone_before + (before - k);
      const float2 xp = xpre[i];
      tmp[dst] = make_float4(xc.x, xc.y, xp.x, xp.y);
      tmpw[dst] = w[i];
    }
    run_kept += tot;
  }
}

__global__ __launch_bounds__(256) void k_prune_copy(int n, const float4* tmp, const float* tmpw, float2* xcur,
                                                   float2* xpre, float* w) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 t = tmp[i];
  xcur[i] = make_float2(t.x, t.y);
  xpre[i] = make_float2(t.z, t.w);
  w[i] = tmpw[i];
}

hipError_t launch_prune(int n, float4 lim, float2* xcur, float2* xpre, float* w, uint32_t* tcount, float4* tmp,
                        float* tmpw, int* kept, hipStream_t s) {
  const int ntiles = cdiv(n, kPruneTile);
  hipLaunchKernelGGL(k_prune_count, dim3(ntiles), dim3(kPruneThreads), 0, s, n, lim, xcur, tcount);
  hipLaunchKernelGGL(k_prune_scan, dim3(1), dim3(kPruneThreads), 0, s, ntiles, tcount, kept);
  hipLaunchKernelGGL(k_prune_scatter, dim3(ntiles), dim3(kPruneThreads), 0, s, n, lim, xcur, xpre, w, tcount,
                     kept, tmp, tmpw);
  hipLaunchKernelGGL(k_prune_copy, dim3(cdiv(n, 256)), dim3(256), 0, s, n, tmp, tmpw, xcur, xpre, w);
  return hipGetLastError();
}

int prune_tiles(int n) { return cdiv(n, kPruneTile); }

}  // namespace nbfmm_amd
