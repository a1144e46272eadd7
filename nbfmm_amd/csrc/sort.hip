// Binning and stable sort of particles by leaf cell (the predo stage).
//
// Reference: source/nbfmm/core/solver/predo.cu:26-135.
//   computeParticleIndex  -> k_bin (same IEEE expression, plus clamp + flag)
//   thrust::sort_by_key   -> two LSD radix passes (digit ix, then digit iy)
//   extractHead           -> k_cscan_* (exclusive scan of the leaf histogram)
//   permuteInputVector    -> fused into the last radix pass
//
// The sort is a least-significant-digit radix sort on key = iy*B + ix with
// radix B: pass 0 orders by ix, pass 1 by iy.  Each pass is stable, so the
// result is the stable argsort by (iy, ix) that thrust::sort_by_key gives
// (row-major int2 operator<, include/nbfmm/utility/operator.hpp:107-109).
// Per pass: k_radix_up builds per-tile digit histograms in LDS, k_radix_scan
// turns them into per-(digit, tile) offsets, and k_radix_down ranks each
// element inside its tile with wave ballots (a 64-lane match of equal
// digits) and scatters it.
#include "common.hpp"
#include "launch.hpp"

namespace nbfmm_amd {

// predo.cu:37-38: floorf((p - lo) / cell) with IEEE division.  Particles
// outside the box (or NaN) are clamped into the grid and flagged; the
// reference has no clamp (a particle at x_max indexes cell B).
__global__ __launch_bounds__(256) void k_bin(SysView v) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.n) return;
  const float2 p = v.pos0[i];
  const float fx = floorf((p.x - v.xmin) / v.cw);
  const float fy = floorf((p.y - v.ymin) / v.ch);
  int ix, iy;
  bool bad = false;
  if (!(fx >= 0.0f)) { ix = 0; bad = true; }
  else if (fx >= (float)v.B) { ix = v.B - 1; bad = true; }
  else ix = (int)fx;
  if (!(fy >= 0.0f)) { iy = 0; bad = true; }
  else if (fy >= (float)v.B) { iy = v.B - 1; bad = true; }
  else iy = (int)fy;
  const uint32_t key = (uint32_t)(iy * v.B + ix);
  v.key0[i] = key;
  atomicAdd(&v.count[key], 1u);
  if (bad) atomicOr(v.flags, 1u);
}

// Leaf histogram -> head (exclusive scan, head[B*B] = N) and the compacted
// list of occupied leaves.  Three kernels: block sums, scan of the block
// sums, block-local scans.  The last one also clears count for the next
// solve.
__global__ __launch_bounds__(kScanThreads) void k_cscan_reduce(SysView v) {
  __shared__ uint32_t sh[kScanThreads / kWave];
  const int C = v.B * v.B;
  const int base = blockIdx.x * kScanChunk;
  uint32_t cnt = 0, occ = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int c = base + j * kScanThreads + threadIdx.x;
    if (c < C) {
      const uint32_t k = v.count[c];
      cnt += k;
      occ += k ? 1u : 0u;
    }
  }
  uint32_t tc, to;
  block_exclusive_scan<kScanThreads>(cnt, sh, &tc);
  block_exclusive_scan<kScanThreads>(occ, sh, &to);
  if (threadIdx.x == 0) {
    v.part_cnt[blockIdx.x] = tc;
    v.part_occ[blockIdx.x] = to;
  }
}

__global__ __launch_bounds__(kScanThreads) void k_cscan_top(SysView v, int nblk) {
  __shared__ uint32_t sh[kScanThreads / kWave];
  __shared__ uint32_t carry[2];
  if (threadIdx.x == 0) { carry[0] = 0; carry[1] = 0; }
  __syncthreads();
  for (int b0 = 0; b0 < nblk; b0 += kScanThreads) {
    const int b = b0 + threadIdx.x;
    const uint32_t c = b < nblk ? v.part_cnt[b] : 0u;
    const uint32_t o = b < nblk ? v.part_occ[b] : 0u;
    uint32_t tc, to;
    const uint32_t ec = block_exclusive_scan<kScanThreads>(c, sh, &tc);
    const uint32_t eo = block_exclusive_scan<kScanThreads>(o, sh, &to);
    const uint32_t c0 = carry[0], o0 = carry[1];
    if (b < nblk) {
      v.part_cnt[b] = c0 + ec;
      v.part_occ[b] = o0 + eo;
    }
    __syncthreads();
    if (threadIdx.x == 0) { carry[0] = c0 + tc; carry[1] = o0 + to; }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    v.head[v.B * v.B] = (int)carry[0];
    *v.nocc = (int)carry[1];
  }
}

__global__ __launch_bounds__(kScanThreads) void k_cscan_down(SysView v) {
  __shared__ uint32_t sh[kScanThreads / kWave];
  const int C = v.B * v.B;
  // each thread owns kScanItems consecutive leaves
  const int c0 = blockIdx.x * kScanChunk + threadIdx.x * kScanItems;
  uint32_t k[kScanItems];
  uint32_t cnt = 0, occ = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int c = c0 + j;
    k[j] = c < C ? v.count[c] : 0u;
    cnt += k[j];
    occ += k[j] ? 1u : 0u;
  }
  uint32_t tc, to;
  uint32_t ec = block_exclusive_scan<kScanThreads>(cnt, sh, &tc) + v.part_cnt[blockIdx.x];
  uint32_t eo = block_exclusive_scan<kScanThreads>(occ, sh, &to) + v.part_occ[blockIdx.x];
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const int c = c0 + j;
    if (c < C) {
      v.head[c] = (int)ec;
      if (k[j]) v.occ[eo++] = c;
      ec += k[j];
      v.count[c] = 0u;
    }
  }
}

// Occupied-leaf list rebuilt from head alone (after head was written through
// the buffer API).  Single block; only used by the stage API.
__global__ __launch_bounds__(kScanThreads) void k_occ_from_head(SysView v) {
  __shared__ uint32_t sh[kScanThreads / kWave];
  __shared__ uint32_t carry;
  const int C = v.B * v.B;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int b0 = 0; b0 < C; b0 += kScanThreads) {
    const int c = b0 + threadIdx.x;
    const uint32_t o = (c < C && v.head[c + 1] > v.head[c]) ? 1u : 0u;
    uint32_t tot;
    const uint32_t e = block_exclusive_scan<kScanThreads>(o, sh, &tot);
    const uint32_t base = carry;
    if (o) v.occ[base + e] = c;
    __syncthreads();
    if (threadIdx.x == 0) carry = base + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *v.nocc = (int)carry;
}

__device__ __forceinline__ uint32_t digit_of(uint32_t key, int pass, const SysView& v) {
  return pass == 0 ? (key & (uint32_t)(v.B - 1)) : (key >> v.lgB);
}

// Per-tile digit histogram, stored digit-major: hist[d * ntiles + tile].
template <int PASS>
__global__ __launch_bounds__(kSortThreads) void k_radix_up(SysView v, int ntiles) {
  __shared__ uint32_t h[512];
  const int R = v.B;
  for (int d = threadIdx.x; d < R; d += kSortThreads) h[d] = 0;
  __syncthreads();
  const int base = blockIdx.x * kSortTile;
  const uint32_t* keys = PASS == 0 ? v.key0 : v.keyA;
#pragma unroll
  for (int j = 0; j < kSortWaveItems; ++j) {
    const int i = base + j * kSortThreads + threadIdx.x;
    if (i < v.n) atomicAdd(&h[digit_of(keys[i], PASS, v)], 1u);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < R; d += kSortThreads) v.hist[(size_t)d * ntiles + blockIdx.x] = h[d];
}

// One block per digit: exclusive scan of that digit's tile counts, in
// place; the digit total goes to dtotal[d].
__global__ __launch_bounds__(kScanThreads) void k_radix_scan(SysView v, int ntiles) {
  __shared__ uint32_t sh[kScanThreads / kWave];
  __shared__ uint32_t carry;
  const int d = blockIdx.x;
  uint32_t* row = v.hist + (size_t)d * ntiles;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int t0 = 0; t0 < ntiles; t0 += kScanThreads) {
    const int t = t0 + threadIdx.x;
    const uint32_t c = t < ntiles ? row[t] : 0u;
    uint32_t tot;
    const uint32_t e = block_exclusive_scan<kScanThreads>(c, sh, &tot);
    const uint32_t base = carry;
    if (t < ntiles) row[t] = base + e;
    __syncthreads();
    if (threadIdx.x == 0) carry = base + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) v.dtotal[d] = carry;
}

// Stable scatter of one tile.  Wave w owns elements [w*512, w*512+512) of
// the tile as 8 chunks of 64 in input order; inside a chunk, lanes with equal
// digits are found with one ballot per digit bit, and a per-wave LDS counter
// per digit carries the running rank across chunks.  Per-wave counts are then
// prefix-summed over waves, so the final order inside a digit is the input
// order (stable).
template <int PASS>
__global__ __launch_bounds__(kSortThreads) void k_radix_down(SysView v, int ntiles) {
  __shared__ uint32_t s_base[512];
  __shared__ uint32_t s_wcnt[kSortThreads / kWave][512];
  __shared__ uint32_t sh[kSortThreads / kWave];
  const int R = v.B;
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
  const int tile = blockIdx.x;

  // digit bases: exclusive scan of the digit totals (R <= 512, two per thread)
  {
    const int d0 = 2 * threadIdx.x, d1 = d0 + 1;
    const uint32_t t0 = d0 < R ? v.dtotal[d0] : 0u;
    const uint32_t t1 = d1 < R ? v.dtotal[d1] : 0u;
    uint32_t tot;
    const uint32_t e = block_exclusive_scan<kSortThreads>(t0 + t1, sh, &tot);
    if (d0 < R) s_base[d0] = e + v.hist[(size_t)d0 * ntiles + tile];
    if (d1 < R) s_base[d1] = e + t0 + v.hist[(size_t)d1 * ntiles + tile];
  }
  for (int d = threadIdx.x; d < R; d += kSortThreads) {
#pragma unroll
    for (int w = 0; w < kSortThreads / kWave; ++w) s_wcnt[w][d] = 0;
  }
  __syncthreads();

  const uint32_t* keys = PASS == 0 ? v.key0 : v.keyA;
  const int wbase = tile * kSortTile + wave * (kWave * kSortWaveItems);
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
  uint32_t k[kSortWaveItems], val[kSortWaveItems], rank[kSortWaveItems];
  int nbits = v.lgB;
#pragma unroll
  for (int j = 0; j < kSortWaveItems; ++j) {
    const int i = wbase + j * kWave + lane;
    const bool valid = i < v.n;
    k[j] = valid ? keys[i] : 0u;
    val[j] = valid ? (PASS == 0 ? (uint32_t)i : v.valA[i]) : 0u;
    const uint32_t d = digit_of(k[j], PASS, v);
    uint64_t m = __ballot(valid);
    for (int b = 0; b < nbits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(valid && bit);
      m &= bit ? bal : ~bal;
    }
    const uint32_t r = (uint32_t)__popcll(m & lt_mask);
    const uint32_t cnt = (uint32_t)__popcll(m);
    uint32_t prior = 0;
    if (valid) prior = s_wcnt[wave][d];
    rank[j] = prior + r;
    __syncthreads();
    if (valid && r == 0) s_wcnt[wave][d] = prior + cnt;
    __syncthreads();
  }
  // exclusive prefix of the per-wave counts, per digit
  for (int d = threadIdx.x; d < R; d += kSortThreads) {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < kSortThreads / kWave; ++w) {
      const uint32_t c = s_wcnt[w][d];
      s_wcnt[w][d] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortWaveItems; ++j) {
    const int i = wbase + j * kWave + lane;
    if (i >= v.n) continue;
    const uint32_t d = digit_of(k[j], PASS, v);
    const uint32_t p = s_base[d] + s_wcnt[wave][d] + rank[j];
    if (PASS == 0) {
      v.keyA[p] = k[j];
      v.valA[p] = val[j];
    } else {
      // permuteInputVector (predo.cu:86-100) fused: gather position and weight
      v.key[p] = k[j];
      const float2 q = v.pos0[val[j]];
      v.part[p] = make_float4(q.x, q.y, v.w0[val[j]], __uint_as_float(val[j]));
    }
  }
}

// ---------------------------------------------------------------- launchers

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

hipError_t launch_predo(const SysView& v, hipStream_t s) {
  const int C = v.B * v.B;
  if (v.n > 0) {
    hipLaunchKernelGGL(k_bin, dim3(cdiv(v.n, 256)), dim3(256), 0, s, v);
  }
  const int nblk = cdiv(C, kScanChunk);
  hipLaunchKernelGGL(k_cscan_reduce, dim3(nblk), dim3(kScanThreads), 0, s, v);
  hipLaunchKernelGGL(k_cscan_top, dim3(1), dim3(kScanThreads), 0, s, v, nblk);
  hipLaunchKernelGGL(k_cscan_down, dim3(nblk), dim3(kScanThreads), 0, s, v);
  if (v.n > 0) {
    const int ntiles = cdiv(v.n, kSortTile);
    hipLaunchKernelGGL(k_radix_up<0>, dim3(ntiles), dim3(kSortThreads), 0, s, v, ntiles);
    hipLaunchKernelGGL(k_radix_scan, dim3(v.B), dim3(kScanThreads), 0, s, v, ntiles);
    hipLaunchKernelGGL(k_radix_down<0>, dim3(ntiles), dim3(kSortThreads), 0, s, v, ntiles);
    hipLaunchKernelGGL(k_radix_up<1>, dim3(ntiles), dim3(kSortThreads), 0, s, v, ntiles);
    hipLaunchKernelGGL(k_radix_scan, dim3(v.B), dim3(kScanThreads), 0, s, v, ntiles);
    hipLaunchKernelGGL(k_radix_down<1>, dim3(ntiles), dim3(kSortThreads), 0, s, v, ntiles);
  }
  return hipGetLastError();
}

hipError_t launch_occ_from_head(const SysView& v, hipStream_t s) {
  hipLaunchKernelGGL(k_occ_from_head, dim3(1), dim3(kScanThreads), 0, s, v);
  return hipGetLastError();
}

int scan_blocks(int B) { return cdiv(B * B, kScanChunk); }

}  // namespace nbfmm_amd
