// Binning and stable sort of particles by leaf cell (the predo stage).
//
// Reference: source/nbfmm/core/solver/predo.cu:26-135.
//   computeParticleIndex  -> bin_key (same IEEE expression, plus clamp + flag)
//   thrust::sort_by_key   -> a stable counting sort by row (iy), then a stable
//                            counting sort of every row by column (ix)
//   extractHead           -> written by the row sort (exclusive scan of the
//                            per-row column histogram, plus the row base)
//   permuteInputVector    -> position and weight travel with the keys
//
// The result is the stable argsort by (iy, ix) that thrust::sort_by_key gives
// (row-major int2 operator<, include/nbfmm/utility/operator.hpp:107-109): the
// row pass keeps input order inside a row, and the column pass keeps row-pass
// order inside a cell.
//
// Row pass (tiles of 2048 particles in input order): k_row_up builds each
// tile's row histogram in LDS, k_row_scan turns them into per-(row, tile)
// offsets, and k_row_down ranks every particle inside its tile (wave ballots
// find the lanes with equal rows) and writes {key} and {x, y, m, index} to
// the row-sorted scratch.  Reads of the caller's arrays are coalesced; no
// gather.
// Column pass: one block per row sorts the row's contiguous segment by
// column with an LDS histogram, ranking groups of particles the same way,
// and writes the final sorted particles, keys, head for the row's cells and
// the row's occupied cells.
#include "common.hpp"
#include "launch.hpp"

namespace nbfmm_amd {

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// ------------------------------------------------------------- row pass

// Row histogram of one tile of 2048 particles in input order, by a block of
// NT threads (h: 512 words of LDS).  All threads of the block call it.
template <int NT>
__device__ __forceinline__ void row_up_tile(const SysView& v, const CallIO& io, int tile, int ntiles, uint32_t* h) {
  const float2* pos0 = io.pos0 + v.part_off;
  for (int d = threadIdx.x; d < v.B; d += NT) h[d] = 0;
  __syncthreads();
  const int base = tile * kSortTile;
  bool any_bad = false;
#pragma unroll
  for (int j = 0; j < kSortTile / NT; ++j) {
    const int i = base + j * NT + threadIdx.x;
    if (i < v.n) {
      bool bad;
      const uint32_t key = bin_key(v, pos0[i], bad);
      any_bad |= bad;
      atomicAdd(&h[key >> v.lgB], 1u);
    }
  }
  if (any_bad) atomicOr(v.flags, 1u);
  __syncthreads();
  for (int d = threadIdx.x; d < v.B; d += NT) v.hist[(size_t)d * ntiles + tile] = h[d];
  __syncthreads();
}

template <class Sys>
__global__ __launch_bounds__(kSortThreads) void k_row_up(Sys sys, CallIO io) {
  __shared__ uint32_t h[512];
  int tile, ntiles;
  const SysView v = sys.get(tile, ntiles);
  row_up_tile<kSortThreads>(v, io, tile, ntiles, h);
}

// Exclusive scan of row d's tile counts, in place, by a block of NT
// threads; the row total goes to dtotal[d].  Row 0 also clears
// the occupied-cell counter that the column pass fills.  All threads of the
// block call it.
template <int NT>
__device__ __forceinline__ void row_scan_row(const SysView& v, int d, uint32_t* sh, uint32_t* carry) {
  const int ntiles = (v.n + kSortTile - 1) / kSortTile;
  uint32_t* row = v.hist + (size_t)d * ntiles;
  if (threadIdx.x == 0) {
    *carry = 0;
    if (d == 0) *v.nocc = 0;
  }
  __syncthreads();
  for (int t0 = 0; t0 < ntiles; t0 += NT) {
    const int t = t0 + threadIdx.x;
    const uint32_t c = t < ntiles ? row[t] : 0u;
    uint32_t tot;
    const uint32_t e = block_exclusive_scan<NT>(c, sh, &tot);
    const uint32_t base = *carry;
    if (t < ntiles) row[t] = base + e;
    __syncthreads();
    if (threadIdx.x == 0) *carry = base + tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) v.dtotal[d] = *carry;
  __syncthreads();
}

// One block per row.
template <class Sys>
__global__ __launch_bounds__(kScanThreads) void k_row_scan(Sys sys) {
  __shared__ uint32_t sh[kScanThreads / kWave];
  __shared__ uint32_t carry;
  int d, nrows;
  const SysView v = sys.get(d, nrows);
  row_scan_row<kScanThreads>(v, d, sh, &carry);
}

// Stable scatter of one tile by row, by a block of kSortThreads threads.
// Wave w owns elements [w*512, w*512+512) of the tile as 8 chunks of 64 in
// input order; a per-wave LDS counter per row carries the running rank
// across chunks; per-wave counts are then prefix-summed over waves, so rows
// keep input order (stable).  All threads of the block call it.
struct RowDownLds {
  uint32_t base[512];
  uint32_t wcnt[kSortThreads / kWave][512];
  uint32_t sh[kSortThreads / kWave];
};

__device__ __forceinline__ void row_down_tile(const SysView& v, const CallIO& io, int tile, int ntiles,
                                              RowDownLds& L) {
  uint32_t* s_base = L.base;
  auto& s_wcnt = L.wcnt;
  uint32_t* sh = L.sh;
  const float2* pos0 = io.pos0 + v.part_off;
  const float* w0 = io.w0 + v.part_off;
  const int R = v.B;
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
  {
    // row bases: exclusive scan of the row totals (R <= 512, two per thread)
    const int d0 = 2 * threadIdx.x, d1 = d0 + 1;
    const uint32_t t0 = d0 < R ? v.dtotal[d0] : 0u;
    const uint32_t t1 = d1 < R ? v.dtotal[d1] : 0u;
    uint32_t tot;
    const uint32_t e = block_exclusive_scan<kSortThreads>(t0 + t1, sh, &tot);
    if (d0 < R) s_base[d0] = e + v.hist[(size_t)d0 * ntiles + tile];
    if (d1 < R) s_base[d1] = e + t0 + v.hist[(size_t)d1 * ntiles + tile];
    if (tile == 0) {
      if (d0 < R) v.rowstart[d0] = (int)e;
      if (d1 < R) v.rowstart[d1] = (int)(e + t0);
      if (threadIdx.x == 0) v.rowstart[R] = (int)tot;
    }
  }
  for (int d = threadIdx.x; d < R; d += kSortThreads) {
#pragma unroll
    for (int w = 0; w < kSortThreads / kWave; ++w) s_wcnt[w][d] = 0;
  }
  __syncthreads();

  const int wbase = tile * kSortTile + wave * (kWave * kSortWaveItems);
  const uint64_t below = lanes_below();
  uint32_t key[kSortWaveItems], rank[kSortWaveItems];
  float2 q[kSortWaveItems];
  // all loads first, so they are in flight together
#pragma unroll
  for (int j = 0; j < kSortWaveItems; ++j) {
    const int i = wbase + j * kWave + lane;
    q[j] = i < v.n ? pos0[i] : make_float2(0.0f, 0.0f);
  }
#pragma unroll
  for (int j = 0; j < kSortWaveItems; ++j) {
    const int i = wbase + j * kWave + lane;
    const bool valid = i < v.n;
    bool bad;
    key[j] = valid ? bin_key(v, q[j], bad) : 0u;
    const uint32_t d = key[j] >> v.lgB;
    const uint64_t m = match_mask(d, valid, v.lgB);
    const uint32_t r = (uint32_t)__popcll(m & below);
    uint32_t prior = 0;
    if (valid) prior = s_wcnt[wave][d];
    rank[j] = prior + r;
    wave_lds_order();
    if (valid && r == 0) s_wcnt[wave][d] = prior + (uint32_t)__popcll(m);
    wave_lds_order();
  }
  __syncthreads();
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
    const uint32_t d = key[j] >> v.lgB;
    const uint32_t p = s_base[d] + s_wcnt[wave][d] + rank[j];
    v.tkey[p] = key[j];
    v.tpart[p] = make_float4(q[j].x, q[j].y, w0[i], __uint_as_float((uint32_t)i));
  }
  __syncthreads();
}

// row_down_tile with each item's key and rank kept in LDS instead of
// registers (positions are loaded again for the scatter), so that a kernel
// that carries it as a rarely taken branch keeps its register count.
struct RowDownLeanLds {
  RowDownLds d;
  uint32_t kr[kSortTile];  // key | rank << 18 (key < 2^18, rank < 512)
};

__device__ __forceinline__ void row_down_tile_lean(const SysView& v, const CallIO& io, int tile, int ntiles,
                                                   RowDownLeanLds& L) {
  uint32_t* s_base = L.d.base;
  auto& s_wcnt = L.d.wcnt;
  const float2* pos0 = io.pos0 + v.part_off;
  const float* w0 = io.w0 + v.part_off;
  const int R = v.B;
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
  {
    const int d0 = 2 * threadIdx.x, d1 = d0 + 1;
    const uint32_t t0 = d0 < R ? v.dtotal[d0] : 0u;
    const uint32_t t1 = d1 < R ? v.dtotal[d1] : 0u;
    uint32_t tot;
    const uint32_t e = block_exclusive_scan<kSortThreads>(t0 + t1, L.d.sh, &tot);
    if (d0 < R) s_base[d0] = e + v.hist[(size_t)d0 * ntiles + tile];
    if (d1 < R) s_base[d1] = e + t0 + v.hist[(size_t)d1 * ntiles + tile];
    if (tile == 0) {
      if (d0 < R) v.rowstart[d0] = (int)e;
      if (d1 < R) v.rowstart[d1] = (int)(e + t0);
      if (threadIdx.x == 0) v.rowstart[R] = (int)tot;
    }
  }
  for (int d = threadIdx.x; d < R; d += kSortThreads) {
#pragma unroll
    for (int w = 0; w < kSortThreads / kWave; ++w) s_wcnt[w][d] = 0;
  }
  __syncthreads();
  const int wbase = tile * kSortTile + wave * (kWave * kSortWaveItems);
  const uint64_t below = lanes_below();
#pragma unroll 1
  for (int j = 0; j < kSortWaveItems; ++j) {
    const int i = wbase + j * kWave + lane;
    const bool valid = i < v.n;
    bool bad;
    const uint32_t key = valid ? bin_key(v, pos0[i], bad) : 0u;
    const uint32_t d = key >> v.lgB;
    const uint64_t m = match_mask(d, valid, v.lgB);
    const uint32_t r = (uint32_t)__popcll(m & below);
    uint32_t prior = 0;
    if (valid) prior = s_wcnt[wave][d];
    L.kr[wave * (kWave * kSortWaveItems) + j * kWave + lane] = key | ((prior + r) << 18);
    wave_lds_order();
    if (valid && r == 0) s_wcnt[wave][d] = prior + (uint32_t)__popcll(m);
    wave_lds_order();
  }
  __syncthreads();
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
#pragma unroll 1
  for (int j = 0; j < kSortWaveItems; ++j) {
    const int i = wbase + j * kWave + lane;
    if (i >= v.n) continue;
    const uint32_t kr = L.kr[wave * (kWave * kSortWaveItems) + j * kWave + lane];
    const uint32_t key = kr & 0x3ffffu;
    const uint32_t d = key >> v.lgB;
    const uint32_t p = s_base[d] + s_wcnt[wave][d] + (kr >> 18);
    const float2 q = pos0[i];
    v.tkey[p] = key;
    v.tpart[p] = make_float4(q.x, q.y, w0[i], __uint_as_float((uint32_t)i));
  }
  __syncthreads();
}

template <class Sys>
__global__ __launch_bounds__(kSortThreads) void k_row_down(Sys sys, CallIO io) {
  __shared__ RowDownLds L;
  int tile, ntiles;
  const SysView v = sys.get(tile, ntiles);
  row_down_tile(v, io, tile, ntiles, L);
}

// ---------------------------------------------------------- column pass

constexpr int kRowThreads = 512;                         // 8 waves per row block
constexpr int kRowWaves = kRowThreads / kWave;
constexpr int kRowGroup = kRowThreads * kSortWaveItems;  // 4096 particles per group

// Column pass of row y by a block of NT threads (NT <= 512): stable counting
// sort of the row's segment of the row-sorted scratch by column, writing the
// sorted keys and particles, head of the row's cells and the row's occupied
// cells.  zero_empty: also clear the level-0 multipole of the row's empty
// cells (the fused solve's P2M then visits occupied cells only).  Each thread
// takes 512 / NT consecutive cells in the scans.  All threads of the block
// call it.
template <int NT>
struct RowSortLds {
  uint32_t hist[512];
  uint32_t start[512];
  uint32_t run[512];
  uint32_t wcnt[NT / kWave][512];
  uint32_t sh[NT / kWave];
  uint32_t occ_base;
};

template <int NT>
__device__ __forceinline__ void row_sort_row(const SysView& v, int y, int zero_empty, RowSortLds<NT>& L) {
  constexpr int NW = NT / kWave;
  constexpr int LPT = 512 / NT;        // cells per thread in the scans
  constexpr int GROUP = NT * kSortWaveItems;
  const int B = v.B;
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
  const int rs = v.rowstart[y], re = v.rowstart[y + 1];
  const uint32_t cmask = (uint32_t)(B - 1);

  for (int x = threadIdx.x; x < B; x += NT) { L.hist[x] = 0; L.run[x] = 0; }
  __syncthreads();
  for (int p0 = rs + (int)threadIdx.x; p0 < re; p0 += NT * kSortWaveItems) {
    uint32_t k[kSortWaveItems];
#pragma unroll
    for (int j = 0; j < kSortWaveItems; ++j) {
      const int p = p0 + j * NT;
      k[j] = p < re ? v.tkey[p] : 0xffffffffu;
    }
#pragma unroll
    for (int j = 0; j < kSortWaveItems; ++j)
      if (k[j] != 0xffffffffu) atomicAdd(&L.hist[k[j] & cmask], 1u);
  }
  __syncthreads();
  {
    // cell starts inside the row, head, occupied cells
    const int x0 = threadIdx.x * LPT;
    uint32_t c[LPT], csum = 0, osum = 0;
#pragma unroll
    for (int j = 0; j < LPT; ++j) {
      c[j] = x0 + j < B ? L.hist[x0 + j] : 0u;
      csum += c[j];
      osum += c[j] > 0 ? 1u : 0u;
    }
    uint32_t tot;
    uint32_t e = block_exclusive_scan<NT>(csum, L.sh, &tot);
    uint32_t nocc;
    uint32_t eo = block_exclusive_scan<NT>(osum, L.sh, &nocc);
    if (threadIdx.x == 0) L.occ_base = nocc ? atomicAdd((uint32_t*)v.nocc, nocc) : 0u;
#pragma unroll
    for (int j = 0; j < LPT; ++j) {
      const int x = x0 + j;
      if (x < B) {
        L.start[x] = e;
        v.head[y * B + x] = rs + (int)e;
        if (zero_empty && c[j] == 0) v.mpole[y * B + x] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
      e += c[j];
    }
    if (y == B - 1 && threadIdx.x == 0) v.head[B * B] = re;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < LPT; ++j) {
      if (x0 + j < B && c[j] > 0) {
        v.occ[L.occ_base + eo] = y * B + x0 + j;
        ++eo;
      }
    }
  }
  const uint64_t below = lanes_below();
  for (int g = rs; g < re; g += GROUP) {
    for (int x = threadIdx.x; x < B; x += NT) {
#pragma unroll
      for (int w = 0; w < NW; ++w) L.wcnt[w][x] = 0;
    }
    __syncthreads();
    const int wbase = g + wave * (kWave * kSortWaveItems);
    uint32_t key[kSortWaveItems], rank[kSortWaveItems];
#pragma unroll
    for (int j = 0; j < kSortWaveItems; ++j) {
      const int e = wbase + j * kWave + lane;
      key[j] = e < re ? v.tkey[e] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kSortWaveItems; ++j) {
      const int e = wbase + j * kWave + lane;
      const bool valid = e < re;
      const uint32_t d = key[j] & cmask;
      const uint64_t m = match_mask(d, valid, v.lgB);
      const uint32_t r = (uint32_t)__popcll(m & below);
      uint32_t prior = 0;
      if (valid) prior = L.wcnt[wave][d];
      rank[j] = prior + r;
      wave_lds_order();
      if (valid && r == 0) L.wcnt[wave][d] = prior + (uint32_t)__popcll(m);
      wave_lds_order();
    }
    __syncthreads();
    // per cell: offset of each wave = earlier groups + earlier waves
    for (int x = threadIdx.x; x < B; x += NT) {
      uint32_t run = L.run[x];
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        const uint32_t cc = L.wcnt[w][x];
        L.wcnt[w][x] = run;
        run += cc;
      }
      L.run[x] = run;
    }
    __syncthreads();
    float4 pv[kSortWaveItems];
#pragma unroll
    for (int j = 0; j < kSortWaveItems; ++j) {
      const int e = wbase + j * kWave + lane;
      pv[j] = e < re ? v.tpart[e] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
#pragma unroll
    for (int j = 0; j < kSortWaveItems; ++j) {
      const int e = wbase + j * kWave + lane;
      if (e >= re) continue;
      const uint32_t d = key[j] & cmask;
      const int p = rs + (int)(L.start[d] + L.wcnt[wave][d] + rank[j]);
      v.key[p] = key[j];
      v.part[p] = pv[j];
    }
    __syncthreads();
  }
}

// row_sort_row for a block of NT threads with each item's key and rank kept
// in LDS and one item in registers at a time, so that a kernel that carries
// it as a rarely taken branch keeps its register count.
template <int NT>
struct RowSortLeanLds {
  RowSortLds<NT> s;
  uint32_t kr[NT * kSortWaveItems];  // column | rank << 16 (column < 512, rank < 512)
};

template <int NT>
__device__ __forceinline__ void row_sort_row_lean(const SysView& v, int y, int zero_empty, RowSortLeanLds<NT>& LL) {
  RowSortLds<NT>& L = LL.s;
  constexpr int NW = NT / kWave;
  constexpr int LPT = 512 / NT;
  constexpr int GROUP = NT * kSortWaveItems;
  const int B = v.B;
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
  const int rs = v.rowstart[y], re = v.rowstart[y + 1];
  const uint32_t cmask = (uint32_t)(B - 1);
  for (int x = threadIdx.x; x < B; x += NT) { L.hist[x] = 0; L.run[x] = 0; }
  __syncthreads();
#pragma unroll 1
  for (int p = rs + (int)threadIdx.x; p < re; p += NT) atomicAdd(&L.hist[v.tkey[p] & cmask], 1u);
  __syncthreads();
  {
    const int x0 = threadIdx.x * LPT;
    uint32_t csum = 0, osum = 0;
    for (int j = 0; j < LPT; ++j) {
      const uint32_t c = x0 + j < B ? L.hist[x0 + j] : 0u;
      csum += c;
      osum += c > 0 ? 1u : 0u;
    }
    uint32_t tot;
    uint32_t e = block_exclusive_scan<NT>(csum, L.sh, &tot);
    uint32_t nocc;
    uint32_t eo = block_exclusive_scan<NT>(osum, L.sh, &nocc);
    if (threadIdx.x == 0) L.occ_base = nocc ? atomicAdd((uint32_t*)v.nocc, nocc) : 0u;
    for (int j = 0; j < LPT; ++j) {
      const int x = x0 + j;
      const uint32_t c = x < B ? L.hist[x] : 0u;
      if (x < B) {
        L.start[x] = e;
        v.head[y * B + x] = rs + (int)e;
        if (zero_empty && c == 0) v.mpole[y * B + x] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      }
      e += c;
    }
    if (y == B - 1 && threadIdx.x == 0) v.head[B * B] = re;
    __syncthreads();
    for (int j = 0; j < LPT; ++j) {
      const int x = x0 + j;
      if (x < B && L.hist[x] > 0) {
        v.occ[L.occ_base + eo] = y * B + x;
        ++eo;
      }
    }
  }
  const uint64_t below = lanes_below();
  for (int g = rs; g < re; g += GROUP) {
    for (int x = threadIdx.x; x < B; x += NT) {
      for (int w = 0; w < NW; ++w) L.wcnt[w][x] = 0;
    }
    __syncthreads();
    const int wbase = g + wave * (kWave * kSortWaveItems);
#pragma unroll 1
    for (int j = 0; j < kSortWaveItems; ++j) {
      const int e = wbase + j * kWave + lane;
      const bool valid = e < re;
      const uint32_t d = valid ? v.tkey[e] & cmask : 0u;
      const uint64_t m = match_mask(d, valid, v.lgB);
      const uint32_t r = (uint32_t)__popcll(m & below);
      uint32_t prior = 0;
      if (valid) prior = L.wcnt[wave][d];
      LL.kr[wave * (kWave * kSortWaveItems) + j * kWave + lane] = d | ((prior + r) << 16);
      wave_lds_order();
      if (valid && r == 0) L.wcnt[wave][d] = prior + (uint32_t)__popcll(m);
      wave_lds_order();
    }
    __syncthreads();
    for (int x = threadIdx.x; x < B; x += NT) {
      uint32_t run = L.run[x];
      for (int w = 0; w < NW; ++w) {
        const uint32_t cc = L.wcnt[w][x];
        L.wcnt[w][x] = run;
        run += cc;
      }
      L.run[x] = run;
    }
    __syncthreads();
#pragma unroll 1
    for (int j = 0; j < kSortWaveItems; ++j) {
      const int e = wbase + j * kWave + lane;
      if (e >= re) continue;
      const uint32_t kr = LL.kr[wave * (kWave * kSortWaveItems) + j * kWave + lane];
      const uint32_t d = kr & 0xffffu;
      const int p = rs + (int)(L.start[d] + L.wcnt[wave][d] + (kr >> 16));
      v.key[p] = (uint32_t)(y * B) + d;
      v.part[p] = v.tpart[e];
    }
    __syncthreads();
  }
}

// One block per row y.
template <class Sys>
__global__ __launch_bounds__(kRowThreads) void k_row_sort(Sys sys, int zero_empty) {
  __shared__ RowSortLds<kRowThreads> L;
  int y, nrows;
  const SysView v = sys.get(y, nrows);
  row_sort_row<kRowThreads>(v, y, zero_empty, L);
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

// An empty system: head is all zero and nothing is occupied.
__global__ __launch_bounds__(256) void k_empty(SysView v) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c <= v.B * v.B) v.head[c] = 0;
  if (c == 0) *v.nocc = 0;
}

// --------------------------------------------------------- warm start
//
// When key/part/head hold the stable sort of an earlier call with the same
// N, that order is the starting point.  Particles are gathered in the old
// sorted order; one whose leaf is unchanged (a stayer) keeps its place
// relative to the other stayers of its leaf, because both orders are by
// (leaf, original index).  Only the movers, whose leaf changed, need
// placing: a leaf's new content is its stayers merged with its incoming
// movers by original index.  The result is the same stable argsort as the
// cold path, bit for bit.
//   k_warm_gather  gather, bin, mark movers and the leaves they touch (inside
//                  a run, the previous step's P2P does this pass instead:
//                  warm_first_pass)
//   k_warm_rows    new leaf counts, row-local offsets (all rows, one block each)
//   k_warm_finish  head, occupied-leaf list, movers' group offsets, list
//                  of touched leaves
//   k_warm_place   stayers of untouched leaves go straight to their slot;
//                  movers are grouped by new leaf
//   k_warm_merge   one wave per touched leaf merges its stayers and movers
// Each pass clears the counters it is the last to read.  Per-particle
// atomics are per leaf; an atomic that the whole grid shares serialises
// across the chip, so the occupied and touched leaf lists take their row
// bases from scans of per-row counts instead of shared counters.  A leaf that
// gains more than kMergeLdsMovers movers switches the sort to the cold passes:
// the first pass raises the request when a leaf's incoming count passes that
// limit (the one returning atomic of the wave's count), and the warm kernels
// then make the cold sort's four passes in their place (row histograms in
// k_warm_rows, row scan in k_warm_finish, row scatter in k_warm_place, column
// pass in k_warm_merge), with no launch of their own, so such a step costs
// about one cold sort, whatever the number of movers.

template <class Sys>
__global__ __launch_bounds__(256) void k_warm_gather(Sys sys, CallIO io) {
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int i = blk * 256 + threadIdx.x;
  const bool valid = i < v.n;
  float4 old = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  uint32_t kold = 0;
  int hold = 0;
  float2 p = make_float2(0.0f, 0.0f);
  float m = 0.0f;
  if (valid) {
    old = v.part[i];
    const uint32_t o = __float_as_uint(old.w);
    kold = v.key[i];
    hold = v.head[kold];
    p = io.pos0[v.part_off + o];
    m = io.w0[v.part_off + o];
  }
  const bool bad = warm_first_pass(v, valid, i, kold, hold, p, m, old.w);
  if (__any(bad) && lane_id() == 0) atomicOr(v.flags, 1u);
}

// One block per row: new count of every leaf (old count - out + in), its
// row-local exclusive offset, and the row-local offset of its movers.  A
// leaf that movers enter or leave is marked touched, with its old slot
// range.  The row's totals (particles, movers, occupied leaves, touched
// leaves) go to wrowtot.  Clears wcin/wcout.
template <int NT, class Sys>
__global__ __launch_bounds__(NT) void k_warm_rows(Sys sys, CallIO io) {
  __shared__ uint32_t sh[NT / kWave];
  __shared__ uint32_t s_hist[512];
  int y, nrows;
  const SysView v = sys.get(y, nrows);
  const int B = v.B;
  const int x = threadIdx.x;
  const int c = y * B + x;
  const int cold = v.wstate[kWarmColdReq];
  uint32_t cnt = 0, cin = 0;
  bool touched = false;
  if (x < B) {
    const int ci = v.wcin[c], co = v.wcout[c];
    const int h0 = v.head[c], h1 = v.head[c + 1];
    touched = (ci | co) != 0;
    if (touched) {
      v.wcin[c] = 0;
      v.wcout[c] = 0;
      v.wtouch[c] = 1;
      v.wtold[c] = make_int2(h0, h1 - h0);
    }
    cnt = (uint32_t)(h1 - h0 - co + ci);
    cin = (uint32_t)ci;
  }
  const int nocc = __syncthreads_count(cnt > 0);
  const int ntouch = __syncthreads_count(touched);
  uint32_t tot, itot;
  const uint32_t e = block_exclusive_scan<NT>(cnt, sh, &tot);
  const uint32_t ei = block_exclusive_scan<NT>(cin, sh, &itot);
  if (x < B) {
    v.wcnew[c] = (int)e;
    v.winoff[c] = (int)ei;
  }
  if (x == 0) {
    v.wrowtot[y] = (int)tot;
    v.wrowtot[B + 1 + y] = (int)itot;
    v.wrowtot[2 * (B + 1) + y] = nocc;
    v.wrowtot[3 * (B + 1) + y] = ntouch;
    v.wstate[kWarmCold] = cold;  // every row block writes the same value
  }
  if (cold) {
    // cold switch: the row histograms of the cold row pass (k_row_up), the
    // system's tiles shared among its row blocks
    const int ntiles = (v.n + kSortTile - 1) / kSortTile;
    for (int t = y; t < ntiles; t += nrows) row_up_tile<NT>(v, io, t, ntiles, s_hist);
  }
}

// One block per system: exclusive scans of the row totals of the new
// counts (-> rowstart, head[B*B]), of the movers, the occupied and the
// touched leaves (in place in wrowtot; the last two totals -> nocc and the
// touched-leaf count), when k_warm_finish does not scan them itself
// (batches: one scan per system instead of one per row block).
template <int NT, class Sys>
__global__ __launch_bounds__(NT) void k_warm_rowscan(Sys sys) {
  __shared__ uint32_t sh[NT / kWave];
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int B = v.B;
  const int x = threadIdx.x;
  const uint32_t t = x < B ? (uint32_t)v.wrowtot[x] : 0u;
  const uint32_t ti = x < B ? (uint32_t)v.wrowtot[B + 1 + x] : 0u;
  const uint32_t to = x < B ? (uint32_t)v.wrowtot[2 * (B + 1) + x] : 0u;
  const uint32_t tt = x < B ? (uint32_t)v.wrowtot[3 * (B + 1) + x] : 0u;
  uint32_t tot, itot, otot, ttot;
  const uint32_t e = block_exclusive_scan<NT>(t, sh, &tot);
  const uint32_t ei = block_exclusive_scan<NT>(ti, sh, &itot);
  const uint32_t eo = block_exclusive_scan<NT>(to, sh, &otot);
  const uint32_t et = block_exclusive_scan<NT>(tt, sh, &ttot);
  if (x < B) {
    v.rowstart[x] = (int)e;
    v.wrowtot[B + 1 + x] = (int)ei;
    v.wrowtot[2 * (B + 1) + x] = (int)eo;
    v.wrowtot[3 * (B + 1) + x] = (int)et;
  }
  if (x == 0) {
    v.rowstart[B] = (int)tot;
    v.head[B * B] = (int)tot;
    *v.nocc = (int)otot;
    v.wstate[kWarmTouched] = (int)ttot;
  }
}

// One block per row: the row's start and its bases in the occupied and the
// touched leaf lists (scanned here from the B row totals, or read from
// rowstart and wrowtot when k_warm_rowscan ran), then head and the movers'
// group offsets for the row; writes the row's occupied leaves to occ and its
// touched leaves to wtlist, both in row-major order.  zero_empty: also clear
// the level-0 multipole of the row's empty leaves.
template <int NT, class Sys>
__global__ __launch_bounds__(NT) void k_warm_finish(Sys sys, CallIO io, int zero_empty, int prescanned) {
  __shared__ uint32_t sh[NT / kWave];
  __shared__ uint32_t s_base, s_ibase, s_occ_base, s_touch_base, s_carry;
  int y, nrows;
  const SysView v = sys.get(y, nrows);
  const int B = v.B;
  const int x = threadIdx.x;
  const int c = y * B + x;
  // the row's own loads first, so that their latency overlaps the scans
  const int cold = v.wstate[kWarmCold];
  int e = 0, e1 = 0, win = 0;
  bool tch = false;
  if (x < B) {
    e = v.wcnew[c];
    e1 = x + 1 < B ? v.wcnew[c + 1] : v.wrowtot[y];
    win = v.winoff[c];
    tch = v.wtouch[c] != 0;
  }
  if (prescanned) {
    if (x == 0) {
      s_base = (uint32_t)v.rowstart[y];
      s_ibase = (uint32_t)v.wrowtot[B + 1 + y];
      s_occ_base = (uint32_t)v.wrowtot[2 * (B + 1) + y];
      s_touch_base = (uint32_t)v.wrowtot[3 * (B + 1) + y];
    }
  } else {
    // exclusive scans of the row totals, up to row y
    const uint32_t t = x < B ? (uint32_t)v.wrowtot[x] : 0u;
    const uint32_t ti = x < B ? (uint32_t)v.wrowtot[B + 1 + x] : 0u;
    const uint32_t to = x < B ? (uint32_t)v.wrowtot[2 * (B + 1) + x] : 0u;
    const uint32_t tt = x < B ? (uint32_t)v.wrowtot[3 * (B + 1) + x] : 0u;
    uint32_t tot, itot, otot, ttot;
    const uint32_t e = block_exclusive_scan<NT>(t, sh, &tot);
    const uint32_t ei = block_exclusive_scan<NT>(ti, sh, &itot);
    const uint32_t eo = block_exclusive_scan<NT>(to, sh, &otot);
    const uint32_t et = block_exclusive_scan<NT>(tt, sh, &ttot);
    if (x == y) {
      s_base = e;
      s_ibase = ei;
      s_occ_base = eo;
      s_touch_base = et;
      v.rowstart[y] = (int)e;
    }
    if (y == B - 1 && x == 0) {
      v.rowstart[B] = (int)tot;
      v.head[B * B] = (int)tot;
      // with the cold switch on, the column pass (next kernel) counts the
      // occupied leaves from zero: row 0's block clears the count in this
      // kernel too (row_scan_row), so both blocks must store 0, or the
      // column pass's list bases start past the count and its writes run
      // off the end of occ
      *v.nocc = cold ? 0 : (int)otot;
      v.wstate[kWarmTouched] = (int)ttot;
    }
  }
  __syncthreads();
  const uint32_t cnt = (uint32_t)(e1 - e);  // 0 past the row
  if (x < B) {
    v.head[c] = (int)s_base + e;
    v.winoff[c] = win + (int)s_ibase;
    if (zero_empty && cnt == 0) v.mpole[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  }
  const uint32_t occ = cnt > 0 ? 1u : 0u;
  const uint32_t touched = tch ? 1u : 0u;
  uint32_t nocc, ntouch;
  const uint32_t eo = block_exclusive_scan<NT>(occ, sh, &nocc);
  const uint32_t et = block_exclusive_scan<NT>(touched, sh, &ntouch);
  if (occ) v.occ[s_occ_base + eo] = c;
  if (touched) v.wtlist[s_touch_base + et] = c;
  // k_warm_rows has read the first pass's request: clear it for the next sort
  if (y == 0 && x == 0) v.wstate[kWarmColdReq] = 0;
  if (cold) {
    // cold switch: the row scan of the cold row pass (k_row_scan), the rows
    // shared among the system's row blocks
    for (int d = y; d < B; d += nrows) row_scan_row<NT>(v, d, sh, &s_carry);
  }
}

// Small systems: k_warm_rows and k_warm_finish (and a batch's
// k_warm_rowscan) as one block per system, each thread taking `per`
// consecutive leaves in row-major order, so that the scans over all leaves
// need no second launch.  It writes what those kernels leave for the passes
// after them: the same head, rowstart, movers' group offsets, occupied and
// touched leaf lists, counts and cleared counters (the row-local offsets and
// row totals that only pass between them are not kept).  With the cold switch
// on it also makes the cold row histograms, every tile's at once in LDS, and
// the row scans, a row per wave.  Bounds: at most kWarmOneMaxPer leaves per
// thread (64 x 64 leaves) and kWarmOneMaxTiles tiles of kSortTile particles.
// A solver takes it up to 32 x 32 leaves and 65536 particles (one leaf per
// thread: the 16384-particle disk's step 49.5 -> 47.3 us; with 16 consecutive
// leaves per thread on 128 x 128 leaves one block's scattered loads took
// longer than the two kernels, the Gaussian disk's 50 steps 3.67 -> 6.27 ms),
// a batch whenever every system is within the bounds.
constexpr int kWarmOneThreads = 1024;
constexpr int kWarmOneMaxPer = 4;
constexpr int kWarmOneMaxTiles = 64;
constexpr int kWarmOneMaxN = 65536;  // a solver's limit (a batch's: kWarmOneMaxTiles * kSortTile)

template <class Sys>
__global__ __launch_bounds__(kWarmOneThreads) void k_warm_rows_finish_one(Sys sys, CallIO io, int zero_empty) {
  __shared__ uint32_t sh[kWarmOneThreads / kWave];
  __shared__ uint32_t s_hist[kWarmOneMaxTiles * kWave];  // cold switch: [tile][row] counts, B <= 64
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int B = v.B, C = B * B;
  const int per = C > kWarmOneThreads ? C / kWarmOneThreads : 1;
  const int c0 = (int)threadIdx.x * per;
  const int cold = v.wstate[kWarmColdReq];
  // k_warm_rows: new counts; a leaf that movers enter or leave is touched
  // and keeps its old slot range
  int cnt[kWarmOneMaxPer], cin[kWarmOneMaxPer];
  uint32_t tmask = 0, scnt = 0, scin = 0, socc = 0, stch = 0;
#pragma unroll
  for (int j = 0; j < kWarmOneMaxPer; ++j) {
    cnt[j] = 0;
    cin[j] = 0;
    const int c = c0 + j;
    if (j < per && c < C) {
      const int ci = v.wcin[c], co = v.wcout[c];
      const int h0 = v.head[c], h1 = v.head[c + 1];
      if ((ci | co) != 0) {
        v.wcin[c] = 0;
        v.wcout[c] = 0;
        v.wtouch[c] = 1;
        v.wtold[c] = make_int2(h0, h1 - h0);
        tmask |= 1u << j;
        ++stch;
      }
      cnt[j] = h1 - h0 - co + ci;
      cin[j] = ci;
      scnt += (uint32_t)cnt[j];
      scin += (uint32_t)ci;
      socc += cnt[j] > 0 ? 1u : 0u;
    }
  }
  // k_warm_finish: exclusive scans over all leaves (their barriers also
  // order every read of the old head above before the writes below)
  uint32_t tot, itot, otot, ttot;
  uint32_t e = block_exclusive_scan<kWarmOneThreads>(scnt, sh, &tot);
  uint32_t ei = block_exclusive_scan<kWarmOneThreads>(scin, sh, &itot);
  uint32_t eo = block_exclusive_scan<kWarmOneThreads>(socc, sh, &otot);
  uint32_t et = block_exclusive_scan<kWarmOneThreads>(stch, sh, &ttot);
#pragma unroll
  for (int j = 0; j < kWarmOneMaxPer; ++j) {
    const int c = c0 + j;
    if (j < per && c < C) {
      v.head[c] = (int)e;
      v.winoff[c] = (int)ei;
      if ((c & (B - 1)) == 0) v.rowstart[c >> v.lgB] = (int)e;
      if (cnt[j] > 0)
        v.occ[eo++] = c;
      else if (zero_empty)
        v.mpole[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if ((tmask >> j) & 1u) v.wtlist[et++] = c;
      e += (uint32_t)cnt[j];
      ei += (uint32_t)cin[j];
    }
  }
  if (threadIdx.x == 0) {
    v.rowstart[B] = (int)tot;
    v.head[C] = (int)tot;
    // with the cold switch on, the column pass counts the occupied leaves
    // from zero
    *v.nocc = cold ? 0 : (int)otot;
    v.wstate[kWarmTouched] = (int)ttot;
    v.wstate[kWarmCold] = cold;
    v.wstate[kWarmColdReq] = 0;  // read by every thread before the scans
  }
  if (cold) {
    // the cold row pass's histograms (k_row_up) of every tile at once, then
    // its scans (k_row_scan): one wave per row takes the row's tile counts to
    // exclusive offsets (hist) and its total (dtotal)
    const int ntiles = (v.n + kSortTile - 1) / kSortTile;
    for (int k = (int)threadIdx.x; k < ntiles * B; k += kWarmOneThreads) s_hist[k] = 0;
    __syncthreads();
    const float2* pos0 = io.pos0 + v.part_off;
    bool any_bad = false;
    for (int i = (int)threadIdx.x; i < v.n; i += kWarmOneThreads) {
      bool bad;
      const uint32_t key = bin_key(v, pos0[i], bad);
      any_bad |= bad;
      atomicAdd(&s_hist[(i / kSortTile) * B + (int)(key >> v.lgB)], 1u);
    }
    if (any_bad) atomicOr(v.flags, 1u);
    __syncthreads();
    const int lane = lane_id();
    for (int d = (int)threadIdx.x / kWave; d < B; d += kWarmOneThreads / kWave) {
      const uint32_t k = lane < ntiles ? s_hist[lane * B + d] : 0u;
      const uint32_t inc = wave_inclusive_scan(k);
      if (lane < ntiles) v.hist[(size_t)d * ntiles + lane] = inc - k;
      if (lane == kWave - 1) v.dtotal[d] = inc;
    }
  }
}

bool warm_rows_one_fits(int B, int n) {
  return B * B <= kWarmOneThreads * kWarmOneMaxPer && B <= kWave && n <= kWarmOneMaxTiles * kSortTile;
}

// Stayers of untouched leaves go to head[leaf] + their old rank; movers are
// put into their new leaf's group (slots of the group in any order).
template <class Sys>
__global__ __launch_bounds__(256) void k_warm_place(Sys sys, CallIO io) {
  __shared__ RowDownLeanLds lds;
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int i = blk * 256 + threadIdx.x;
  const bool valid = i < v.n;
  // the particle's own entries are loaded together with its key (a stayer,
  // nearly every particle, needs them; a mover's are simply unused) and the
  // cold switch
  const int cold = v.wstate[kWarmCold];
  const uint32_t t = valid ? v.tkey[i] : 0u;
  const int rel = valid ? v.wrel[i] : 0;
  const float4 p = valid ? v.tpart[i] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (cold) {
    // cold switch: the scatter of the cold row pass (k_row_down), the tiles
    // shared among the system's blocks
    const int ntiles = (v.n + kSortTile - 1) / kSortTile;
    for (int tt = blk; tt < ntiles; tt += nblk) row_down_tile_lean(v, io, tt, ntiles, lds);
    return;
  }
  if (blk * 256 >= v.n) return;  // whole waves below stay for the movers' slots
  const bool mover = valid && (t & kMoverBit);
  const uint32_t k = t & ~kMoverBit;
  // a slot in the new leaf's group (slots of a group in any order); a leaf
  // takes at most kMergeLdsMovers movers here (more switch the sort to cold),
  // so at most that many atomics meet on one address
  if (mover) v.wmgrp[v.winoff[k] + atomicAdd(&v.wfill[k], 1)] = i;
  if (valid && !mover && !v.wtouch[t]) {
    const int dst = v.head[t] + rel;
    v.key[dst] = t;
    v.part[dst] = p;
  }
}

// Original index of the particle a tpart entry holds.
__device__ __forceinline__ uint32_t orig_of(const float4& p) { return __float_as_uint(p.w); }

// One wave per touched leaf (grid-stride over the list): the leaf's stayers
// (old slots [first, first+count) whose tkey is still this leaf, already in
// original-index order) and its incoming movers (grouped at wmgrp[winoff])
// are merged by original index; indices are unique, so counting the smaller
// ones gives each particle's exact slot.  Up to 64 movers (one per lane) are
// ranked in one pass over the stayers.  Up to kMergeLdsMovers movers are
// sorted by original index in the wave's LDS (bitonic); each stayer then
// finds the movers below it by binary search (p = their count) and counts
// itself in H[p], so the stayers below the mover of rank r are
// H[0] + ... + H[r].  A leaf with more movers never gets here: the sort
// switched to the cold passes, and then this kernel only clears the touched
// leaves' flags and makes the cold sort's column pass (k_row_sort).
// Clears the leaf's touch flag and group count.
union MergeLds {
  struct {
    uint2 mv[256 / kWave][kMergeLdsMovers];  // (original index, tpart slot), sorted
    uint32_t h[256 / kWave][kMergeLdsMovers + 1];
  } m;
  RowSortLeanLds<256> r;
};

template <class Sys>
__global__ __launch_bounds__(256) void k_warm_merge(Sys sys) {
  __shared__ MergeLds lds;
  auto& s_mv = lds.m.mv;
  auto& s_h = lds.m.h;
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int lane = lane_id();
  const int wave = (int)(threadIdx.x / kWave);
  const int waves = nblk * (256 / kWave);
  const int nt = v.wstate[kWarmTouched];
  const int cold = v.wstate[kWarmCold];
  const uint64_t below = lanes_below();
  const int cmax = v.B * v.B - 1;  // wtlist holds at most B * B leaves
  int d = blk * (256 / kWave) + wave;
  // the first list entry is loaded together with the count and the cold
  // switch (clamped; used only when d < nt), the next one while this leaf is
  // merged
  uint32_t cnext = (uint32_t)v.wtlist[min(d, cmax)];
  if (cold) {
    for (; d < nt; d += waves) {
      const int c = v.wtlist[d];
      if (lane == 0) {
        v.wtouch[c] = 0;
        v.wfill[c] = 0;
      }
    }
    for (int y = blk; y < v.B; y += nblk) row_sort_row_lean<256>(v, y, 1, lds.r);
    return;
  }
  for (; d < nt; d += waves) {
    const uint32_t c = cnext;
    cnext = (uint32_t)v.wtlist[min(d + waves, cmax)];
    const int2 od = v.wtold[c];
    const int h = v.head[c];
    const int ni = v.wfill[c];
    const int* grp = v.wmgrp + v.winoff[c];
    if (ni > kWave) {  // ni <= kMergeLdsMovers: a leaf with more switched the sort to cold
      uint2* mv = s_mv[wave];
      uint32_t* H = s_h[wave];
      int P = kWave;
      while (P < ni) P <<= 1;
      for (int k = lane; k < P; k += kWave) {
        const int ms = k < ni ? grp[k] : 0;
        mv[k] = make_uint2(k < ni ? orig_of(v.tpart[ms]) : 0xffffffffu, (uint32_t)ms);
      }
      for (int k = lane; k <= ni; k += kWave) H[k] = 0u;
      wave_lds_order();
      // bitonic sort of P entries by original index (padding sorts last)
      for (int kk = 2; kk <= P; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
          for (int t = lane; t < P / 2; t += kWave) {
            const int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
            const int l = i | j;
            const uint2 a = mv[i], b = mv[l];
            if ((a.x > b.x) == ((i & kk) == 0)) {
              mv[i] = b;
              mv[l] = a;
            }
          }
          wave_lds_order();
        }
      }
      // stayers: rank among the stayers plus the movers below; the next
      // chunk's loads are in flight while this one is placed (indices
      // clamped to the old range's last slot, od.y > 0 inside the loop)
      int before = 0;
      const int olast = od.x + od.y - 1;
      uint32_t tk = 0u;
      float4 tp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (od.y > 0) {
        tk = v.tkey[min(od.x + lane, olast)];
        tp = v.tpart[min(od.x + lane, olast)];
      }
      for (int b = 0; b < od.y; b += kWave) {
        const uint32_t tkc = tk;
        const float4 tpc = tp;
        tk = v.tkey[min(od.x + b + kWave + lane, olast)];
        tp = v.tpart[min(od.x + b + kWave + lane, olast)];
        const bool stay = b + lane < od.y && tkc == c;
        const uint32_t os = stay ? orig_of(tpc) : 0u;
        int p = 0;
        for (int step = P; step >= 1; step >>= 1)
          if (p + step <= ni && mv[p + step - 1].x < os) p += step;
        const uint64_t sm = __ballot(stay);
        if (stay) {
          atomicAdd(&H[p], 1u);
          const int dst = h + before + __popcll(sm & below) + p;
          v.key[dst] = c;
          v.part[dst] = tpc;
        }
        before += __popcll(sm);
      }
      wave_lds_order();
      // movers in rank order: rank plus the stayers below (prefix of H)
      uint32_t carry = 0;
      for (int r0 = 0; r0 < ni; r0 += kWave) {
        const int r = r0 + lane;
        const uint32_t inc = wave_inclusive_scan(r < ni ? H[r] : 0u) + carry;
        if (r < ni) {
          const int dst = h + r + (int)inc;
          v.key[dst] = c;
          v.part[dst] = v.tpart[mv[r].y];
        }
        carry = __shfl(inc, kWave - 1, kWave);
      }
      wave_lds_order();
    } else {
      const int ms = lane < ni ? grp[lane] : 0;
      const uint32_t om = lane < ni ? orig_of(v.tpart[ms]) : 0xffffffffu;
      int rank_m = 0, cnt_m = 0;
      for (int l = 0; l < ni; ++l) rank_m += __shfl(om, l, kWave) < om ? 1 : 0;
      int before = 0;
      for (int b = 0; b < od.y; b += kWave) {
        const int slot = od.x + b + lane;
        const bool stay = b + lane < od.y && v.tkey[slot] == c;
        const uint32_t os = stay ? orig_of(v.tpart[slot]) : 0u;
        int cnt_s = 0;
        for (int l = 0; l < ni; ++l) {
          const uint32_t o = __shfl(om, l, kWave);
          cnt_s += o < os ? 1 : 0;
          const int lt = __popcll(__ballot(stay && os < o));
          if (lane == l) cnt_m += lt;
        }
        const uint64_t sm = __ballot(stay);
        if (stay) {
          const int dst = h + before + __popcll(sm & below) + cnt_s;
          v.key[dst] = c;
          v.part[dst] = v.tpart[slot];
        }
        before += __popcll(sm);
      }
      if (lane < ni) {
        const int dst = h + rank_m + cnt_m;
        v.key[dst] = c;
        v.part[dst] = v.tpart[ms];
      }
    }
    if (lane == 0) {
      v.wtouch[c] = 0;
      v.wfill[c] = 0;
    }
  }
}

// ---------------------------------------------------------------- launchers

hipError_t launch_predo(const SysView& v, const CallIO& io, int zero_empty, hipStream_t s) {
  if (v.n <= 0) {
    hipLaunchKernelGGL(k_empty, dim3(cdiv(v.B * v.B + 1, 256)), dim3(256), 0, s, v);
    return hipGetLastError();
  }
  const OneSys one{v};
  const int ntiles = cdiv(v.n, kSortTile);
  hipLaunchKernelGGL(k_row_up<OneSys>, dim3(ntiles), dim3(kSortThreads), 0, s, one, io);
  hipLaunchKernelGGL(k_row_scan<OneSys>, dim3(v.B), dim3(kScanThreads), 0, s, one);
  hipLaunchKernelGGL(k_row_down<OneSys>, dim3(ntiles), dim3(kSortThreads), 0, s, one, io);
  hipLaunchKernelGGL(k_row_sort<OneSys>, dim3(v.B), dim3(kRowThreads), 0, s, one, zero_empty);
  return hipGetLastError();
}

hipError_t launch_predo_batch(const BatchLaunch& b, const CallIO& io, hipStream_t s) {
  hipLaunchKernelGGL(k_row_up<ManySys>, dim3(b.tiles.blocks), dim3(kSortThreads), 0, s, b.tiles.sys, io);
  hipLaunchKernelGGL(k_row_scan<ManySys>, dim3(b.rows.blocks), dim3(kScanThreads), 0, s, b.rows.sys);
  hipLaunchKernelGGL(k_row_down<ManySys>, dim3(b.tiles.blocks), dim3(kSortThreads), 0, s, b.tiles.sys, io);
  hipLaunchKernelGGL(k_row_sort<ManySys>, dim3(b.rows.blocks), dim3(kRowThreads), 0, s, b.rows.sys, 1);
  return hipGetLastError();
}

// The warm row kernels take one thread per leaf of a row: blocks of the
// smallest of 64..512 threads that covers the widest row launched.
static int row_threads(int maxB) {
  int t = kWave;
  while (t < maxB) t <<= 1;
  return t;
}

template <class Sys>
static void launch_warm_rows(int nt, dim3 g, const Sys& sys, const CallIO& io, hipStream_t s) {
  switch (nt) {
    case 64: hipLaunchKernelGGL((k_warm_rows<64, Sys>), g, dim3(64), 0, s, sys, io); break;
    case 128: hipLaunchKernelGGL((k_warm_rows<128, Sys>), g, dim3(128), 0, s, sys, io); break;
    case 256: hipLaunchKernelGGL((k_warm_rows<256, Sys>), g, dim3(256), 0, s, sys, io); break;
    default: hipLaunchKernelGGL((k_warm_rows<512, Sys>), g, dim3(512), 0, s, sys, io); break;
  }
}

template <class Sys>
static void launch_warm_rowscan(int nt, dim3 g, const Sys& sys, hipStream_t s) {
  switch (nt) {
    case 64: hipLaunchKernelGGL((k_warm_rowscan<64, Sys>), g, dim3(64), 0, s, sys); break;
    case 128: hipLaunchKernelGGL((k_warm_rowscan<128, Sys>), g, dim3(128), 0, s, sys); break;
    case 256: hipLaunchKernelGGL((k_warm_rowscan<256, Sys>), g, dim3(256), 0, s, sys); break;
    default: hipLaunchKernelGGL((k_warm_rowscan<512, Sys>), g, dim3(512), 0, s, sys); break;
  }
}

template <class Sys>
static void launch_warm_finish(int nt, dim3 g, const Sys& sys, const CallIO& io, int zero_empty, int prescanned,
                               hipStream_t s) {
  switch (nt) {
    case 64: hipLaunchKernelGGL((k_warm_finish<64, Sys>), g, dim3(64), 0, s, sys, io, zero_empty, prescanned); break;
    case 128:
      hipLaunchKernelGGL((k_warm_finish<128, Sys>), g, dim3(128), 0, s, sys, io, zero_empty, prescanned);
      break;
    case 256:
      hipLaunchKernelGGL((k_warm_finish<256, Sys>), g, dim3(256), 0, s, sys, io, zero_empty, prescanned);
      break;
    default:
      hipLaunchKernelGGL((k_warm_finish<512, Sys>), g, dim3(512), 0, s, sys, io, zero_empty, prescanned);
      break;
  }
}

// Warm-started predo (key/part/head hold an earlier sort of the same n).
hipError_t launch_predo_warm(const SysView& v, const CallIO& io, int zero_empty, hipStream_t s, bool gathered) {
  if (v.n <= 0) return launch_predo(v, io, zero_empty, s);
  const OneSys one{v};
  const int nb = cdiv(v.n, 256);
  if (!gathered) hipLaunchKernelGGL(k_warm_gather<OneSys>, dim3(nb), dim3(256), 0, s, one, io);
  if (v.B * v.B <= kWarmOneThreads && v.n <= kWarmOneMaxN && warm_rows_one_fits(v.B, v.n)) {
    hipLaunchKernelGGL(k_warm_rows_finish_one<OneSys>, dim3(1), dim3(kWarmOneThreads), 0, s, one, io, zero_empty);
  } else {
    launch_warm_rows(row_threads(v.B), dim3(v.B), one, io, s);
    launch_warm_finish(row_threads(v.B), dim3(v.B), one, io, zero_empty, 0, s);
  }
  hipLaunchKernelGGL(k_warm_place<OneSys>, dim3(nb), dim3(256), 0, s, one, io);
  hipLaunchKernelGGL(k_warm_merge<OneSys>, dim3(cdiv(std::min(kWarmMergeWaves, v.B * v.B), 256 / kWave)),
                     dim3(256), 0, s, one);
  return hipGetLastError();
}

hipError_t launch_predo_warm_batch(const BatchLaunch& b, const CallIO& io, hipStream_t s, bool gathered) {
  if (!gathered)
    hipLaunchKernelGGL(k_warm_gather<ManySys>, dim3(b.adj_sort.blocks), dim3(256), 0, s, b.adj_sort.sys, io);
  if (b.warm_one) {
    hipLaunchKernelGGL(k_warm_rows_finish_one<ManySys>, dim3(b.one.blocks), dim3(kWarmOneThreads), 0, s, b.one.sys, io,
                       1);
  } else {
    const int nt = row_threads(b.max_B);
    launch_warm_rows(nt, dim3(b.rows.blocks), b.rows.sys, io, s);
    launch_warm_rowscan(nt, dim3(b.one.blocks), b.one.sys, s);
    launch_warm_finish(nt, dim3(b.rows.blocks), b.rows.sys, io, 1, 1, s);
  }
  hipLaunchKernelGGL(k_warm_place<ManySys>, dim3(b.adj_sort.blocks), dim3(256), 0, s, b.adj_sort.sys, io);
  hipLaunchKernelGGL(k_warm_merge<ManySys>, dim3(b.warm_merge.blocks), dim3(256), 0, s, b.warm_merge.sys);
  return hipGetLastError();
}

hipError_t launch_occ_from_head(const SysView& v, hipStream_t s) {
  hipLaunchKernelGGL(k_occ_from_head, dim3(1), dim3(kScanThreads), 0, s, v);
  return hipGetLastError();
}

}  // namespace nbfmm_amd
