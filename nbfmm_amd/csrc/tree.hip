// Quadtree passes of the monopole FMM: P2M, M2M, M2L, L2L.
//
// Multipoles are float4 {COM x, COM y, mass, particles in the cell}: the
// count (not in the reference, which keeps COM and mass) tells the fused M2L
// which cells hold particles.
//
// Reference: source/nbfmm/core/solver/{p2m,m2m,m2l,l2l}.cu.  The reference
// keeps every level in a sparse B x B x L cube; here each level l is a dense
// D_l x D_l array (D_l = B >> l), so every pass is a coalesced stream
// (multipoles float4 {COM x, COM y, mass, 0}, locals float2).  The reference
// cube layout is produced only by the buffer API (k_cube_read/k_cube_write).
//
// All tree arithmetic uses IEEE-rounded division and square root with FMA
// contraction off, in the same expression order as the reference, so these
// passes reproduce the CPU oracle bit for bit given the same inputs.
#include "common.hpp"
#include "launch.hpp"

namespace nbfmm_amd {

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// kernel_function.hpp:29-40 with correctly rounded sqrt and division.
__device__ __forceinline__ void kernel_accum_ieee(float tx, float ty, float sx, float sy, float m, float eps,
                                                  float& ax, float& ay) {
#pragma clang fp contract(off)
  const float dx = sx - tx;
  const float dy = sy - ty;
  const float r = sqrtf(dx * dx + dy * dy);
  const float tmp = m / (r * r * r + eps);
  ax += dx * tmp;
  ay += dy * tmp;
}

// P2M (p2m.cu:24-110, intended semantics): every leaf gets the centre of mass
// sum(m*x)/sum(m) and mass sum(m) of its particles, empty leaves (0,0) and 0.
// One thread per leaf, particles summed in sorted order.
__device__ __forceinline__ float4 p2m_leaf(const SysView& v, int c);
__device__ __forceinline__ float4 m2m_combine(float4 c00, float4 c10, float4 c01, float4 c11);

__global__ __launch_bounds__(256) void k_p2m(SysView v) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= v.B * v.B) return;
  v.mpole[c] = p2m_leaf(v, c);
}

// M2M (m2m.cu:22-54): parent COM = (sum of child COM*mass) / (sum of masses),
// children (0,0),(1,0),(0,1),(1,1) in that order; a massless parent gets
// COM (0,0).  One thread per parent cell of level l.
__global__ __launch_bounds__(256) void k_m2m(SysView v, int l) {
  const int D = v.B >> l;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= D * D) return;
  const int X = c % D, Y = c / D;
  const int Dc = D * 2;
  const float4* ch = v.mpole + v.lvl_off(l - 1);
  const float4 c00 = ch[(2 * Y) * Dc + 2 * X];
  const float4 c10 = ch[(2 * Y) * Dc + 2 * X + 1];
  const float4 c01 = ch[(2 * Y + 1) * Dc + 2 * X];
  const float4 c11 = ch[(2 * Y + 1) * Dc + 2 * X + 1];
  v.mpole[v.lvl_off(l) + c] = m2m_combine(c00, c10, c01, c11);
}

// Parent multipole from its four children (m2m.cu:40-53), the expression
// shared by k_m2m and the fused upward pass.
__device__ __forceinline__ float4 m2m_combine(float4 c00, float4 c10, float4 c01, float4 c11) {
#pragma clang fp contract(off)
  const float px = c00.x * c00.z + c10.x * c10.z + c01.x * c01.z + c11.x * c11.z;
  const float py = c00.y * c00.z + c10.y * c10.z + c01.y * c01.z + c11.y * c11.z;
  const float pw = c00.z + c10.z + c01.z + c11.z;
  // w: particles in the cell (counts are exact in float up to 2^24, and a
  // sum of nonnegative counts is nonzero exactly when one of them is)
  float4 out = make_float4(0.0f, 0.0f, pw, c00.w + c10.w + c01.w + c11.w);
  if (pw != 0.0f) {
    out.x = px / pw;
    out.y = py / pw;
  }
  return out;
}

// Leaf multipole from its sorted particle range (p2m.cu:24-74, intended):
// the range is cut at the multiples of kP2MSeg, each segment summed in order
// from 0, the segment sums added in order from 0 (k_p2m_occ's order; a leaf
// within one segment is summed plainly).
__device__ __forceinline__ float4 p2m_leaf(const SysView& v, int c) {
#pragma clang fp contract(off)
  const int a = v.head[c], b = v.head[c + 1];
  float sx = 0.0f, sy = 0.0f, sw = 0.0f;
  for (int s = a; s < b;) {
    const int e = min(b, (s / kP2MSeg + 1) * kP2MSeg);
    float px = 0.0f, py = 0.0f, pw = 0.0f;
    for (int i = s; i < e; ++i) {
      const float4 p = v.part[i];
      px += p.x * p.z;
      py += p.y * p.z;
      pw += p.z;
    }
    sx += px;
    sy += py;
    sw += pw;
    s = e;
  }
  float4 out = make_float4(0.0f, 0.0f, sw, (float)(b - a));
  if (b != a) {
    out.x = sx / sw;
    out.y = sy / sw;
  }
  return out;
}


// P2M of the occupied leaves (the sort already cleared the empty ones), in
// p2m_leaf's order.  The grid covers every leaf and reads its range from
// head directly (an empty leaf's lane does nothing): going through the
// occupied-leaf list would add two dependent loads (its length, then its
// entry) before the particle loads.  One lane per leaf walks its segments in order, one
// segment per step with the next segment's loads already in flight (indices
// clamped inside the segment, only the adds predicated).  A leaf of more
// than kP2MCoop particles (a dense cluster, whose serial walk would be the
// kernel's time) is listed in LDS instead, and the block's waves take the
// listed leaves one each: lane t sums segment t and the wave adds the
// segment sums in order.  Which wave takes a leaf does not change its bits.
constexpr int kP2MCoop = 256;

__device__ __forceinline__ void p2m_segment(const SysView& v, int s, int e, float& px, float& py, float& pw) {
#pragma clang fp contract(off)
  float4 p[kP2MSeg];
#pragma unroll
  for (int k = 0; k < kP2MSeg; ++k) p[k] = v.part[min(s + k, e - 1)];
  px = 0.0f;
  py = 0.0f;
  pw = 0.0f;
#pragma unroll
  for (int k = 0; k < kP2MSeg; ++k) {
    if (s + k < e) {
      px += p[k].x * p[k].z;
      py += p[k].y * p[k].z;
      pw += p[k].z;
    }
  }
}

// Multipole of an occupied leaf [a, b) of at most kP2MCoop particles in
// p2m_leaf's order: one lane walks its segments in order, one segment per
// step with the next segment's loads already in flight (indices clamped
// inside the segment, only the adds predicated).
__device__ __forceinline__ float4 p2m_leaf_lane(const SysView& v, int a, int b) {
#pragma clang fp contract(off)
  const int first = (a / kP2MSeg + 1) * kP2MSeg;
  float sx = 0.0f, sy = 0.0f, sw = 0.0f;
  int s = a, e = min(b, first);
  float4 next[kP2MSeg];
#pragma unroll
  for (int k = 0; k < kP2MSeg; ++k) next[k] = v.part[min(s + k, e - 1)];
  while (s < b) {
    float4 cur[kP2MSeg];
#pragma unroll
    for (int k = 0; k < kP2MSeg; ++k) cur[k] = next[k];
    const int s2 = e, e2 = min(b, e + kP2MSeg);
#pragma unroll
    for (int k = 0; k < kP2MSeg; ++k) next[k] = v.part[min(s2 + k, e2 - 1)];  // past the leaf: b - 1
    float px = 0.0f, py = 0.0f, pw = 0.0f;
#pragma unroll
    for (int k = 0; k < kP2MSeg; ++k) {
      if (s + k < e) {
        px += cur[k].x * cur[k].z;
        py += cur[k].y * cur[k].z;
        pw += cur[k].z;
      }
    }
    sx += px;
    sy += py;
    sw += pw;
    s = s2;
    e = e2;
  }
  return make_float4(sx / sw, sy / sw, sw, (float)(b - a));
}

// Multipole of a leaf [ha, hb) of more than kP2MSeg particles by a whole
// wave (wave-uniform arguments): lane t sums segment t, the wave adds the
// segment sums in order (p2m_leaf's order).  The result is in every lane.
__device__ __forceinline__ float4 p2m_leaf_wave(const SysView& v, int ha, int hb, int lane) {
#pragma clang fp contract(off)
  const int hf = (ha / kP2MSeg + 1) * kP2MSeg;
  const int nseg = 1 + (hb - hf + kP2MSeg - 1) / kP2MSeg;
  float sx = 0.0f, sy = 0.0f, sw = 0.0f;
  for (int t0 = 0; t0 < nseg; t0 += kWave) {
    const int t = t0 + lane;
    const int s = t == 0 ? ha : hf + (t - 1) * kP2MSeg;
    const int e = t == 0 ? hf : min(hb, s + kP2MSeg);
    float px = 0.0f, py = 0.0f, pw = 0.0f;
    if (t < nseg) p2m_segment(v, s, e, px, py, pw);
    const int cnt = min(kWave, nseg - t0);
    for (int u = 0; u < cnt; ++u) {
      sx += __shfl(px, u, kWave);
      sy += __shfl(py, u, kWave);
      sw += __shfl(pw, u, kWave);
    }
  }
  return make_float4(sx / sw, sy / sw, sw, (float)(hb - ha));
}

// Upward pass of the fused paths: one block per kUpTile x kUpTile tile of
// leaves (smaller grids: the whole grid), one thread per leaf.  P2M of the
// tile's occupied leaves (the sort already cleared the empty ones; leaves
// above kP2MCoop particles, a dense cluster whose serial walk would be the
// kernel's time, are listed in LDS and taken by the block's waves one each,
// which does not change their bits), then M2M levels 1..min(L-1, lgT) in LDS
// from those multipoles (m2m.cu:22-54), every level written out.  The levels
// above the tiles are made by the fused M2L (or k_m2m_top).
template <class Sys>
__global__ __launch_bounds__(kUpTile * kUpTile) void k_upward(Sys sys) {
  __shared__ float4 lv[2][kUpTile * kUpTile];
  __shared__ int s_coop[kUpTile * kUpTile];
  __shared__ int s_ncoop;
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int B = v.B;
  const int T = up_tile_dim(B);
  const int tiles = B / T;
  const int X0 = (blk % tiles) * T, Y0 = (blk / tiles) * T;
  const int t = threadIdx.x;
  const int lane = lane_id();
  if (t == 0) s_ncoop = 0;
  __syncthreads();
  const bool in = t < T * T;
  const int c = (Y0 + t / T) * B + X0 + t % T;
  int a = 0, b = 0;
  if (in) {
    a = v.head[c];
    b = v.head[c + 1];
  }
  const bool valid = b > a;
  const bool coop = valid && b - a > kP2MCoop;
  float4 m = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (valid && !coop) {
    m = p2m_leaf_lane(v, a, b);
    v.mpole[c] = m;
  }
  if (in) lv[0][t] = m;
  if (coop) s_coop[atomicAdd(&s_ncoop, 1)] = t;
  __syncthreads();
  const int ncoop = s_ncoop;
  for (int q = (int)(threadIdx.x / kWave); q < ncoop; q += (kUpTile * kUpTile) / kWave) {
    const int tq = __builtin_amdgcn_readfirstlane(s_coop[q]);
    const int cq = (Y0 + tq / T) * B + X0 + tq % T;
    const int ha = __builtin_amdgcn_readfirstlane(v.head[cq]), hb = __builtin_amdgcn_readfirstlane(v.head[cq + 1]);
    const float4 mq = p2m_leaf_wave(v, ha, hb, lane);
    if (lane == 0) {
      v.mpole[cq] = mq;
      lv[0][tq] = mq;
    }
  }
  __syncthreads();
  int lgT = 0;
  while ((1 << lgT) < T) ++lgT;
  const int top = (v.L - 1) < lgT ? (v.L - 1) : lgT;
  int cur = 0;
  for (int l = 1; l <= top; ++l) {
    const int Dc = T >> (l - 1), Dl = T >> l;
    if (t < Dl * Dl) {
      const int x = t % Dl, y = t / Dl;
      const float4* ch = lv[cur];
      const float4 mm = m2m_combine(ch[(2 * y) * Dc + 2 * x], ch[(2 * y) * Dc + 2 * x + 1],
                                    ch[(2 * y + 1) * Dc + 2 * x], ch[(2 * y + 1) * Dc + 2 * x + 1]);
      lv[cur ^ 1][t] = mm;
      v.mpole[v.lvl_off(l) + ((Y0 >> l) + y) * (B >> l) + (X0 >> l) + x] = mm;
    }
    cur ^= 1;
    __syncthreads();
  }
}

// Front of a small tree's step in one block (1 <= L, B <= kFrontMaxB,
// n <= kFrontMaxN): the cold sort (keys, particles, head, row starts,
// occupied leaves; out-of-box particles clamped and flagged as bin_key does)
// and the multipoles of every level, as the sort, the upward pass and the
// fused M2L's top levels write them, bit for bit (the same P2M and M2M
// expressions on the same sorted particles).  At these sizes each of those
// kernels is a few dependent memory round trips on an otherwise idle GPU;
// here one block runs them back to back, its barriers in place of kernel
// boundaries (predo.cu:26-135, p2m.cu:24-110, m2m.cu:22-54).
//
// Sort: stable counting sort by leaf.  Wave w owns the particles
// [w P, w P + P) in input order.  Pass 1 counts them per leaf in the wave's
// LDS histogram; the scan turns the histograms into each wave's first slot
// in every leaf (leaf start + the counts of the waves before it); pass 2
// walks the wave's particles again in order, 64 at a time, ranking equal
// leaves inside the 64 with ballots, and scatters them.
constexpr int kFrontThreads = 1024;
constexpr int kFrontWaves = kFrontThreads / kWave;
constexpr int kFrontMaxB = 32;
constexpr int kFrontMaxC = kFrontMaxB * kFrontMaxB;
// Measured crossover (10-step graph runs of a uniform disk, 32 x 32
// leaves): the block is faster up to about 6000 particles (29.9 against
// 36.0 us per step at 512, 36.8 against 40.2 at 4096) and slower beyond
// (72 against 53 at 16384), where one CU's memory pipeline spends about 64
// cycles on each wave access to 64 scattered addresses (the permutation's
// stores, the per-leaf P2M reads) that the general kernels spread over the
// whole GPU.
constexpr int kFrontMaxN = 4096;
constexpr int kFrontBatch = 4;  // 64-particle groups whose loads are issued together

struct FrontLds {
  union {
    uint32_t cnt[kFrontWaves][kFrontMaxC];   // sort: per-wave leaf counts, then slots
    float4 mp[(4 * kFrontMaxC) / 3 + 1];     // every level's multipoles (lvl_off layout)
  } u;
  int start[kFrontMaxC + 1];                 // head
  int coop[kFrontMaxC];                      // leaves above kP2MCoop particles
  int ncoop;
  uint32_t sh[kFrontWaves];
};

__global__ __launch_bounds__(kFrontThreads) void k_small_front(SysView v, CallIO io) {
#pragma clang fp contract(off)
  __shared__ FrontLds S;
  const int C = v.B * v.B;
  const int n = v.n;
  const int t = threadIdx.x;
  const int lane = lane_id();
  const int wave = t / kWave;
  const float2* pos0 = io.pos0 + v.part_off;
  const float* w0 = io.w0 + v.part_off;
  constexpr int kSpan = kFrontWaves * kWave;
  const int P = ((n + kSpan - 1) / kSpan) * kWave;
  const int lo = min(wave * P, n), hi = min(lo + P, n);
  if (t < C) {
#pragma unroll
    for (int w = 0; w < kFrontWaves; ++w) S.u.cnt[w][t] = 0;
  }
  if (t == 0) S.ncoop = 0;
  __syncthreads();
  // pass 1: leaf counts of the wave's particles
  bool bad_any = false;
  for (int i0 = lo; i0 < hi; i0 += kFrontBatch * kWave) {
    float2 q[kFrontBatch];
#pragma unroll
    for (int k = 0; k < kFrontBatch; ++k) {
      const int i = i0 + k * kWave + lane;
      q[k] = i < hi ? pos0[i] : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int k = 0; k < kFrontBatch; ++k) {
      const int i = i0 + k * kWave + lane;
      if (i < hi) {
        bool bad;
        const uint32_t key = bin_key(v, q[k], bad);
        bad_any |= bad;
        atomicAdd(&S.u.cnt[wave][key], 1u);
      }
    }
  }
  if (bad_any) atomicOr(v.flags, 1u);
  __syncthreads();
  // scan: the waves' slots in leaf t, head, row starts, occupied leaves
  {
    uint32_t tot = 0;
    if (t < C) {
#pragma unroll
      for (int w = 0; w < kFrontWaves; ++w) {
        const uint32_t k = S.u.cnt[w][t];
        S.u.cnt[w][t] = tot;
        tot += k;
      }
    }
    uint32_t all, nocc;
    const uint32_t e = block_exclusive_scan<kFrontThreads>(tot, S.sh, &all);
    const uint32_t eo = block_exclusive_scan<kFrontThreads>(tot > 0 ? 1u : 0u, S.sh, &nocc);
    if (t < C) {
#pragma unroll
      for (int w = 0; w < kFrontWaves; ++w) S.u.cnt[w][t] += e;
      S.start[t] = (int)e;
      v.head[t] = (int)e;
      if ((t & (v.B - 1)) == 0) v.rowstart[t >> v.lgB] = (int)e;
      if (tot > 0) v.occ[eo] = t;
    }
    if (t == 0) {
      S.start[C] = n;
      v.head[C] = n;
      v.rowstart[v.B] = n;
      *v.nocc = (int)nocc;
    }
  }
  __syncthreads();
  // pass 2: stable scatter
  {
    const uint64_t below = lanes_below();
    for (int i0 = lo; i0 < hi; i0 += kFrontBatch * kWave) {
      float2 q[kFrontBatch];
      float m[kFrontBatch];
#pragma unroll
      for (int k = 0; k < kFrontBatch; ++k) {
        const int i = i0 + k * kWave + lane;
        q[k] = i < hi ? pos0[i] : make_float2(0.0f, 0.0f);
        m[k] = i < hi ? w0[i] : 0.0f;
      }
#pragma unroll
      for (int k = 0; k < kFrontBatch; ++k) {
        const int i = i0 + k * kWave + lane;
        const bool valid = i < hi;
        bool bad;
        const uint32_t key = valid ? bin_key(v, q[k], bad) : 0u;
        const uint64_t mm = match_mask(key, valid, 2 * v.lgB);
        const uint32_t r = (uint32_t)__popcll(mm & below);
        uint32_t prior = 0;
        if (valid) prior = S.u.cnt[wave][key];
        wave_lds_order();
        if (valid && r == 0) S.u.cnt[wave][key] = prior + (uint32_t)__popcll(mm);
        wave_lds_order();
        if (valid) {
          const uint32_t p = prior + r;
          v.key[p] = key;
          v.part[p] = make_float4(q[k].x, q[k].y, m[k], __uint_as_float((uint32_t)i));
        }
      }
    }
  }
  __syncthreads();
  // P2M (the counts are dead: their LDS now holds the multipoles)
  if (t < C) {
    const int a = S.start[t], b = S.start[t + 1];
    if (b - a > kP2MCoop) {
      S.coop[atomicAdd(&S.ncoop, 1)] = t;
    } else {
      const float4 mt = b > a ? p2m_leaf_lane(v, a, b) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      S.u.mp[t] = mt;
      v.mpole[t] = mt;
    }
  }
  __syncthreads();
  const int ncoop = S.ncoop;
  for (int q = wave; q < ncoop; q += kFrontWaves) {
    const int cq = __builtin_amdgcn_readfirstlane(S.coop[q]);
    const int ha = __builtin_amdgcn_readfirstlane(S.start[cq]);
    const int hb = __builtin_amdgcn_readfirstlane(S.start[cq + 1]);
    const float4 mq = p2m_leaf_wave(v, ha, hb, lane);
    if (lane == 0) {
      S.u.mp[cq] = mq;
      v.mpole[cq] = mq;
    }
  }
  __syncthreads();
  // M2M, levels 1..L-1
  for (int l = 1; l < v.L; ++l) {
    const int D = v.B >> l, Dc = 2 * D;
    const float4* ch = S.u.mp + v.lvl_off(l - 1);
    if (t < D * D) {
      const int x = t % D, y = t / D;
      const float4 mm = m2m_combine(ch[(2 * y) * Dc + 2 * x], ch[(2 * y) * Dc + 2 * x + 1],
                                    ch[(2 * y + 1) * Dc + 2 * x], ch[(2 * y + 1) * Dc + 2 * x + 1]);
      S.u.mp[v.lvl_off(l) + t] = mm;
      v.mpole[v.lvl_off(l) + t] = mm;
    }
    __syncthreads();
  }
}

bool small_front_fits(int L, int n) { return L >= 1 && (2 << L) <= kFrontMaxB && n > 0 && n <= kFrontMaxN; }

hipError_t launch_small_front(const SysView& v, const CallIO& io, hipStream_t s) {
  if (!small_front_fits(v.L, v.n)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_small_front, dim3(1), dim3(kFrontThreads), 0, s, v, io);
  return hipGetLastError();
}

// The M2M levels above the tiles (levels lgT+1..L-1), one block.
template <class Sys>
__global__ __launch_bounds__(1024) void k_m2m_top(Sys sys) {
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  int first = 1;
  while ((1 << (first - 1)) < up_tile_dim(v.B)) ++first;
  for (int l = first; l < v.L; ++l) {
    const int D = v.B >> l, Dc = D * 2;
    const float4* ch = v.mpole + v.lvl_off(l - 1);
    for (int c = threadIdx.x; c < D * D; c += blockDim.x) {
      const int x = c % D, y = c / D;
      v.mpole[v.lvl_off(l) + c] = m2m_combine(ch[(2 * y) * Dc + 2 * x], ch[(2 * y) * Dc + 2 * x + 1],
                                              ch[(2 * y + 1) * Dc + 2 * x], ch[(2 * y + 1) * Dc + 2 * x + 1]);
    }
    __syncthreads();
  }
}

// M2L (m2l.cu:24-60), all levels in one launch: each target cell sums the
// kernel at its own COM over the children of its parent's 3x3 neighbourhood
// (window [p-2, p+4) per axis) that are not adjacent to it; y outer, x inner.
// Consecutive threads take consecutive cells of a row, so a wave shares the
// window rows.  The six cells of a window row are loaded together (clamped
// addresses, used only where the window cell is in the box and not
// adjacent), so a row costs one wait instead of one per interaction.
template <class Sys>
__global__ __launch_bounds__(256) void k_m2l(Sys sys) {
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int t = blk * blockDim.x + threadIdx.x;
  if (t >= v.lvl_off(v.L)) return;
  int l = 0;
  while (t >= v.lvl_off(l + 1)) ++l;
  const int D = v.B >> l;
  const int c = t - v.lvl_off(l);
  const int tx = c % D, ty = c / D;
  const float4* mp = v.mpole + v.lvl_off(l);
  const float4 tc = mp[c];
  const int px = tx & ~1, py = ty & ~1;
  const float eps = v.eps;
  float ax = 0.0f, ay = 0.0f;
  // window row wy + 1 is loaded while row wy is summed
  float4 nxt[6];
  {
    const int yc = min(max(py - 2, 0), D - 1);
#pragma clang loop unroll(full)
    for (int k = 0; k < 6; ++k) nxt[k] = mp[yc * D + min(max(px - 2 + k, 0), D - 1)];
  }
#pragma clang loop unroll(full)
  for (int wy = -2; wy < 4; ++wy) {
    const int y = py + wy;
    const bool row_in = y >= 0 && y < D;
    float4 sc[6];
#pragma clang loop unroll(full)
    for (int k = 0; k < 6; ++k) sc[k] = nxt[k];
    if (wy + 1 < 4) {
      const int yn = min(max(y + 1, 0), D - 1);
#pragma clang loop unroll(full)
      for (int k = 0; k < 6; ++k) nxt[k] = mp[yn * D + min(max(px - 2 + k, 0), D - 1)];
    }
#pragma clang loop unroll(full)
    for (int k = 0; k < 6; ++k) {
      const int x = px - 2 + k;
      if (row_in && x >= 0 && x < D && (abs(x - tx) > 1 || abs(y - ty) > 1))
        kernel_accum_ieee(tc.x, tc.y, sc[k].x, sc[k].y, sc[k].z, eps, ax, ay);
    }
  }
  v.local[t] = make_float2(ax, ay);
}


// One target's M2L sum over its 6 x 6 window in k_m2l's order (rows, then
// columns), over the 27 cells not adjacent to the target.  The three window
// rows that meet the target's 3 x 3 neighbourhood (rows 1-3 or 2-4 by the
// target's row parity; the choice is the same for a whole wave on levels at
// least 64 cells wide) evaluate only their three other cells: columns 0 and
// 5 and, by the target's column parity, 4 or 1.  A remaining cell that is
// outside the grid or empty enters as a massless source at the target
// itself, whose term is exactly +0 (d = 0, 0 / eps = 0), so the sum has
// k_m2l's bits (its running sum is never -0) while the evaluations of a row
// can overlap.  Rows are loaded one ahead.
__device__ __forceinline__ float2 m2l_window(const float4* mp, int D, int tx, int ty, float4 tc, float eps) {
  const int px = tx & ~1, py = ty & ~1;
  const int mid = (tx & 1) ? 1 : 4;  // the non-adjacent column between 0 and 5 in a neighbouring row
  float ax = 0.0f, ay = 0.0f;
  float4 nxt[6];
  {
    const int yc = min(max(py - 2, 0), D - 1);
#pragma clang loop unroll(full)
    for (int k = 0; k < 6; ++k) nxt[k] = mp[yc * D + min(max(px - 2 + k, 0), D - 1)];
  }
  auto term = [&](const float4& sc, int x, bool row_in) {
    const bool use = row_in && x >= 0 && x < D && sc.w != 0.0f;
    kernel_accum_ieee(tc.x, tc.y, use ? sc.x : tc.x, use ? sc.y : tc.y, use ? sc.z : 0.0f, eps, ax, ay);
  };
#pragma clang loop unroll(full)
  for (int r = 0; r < 6; ++r) {
    const int y = py - 2 + r;
    const bool row_in = y >= 0 && y < D;
    float4 sc[6];
#pragma clang loop unroll(full)
    for (int k = 0; k < 6; ++k) sc[k] = nxt[k];
    if (r + 1 < 6) {
      const int yn = min(max(y + 1, 0), D - 1);
#pragma clang loop unroll(full)
      for (int k = 0; k < 6; ++k) nxt[k] = mp[yn * D + min(max(px - 2 + k, 0), D - 1)];
    }
    if ((unsigned)(r - 1 - (ty & 1)) < 3u) {
      term(sc[0], px - 2, row_in);
      term(mid == 1 ? sc[1] : sc[4], px - 2 + mid, row_in);
      term(sc[5], px + 3, row_in);
    } else {
#pragma clang loop unroll(full)
      for (int k = 0; k < 6; ++k) term(sc[k], px - 2 + k, row_in);
    }
  }
  return make_float2(ax, ay);
}

// M2L of the fused paths (solve, step, batches).  Cells without particles
// get (0, 0) (the P2P reads the locals of occupied leaves and of their
// ancestors only) and empty source cells are skipped: the other cells' bits
// are k_m2l's.  When the tree has levels above the upward tiles, the
// system's first kM2LTopBlocks blocks each make them in LDS from the tiles'
// top level (m2m_combine; block 0 writes them out as k_m2m_top would) and
// share their cells' M2L; the other blocks take levels 0..lgT as k_m2l does,
// one thread per cell.
template <class Sys>
__global__ __launch_bounds__(256) void k_m2l_fused(Sys sys) {
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int lgT = up_tile_lg(v.B);
  const bool has_top = v.L - 1 > lgT;
  const float eps = v.eps;
  if (has_top && blk < kM2LTopBlocks) {
    // LDS: level lgT + 1 at 0 (made from level lgT in global memory), each
    // level above right after the one below; every top block makes them, and
    // block 0 writes them out
    __shared__ float4 lv[kTopLds];
    int D = v.B >> (lgT + 1);
    {
      const float4* ch = v.mpole + v.lvl_off(lgT);
      const int Dp = 2 * D;
      for (int c = threadIdx.x; c < D * D; c += blockDim.x) {
        const int x = c % D, y = c / D;
        const float4 m = m2m_combine(ch[(2 * y) * Dp + 2 * x], ch[(2 * y) * Dp + 2 * x + 1],
                                     ch[(2 * y + 1) * Dp + 2 * x], ch[(2 * y + 1) * Dp + 2 * x + 1]);
        lv[c] = m;
        if (blk == 0) v.mpole[v.lvl_off(lgT + 1) + c] = m;
      }
    }
    __syncthreads();
    int po = 0;
    for (int l = lgT + 2; l < v.L; ++l) {
      const int Dp = D, o = po + Dp * Dp;
      D >>= 1;
      for (int c = threadIdx.x; c < D * D; c += blockDim.x) {
        const int x = c % D, y = c / D;
        const float4* ch = lv + po;
        const float4 m = m2m_combine(ch[(2 * y) * Dp + 2 * x], ch[(2 * y) * Dp + 2 * x + 1],
                                     ch[(2 * y + 1) * Dp + 2 * x], ch[(2 * y + 1) * Dp + 2 * x + 1]);
        lv[o + c] = m;
        if (blk == 0) v.mpole[v.lvl_off(l) + c] = m;
      }
      po = o;
      __syncthreads();
    }
    // M2L of the top levels' cells, numbered across the levels: block b takes
    // cells b * blockDim.x + threadIdx.x + k * kM2LTopBlocks * blockDim.x
    const int ntop = v.lvl_off(v.L) - v.lvl_off(lgT + 1);
    for (int f = blk * blockDim.x + threadIdx.x; f < ntop; f += kM2LTopBlocks * blockDim.x) {
      int l = lgT + 1, off = 0;
      D = v.B >> l;
      while (f - off >= D * D) {
        off += D * D;
        ++l;
        D >>= 1;
      }
      const int c = f - off;
      const float4* mp = lv + off;
      v.local[v.lvl_off(l) + c] =
          mp[c].w != 0.0f ? m2l_window(mp, D, c % D, c / D, mp[c], eps) : make_float2(0.0f, 0.0f);
    }
    return;
  }
  const int t = (blk - (has_top ? kM2LTopBlocks : 0)) * blockDim.x + threadIdx.x;
  if (t >= v.lvl_off(has_top ? lgT + 1 : v.L)) return;
  int l = 0;
  while (t >= v.lvl_off(l + 1)) ++l;
  const int D = v.B >> l;
  const int c = t - v.lvl_off(l);
  const int tx = c % D, ty = c / D;
  const float4* mp = v.mpole + v.lvl_off(l);
  const float4 tc = mp[c];
  if (tc.w == 0.0f) {
    v.local[t] = make_float2(0.0f, 0.0f);
    return;
  }
  v.local[t] = m2l_window(mp, D, tx, ty, tc, eps);
}

// L2L (l2l.cu:21-36): level l cell += its parent's accumulated effect.
__global__ __launch_bounds__(256) void k_l2l(SysView v, int l) {
#pragma clang fp contract(off)
  const int D = v.B >> l;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= D * D) return;
  const int X = c % D, Y = c / D;
  const float2 p = v.local[v.lvl_off(l + 1) + (Y >> 1) * (D >> 1) + (X >> 1)];
  float2 e = v.local[v.lvl_off(l) + c];
  e.x = e.x + p.x;
  e.y = e.y + p.y;
  v.local[v.lvl_off(l) + c] = e;
}

// Reference cube [levels][B][B] <-> dense levels.
__global__ __launch_bounds__(256) void k_cube_read(SysView v, int which, void* dst, int levels) {
  const int B = v.B;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= levels * B * B) return;
  const int l = i / (B * B);
  const int rem = i - l * B * B;
  const int row = rem / B, col = rem % B;
  const int s = 1 << l;
  const bool slot = l < v.L && (row % s) == 0 && (col % s) == 0;
  const int D = B >> l;
  const int d = slot ? v.lvl_off(l) + (row / s) * D + (col / s) : 0;
  if (which == 0) {
    const float4 m = slot ? v.mpole[d] : make_float4(0, 0, 0, 0);
    ((float2*)dst)[i] = make_float2(m.x, m.y);
  } else if (which == 1) {
    ((float*)dst)[i] = slot ? v.mpole[d].z : 0.0f;
  } else {
    ((float2*)dst)[i] = slot ? v.local[d] : make_float2(0, 0);
  }
}

__global__ __launch_bounds__(256) void k_cube_write(SysView v, int which, const void* src, int levels) {
  const int B = v.B;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= levels * B * B) return;
  const int l = i / (B * B);
  const int rem = i - l * B * B;
  const int row = rem / B, col = rem % B;
  const int s = 1 << l;
  if (!(l < v.L && (row % s) == 0 && (col % s) == 0)) return;
  const int D = B >> l;
  const int d = v.lvl_off(l) + (row / s) * D + (col / s);
  if (which == 0) {
    const float2 p = ((const float2*)src)[i];
    v.mpole[d].x = p.x;
    v.mpole[d].y = p.y;
  } else if (which == 1) {
    v.mpole[d].z = ((const float*)src)[i];
  } else {
    v.local[d] = ((const float2*)src)[i];
  }
}

hipError_t launch_p2m(const SysView& v, hipStream_t s) {
  if (v.L <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_p2m, dim3(cdiv(v.B * v.B, 256)), dim3(256), 0, s, v);
  return hipGetLastError();
}

int up_blocks(int L, int B) {
  if (L <= 0) return 0;
  const int T = up_tile_dim(B);
  return (B / T) * (B / T);
}

int upward_top_blocks(int L, int B) { return (L - 1 > up_tile_lg(B)) ? 1 : 0; }

int m2l_blocks(const SysView& v) { return v.L > 0 ? cdiv(v.lvl_off(v.L), 256) : 0; }

int m2l_fused_blocks(const SysView& v) {
  if (v.L <= 0) return 0;
  const int lgT = up_tile_lg(v.B);
  if (v.L - 1 > lgT) return kM2LTopBlocks + cdiv(v.lvl_off(lgT + 1), 256);
  return cdiv(v.lvl_off(v.L), 256);
}


hipError_t launch_upward(const SysView& v, hipStream_t s, bool with_top) {
  if (v.L <= 0) return hipSuccess;
  const OneSys one{v};
  hipLaunchKernelGGL(k_upward<OneSys>, dim3(up_blocks(v.L, v.B)), dim3(kUpTile * kUpTile), 0, s, one);
  if (with_top && upward_top_blocks(v.L, v.B)) hipLaunchKernelGGL(k_m2m_top<OneSys>, dim3(1), dim3(1024), 0, s, one);
  return hipGetLastError();
}

hipError_t launch_m2l_fused(const SysView& v, hipStream_t s) {
  if (v.L <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_m2l_fused<OneSys>, dim3(m2l_fused_blocks(v)), dim3(256), 0, s, OneSys{v});
  return hipGetLastError();
}

hipError_t launch_tree_batch(const BatchLaunch& b, hipStream_t s, bool forward) {
  if (b.up.blocks) hipLaunchKernelGGL(k_upward<ManySys>, dim3(b.up.blocks), dim3(kUpTile * kUpTile), 0, s, b.up.sys);
  if (forward) {
    if (b.m2l.blocks) hipLaunchKernelGGL(k_m2l_fused<ManySys>, dim3(b.m2l.blocks), dim3(256), 0, s, b.m2l.sys);
  } else if (b.m2m_top.blocks) {
    hipLaunchKernelGGL(k_m2m_top<ManySys>, dim3(b.m2m_top.blocks), dim3(1024), 0, s, b.m2m_top.sys);
  }
  return hipGetLastError();
}

hipError_t launch_m2m(const SysView& v, hipStream_t s) {
  for (int l = 1; l < v.L; ++l) {
    const int D = v.B >> l;
    hipLaunchKernelGGL(k_m2m, dim3(cdiv(D * D, 256)), dim3(256), 0, s, v, l);
  }
  return hipGetLastError();
}

hipError_t launch_m2l(const SysView& v, hipStream_t s) {
  if (v.L <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_m2l<OneSys>, dim3(m2l_blocks(v)), dim3(256), 0, s, OneSys{v});
  return hipGetLastError();
}

hipError_t launch_l2l(const SysView& v, hipStream_t s) {
  for (int l = v.L - 2; l >= 0; --l) {
    const int D = v.B >> l;
    hipLaunchKernelGGL(k_l2l, dim3(cdiv(D * D, 256)), dim3(256), 0, s, v, l);
  }
  return hipGetLastError();
}

static int cube_levels(int L) { return L < 1 ? 1 : L; }

hipError_t launch_cube_read(const SysView& v, int which, void* dst, hipStream_t s) {
  const int levels = cube_levels(v.L);
  hipLaunchKernelGGL(k_cube_read, dim3(cdiv(levels * v.B * v.B, 256)), dim3(256), 0, s, v, which, dst, levels);
  return hipGetLastError();
}

hipError_t launch_cube_write(const SysView& v, int which, const void* src, hipStream_t s) {
  const int levels = cube_levels(v.L);
  hipLaunchKernelGGL(k_cube_write, dim3(cdiv(levels * v.B * v.B, 256)), dim3(256), 0, s, v, which, src, levels);
  return hipGetLastError();
}

}  // namespace nbfmm_amd
