// Quadtree passes of the monopole FMM: P2M, M2M, M2L, L2L.
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
  float4 out = make_float4(0.0f, 0.0f, pw, 0.0f);
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
  float4 out = make_float4(0.0f, 0.0f, sw, 0.0f);
  if (b != a) {
    out.x = sx / sw;
    out.y = sy / sw;
  }
  return out;
}

constexpr int kUpTile = 16;  // leaves per side of one M2M tile block

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
constexpr int kP2MThreads = 256;  // the adjoint's leaf kernel shares this grid (3 blocks per CU at 140 VGPRs)

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

template <class Sys>
__global__ __launch_bounds__(kP2MThreads) void k_p2m_occ(Sys sys) {
#pragma clang fp contract(off)
  __shared__ int s_coop[kP2MThreads];
  __shared__ int s_ncoop;
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int lane = lane_id();
  if (threadIdx.x == 0) s_ncoop = 0;
  __syncthreads();
  // one thread per leaf; empty leaves (cleared by the sort) are skipped
  const int c = blk * kP2MThreads + (int)threadIdx.x;
  int a = 0, b = 0;
  if (c < v.B * v.B) {
    a = v.head[c];
    b = v.head[c + 1];
  }
  const bool valid = b > a;
  const bool coop = valid && b - a > kP2MCoop;
  // the segments of [a, b): [a, first), then kP2MSeg at a time
  const int first = (a / kP2MSeg + 1) * kP2MSeg;
  if (valid && !coop) {
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
    v.mpole[c] = make_float4(sx / sw, sy / sw, sw, 0.0f);
  }
  if (coop) s_coop[atomicAdd(&s_ncoop, 1)] = c;
  __syncthreads();
  // the block's large leaves, one per wave at a time, every lane taking a
  // segment
  const int ncoop = s_ncoop;
  for (int q = (int)(threadIdx.x / kWave); q < ncoop; q += kP2MThreads / kWave) {
    const int hc = __builtin_amdgcn_readfirstlane(s_coop[q]);
    const int ha = __builtin_amdgcn_readfirstlane(v.head[hc]), hb = __builtin_amdgcn_readfirstlane(v.head[hc + 1]);
    const int hf = (ha / kP2MSeg + 1) * kP2MSeg;
    const int nseg = 1 + (hb - hf + kP2MSeg - 1) / kP2MSeg;  // hb > hf: more than kP2MSeg particles
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
    if (lane == 0) v.mpole[hc] = make_float4(sx / sw, sy / sw, sw, 0.0f);
  }
}

// M2M levels 1..min(L-1, 4) for one kUpTile x kUpTile tile of leaves, built
// in LDS from the tile's level-0 multipoles.
template <class Sys>
__global__ __launch_bounds__(kUpTile * kUpTile) void k_m2m_tile(Sys sys) {
  __shared__ float4 lv[2][kUpTile * kUpTile];
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int B = v.B;
  const int T = B < kUpTile ? B : kUpTile;
  const int tiles = B / T;
  const int X0 = (blk % tiles) * T, Y0 = (blk / tiles) * T;
  const int t = threadIdx.x;
  int lgT = 0;
  while ((1 << lgT) < T) ++lgT;
  if (t < T * T) lv[0][t] = v.mpole[(Y0 + t / T) * B + X0 + t % T];
  __syncthreads();
  const int top = (v.L - 1) < lgT ? (v.L - 1) : lgT;
  int cur = 0;
  for (int l = 1; l <= top; ++l) {
    const int Dc = T >> (l - 1), Dl = T >> l;
    if (t < Dl * Dl) {
      const int x = t % Dl, y = t / Dl;
      const float4* ch = lv[cur];
      const float4 m = m2m_combine(ch[(2 * y) * Dc + 2 * x], ch[(2 * y) * Dc + 2 * x + 1],
                                   ch[(2 * y + 1) * Dc + 2 * x], ch[(2 * y + 1) * Dc + 2 * x + 1]);
      lv[cur ^ 1][t] = m;
      v.mpole[v.lvl_off(l) + ((Y0 >> l) + y) * (B >> l) + (X0 >> l) + x] = m;
    }
    cur ^= 1;
    __syncthreads();
  }
}

// The M2M levels above the tiles (levels lgT+1..L-1), one block.
template <class Sys>
__global__ __launch_bounds__(1024) void k_m2m_top(Sys sys) {
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  int first = 1;
  while ((1 << (first - 1)) < (v.B < kUpTile ? v.B : kUpTile)) ++first;
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

int upward_tile_blocks(int L, int B) {
  if (L <= 1) return 0;
  const int T = B < kUpTile ? B : kUpTile;
  return (B / T) * (B / T);
}

int upward_top_blocks(int L, int B) {
  const int T = B < kUpTile ? B : kUpTile;
  int lgT = 0;
  while ((1 << lgT) < T) ++lgT;
  return (L - 1 > lgT) ? 1 : 0;
}

int m2l_blocks(const SysView& v) { return v.L > 0 ? cdiv(v.lvl_off(v.L), 256) : 0; }

int p2m_occ_blocks(int max_occ) { return cdiv(max_occ, kP2MThreads); }
int p2m_leaf_blocks(int B) { return cdiv(B * B, kP2MThreads); }

hipError_t launch_upward(const SysView& v, int max_occ, hipStream_t s) {
  if (v.L <= 0) return hipSuccess;
  const OneSys one{v};
  hipLaunchKernelGGL(k_p2m_occ<OneSys>, dim3(p2m_leaf_blocks(v.B)), dim3(kP2MThreads), 0, s, one);
  const int tb = upward_tile_blocks(v.L, v.B);
  if (tb) hipLaunchKernelGGL(k_m2m_tile<OneSys>, dim3(tb), dim3(kUpTile * kUpTile), 0, s, one);
  if (upward_top_blocks(v.L, v.B)) hipLaunchKernelGGL(k_m2m_top<OneSys>, dim3(1), dim3(1024), 0, s, one);
  return hipGetLastError();
}

hipError_t launch_tree_batch(const BatchLaunch& b, hipStream_t s) {
  if (b.p2m.blocks) hipLaunchKernelGGL(k_p2m_occ<ManySys>, dim3(b.p2m.blocks), dim3(kP2MThreads), 0, s, b.p2m.sys);
  if (b.m2m_tile.blocks)
    hipLaunchKernelGGL(k_m2m_tile<ManySys>, dim3(b.m2m_tile.blocks), dim3(kUpTile * kUpTile), 0, s, b.m2m_tile.sys);
  if (b.m2m_top.blocks)
    hipLaunchKernelGGL(k_m2m_top<ManySys>, dim3(b.m2m_top.blocks), dim3(1024), 0, s, b.m2m_top.sys);
  if (b.m2l.blocks) hipLaunchKernelGGL(k_m2l<ManySys>, dim3(b.m2l.blocks), dim3(256), 0, s, b.m2l.sys);
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
