// Adjoint (vector-Jacobian product) of the force evaluation, and the
// adjoint of one Verlet step.  No reference counterpart: the reference only
// evaluates forces (source/nbfmm/core/solver/solve.cu:11-27).
//
// Forward being differentiated, per sorted particle i in leaf c:
//   a_i = sum_{j in 3x3 leaves} K(x_i, x_j, m_j) + sum_l E_l(anc_l(c)),
//   E_l(t) = sum_{s in IL(t)} K(X_t, X_s, M_s),  K(t, s, m) = m d f(|d|),
//   d = s - t, f(r) = 1/(r^3 + eps);
// cell multipoles X (centre of mass) and M (mass) built by P2M/M2M.  The
// cell assignment is piecewise constant and is held fixed.
//
// With upstream g_i = dloss/da_i the gradients are, in gather form (the
// neighbour relation and the interaction lists are symmetric):
//   P2P^T:  gx_i += sum_j f h - 3 r f^2 d (d.h),  h = m_i g_j - m_j g_i,
//           gm_i += sum_j -f d.g_j                        (d = x_j - x_i)
//   M2L^T:  the same with (X, M, G) per cell, G = sum of g over the cell's
//           particles (L2P^T and L2L^T), giving dloss/dX_t, dloss/dM_t
//   M2M^T:  top-down, child c of parent p:
//           gX_c += (M_c/M_p) gX_p,  gM_c += gM_p + (X_c - X_p).gX_p / M_p
//   P2M^T:  gx_i += (m_i/M_c) gX_c,  gm_i += gM_c + (x_i - X_c).gX_c / M_c
// A self pair contributes exactly 0 (d = 0 and h = m_i g_i - m_i g_i = 0 with
// both products rounded).  In P2P^T a target of nonzero mass factors h as
// m_i (g_j - m_j g_i / m_i): fewer operations per pair, m_i applied once.
#include "common.hpp"
#include "launch.hpp"

namespace nbfmm_amd {

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// One pair of the gather-form adjoint; target (t, mt, gt), source (s, ms, gs).
// The 3 r f^2 (d.h) d term is summed apart, in (cx, cy), and folded in once
// per target by adjoint_fold: one multiply fewer per pair.
__device__ __forceinline__ void pair_adjoint(float tx, float ty, float mt, float gtx, float gty, float sx, float sy,
                                             float ms, float gsx, float gsy, float eps, float& gx, float& gy,
                                             float& cx, float& cy, float& gm) {
#pragma clang fp contract(off)
  const float dx = sx - tx;
  const float dy = sy - ty;
  const float r2 = __builtin_fmaf(dx, dx, dy * dy);
  const float r = __builtin_amdgcn_sqrtf(r2);
  const float f = __builtin_amdgcn_rcpf(__builtin_fmaf(r2, r, eps));
  const float hx = mt * gsx - ms * gtx;
  const float hy = mt * gsy - ms * gty;
  const float dh = __builtin_fmaf(dx, hx, dy * hy);
  const float c = (r * f) * (f * dh);
  gx = __builtin_fmaf(f, hx, gx);
  gy = __builtin_fmaf(f, hy, gy);
  cx = __builtin_fmaf(c, dx, cx);
  cy = __builtin_fmaf(c, dy, cy);
  const float dg = __builtin_fmaf(dx, gsx, dy * gsy);
  gm = __builtin_fmaf(-f, dg, gm);
}

// The same pair for a target of nonzero mass mt, given b = g_t / m_t:
// h = m_t (g_s - m_s b) = m_t h', and every term with h carries the factor
// m_t, which the caller applies once per target; two fused multiply-adds
// form h' where h takes six operations.  Here the self pair gives
// f_self h'_self (h'_self = g_t - m_t b is the rounding residue of b, not
// exactly 0), which the caller takes out again (adjoint_unself).
__device__ __forceinline__ void pair_adjoint_scaled(float tx, float ty, float bx, float by, float sx, float sy,
                                                    float ms, float gsx, float gsy, float eps, float& gx, float& gy,
                                                    float& cx, float& cy, float& gm) {
#pragma clang fp contract(off)
  const float dx = sx - tx;
  const float dy = sy - ty;
  const float r2 = __builtin_fmaf(dx, dx, dy * dy);
  const float r = __builtin_amdgcn_sqrtf(r2);
  const float f = __builtin_amdgcn_rcpf(__builtin_fmaf(r2, r, eps));
  const float hx = __builtin_fmaf(-ms, bx, gsx);
  const float hy = __builtin_fmaf(-ms, by, gsy);
  const float dh = __builtin_fmaf(dx, hx, dy * hy);
  const float c = (r * f) * (f * dh);
  gx = __builtin_fmaf(f, hx, gx);
  gy = __builtin_fmaf(f, hy, gy);
  cx = __builtin_fmaf(c, dx, cx);
  cy = __builtin_fmaf(c, dy, cy);
  const float dg = __builtin_fmaf(dx, gsx, dy * gsy);
  gm = __builtin_fmaf(-f, dg, gm);
}

// Takes the self pair's term out of the sums of pair_adjoint_scaled: the
// same operations as the pair with d = 0 (f = 1/eps, c = 0, dg = 0).
__device__ __forceinline__ void adjoint_unself(float mt, float gtx, float gty, float bx, float by, float eps, float& gx,
                                               float& gy) {
  const float f = __builtin_amdgcn_rcpf(__builtin_fmaf(0.0f, 0.0f, eps));
  gx = __builtin_fmaf(-f, __builtin_fmaf(-mt, bx, gtx), gx);
  gy = __builtin_fmaf(-f, __builtin_fmaf(-mt, by, gty), gy);
}

// gx -= 3 cx, gy -= 3 cy: the term pair_adjoint sums apart.
__device__ __forceinline__ void adjoint_fold(float& gx, float& gy, float cx, float cy) {
  gx = __builtin_fmaf(-3.0f, cx, gx);
  gy = __builtin_fmaf(-3.0f, cy, gy);
}

// Upstream gradient into sorted order, scaled: one thread per particle.
template <class Sys>
__global__ __launch_bounds__(256) void k_adj_sort(Sys sys, CallIO io) {
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int i = blk * blockDim.x + threadIdx.x;
  if (i >= v.n) return;
  const float2 g0 = io.gin[v.part_off + __float_as_uint(v.part[i].w)];
  // gscale < 0: the system's own tick^2 (Verlet step adjoint of a batch)
  const float sc = io.gscale < 0.0f ? v.tick * v.tick : io.gscale;
  v.gsort[i] = make_float2(sc * g0.x, sc * g0.y);
}

// Sum of the sorted upstream gradient over leaf c (L2P^T), in sorted order
// (0 for an empty leaf).  As in P2M, the time is the chain of the heaviest
// leaf, so the thread keeps the next kAdjLeafDepth values in flight while it
// adds the current ones; loads are unconditional (indices clamped to the
// leaf's last particle), only the adds are predicated.
constexpr int kAdjLeafDepth = 16;

__device__ __forceinline__ float2 adj_leaf_sum(const SysView& v, int c) {
#pragma clang fp contract(off)
  const int a = v.head[c], b = v.head[c + 1];
  float sx = 0.0f, sy = 0.0f;
  if (b > a) {
    const int last = b - 1;
    float2 next[kAdjLeafDepth];
#pragma unroll
    for (int k = 0; k < kAdjLeafDepth; ++k) next[k] = v.gsort[min(a + k, last)];
    for (int s = a; s < b; s += kAdjLeafDepth) {
      float2 cur[kAdjLeafDepth];
#pragma unroll
      for (int k = 0; k < kAdjLeafDepth; ++k) cur[k] = next[k];
#pragma unroll
      for (int k = 0; k < kAdjLeafDepth; ++k) next[k] = v.gsort[min(s + kAdjLeafDepth + k, last)];
      const int m = b - s;
#pragma unroll
      for (int k = 0; k < kAdjLeafDepth; ++k) {
        if (k < m) {
          sx += cur[k].x;
          sy += cur[k].y;
        }
      }
    }
  }
  return make_float2(sx, sy);
}

// Subtree sum of a parent from its four children (x0y0, x1y0, x0y1, x1y1).
__device__ __forceinline__ float2 adj_up_combine(float2 a, float2 b, float2 c, float2 d) {
#pragma clang fp contract(off)
  return make_float2(a.x + b.x + c.x + d.x, a.y + b.y + c.y + d.y);
}

// Subtree sums of the gradient, levels 0..lgT (L2P^T and L2L^T): one block
// per kUpTile x kUpTile tile of leaves (the upward pass's grid), one thread
// per leaf summing its particles, then the tile's levels in LDS.  The levels
// above the tiles are made by k_adj_m2l_fused.
template <class Sys>
__global__ __launch_bounds__(kUpTile * kUpTile) void k_adj_up(Sys sys) {
  __shared__ float2 lv[2][kUpTile * kUpTile];
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int B = v.B;
  const int T = adj_tile_dim(B);
  const int tiles = B / T;
  const int X0 = (blk % tiles) * T, Y0 = (blk / tiles) * T;
  const int t = threadIdx.x;
  if (t < T * T) {
    const int c = (Y0 + t / T) * B + X0 + t % T;
    const float2 g = adj_leaf_sum(v, c);
    v.gup[c] = g;
    lv[0][t] = g;
  }
  __syncthreads();
  const int lgT = adj_tile_lg(B);
  const int top = (v.L - 1) < lgT ? (v.L - 1) : lgT;
  int cur = 0;
  for (int l = 1; l <= top; ++l) {
    const int Dc = T >> (l - 1), Dl = T >> l;
    if (t < Dl * Dl) {
      const int x = t % Dl, y = t / Dl;
      const float2* ch = lv[cur];
      const float2 m = adj_up_combine(ch[(2 * y) * Dc + 2 * x], ch[(2 * y) * Dc + 2 * x + 1],
                                      ch[(2 * y + 1) * Dc + 2 * x], ch[(2 * y + 1) * Dc + 2 * x + 1]);
      lv[cur ^ 1][t] = m;
      v.gup[v.lvl_off(l) + ((Y0 >> l) + y) * (B >> l) + (X0 >> l) + x] = m;
    }
    cur ^= 1;
    __syncthreads();
  }
}

// M2L^T of cell c of level l: gradient wrt its COM and mass from the M2L
// interactions it takes part in (as target and as source); gu holds the
// level's subtree sums (global memory or a copy in LDS).  The six cells of a
// window row are loaded together (clamped addresses, used only where the
// window cell is in the box and not adjacent).
__device__ __forceinline__ void adj_m2l_cell(const SysView& v, int l, int c, const float2* gu) {
  const int D = v.B >> l;
  const int tx = c % D, ty = c / D;
  const float4* mp = v.mpole + v.lvl_off(l);
  const float4 tc = mp[c];
  const float2 tg = gu[c];
  const int px = tx & ~1, py = ty & ~1;
  const float eps = v.eps;
  float gx = 0.0f, gy = 0.0f, gm = 0.0f, cx = 0.0f, cy = 0.0f;
#pragma clang loop unroll(full)
  for (int wy = -2; wy < 4; ++wy) {
    const int y = py + wy;
    const bool row_in = y >= 0 && y < D;
    const int yc = min(max(y, 0), D - 1);
    float4 sc[6];
    float2 sg[6];
#pragma clang loop unroll(full)
    for (int k = 0; k < 6; ++k) {
      const int a = yc * D + min(max(px - 2 + k, 0), D - 1);
      sc[k] = mp[a];
      sg[k] = gu[a];
    }
#pragma clang loop unroll(full)
    for (int k = 0; k < 6; ++k) {
      const int x = px - 2 + k;
      if (row_in && x >= 0 && x < D && (abs(x - tx) > 1 || abs(y - ty) > 1))
        pair_adjoint(tc.x, tc.y, tc.z, tg.x, tg.y, sc[k].x, sc[k].y, sc[k].z, sg[k].x, sg[k].y, eps, gx, gy, cx, cy,
                     gm);
    }
  }
  adjoint_fold(gx, gy, cx, cy);
  v.gmp[v.lvl_off(l) + c] = make_float4(gx, gy, gm, 0.0f);
}

// M2L^T, all levels in one launch (the fused M2L's grid): when the tree has
// levels above the upward tiles, each of the system's first kM2LTopBlocks
// blocks makes their subtree sums in LDS, level by level from the tiles' top
// level (block 0 writes them out), and takes its share of their cells; the
// other blocks take the cells of levels 0..lgT, one thread per cell,
// consecutive threads on consecutive cells of a row.
template <class Sys>
__global__ __launch_bounds__(256) void k_adj_m2l_fused(Sys sys) {
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int lgT = adj_tile_lg(v.B);
  const bool has_top = v.L - 1 > lgT;
  if (has_top && blk < kM2LTopBlocks) {
    __shared__ float2 lv[kTopLds];  // level lgT + 1 at 0, each level above right after the one below
    int D = v.B >> (lgT + 1);
    {
      const float2* ch = v.gup + v.lvl_off(lgT);
      const int Dp = 2 * D;
      for (int c = threadIdx.x; c < D * D; c += blockDim.x) {
        const int x = c % D, y = c / D;
        const float2 m = adj_up_combine(ch[(2 * y) * Dp + 2 * x], ch[(2 * y) * Dp + 2 * x + 1],
                                        ch[(2 * y + 1) * Dp + 2 * x], ch[(2 * y + 1) * Dp + 2 * x + 1]);
        lv[c] = m;
        if (blk == 0) v.gup[v.lvl_off(lgT + 1) + c] = m;
      }
    }
    __syncthreads();
    int po = 0;
    for (int l = lgT + 2; l < v.L; ++l) {
      const int Dp = D, o = po + Dp * Dp;
      D >>= 1;
      for (int c = threadIdx.x; c < D * D; c += blockDim.x) {
        const int x = c % D, y = c / D;
        const float2* ch = lv + po;
        const float2 m = adj_up_combine(ch[(2 * y) * Dp + 2 * x], ch[(2 * y) * Dp + 2 * x + 1],
                                        ch[(2 * y + 1) * Dp + 2 * x], ch[(2 * y + 1) * Dp + 2 * x + 1]);
        lv[o + c] = m;
        if (blk == 0) v.gup[v.lvl_off(l) + c] = m;
      }
      po = o;
      __syncthreads();
    }
    // the top levels' cells, numbered across the levels: block b takes cells
    // b * blockDim.x + threadIdx.x + k * kM2LTopBlocks * blockDim.x
    const int ntop = v.lvl_off(v.L) - v.lvl_off(lgT + 1);
    for (int f = blk * blockDim.x + threadIdx.x; f < ntop; f += kM2LTopBlocks * blockDim.x) {
      int l = lgT + 1, off = 0;
      D = v.B >> l;
      while (f - off >= D * D) {
        off += D * D;
        ++l;
        D >>= 1;
      }
      adj_m2l_cell(v, l, f - off, lv + off);
    }
    return;
  }
  const int t = (blk - (has_top ? kM2LTopBlocks : 0)) * blockDim.x + threadIdx.x;
  if (t >= v.lvl_off(has_top ? lgT + 1 : v.L)) return;
  int l = 0;
  while (t >= v.lvl_off(l + 1)) ++l;
  adj_m2l_cell(v, l, t - v.lvl_off(l), v.gup + v.lvl_off(l));
}

// Gradient wrt the leaf's (X, M): the M2M^T chain from the top level down,
// per lane (lanes of one leaf load the same cells).
__device__ __forceinline__ float4 leaf_multipole_grad(const SysView& v, int cx, int cy) {
  int l = v.L - 1;
  float4 T = v.gmp[v.lvl_off(l) + (cy >> l) * (v.B >> l) + (cx >> l)];
  for (l = v.L - 2; l >= 0; --l) {
    const int ci = v.lvl_off(l) + (cy >> l) * (v.B >> l) + (cx >> l);
    const int pi = v.lvl_off(l + 1) + (cy >> (l + 1)) * (v.B >> (l + 1)) + (cx >> (l + 1));
    const float4 mc = v.mpole[ci], mp = v.mpole[pi];
    const float4 d = v.gmp[ci];
    float4 n = make_float4(d.x, d.y, d.z + T.z, 0.0f);
    if (mp.z != 0.0f) {
      const float ratio = mc.z / mp.z;
      n.x = __builtin_fmaf(ratio, T.x, d.x);
      n.y = __builtin_fmaf(ratio, T.y, d.y);
      n.z += ((mc.x - mp.x) * T.x + (mc.y - mp.y) * T.y) / mp.z;
    }
    T = n;
  }
  return T;
}

// P2P^T plus P2M^T, over the same runs as the forward P2P (64 consecutive
// sorted targets of one row per wavefront, one per lane; each lane loops
// over its own sources inside the staged union of the run's ranges, in
// sorted order), with the epilogue of the mode:
//   kAdjSolve: gpos[o] = dloss/dx_o, gw[o] = dloss/dm_o
//   kAdjStep:  Verlet adjoint; with lambda = gpos, pi = gpre:
//              lambda_k = 2 lambda + tick^2 (da/dx)^T lambda + pi,
//              pi_k = -lambda,  gw += tick^2 (da/dm)^T lambda
//              (the tick^2 scale was applied to the sorted upstream).
enum AdjMode : int { kAdjSolve = 0, kAdjStep = 1 };

// Three blocks of kAdjP2PWaves waves fit a CU's LDS, so registers are held
// to what six waves per SIMD allow (80 VGPRs; unbounded it takes 83).
template <int MODE, class Sys>
__global__ __launch_bounds__(kAdjP2PWaves * kWave) __attribute__((amdgpu_waves_per_eu(6))) void k_adj_p2p(Sys sys,
                                                                                                       CallIO io) {
  __shared__ float4 s_pos[kAdjP2PWaves][kP2PStage];
  __shared__ float2 s_g[kAdjP2PWaves][kP2PStage];
  __shared__ int s_runs[(1 << (kMaxLevel + 1)) + 1];
  __shared__ uint32_t sh[kAdjP2PWaves];
  __shared__ int s_next;
  int blk, nblk;
  const SysView v = sys.get(blk, nblk);
  const int lane = lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  float4* sp = s_pos[wave];
  float2* sg = s_g[wave];
  float2* gpos = io.gpos + v.part_off;
  float2* gpre = io.gpre + v.part_off;
  float* gw = io.gw + v.part_off;
  const int B = v.B;
  const float eps = v.eps;
  row_runs<kAdjP2PWaves * kWave>(v, s_runs, sh);
  if (threadIdx.x == 0) s_next = 0;
  __syncthreads();
  const int total = s_runs[B];
  const int lblk = xcd_block(blk, nblk);
  for (int j = lblk + block_take(&s_next) * nblk; j < total; j = lblk + block_take(&s_next) * nblk) {
    const int y = __builtin_amdgcn_readfirstlane(run_row_wave(s_runs, B, j, lane));
    const int T0 = v.head[y * B] + kWave * (j - s_runs[y]);
    const int nt = min(kWave, v.head[(y + 1) * B] - T0);
    const bool valid = lane < nt;
    const int ti = T0 + (valid ? lane : 0);
    const float4 pt = v.part[ti];
    const float2 gt = v.gsort[ti];
    const int xt = (int)(v.key[ti] & (uint32_t)(B - 1));
    const int xf = __shfl(xt, 0, kWave);
    const int xl = __shfl(xt, nt - 1, kWave);
    // a wave whose targets all have nonzero mass takes the scaled pair
    // (b = g_t / m_t, the factor m_t applied below); otherwise every lane
    // takes the plain one
    const bool scaled = __builtin_amdgcn_ballot_w64(valid && !(pt.z != 0.0f)) == 0;
    const float2 bt = scaled && valid ? make_float2(gt.x / pt.z, gt.y / pt.z) : make_float2(0.0f, 0.0f);
    float gx = 0.0f, gy = 0.0f, gm = 0.0f, cx = 0.0f, cy = 0.0f;
#pragma unroll
    for (int dr = -1; dr <= 1; ++dr) {
      const int r = y + dr;
      if (r < 0 || r >= B) continue;
      const int* hr = v.head + r * B;
      const int us = hr[xf > 0 ? xf - 1 : 0];
      const int ulen = hr[(xl < B - 1 ? xl + 1 : B - 1) + 1] - us;
      const int a = valid ? hr[xt > 0 ? xt - 1 : 0] - us : 0;
      const int b = valid ? hr[(xt < B - 1 ? xt + 1 : B - 1) + 1] - us : 0;
      for (int w0 = 0; w0 < ulen; w0 += kP2PStage) {
        const int cnt = min(kP2PStage, ulen - w0);
        // as the forward P2P: every load of the chunk in flight at once
        // (indices clamped to its last source, cnt >= 1), then unconditional
        // LDS writes; slots from cnt on are never read
        {
          float4 stp[kP2PStage / kWave];
          float2 stg[kP2PStage / kWave];
#pragma unroll
          for (int q = 0; q < kP2PStage / kWave; ++q) {
            const int k = us + w0 + min(q * kWave + lane, cnt - 1);
            stp[q] = v.part[k];
            stg[q] = v.gsort[k];
          }
#pragma unroll
          for (int q = 0; q < kP2PStage / kWave; ++q) {
            sp[q * kWave + lane] = stp[q];
            sg[q * kWave + lane] = stg[q];
          }
        }
        wave_lds_order();
        const int lo = min(max(a - w0, 0), cnt);
        const int hi = min(max(b - w0, 0), cnt);
        // four sources per iteration, added in the same order as one at a
        // time; loop control on the LDS address alone (pfull >= p, so the
        // unsigned address compare never sees a bound below the buffer)
        const float4* p = sp + lo;
        const float2* q = sg + lo;
        const float4* const pend = sp + hi;
        const float4* const pfull = sp + max(hi - 3, lo);
        if (scaled) {
          for (; p < pfull; p += 4, q += 4) {
            float4 s4[4];
            float2 g4[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              s4[u] = p[u];
              g4[u] = q[u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
              pair_adjoint_scaled(pt.x, pt.y, bt.x, bt.y, s4[u].x, s4[u].y, s4[u].z, g4[u].x, g4[u].y, eps, gx, gy,
                                  cx, cy, gm);
          }
          for (; p < pend; ++p, ++q) {
            const float4 s1 = *p;
            const float2 g1 = *q;
            pair_adjoint_scaled(pt.x, pt.y, bt.x, bt.y, s1.x, s1.y, s1.z, g1.x, g1.y, eps, gx, gy, cx, cy, gm);
          }
        } else {
          for (; p < pfull; p += 4, q += 4) {
            float4 s4[4];
            float2 g4[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              s4[u] = p[u];
              g4[u] = q[u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
              pair_adjoint(pt.x, pt.y, pt.z, gt.x, gt.y, s4[u].x, s4[u].y, s4[u].z, g4[u].x, g4[u].y, eps, gx, gy,
                           cx, cy, gm);
          }
          for (; p < pend; ++p, ++q) {
            const float4 s1 = *p;
            const float2 g1 = *q;
            pair_adjoint(pt.x, pt.y, pt.z, gt.x, gt.y, s1.x, s1.y, s1.z, g1.x, g1.y, eps, gx, gy, cx, cy, gm);
          }
        }
        wave_lds_order();
      }
    }
    if (scaled) adjoint_unself(pt.z, gt.x, gt.y, bt.x, bt.y, eps, gx, gy);
    adjoint_fold(gx, gy, cx, cy);
    if (scaled) {
      gx = pt.z * gx;
      gy = pt.z * gy;
    }
    if (valid) {
      if (v.L > 0) {
        // P2M^T: X_c = sum m x / M_c,  M_c = sum m
        const float4 leafg = leaf_multipole_grad(v, xt, y);
        const float4 leafm = v.mpole[y * B + xt];
        const float rr = pt.z / leafm.z;
        gx = __builtin_fmaf(rr, leafg.x, gx);
        gy = __builtin_fmaf(rr, leafg.y, gy);
        gm += leafg.z + ((pt.x - leafm.x) * leafg.x + (pt.y - leafm.y) * leafg.y) / leafm.z;
      }
      const uint32_t o = __float_as_uint(pt.w);
      if (MODE == kAdjSolve) {
        gpos[o] = make_float2(gx, gy);
        gw[o] = gm;
      } else {
        const float2 lam = gpos[o];
        const float2 pi = gpre[o];
        gpos[o] = make_float2(2.0f * lam.x + gx + pi.x, 2.0f * lam.y + gy + pi.y);
        gpre[o] = make_float2(-lam.x, -lam.y);
        gw[o] += gm;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_axpy(int n, float a, const float* x, float* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = __builtin_fmaf(a, x[i], y[i]);
}

// ---------------------------------------------------------------- launchers

__global__ __launch_bounds__(256) void k_fill(int n, float a, float* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a;
}

hipError_t launch_fill(int n, float a, float* y, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, y);
  return hipGetLastError();
}

hipError_t launch_axpy(int n, float a, const float* x, float* y, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpy, dim3(cdiv(n, 256)), dim3(256), 0, s, n, a, x, y);
  return hipGetLastError();
}

// Device-to-device copies of up to kCopyRegions regions in one launch:
// blockIdx.y picks the region, the x blocks stride over it 16 bytes at a
// time when both ends are 16-byte aligned (4 bytes at a time otherwise).
__global__ __launch_bounds__(256) void k_copy_regions(CopyRegions c) {
  const int r = blockIdx.y;
  const uint32_t* src = c.src[r];
  uint32_t* dst = c.dst[r];
  const size_t n = c.words[r];
  const size_t first = (size_t)blockIdx.x * 256 + threadIdx.x, stride = (size_t)gridDim.x * 256;
  size_t done = 0;
  if (((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0) {
    const size_t n4 = n / 4;
    const uint4* s4 = reinterpret_cast<const uint4*>(src);
    uint4* d4 = reinterpret_cast<uint4*>(dst);
    for (size_t i = first; i < n4; i += stride) d4[i] = s4[i];
    done = n4 * 4;
  }
  for (size_t i = done + first; i < n; i += stride) dst[i] = src[i];
}

hipError_t launch_copy_regions(const CopyRegions& c, int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (count > kCopyRegions) return hipErrorInvalidValue;
  size_t most = 0;
  for (int r = 0; r < count; ++r) most = c.words[r] > most ? c.words[r] : most;
  if (most == 0) return hipSuccess;
  const size_t blocks = (most / 4 + 255) / 256;
  hipLaunchKernelGGL(k_copy_regions, dim3((unsigned)(blocks < 2048 ? (blocks ? blocks : 1) : 2048), count),
                     dim3(256), 0, s, c);
  return hipGetLastError();
}

int adj_up_blocks(int L, int B) {
  if (L <= 0) return 0;
  const int T = adj_tile_dim(B);
  return (B / T) * (B / T);
}

int adj_m2l_fused_blocks(const SysView& v) {
  if (v.L <= 0) return 0;
  const int lgT = adj_tile_lg(v.B);
  if (v.L - 1 > lgT) return kM2LTopBlocks + cdiv(v.lvl_off(lgT + 1), 256);
  return cdiv(v.lvl_off(v.L), 256);
}

hipError_t launch_adjoint(const SysView& v, const CallIO& io, int mode, int grid_blocks, hipStream_t s) {
  if (v.n <= 0) return hipSuccess;
  const OneSys one{v};
  hipLaunchKernelGGL(k_adj_sort<OneSys>, dim3(cdiv(v.n, 256)), dim3(256), 0, s, one, io);
  if (v.L > 0) {
    hipLaunchKernelGGL(k_adj_up<OneSys>, dim3(adj_up_blocks(v.L, v.B)), dim3(kUpTile * kUpTile), 0, s, one);
    hipLaunchKernelGGL(k_adj_m2l_fused<OneSys>, dim3(adj_m2l_fused_blocks(v)), dim3(256), 0, s, one);
  }
  if (mode == kAdjSolve)
    hipLaunchKernelGGL((k_adj_p2p<kAdjSolve, OneSys>), dim3(grid_blocks), dim3(kAdjP2PWaves * kWave), 0, s, one, io);
  else
    hipLaunchKernelGGL((k_adj_p2p<kAdjStep, OneSys>), dim3(grid_blocks), dim3(kAdjP2PWaves * kWave), 0, s, one, io);
  return hipGetLastError();
}

hipError_t launch_adjoint_batch(const BatchLaunch& bl, const CallIO& io, int mode, hipStream_t s) {
  hipLaunchKernelGGL(k_adj_sort<ManySys>, dim3(bl.adj_sort.blocks), dim3(256), 0, s, bl.adj_sort.sys, io);
  if (bl.adj_up.blocks)
    hipLaunchKernelGGL(k_adj_up<ManySys>, dim3(bl.adj_up.blocks), dim3(kUpTile * kUpTile), 0, s, bl.adj_up.sys);
  if (bl.adj_m2l.blocks)
    hipLaunchKernelGGL(k_adj_m2l_fused<ManySys>, dim3(bl.adj_m2l.blocks), dim3(256), 0, s, bl.adj_m2l.sys);
  if (mode == kAdjSolve)
    hipLaunchKernelGGL((k_adj_p2p<kAdjSolve, ManySys>), dim3(bl.adj_p2p.blocks), dim3(kAdjP2PWaves * kWave), 0, s, bl.adj_p2p.sys,
                       io);
  else
    hipLaunchKernelGGL((k_adj_p2p<kAdjStep, ManySys>), dim3(bl.adj_p2p.blocks), dim3(kAdjP2PWaves * kWave), 0, s, bl.adj_p2p.sys,
                       io);
  return hipGetLastError();
}

}  // namespace nbfmm_amd
