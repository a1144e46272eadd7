// Shared device-side definitions of the nbfmm_amd kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nbfmm_amd {

constexpr int kWave = 64;
// Polls (each a short sleep) before a P2P wave stops waiting for another wave
// of its block; nbfmm_set_p2p_wait_polls changes it per solver or batch.
constexpr int kP2PWaitPolls = 1 << 20;
constexpr int kMaxLevel = 8;          // base_dim <= 512 (source/nbfmm/core/solver.cu:25)
constexpr int kSortThreads = 256;     // 4 waves per sort tile
constexpr int kSortWaveItems = 8;     // 64-element chunks per wave
constexpr int kSortTile = kSortThreads * kSortWaveItems;  // 2048 elements
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanChunk = kScanThreads * kScanItems;     // 2048 cells
constexpr int kP2PStage = 256;        // adjoint P2P: sources staged in LDS per wave and pass
constexpr int kAdjP2PWaves = 8;       // adjoint P2P: waves per block (48 KiB of LDS)
constexpr int kAdjP2PBlocksPerCU = 3; // adjoint P2P: resident blocks per CU (LDS bound)
// Batched adjoint P2P: a few rounds of residency, so that blocks ending
// early are replaced while others still work through heavy runs.  Measured
// on the 64-collision trajectory backward, whose adjoint steps share the GPU
// with the recompute of the segment before: 4, 6, 8, 12, 16 and 32 blocks
// per CU gave 538, 516, 497-507, 523, 519 and 567 ms.
constexpr int kAdjP2PBatchBlocksPerCU = 8;

// P2P output modes.
enum P2PMode : int {
  kP2POnly = 0,     // sorted effect = P2P sum (stage p2p)
  kP2PFinal = 1,    // + leaf local (L2P), scattered to original order (postdo)
  kP2PStep = 2,     // as kP2PFinal, then the Verlet update of that particle
  kP2PStepSort = 3, // as kP2PStep, and the first pass of the next step's warm-started sort
};

// Segments of P2P runs: a run whose staged union holds more than S_min
// sources is summed in K = min(kP2PMaxSeg, ceil(U / S_len)) segments of the
// union: each lane sums its sources of every segment from zero and adds the
// segment sums in segment order.  K and the segment bounds depend on U and
// the system's particle count alone.  Whether the segments are then shared
// by the waves of the block (a split run) or summed one after another by one
// wave in one pass (folded) changes no bits, so every kernel configuration,
// the solver, the batch and the allocation-free solve give the same bits.
// Large systems segment only heavy runs (dense clusters, whose serial sum
// would be the kernel's time), and split them; a system of at most
// kP2PSmallN particles has about n / 64 runs for the GPU's 1024 SIMDs, so its
// runs are segmented from kP2PSmallSplitMin sources into segments of
// kP2PSmallSegLen (S_min, S_len; kP2PSplitMin and kP2PSegLen otherwise) and,
// when it runs alone (SysView::split_small), split: the block's idle waves
// share each run and the chain of a wave is a few times shorter.  In a batch
// the other systems fill the GPU, so a small system's runs are folded and
// only runs above kP2PSplitMin sources are split.
constexpr int kP2PSplitMin = 2048;
constexpr int kP2PSegLen = 1024;
constexpr int kP2PSmallN = 1 << 16;
constexpr int kP2PSmallSplitMin = 256;
constexpr int kP2PSmallSegLen = 128;

// Everything one system's kernels need: sizes, box and the device arrays.
// Passed by value to every kernel.
struct SysView {
  int n;            // particles
  int L;            // reference num_level
  int B;            // base_dim = 2^(L+1)
  int lgB;          // L + 1
  float xmin, ymin; // lower box corner
  float cw, ch;     // leaf cell size, computed on the host as (max-min)/B in float
  float eps;        // kernel softening
  float tick;       // Verlet step (dt*sqrt(G)), step mode only
  int part_off;     // first particle of this system in the caller's arrays
  int wait_polls = kP2PWaitPolls;  // P2P: bound on every wait for another wave of the block
  int split_small = 1;             // P2P: split the segmented runs of a small system (0 in batches)
  // particle scratch (capacity max_n)
  uint32_t* tkey;   // row-sorted keys (row pass output)
  float4* tpart;    // row-sorted particles: x, y, m, original index (bits)
  uint32_t* key;    // sorted cell keys iy*B + ix
  float4* part;     // sorted particles: x, y, m, original index (bits)
  float2* eff;      // sorted effects (stage API)
  // cell scratch
  int* head;        // B*B+1
  int* rowstart;    // B+1: first sorted slot of each row
  int* occ;         // occupied leaves (row by row)
  int* nocc;        // number of occupied leaves
  uint32_t* hist;   // row-pass tile histograms [B][ntiles]
  uint32_t* dtotal; // row totals [B]
  float4* mpole;    // per level, dense D_l x D_l: COM x, COM y, mass, 0
  float2* local;    // per level, dense D_l x D_l: M2L effect (L2L accumulates)
  // adjoint scratch
  float2* gsort;    // upstream gradient of the effects, sorted order
  float2* gup;      // per level: sum of the upstream gradient over the cell's particles
  float4* gmp;      // per level: gradient wrt the cell's COM (x, y) and mass (z) from M2L
  uint32_t* flags;  // NBFMM_FLAG_* bits
  // warm-started sort (the previous sort's key/part/head as the starting
  // order); every array is left cleared by the pass that last uses it
  int* wstate;      // kWarmInts: [kWarmTouched] = touched-leaf count
  int* wrel;        // N: rank of a staying particle inside its old leaf
  int* wmgrp;       // N: old slots of the movers grouped by new leaf
  int* wcin;        // C: movers into each leaf
  int* wcout;       // C: movers out of each leaf
  int* wtouch;      // C: leaf gained or lost a mover
  int2* wtold;      // C: old (first slot, count) of each touched leaf
  int* wtlist;      // C: touched leaves (appended one row at a time)
  int* wcnew;       // C: row-local exclusive offset of the new leaf counts
  int* winoff;      // C: first slot of each leaf's movers in wmgrp
  int* wfill;       // C: movers placed into each leaf's group
  int* wrowtot;     // 4 (B + 1): row totals of the new counts, movers, occupied and touched leaves

  // Start of level l in mpole/local; lvl_off(L) = total.  The dense levels
  // k < max(L, 1) hold (B >> k)^2 = 4^(lgB - k) cells each, so
  //   lvl_off(l) = sum_{k < l} 4^(lgB - k) = (4^(lgB + 1) - 4^(lgB - l + 1)) / 3
  // (exact: 4^a - 4^b is a multiple of 3), with l clamped to max(L, 1).  A
  // function rather than a table, so that no kernel indexes into the view at
  // run time and a view read from device memory stays in registers.
  __host__ __device__ __forceinline__ int lvl_off(int l) const {
    const int levels = L < 1 ? 1 : L;
    const int m = l < levels ? l : levels;
    return (int)((((uint32_t)1 << (2 * lgB + 2)) - ((uint32_t)1 << (2 * (lgB - m) + 2))) / 3u);
  }
};

// P2M: a leaf's sorted range is cut at the multiples of kP2MSeg (its
// segments); segment sums, then the leaf's sum of them, both in order.
constexpr int kP2MSeg = 16;

// Upward passes (forward and adjoint): one block of kUpTile x kUpTile
// threads per T x T tile of leaves (up_tile_dim) makes levels 1..lgT of its
// tile in LDS; a single block makes the levels above (block 0 of the fused
// M2L launch, or k_m2m_top).  Grids of at most kUpTileSmallMaxB leaves a side
// take 8 x 8 tiles, so that their leaves spread over four times as many CUs.
constexpr int kUpTile = 16;
constexpr int kUpTileSmall = 8;
constexpr int kUpTileSmallMaxB = 128;
// The levels above the tiles in a fused M2L launch (forward or adjoint):
// each of the system's first kM2LTopBlocks blocks makes them in LDS (at most
// kTopLds cells: levels lgT+1..7 at B = 512) and takes its share of their
// cells.
constexpr int kM2LTopBlocks = 2;
constexpr int kTopLds = 256 + 64 + 16;
// log2 of the upward tiles' side (lgT).
__device__ __host__ __forceinline__ int up_tile_dim(int B) {
  const int T = B <= kUpTileSmallMaxB ? kUpTileSmall : kUpTile;
  return B < T ? B : T;
}
__device__ __host__ __forceinline__ int up_tile_lg(int B) {
  const int T = up_tile_dim(B);
  int lgT = 0;
  while ((1 << lgT) < T) ++lgT;
  return lgT;
}
// The adjoint's upward and M2L^T passes keep kUpTile x kUpTile tiles at every
// size: there a lane sums a whole leaf's gradient, and 8 x 8 tiles left three
// of a block's four waves idle (the 64-collision backward 496 -> 502 ms).
__device__ __host__ __forceinline__ int adj_tile_dim(int B) { return B < kUpTile ? B : kUpTile; }
__device__ __host__ __forceinline__ int adj_tile_lg(int B) {
  const int T = adj_tile_dim(B);
  int lgT = 0;
  while ((1 << lgT) < T) ++lgT;
  return lgT;
}

// Warm-started sort.
constexpr uint32_t kMoverBit = 0x80000000u;  // tkey flag: the particle changed leaf
// wstate words: [kWarmTouched] touched-leaf count; [kWarmColdReq] set by the
// first pass when a leaf's mover count passes kMergeLdsMovers (cleared by
// k_warm_finish, after k_warm_rows has read it); [kWarmCold] k_warm_rows'
// copy of it, which the later passes of the same sort read: the sort then
// takes the cold passes instead of the merge
constexpr int kWarmTouched = 1, kWarmCold = 2, kWarmColdReq = 3, kWarmInts = 4;
constexpr int kMergeLdsMovers = 512;  // movers of one leaf the merge sorts in LDS

constexpr int kP2PMaxSeg = 16;
constexpr uint32_t kFlagP2PWait = 2u;      // NBFMM_FLAG_P2P_WAIT
constexpr uint32_t kFlagP2PProtocol = 4u;  // NBFMM_FLAG_P2P_LOST
constexpr int kWarmMergeWaves = 16384;       // merge grid: one wave per touched leaf, grid-stride beyond

// Caller arrays of one call, original particle order; a system's particles
// start at its part_off.
struct CallIO {
  const float2* pos0 = nullptr;  // positions (input)
  const float* w0 = nullptr;     // masses (input)
  float2* acc0 = nullptr;        // effects (output of solve/step)
  float2* xcur = nullptr;        // Verlet state (step)
  float2* xpre = nullptr;
  // adjoint
  const float2* gin = nullptr;   // upstream gradient (solve adjoint: d loss / d effect;
                                 // step adjoint: lambda = d loss / d x_{k+1})
  float2* gpos = nullptr;        // solve adjoint: d loss / d position (out);
                                 // step adjoint: lambda, overwritten with lambda_k
  float2* gpre = nullptr;        // step adjoint: pi (in/out)
  float* gw = nullptr;           // d loss / d mass (solve adjoint: out; step adjoint: += )
  float gscale = 0.0f;           // upstream scale (step adjoint: tick^2)
};

// Which system a block works on, and its block index / block count inside
// that system.  OneSys: a single system passed by value.  ManySys: a batch;
// bsys[b] is the system of block b and bstart[s] the first block of system s.
struct OneSys {
  SysView v;
  __device__ __forceinline__ SysView get(int& blk, int& nblk) const {
    blk = blockIdx.x;
    nblk = gridDim.x;
    return v;
  }
};

// One system of a batch: its sizes, its box and where its part of each of
// the batch's concatenated arrays starts.
struct SysDesc {
  int n, L, B, lgB;
  float xmin, ymin, cw, ch, eps, tick;
  int part_off;  // particle arrays
  int head_off;  // head (B*B + 1 per system)
  int row_off;   // rowstart, dtotal (B + 1 per system); wrowtot at four times it
  int cell_off;  // leaf arrays (B*B per system)
  int hist_off;  // row-pass tile histograms
  int tree_off;  // per-level cell arrays
};

// The kernels of a batch see one system per block.  Its view is the batch's
// base view (the kernel argument, whose arrays start at system 0) moved to
// the system's offsets: derived from kernel-argument pointers, every array
// stays in the global address space, and the view in registers (a view read
// whole from device memory would carry generic pointers, i.e. flat accesses).
struct ManySys {
  SysView base;
  const SysDesc* desc;
  const int* bsys;
  const int* bstart;
  __device__ __forceinline__ SysView get(int& blk, int& nblk) const {
    const int b = blockIdx.x;
    const int s = bsys[b];
    const int b0 = bstart[s];
    blk = b - b0;
    nblk = bstart[s + 1] - b0;
    const SysDesc d = desc[s];
    SysView v = base;
    v.n = d.n;
    v.L = d.L;
    v.B = d.B;
    v.lgB = d.lgB;
    v.xmin = d.xmin;
    v.ymin = d.ymin;
    v.cw = d.cw;
    v.ch = d.ch;
    v.eps = d.eps;
    v.tick = d.tick;
    v.part_off = d.part_off;
    v.tkey += d.part_off;
    v.tpart += d.part_off;
    v.key += d.part_off;
    v.part += d.part_off;
    v.eff += d.part_off;
    v.gsort += d.part_off;
    v.wrel += d.part_off;
    v.wmgrp += d.part_off;
    v.head += d.head_off;
    v.rowstart += d.row_off;
    v.dtotal += d.row_off;
    v.wrowtot += 4 * d.row_off;
    v.occ += d.cell_off;
    v.wcin += d.cell_off;
    v.wcout += d.cell_off;
    v.wtouch += d.cell_off;
    v.wtlist += d.cell_off;
    v.wcnew += d.cell_off;
    v.winoff += d.cell_off;
    v.wfill += d.cell_off;
    v.wtold += d.cell_off;
    v.hist += d.hist_off;
    v.mpole += d.tree_off;
    v.local += d.tree_off;
    v.gup += d.tree_off;
    v.gmp += d.tree_off;
    v.nocc += s;
    v.flags += s;
    v.wstate += s * kWarmInts;
    return v;
  }
};

// include/nbfmm/core/kernel_function.hpp:29-40 on the GPU:
//   d = s - t; r = |d|; a += d * m / (r^3 + eps)
// r comes from the hardware square root and 1/(r^3+eps) from the hardware
// reciprocal (v_sqrt_f32, v_rcp_f32), not the IEEE-rounded sequences: the
// result differs from the reference's correctly rounded sqrt and division by
// about one ulp per pair.  A pair with d = 0 (a particle with itself)
// contributes exactly 0 because eps > 0.
// predo.cu:37-38: floorf((p - lo) / cell) with IEEE division.  Particles
// outside the box (or NaN) are clamped into the grid and reported through
// `bad`; the reference has no clamp (a particle at x_max indexes cell B).
__device__ __forceinline__ uint32_t bin_key(const SysView& v, float2 p, bool& bad) {
#pragma clang fp contract(off)
  const float fx = floorf((p.x - v.xmin) / v.cw);
  const float fy = floorf((p.y - v.ymin) / v.ch);
  int ix, iy;
  bad = false;
  if (!(fx >= 0.0f)) { ix = 0; bad = true; }
  else if (fx >= (float)v.B) { ix = v.B - 1; bad = true; }
  else ix = (int)fx;
  if (!(fy >= 0.0f)) { iy = 0; bad = true; }
  else if (fy >= (float)v.B) { iy = v.B - 1; bad = true; }
  else iy = (int)fy;
  return (uint32_t)(iy * v.B + ix);
}

// Lanes of the wave whose `d` (nbits wide) equals this lane's, among the
// lanes where `valid` holds.
__device__ __forceinline__ uint64_t match_mask(uint32_t d, bool valid, int nbits) {
  uint64_t m = __ballot(valid);
  for (int b = 0; b < nbits; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bal = __ballot(valid && bit);
    m &= bit ? bal : ~bal;
  }
  return m;
}

__device__ __forceinline__ uint64_t lanes_below() {
  const int lane = (int)(threadIdx.x & (kWave - 1));
  return lane == 0 ? 0ull : (~0ull >> (kWave - lane));
}


// Adds, for every lane where `act` holds, one to cnt[k] of that lane (k of
// nbits bits).  A wave with few active lanes adds one per lane; one with many
// (a crowd of particles moving into or out of one leaf together, which
// would otherwise serialise on one address) adds once per distinct k, all
// leaders in one instruction.  limit >= 0: the add that takes a count past
// `limit` also sets *over (those atomics return their old value).  All lanes
// of the wave call it.
__device__ __forceinline__ void wave_count_add(int* cnt, uint32_t k, bool act, int nbits, int limit = -1,
                                               int* over = nullptr) {
  const uint64_t any = __ballot(act);
  if (any == 0) return;
  int add = 1;
  bool go = act;
  if (__popcll(any) > 8) {
    const uint64_t m = match_mask(k, act, nbits);  // this lane's group (empty for inactive lanes)
    go = act && (threadIdx.x & (kWave - 1)) == (unsigned)__builtin_ctzll(m);
    add = (int)__popcll(m);
  }
  if (go) {
    if (limit >= 0) {
      const int old = atomicAdd(&cnt[k], add);
      if (old <= limit && old + add > limit) atomicOr(over, 1);
    } else {
      atomicAdd(&cnt[k], add);
    }
  }
}

// First pass of the warm-started sort for the particle at sorted slot i of
// the current sort (key kold), now at position p with mass m: its entry in
// the new order's input (tpart, tkey, wrel), and for a particle that changes
// leaf (a mover) the per-leaf in/out counts.  Lanes with `valid` false take
// part only in the counts (all lanes of the wave call it).  Returns whether
// p was outside the box.  hold = head[kold], the old leaf's first slot.
__device__ __forceinline__ bool warm_first_pass(const SysView& v, bool valid, int i, uint32_t kold, int hold,
                                                float2 p, float m, float w) {
  bool bad = false, mover = false;
  uint32_t k = 0;
  if (valid) {
    k = bin_key(v, p, bad);
    v.tpart[i] = make_float4(p.x, p.y, m, w);
    mover = k != kold;
    if (!mover) {
      v.tkey[i] = k;
      v.wrel[i] = i - hold;
    } else {
      v.tkey[i] = k | kMoverBit;
    }
  }
  // the next pass marks the leaves they touch; a leaf whose incoming count
  // passes kMergeLdsMovers switches this sort to the cold passes
  wave_count_add(v.wcout, kold, mover, 2 * v.lgB);
  wave_count_add(v.wcin, k, mover, 2 * v.lgB, kMergeLdsMovers, &v.wstate[kWarmColdReq]);
  return bad;
}

__device__ __forceinline__ void kernel_accum(float tx, float ty, float sx, float sy, float m, float eps,
                                             float& ax, float& ay) {
  const float dx = sx - tx;
  const float dy = sy - ty;
  const float r2 = __builtin_fmaf(dx, dx, dy * dy);
  const float r = __builtin_amdgcn_sqrtf(r2);
  const float den = __builtin_fmaf(r2, r, eps);
  const float f = m * __builtin_amdgcn_rcpf(den);
  ax = __builtin_fmaf(dx, f, ax);
  ay = __builtin_fmaf(dy, f, ay);
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

// The calling block's place among its system's nblk blocks in the run sweep
// (block blk takes runs blk', blk' + nblk, ...).  Blocks go to the eight XCDs
// round-robin by their global index, each XCD with its own L2; numbered so
// that each XCD takes a contiguous eighth of every band of nblk runs,
// neighbouring runs, which stage mostly the same sources, run on one XCD and
// share its L2.  The global index is blockIdx.x = (the system's first block)
// + blk; rotating blk by that first block's XCD gives each block its own XCD
// in the low three bits, so the map is a bijection of [0, nblk) when nblk is
// a multiple of 8 (otherwise blk itself).
__device__ __forceinline__ int xcd_block(int blk, int nblk) {
  if (nblk & 7) return blk;
  int r = blk + (int)((blockIdx.x - (unsigned)blk) & 7u);
  if (r >= nblk) r -= nblk;
  return (r & 7) * (nblk >> 3) + (r >> 3);
}

// Next position in the calling block's share of a work list: lane 0 bumps
// the block's LDS counter, the whole wave gets the value.
__device__ __forceinline__ int block_take(int* s_next) {
  int i = 0;
  if (lane_id() == 0) i = atomicAdd(s_next, 1);
  return __builtin_amdgcn_readfirstlane(i);
}

// Ordering point for LDS data that only the calling wave touches: the
// hardware executes one wave's LDS instructions in program order, so only
// the compiler must be kept from moving memory accesses across this point.
// (A workgroup-scope fence would also wait for every outstanding global
// load.)
__device__ __forceinline__ void wave_lds_order() { asm volatile("" ::: "memory"); }

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t o = __shfl_up(v, off, kWave);
    if (lane >= off) v += o;
  }
  return v;
}

// Exclusive prefix sum over a block of NT threads; returns the block total
// in *total.  Uses sh[NT/64] of shared scratch.  All threads must call it.
template <int NT>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
  constexpr int NW = NT / kWave;
  const int lane = lane_id();
  const int wave = threadIdx.x / kWave;
  const uint32_t inc = wave_inclusive_scan(v);
  if (lane == kWave - 1) sh[wave] = inc;
  __syncthreads();
  uint32_t wave_off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < NW; ++w) {
    const uint32_t s = sh[w];
    if (w < wave) wave_off += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return wave_off + inc - v;
}

// Target runs: every row of leaves is cut into runs of 64 consecutive sorted
// particles (the last run of a row may be shorter).  s_runs[y] is the number
// of runs in the rows before y; s_runs[B] is the total.  All threads of the
// block call this.
template <int NT>
__device__ __forceinline__ void row_runs(const SysView& v, int* s_runs, uint32_t* sh) {
  const int B = v.B;
  uint32_t carry = 0;
  for (int r0 = 0; r0 < B; r0 += NT) {
    const int r = r0 + (int)threadIdx.x;
    const uint32_t cnt = r < B ? (uint32_t)((v.head[(r + 1) * B] - v.head[r * B] + kWave - 1) / kWave) : 0u;
    uint32_t tot;
    const uint32_t e = block_exclusive_scan<NT>(cnt, sh, &tot);
    if (r < B) s_runs[r] = (int)(carry + e);
    carry += tot;
  }
  if (threadIdx.x == 0) s_runs[B] = (int)carry;
}

// Row of run j (j < total): the last row whose first run is <= j.
__device__ __forceinline__ int run_row(const int* s_runs, int B, int j) {
  int lo = 0, hi = B;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (s_runs[mid] <= j) lo = mid; else hi = mid;
  }
  return lo;
}

// run_row for a whole wave (j wave-uniform): every lane tests one row of a
// stride-S sample (S = ceil(B / 64)), then the first S lanes test the rows
// of the sampled interval that holds the answer; two LDS round trips instead
// of log2(B).  s_runs is nondecreasing, so each test's true lanes form a
// prefix and the answer is the highest of them.
__device__ __forceinline__ int run_row_wave(const int* s_runs, int B, int j, int lane) {
  const int S = (B + kWave - 1) / kWave;
  const int rc = lane * S;
  const uint64_t coarse = __ballot(rc < B && s_runs[rc < B ? rc : 0] <= j);
  const int base = (63 - __builtin_clzll(coarse)) * S;  // lane 0 always holds (s_runs[0] = 0)
  const int rf = base + lane;
  const uint64_t fine = __ballot(lane < S && rf < B && s_runs[rf < B ? rf : 0] <= j);
  return base + 63 - __builtin_clzll(fine);
}

inline int level_dim(int B, int l) { return B >> l; }

}  // namespace nbfmm_amd
