// Host-side launchers of the nbfmm_amd kernels (one translation unit each).
#pragma once

#include <string>

#include "common.hpp"

namespace nbfmm_amd {

// Records `msg` as the calling thread's nbfmm_last_error() text; returns `code`.
int report_error(int code, const std::string& msg);

// Launch map of one kernel over a batch: `blocks` blocks in all, ManySys
// telling each block its system.
struct KernelMap {
  ManySys sys;
  int blocks;
};

struct BatchLaunch {
  KernelMap tiles;     // row pass: cdiv(n_s, kSortTile) per system
  KernelMap rows;      // row scan / column pass: B_s per system
  KernelMap up;        // forward upward pass: one block per tile of leaves (up_blocks)
  KernelMap m2m_top;   // M2M top levels
  KernelMap m2l;       // fused M2L: m2l_fused_blocks
  KernelMap adj_up;    // adjoint upward pass: one block per adjoint tile (adj_up_blocks)
  KernelMap adj_m2l;   // adjoint M2L: adj_m2l_fused_blocks
  KernelMap p2p;       // P2P: share of the P2P grid (blocks of the chosen variant)
  KernelMap adj_p2p;   // adjoint P2P: share of its grid (kAdjP2PWaves-wave blocks)
  KernelMap adj_sort;  // adjoint: upstream gradient to sorted order, cdiv(n_s, 256)
  // warm-started sort: the particle kernels use adj_sort, the row kernels rows
  KernelMap warm_merge;  // one wave per touched leaf (grid-stride): its share of waves
  KernelMap one;       // one block per system
  int max_B = 0;       // widest leaf row among the systems
  bool warm_one = false;  // every system within warm_rows_one_fits: the warm row passes as one launch
};

// sort.hip
// zero_empty: also clear the level-0 multipole of empty leaves (fused solve)
hipError_t launch_predo(const SysView& v, const CallIO& io, int zero_empty, hipStream_t s);
hipError_t launch_predo_batch(const BatchLaunch& b, const CallIO& io, hipStream_t s);
hipError_t launch_occ_from_head(const SysView& v, hipStream_t s);
// warm start from the previous sort of the same n; a leaf gaining more than
// kMergeLdsMovers movers switches it to the cold passes (decided on the device)
// gathered: the previous step's P2P (kP2PStepSort) made the first pass
hipError_t launch_predo_warm(const SysView& v, const CallIO& io, int zero_empty, hipStream_t s, bool gathered = false);
hipError_t launch_predo_warm_batch(const BatchLaunch& b, const CallIO& io, hipStream_t s, bool gathered = false);
// whether a system of B x B leaves and n particles fits the one-block warm
// row passes (k_warm_rows_finish_one)
bool warm_rows_one_fits(int B, int n);

// tree.hip
hipError_t launch_p2m(const SysView& v, hipStream_t s);
// fused paths: P2M of the occupied leaves (empty ones cleared by the sort) and
// the M2M levels of the upward tiles; with_top: also the levels above them
// (the fused M2L makes those itself)
hipError_t launch_upward(const SysView& v, hipStream_t s, bool with_top);
// fused paths' M2L: cells with particles only, and the M2M levels above the tiles
hipError_t launch_m2l_fused(const SysView& v, hipStream_t s);
// small trees (small_front_fits): cold sort + P2M + every M2M level in one
// block, the same arrays and bits as the sort + upward pass + fused M2L's top
bool small_front_fits(int L, int n);
hipError_t launch_small_front(const SysView& v, const CallIO& io, hipStream_t s);
// forward: upward pass + fused M2L; else (adjoints) upward pass + the top M2M levels
hipError_t launch_tree_batch(const BatchLaunch& b, hipStream_t s, bool forward);
int up_blocks(int L, int B);           // the upward pass's tiles
int upward_top_blocks(int L, int B);
int m2l_blocks(const SysView& v);
int m2l_fused_blocks(const SysView& v);
hipError_t launch_m2m(const SysView& v, hipStream_t s);
hipError_t launch_m2l(const SysView& v, hipStream_t s);
hipError_t launch_l2l(const SysView& v, hipStream_t s);
// reference cube <-> dense levels; which: 0 mpole com, 1 mpole mass, 2 local
hipError_t launch_cube_read(const SysView& v, int which, void* dst, hipStream_t s);
hipError_t launch_cube_write(const SysView& v, int which, const void* src, hipStream_t s);

// p2p.hip
hipError_t launch_p2p(const SysView& v, int mode, const CallIO& io, int grid_blocks, int variant, hipStream_t s);
hipError_t launch_p2p_batch(const BatchLaunch& b, int mode, const CallIO& io, int variant, hipStream_t s);
int p2p_num_variants();
// default P2P configurations (measured): a solver runs variant {384 staged,
// unroll 4, 8 waves} on 3 blocks per CU (one round: LDS holds three such
// blocks per CU, so every block starts at once), and one of at most
// kP2PSmallN particles (whose runs are split among the waves of a block)
// variant {384, 4, 4 waves} on 4 blocks per CU: 10-step runs of the
// 16384-particle disk 52.9 -> 51.0 us a step, of the 65536-particle Gaussian
// disk 76.5 -> 69.3 us (2^20 two-disk step: 213 against 230 us).  A batch
// spreads variant {320, 4, 4 waves} on 48 blocks per CU over its systems.
constexpr int kDefaultP2PVariant = 5;
constexpr int kDefaultP2PBlocksPerCU = 3;
constexpr int kDefaultSmallP2PVariant = 0;
constexpr int kDefaultSmallP2PBlocksPerCU = 4;
constexpr int kDefaultBatchP2PVariant = 6;
constexpr int kDefaultBatchP2PBlocksPerCU = 48;
inline void default_p2p_config(int max_n, int* variant, int* blocks_per_cu) {
  const bool small = max_n <= kP2PSmallN;
  *variant = small ? kDefaultSmallP2PVariant : kDefaultP2PVariant;
  *blocks_per_cu = small ? kDefaultSmallP2PBlocksPerCU : kDefaultP2PBlocksPerCU;
}
void p2p_variant_info(int variant, int* stage, int* unroll, int* waves);
hipError_t launch_l2p(const SysView& v, hipStream_t s);

// adjoint.hip; mode 0: solve adjoint, 1: Verlet step adjoint
int adj_up_blocks(int L, int B);  // the adjoint's upward tiles
int adj_m2l_fused_blocks(const SysView& v);
hipError_t launch_adjoint(const SysView& v, const CallIO& io, int mode, int grid_blocks, hipStream_t s);
hipError_t launch_adjoint_batch(const BatchLaunch& b, const CallIO& io, int mode, hipStream_t s);
// y += a * x over n floats
hipError_t launch_axpy(int n, float a, const float* x, float* y, hipStream_t s);
// y = a over n floats (a kernel, so that callers may capture it in a graph)
hipError_t launch_fill(int n, float a, float* y, hipStream_t s);
// the first `count` (<= kCopyRegions) copies of dst[r] = src[r], words[r]
// 32-bit words each, in one launch (regions must not overlap)
constexpr int kCopyRegions = 6;
struct CopyRegions {
  const uint32_t* src[kCopyRegions];
  uint32_t* dst[kCopyRegions];
  size_t words[kCopyRegions];
};
hipError_t launch_copy_regions(const CopyRegions& c, int count, hipStream_t s);
hipError_t launch_postdo(const SysView& v, float2* eff0, hipStream_t s);
hipError_t launch_update(int n, float tick, float2* xcur, float2* xpre, const float2* acc, hipStream_t s);
// particle buffers <-> reference layouts; which = NBFMM_BUF_* id
hipError_t launch_part_read(const SysView& v, int which, int count, void* dst, hipStream_t s);
hipError_t launch_part_write(const SysView& v, int which, int count, const void* src, hipStream_t s);
hipError_t launch_pair_count(const SysView& v, unsigned long long* out, hipStream_t s);

// prune.hip: Stars::prune on (xcur, xpre, w), n > 0; tcount needs prune_tiles(n)
// entries, tmp and tmpw n each; *kept (device) receives the new count
hipError_t launch_prune(int n, float4 lim, float2* xcur, float2* xpre, float* w, uint32_t* tcount, float4* tmp,
                        float* tmpw, int* kept, hipStream_t s);
int prune_tiles(int n);

}  // namespace nbfmm_amd
