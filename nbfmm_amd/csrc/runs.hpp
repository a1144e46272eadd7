// Runs: the unit of work of the P2P kernel.
//
// The sorted particles of every row of leaves are cut into runs of 64
// consecutive targets, one wavefront per run and one target per lane.  A
// target in leaf x of row y needs, in each neighbour row r, the sorted range
// [head[r, x-1], head[r, x+2]) (the row-range idea of
// source/nbfmm/core/solver/p2p.cu:105-118); those ranges lie inside the run's
// union [head[r, x_first-1], head[r, x_last+2]).  A block rebuilds the run
// table (runs before each row) from head and its waves draw runs b, b + nblk,
// ... through an LDS counter (b = xcd_block(block, nblk): the blocks of one
// XCD take neighbouring runs).
//
// A run whose union exceeds the system's segment threshold (a cluster of
// dense leaves; in a small system, most runs) is summed in K segments of the
// union (p2p_segments: K and segments of S sources, a multiple of 64, both
// functions of the union size U and the system's particle count alone):
// each lane's sum of every segment starts from zero and the segment sums are
// added in segment order.  One wave can do this in one pass over the union
// (folded at the segment bounds), or the segments can be shared by the waves
// of the block (split); both give the same bits (the segment sums are added
// in segment order either way).  A run that would be one
// long serial chain on one wave is split (p2p_split).  The wave
// that takes such a run publishes it in the block's slot; every wave of the
// block that is free takes a segment (a ticket from an LDS counter,
// generation and K in its high bits), sums it, waits until the segments
// before it have been added into the slot's running sums, adds its own and
// passes on; the last one runs the epilogue and frees the slot.  The wait
// ends normally (tickets go out in order and segment 0 waits for nothing);
// it is bounded all the same: after a poll limit the waiting wave abandons
// the run, sums its whole union alone in one folded pass and runs the
// epilogue, and the other waves of that run drop their segments (RunSlot
// below).  A wave whose heavy run finds the slot taken keeps it pending,
// helps with the published run and publishes its own once the slot is free
// (after a poll limit it sums the whole union itself in one folded pass).
// Either give-up raises NBFMM_FLAG_P2P_WAIT, which reports the wait only:
// the folded pass gives the split sum's bits.  A wave that finds the run
// queue empty stays to help while another wave holds a run that may still be
// published (the low bits of the queue word count runs from their take until
// they are known light, published, or summed alone) or a split run is
// published.  All of this state is in LDS, and the hardware executes each
// wave's LDS instructions in program order, so the atomics are relaxed and
// only the compiler is kept from reordering (wave_lds_order); no wave reads
// another's global writes here, so no wave waits for its global stores.
// Every kernel configuration gives the same bits.
#pragma once

#include "common.hpp"

namespace nbfmm_amd {

// Geometry of run j: its targets (one per lane) and, per neighbour row
// q = 0, 1, 2 (rows y-1, y, y+1), the run's union of source ranges
// [us, us + ulen) in the sorted particles and this lane's part [a, b) of it
// (relative to us; empty for idle lanes and rows outside the grid).
struct RunGeom {
  int y, i, xt;
  bool valid;
  float4 pt;
  int us[3], ulen[3];  // wave-uniform
  int a[3], b[3];      // this lane's range in each row's union (relative to us)
  int total;           // sum of ulen: the union's size
  int hk;              // head of the target's own leaf (loaded only when asked for)
};

template <bool HK>
__device__ __forceinline__ void run_geom(const SysView& v, const int* s_runs, int j, int lane, RunGeom& g) {
  const int B = v.B;
  // everything but the lane's own target is wave-uniform: scalar registers
  g.y = __builtin_amdgcn_readfirstlane(run_row_wave(s_runs, B, j, lane));
  const int T0 = __builtin_amdgcn_readfirstlane(v.head[g.y * B] + kWave * (j - s_runs[g.y]));
  const int nt = __builtin_amdgcn_readfirstlane(min(kWave, v.head[(g.y + 1) * B] - T0));
  g.valid = lane < nt;
  g.i = T0 + (g.valid ? lane : 0);
  g.pt = v.part[g.i];
  g.xt = (int)(v.key[g.i] & (uint32_t)(B - 1));
  // in flight during the run's sums, for the epilogue
  g.hk = HK ? v.head[g.y * B + g.xt] : 0;
  const int xf = __builtin_amdgcn_readfirstlane(__shfl(g.xt, 0, kWave));
  const int xl = __builtin_amdgcn_readfirstlane(__shfl(g.xt, nt - 1, kWave));
  g.total = 0;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int r = g.y + q - 1;
    if (r < 0 || r >= B) {
      g.us[q] = g.ulen[q] = g.a[q] = g.b[q] = 0;
      continue;
    }
    const int* hr = v.head + r * B;
    g.us[q] = __builtin_amdgcn_readfirstlane(hr[xf > 0 ? xf - 1 : 0]);
    g.ulen[q] = __builtin_amdgcn_readfirstlane(hr[(xl < B - 1 ? xl + 1 : B - 1) + 1]) - g.us[q];
    g.a[q] = g.valid ? hr[g.xt > 0 ? g.xt - 1 : 0] - g.us[q] : 0;
    g.b[q] = g.valid ? hr[(g.xt < B - 1 ? g.xt + 1 : B - 1) + 1] - g.us[q] : 0;
    g.total += g.ulen[q];
  }
}

// Segments of a union of `total` sources in a system of v.n particles: 1 up
// to S_min, else min(kP2PMaxSeg, ceil(total / S_len)) segments of S sources
// (a multiple of 64), the last one shorter; S_min, S_len are the small
// system's rule when v.n <= kP2PSmallN.
__device__ __forceinline__ int p2p_segments(const SysView& v, int total) {
  const bool small = v.n <= kP2PSmallN;
  if (total <= (small ? kP2PSmallSplitMin : kP2PSplitMin)) return 1;
  const int len = small ? kP2PSmallSegLen : kP2PSegLen;
  return min(kP2PMaxSeg, (total + len - 1) / len);
}
__device__ __forceinline__ int p2p_segment_len(int total, int K) { return (((total + K - 1) / K) + 63) & ~63; }
// Whether a run of K segments is split among the waves of its block (else a
// K > 1 run is folded by one wave): a small system alone on the GPU splits
// every segmented run, otherwise only runs above kP2PSplitMin sources.
__device__ __forceinline__ bool p2p_split(const SysView& v, int total, int K) {
  return K > 1 && (v.split_small || total > kP2PSplitMin);
}

// The block's run queue and its slot for a split run, in LDS; NACC running
// sums per lane.
//
// The slot's completion word `done` holds the generation of the published
// run (bits 20..29, the generation of its tickets), a lock bit, an abandon
// bit and the number of segments added so far (bits 0..15).  The wave
// holding segment k adds it only after locking the word at count k (a
// compare-and-swap from the unlocked value), so the running sums are written
// only by the lock holder of the current generation; it then stores count
// k + 1, unlocked.  A wave that has waited its poll limit for the segments
// before its own abandons the run instead (a compare-and-swap from an
// unlocked value of the same generation to the abandoned one): it sums the
// whole union alone, in one pass, runs the epilogue and frees the slot.
// Every other wave of that generation then drops its segment.  Only one of
// "count reaches K - 1 under the last segment's lock" and "abandoned" can
// happen, so the epilogue of a run runs exactly once, and a wave of an old
// generation can neither lock nor abandon a run published later.
constexpr uint32_t kDoneCount = 0xffffu;
constexpr uint32_t kDoneLock = 1u << 16;
constexpr uint32_t kDoneAbandon = 1u << 17;
constexpr int kDoneGenShift = 20;

template <int NACC>
struct RunSlot {
  float acc[kWave][NACC];  // running sums of the segments added so far
  uint32_t queue;          // runs taken (bits 31..4) and runs counted active (bits 3..0)
  uint32_t ticket;         // generation << 22 | K << 16 | next segment
  uint32_t done;           // generation << 20 | abandon | lock | segments added
  int busy, run;
};

template <int NACC>
__device__ __forceinline__ void run_slot_init(RunSlot<NACC>& sl) {
  sl.queue = 0;
  sl.ticket = 0;
  sl.done = 0;
  sl.busy = 0;
  sl.run = 0;
}

// Lane 0's compare-and-swap on an LDS word, its outcome made wave-uniform.
__device__ __forceinline__ bool wave_cas(uint32_t* w, uint32_t expected, uint32_t desired, int lane) {
  int ok = 0;
  if (lane == 0) {
    uint32_t e = expected;
    ok = __hip_atomic_compare_exchange_strong(w, &e, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_WORKGROUP) ? 1 : 0;
  }
  wave_lds_order();
  return __builtin_amdgcn_readfirstlane(ok) != 0;
}

__device__ __forceinline__ uint32_t wave_load(uint32_t* w) {
  wave_lds_order();
  const uint32_t x = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  return __builtin_amdgcn_readfirstlane(x);
}

// The block's waves work through its runs.  seg(g, sa, sb, acc) sums this
// lane's sources among union positions [sa, sb) of run g into acc[NACC]
// (from zero); seg_lean is the same sum in the same order (its bits) with
// fewer registers; fin(g, acc) is the epilogue of run g (HK: g.hk is
// loaded).
// The slot must have been initialised and the block synchronised.
// v.wait_polls bounds every wait on another wave.
template <int NACC, bool HK, class Seg, class SegLean, class Fin>
__device__ __forceinline__ void run_queue(const SysView& v, const int* s_runs, RunSlot<NACC>& sl, int blk, int nblk,
                                          int total, int lane, Seg&& seg, SegLean&& seg_lean, Fin&& fin) {
  RunGeom g;
  const int limit = v.wait_polls;
  const int lblk = xcd_block(blk, nblk);
  bool exhausted = false;  // this wave found the block's run queue empty
  int idle_polls = 0;
  int pend = -1, pend_k = 1, pend_polls = 0;  // a split run waiting for the slot (wave-uniform)
  int redo = -1, redo_k = 1;  // a published run this wave abandoned: summed whole next, slot freed after
  // publishes split run jr of kr segments if the slot is free (lane 0's
  // view, made wave-uniform); the run stops counting as active once published
  auto try_publish = [&](int jr, int kr) -> bool {
    int got = 0;
    if (lane == 0) {
      int expected = 0;
      if (__hip_atomic_compare_exchange_strong(&sl.busy, &expected, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP)) {
        got = 1;
        sl.run = jr;
        const uint32_t gen = ((sl.ticket >> 22) + 1u) & 0x3ffu;
        __hip_atomic_store(&sl.done, gen << kDoneGenShift, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        wave_lds_order();
        // the tickets open last: a wave that draws one sees this run and its done word
        __hip_atomic_exchange(&sl.ticket, (gen << 22) | ((uint32_t)kr << 16), __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_WORKGROUP);
        wave_lds_order();
        __hip_atomic_fetch_add(&sl.queue, 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
    return __builtin_amdgcn_readfirstlane(got) != 0;
  };
  for (;;) {
    // what this iteration sums: segment k0 of run j out of K (a ticket of
    // the published run, generation gen), or all of run j
    int j = -1, K = 1, k0 = 0;
    uint32_t gen = 0;
    bool ticket = false, own_slot = false;
    if (redo >= 0) {
      // 0. the run this wave abandoned (g still holds its geometry)
      j = redo;
      K = redo_k;
      redo = -1;
      own_slot = true;
    }
    // 1. a segment of the published split run, if any is left
    if (!own_slot) {
      const uint32_t t0 = wave_load(&sl.ticket);
      if ((t0 & 0xffffu) < ((t0 >> 16) & 0x3fu)) {
        uint32_t t = 0;
        if (lane == 0) t = __hip_atomic_fetch_add(&sl.ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        wave_lds_order();
        t = __builtin_amdgcn_readfirstlane(t);
        const int k = (int)(t & 0xffffu), Kt = (int)((t >> 16) & 0x3fu);
        if (k < Kt) {
          ticket = true;
          gen = t >> 22;
          j = __builtin_amdgcn_readfirstlane(sl.run);
          K = Kt;
          k0 = k;
          run_geom<HK>(v, s_runs, j, lane, g);
        }
      }
    }
    if (!ticket && !own_slot) {
      if (pend >= 0) {
        // 2. this wave's split run waits for the slot, helping meanwhile;
        // after the poll limit it is summed here alone, in one folded pass
        if (try_publish(pend, pend_k)) {
          pend = -1;
          continue;
        }
        if (++pend_polls < limit) {
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        j = pend;
        K = pend_k;
        pend = -1;
        if (lane == 0) atomicOr(v.flags, kFlagP2PWait);
        run_geom<HK>(v, s_runs, j, lane, g);
      } else {
        // 3. a new run, counted active from its take until it is known light,
        // published or summed alone
        if (exhausted) {
          // leave once no run is counted active and no split run is
          // published; new segments can appear only from those (a wave that
          // stops waiting only stops helping, so this bound is no give-up and
          // does not follow v.wait_polls)
          wave_lds_order();
          if ((__hip_atomic_load(&sl.queue, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) & 0xfu) == 0 &&
              __hip_atomic_load(&sl.busy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
            break;
          if (++idle_polls >= kP2PWaitPolls) break;
          __builtin_amdgcn_s_sleep(2);
          continue;
        }
        uint32_t q = 0;
        if (lane == 0) q = __hip_atomic_fetch_add(&sl.queue, 17u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        wave_lds_order();
        j = lblk + (int)(__builtin_amdgcn_readfirstlane(q) >> 4) * nblk;
        if (j >= total) {
          if (lane == 0) __hip_atomic_fetch_add(&sl.queue, 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          exhausted = true;
          continue;
        }
        run_geom<HK>(v, s_runs, j, lane, g);
        K = p2p_segments(v, g.total);
        const bool split = p2p_split(v, g.total, K);
        // a run summed here can publish nothing: it stops counting at once,
        // so that waves without work need not stay for its end
        if (!split && lane == 0)
          __hip_atomic_fetch_add(&sl.queue, 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (split) {
          if (!try_publish(j, K)) {
            pend = j;
            pend_k = K;
            pend_polls = 0;
          }
          continue;  // this wave takes segments like any other
        }
      }
    }
    // a ticket's one segment of S sources, or all K segments of the run
    // (K = 1: S >= the union); each segment's sum from zero, added in
    // segment order.  A whole run's first segment by seg, the others (of a
    // folded run) by seg_lean: the same sum with fewer registers, since acc
    // is live beside it.  (Separate call sites for a ticket and a whole run:
    // the compiler keeps fewer registers live than with computed bounds.)
    const int S = p2p_segment_len(g.total, K);
    float acc[NACC];
    if (ticket) {
      seg(g, k0 * S, min(k0 * S + S, g.total), acc);
    } else {
      seg(g, 0, min(S, g.total), acc);
      for (int k = 1; k < K; ++k) {
        float part[NACC];
        seg_lean(g, k * S, min(k * S + S, g.total), part);
#pragma unroll
        for (int c = 0; c < NACC; ++c) acc[c] = acc[c] + part[c];
      }
    }
    if (!ticket) {
      fin(g, acc);
      if (own_slot) {
        // the abandoned run is done: the slot is free
        wave_lds_order();
        if (lane == 0) __hip_atomic_store(&sl.busy, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      } else if (p2p_split(v, g.total, K) && lane == 0) {
        // a split run summed here alone (poll limit) counted until now
        __hip_atomic_fetch_add(&sl.queue, 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      continue;
    }
    // a ticket: lock the done word at count k0 (segments 0 .. k0-1 added),
    // or abandon the run after the poll limit
    const uint32_t mine = (gen << kDoneGenShift) | (uint32_t)k0;
    int outcome = 0;  // 1: locked, 2: abandoned by this wave, 3: dropped
    for (int polls = 0; outcome == 0; ++polls) {
      const uint32_t d = wave_load(&sl.done);
      if ((d >> kDoneGenShift) != gen || (d & kDoneAbandon)) {
        outcome = 3;  // abandoned by another wave, or the slot already holds a later run
      } else if (d == mine) {
        if (wave_cas(&sl.done, d, d | kDoneLock, lane)) outcome = 1;
      } else if (!(d & kDoneLock) && polls >= limit) {
        if (wave_cas(&sl.done, d, d | kDoneAbandon, lane)) outcome = 2;
      } else if (polls >= limit + (1 << 16)) {
        // a lock is held only across a few LDS accesses (a few microseconds
        // at most; this is milliseconds): reaching it is a broken protocol,
        // reported rather than waited for
        if (lane == 0) atomicOr(v.flags, kFlagP2PWait | kFlagP2PProtocol);
        outcome = 3;
      } else {
        __builtin_amdgcn_s_sleep(1);
      }
    }
    if (outcome == 3) continue;
    if (outcome == 2) {
      // this wave owns the run: no further tickets; next, the whole union
      // here segment after segment, its epilogue, then the slot is free
      if (lane == 0) {
        atomicOr(v.flags, kFlagP2PWait);
        __hip_atomic_store(&sl.ticket, (gen << 22) | ((uint32_t)K << 16) | (uint32_t)K, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      wave_lds_order();
      redo = j;
      redo_k = K;
      continue;
    }
    // locked: add this segment to the running sums
    if (k0 > 0) {
#pragma unroll
      for (int c = 0; c < NACC; ++c) acc[c] = sl.acc[lane][c] + acc[c];
    }
    if (k0 < K - 1) {
#pragma unroll
      for (int c = 0; c < NACC; ++c) sl.acc[lane][c] = acc[c];
      wave_lds_order();
      if (lane == 0) __hip_atomic_store(&sl.done, mine + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
      wave_lds_order();
      if (lane == 0) {
        __hip_atomic_store(&sl.done, mine + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_store(&sl.busy, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
      fin(g, acc);
    }
  }
}

}  // namespace nbfmm_amd
