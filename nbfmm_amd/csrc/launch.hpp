// Host-side launchers of the nbfmm_amd kernels (one translation unit each).
#pragma once

#include "common.hpp"

namespace nbfmm_amd {

// sort.hip
hipError_t launch_predo(const SysView& v, hipStream_t s);
hipError_t launch_occ_from_head(const SysView& v, hipStream_t s);
int scan_blocks(int B);

// tree.hip
hipError_t launch_p2m(const SysView& v, hipStream_t s);
hipError_t launch_m2m(const SysView& v, hipStream_t s);
hipError_t launch_m2l(const SysView& v, hipStream_t s);
hipError_t launch_l2l(const SysView& v, hipStream_t s);
// reference cube <-> dense levels; which: 0 mpole com, 1 mpole mass, 2 local
hipError_t launch_cube_read(const SysView& v, int which, void* dst, hipStream_t s);
hipError_t launch_cube_write(const SysView& v, int which, const void* src, hipStream_t s);

// p2p.hip
struct StepArgs {
  float2* acc0;     // original-order output effect
  float2* xcur;     // Verlet state (kP2PStep only)
  float2* xpre;
  float tick;
};
hipError_t launch_p2p(const SysView& v, int mode, const StepArgs& a, int grid_blocks, hipStream_t s);
hipError_t launch_l2p(const SysView& v, hipStream_t s);
hipError_t launch_postdo(const SysView& v, float2* eff0, hipStream_t s);
hipError_t launch_update(int n, float tick, float2* xcur, float2* xpre, const float2* acc, hipStream_t s);
// particle buffers <-> reference layouts; which = NBFMM_BUF_* id
hipError_t launch_part_read(const SysView& v, int which, int count, void* dst, hipStream_t s);
hipError_t launch_part_write(const SysView& v, int which, int count, const void* src, hipStream_t s);
hipError_t launch_pair_count(const SysView& v, unsigned long long* out, hipStream_t s);

}  // namespace nbfmm_amd
