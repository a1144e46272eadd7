// nbfmm_solve_noalloc: nbfmm_solve without a solver and without any device
// allocation, for callers that cannot allocate (code under hipGraph capture,
// tight serving loops, harnesses that forbid hipMalloc).  All scratch lives
// in one file-scope device array of kNoAllocScratchBytes per GPU, held until
// the process ends; the caller passes only inputs and outputs.
//
// The kernels are the solver's (cold sort, upward pass, fused M2L, P2P with the
// L2P/un-permute epilogue), on a view whose arrays are carved out of that
// array, so the results equal nbfmm_solve's bit for bit.  The array keeps
// the last call's contents; the cold sort rewrites every array it reads
// (histograms, offsets, head, sorted particles, occupied-leaf count, every
// tree cell), and the flag word is cleared with a fill kernel at the start
// of each call.  Calls that overlap on one GPU (two streams, two threads)
// share the array and race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>

#include "../../include/nbfmm_amd.h"
#include "common.hpp"
#include "launch.hpp"

namespace nbfmm_amd {

constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;  // 16 MiB per GPU
constexpr int kNoAllocMaxDevices = 64;

// The scratch array.  In host code this name is a stand-in, not a device
// address: host code uses the per-device pointer from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

struct DeviceScratch {
  void* ptr = nullptr;  // this device's copy of `scratch`
  int cus = 0;
};
DeviceScratch g_dev[kNoAllocMaxDevices];
std::mutex g_mutex;

// The calling thread's current device's scratch address and CU count,
// looked up once per device with that device current.
int device_scratch(DeviceScratch* out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return report_error(NBFMM_ERR_NO_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kNoAllocMaxDevices) return report_error(NBFMM_ERR_NO_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_mutex);
  DeviceScratch& d = g_dev[dev];
  if (!d.ptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return report_error(NBFMM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess || cus <= 0) cus = 256;
    d.cus = cus;
    d.ptr = p;
  }
  *out = d;
  return NBFMM_OK;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Carves the arrays of a cold solve of n particles on a 2^(L+1) leaf grid
// out of `base` (null: only sizes them); returns the bytes used.
size_t carve(unsigned char* base, int n, int L, SysView& v) {
  const size_t N = (size_t)(n > 0 ? n : 1);
  const int B = 1 << (L + 1);
  const size_t C = (size_t)B * B;
  const int levels = L < 1 ? 1 : L;
  size_t cells = 0;
  for (int l = 0; l < levels; ++l) cells += (size_t)(B >> l) * (B >> l);
  const size_t ntiles = (N + kSortTile - 1) / kSortTile;
  size_t off = 0;
  auto take = [&](size_t bytes) -> void* {
    void* p = base ? (void*)(base + off) : nullptr;
    off += align256(bytes);
    return p;
  };
  v.tkey = (uint32_t*)take(N * 4);
  v.tpart = (float4*)take(N * 16);
  v.key = (uint32_t*)take(N * 4);
  v.part = (float4*)take(N * 16);
  v.head = (int*)take((C + 1) * 4);
  v.rowstart = (int*)take(((size_t)B + 1) * 4);
  v.occ = (int*)take(C * 4);
  v.nocc = (int*)take(4);
  v.hist = (uint32_t*)take((size_t)B * ntiles * 4);
  v.dtotal = (uint32_t*)take((size_t)B * 4);
  v.mpole = (float4*)take(cells * 16);
  v.local = (float2*)take(cells * 8);
  v.flags = (uint32_t*)take(4);
  return off;
}

int cdiv(int a, int b) { return (a + b - 1) / b; }

}  // namespace
}  // namespace nbfmm_amd

using namespace nbfmm_amd;

extern "C" int nbfmm_solve_noalloc(int num_level, int num_particle, const float position_limits[4], float epsilon,
                                   const float* pos0, const float* w0, float* eff0, void* stream) {
  if (num_level < 0 || num_level > kMaxLevel) return report_error(NBFMM_ERR_INVALID_ARGUMENT, "num_level must be in [0, 8]");
  if (num_particle < 0) return report_error(NBFMM_ERR_INVALID_ARGUMENT, "num_particle < 0");
  if (!position_limits) return report_error(NBFMM_ERR_INVALID_ARGUMENT, "position_limits is null");
  if (!(position_limits[0] < position_limits[2] && position_limits[1] < position_limits[3]))
    return report_error(NBFMM_ERR_INVALID_ARGUMENT, "position_limits must satisfy x_min < x_max and y_min < y_max");
  if (!(epsilon > 0.0f) || !(epsilon < INFINITY)) return report_error(NBFMM_ERR_INVALID_ARGUMENT, "epsilon must be > 0");
  if (num_particle > 0 && (!pos0 || !w0 || !eff0)) return report_error(NBFMM_ERR_INVALID_ARGUMENT, "null array");
  SysView v{};
  const size_t need = carve(nullptr, num_particle, num_level, v);
  if (need > kNoAllocScratchBytes)
    return report_error(NBFMM_ERR_CAPACITY, "nbfmm_solve_noalloc: " + std::to_string(num_particle) +
                                                " particles at num_level " + std::to_string(num_level) + " need " +
                                                std::to_string(need) + " bytes of scratch; the static scratch holds " +
                                                std::to_string(kNoAllocScratchBytes));
  if (num_particle == 0) return NBFMM_OK;
  DeviceScratch d;
  if (int rc = device_scratch(&d)) return rc;
  carve((unsigned char*)d.ptr, num_particle, num_level, v);
  v.n = num_particle;
  v.L = num_level;
  v.B = 1 << (num_level + 1);
  v.lgB = num_level + 1;
  v.xmin = position_limits[0];
  v.ymin = position_limits[1];
  // predo.cu:112-113: cell size computed in float on the host
  v.cw = (position_limits[2] - position_limits[0]) / (float)v.B;
  v.ch = (position_limits[3] - position_limits[1]) / (float)v.B;
  v.eps = epsilon;
  CallIO io;
  io.pos0 = (const float2*)pos0;
  io.w0 = w0;
  io.acc0 = (float2*)eff0;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = launch_fill(1, 0.0f, (float*)v.flags, st);
  if (e == hipSuccess) e = launch_predo(v, io, 1, st);
  if (e == hipSuccess) e = launch_upward(v, st, false);
  if (e == hipSuccess) e = launch_m2l_fused(v, st);
  if (e == hipSuccess) {
    // the default P2P configuration and grid of a solver of this n (p2p_grid_for)
    int variant = 0, bpc = 0, stage = 0, unroll = 0, waves = 0;
    default_p2p_config(num_particle, &variant, &bpc);
    p2p_variant_info(variant, &stage, &unroll, &waves);
    const int C = v.B * v.B;
    const int want = std::max(cdiv(std::min(C, num_particle), waves), cdiv(num_particle, kWave) + v.B);
    const int grid = std::max(1, std::min(want, d.cus * bpc));
    e = launch_p2p(v, kP2PFinal, io, grid, variant, st);
  }
  if (e != hipSuccess) return report_error(NBFMM_ERR_HIP, std::string("nbfmm_solve_noalloc: ") + hipGetErrorString(e));
  return NBFMM_OK;
}
