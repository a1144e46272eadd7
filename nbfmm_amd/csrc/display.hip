// Video frames of the stars (source/nbfmm/display/stars/display.cu) and the
// YUV4MPEG2 stream they are written to (main/main.cu:234-250).
//
// A frame is planar YUV 4:2:0: width*height luma bytes, then the U and V
// planes, width*height/2 bytes together, held at 128 (no colour).  Each star
// whose pixel lies in the frame lights a stamp of luma 255 whose size grows
// with sqrt(weight) / size_scale: level s = floor(sqrt(w) / size_scale),
// clamped to [0, 7], lights every pixel (x + dx, y + dy) of the frame with
// dx^2 + dy^2 <= kStampR2[s], i.e. the point, its 4-neighbours, the 3x3
// square, then the lattice disks of radius^2 4, 5, 9, 10 and 16
// (1, 5, 9, 13, 21, 29, 37 and 49 pixels; display.cu:54-211 lists the same
// offsets level by level).  The pixel of a star is
//   x = floor((p.x - x_min) / unit_width), y = floor((y_max - p.y) / unit_height),
// unit = (max - min) / size, in IEEE float as display.cu:45-48 (row 0 at
// the top).  Stamps are clipped to the frame; the reference also lets the
// star's own row reach y == height (display.cu:53) and the upward arms skip
// their bound (display.cu:62, 90, 130, 182), writing outside the luma plane.
//
// Stars light pixels independently, and every store writes the same value,
// so concurrent stores to one pixel need no ordering.
#include <cstdio>
#include <string>

#include "launch.hpp"
#include "../../include/nbfmm_amd.h"

namespace nbfmm_amd {

__constant__ int kStampR2[8] = {0, 1, 2, 4, 5, 9, 10, 16};

// Luma 0, chroma 128: 16 bytes per thread, byte stores for a chunk that is
// unaligned or straddles the luma/chroma boundary or the end.
__global__ void k_frame_clear(size_t luma, size_t total, uint8_t* __restrict__ frame) {
  const size_t b = (size_t)(blockIdx.x * blockDim.x + threadIdx.x) * 16;
  if (b >= total) return;
  const bool aligned = (((uintptr_t)(frame + b)) & 15) == 0;
  if (aligned && b + 16 <= total && (b + 16 <= luma || b >= luma)) {
    const uint32_t v = b < luma ? 0u : 0x80808080u;
    *(uint4*)(frame + b) = make_uint4(v, v, v, v);
    return;
  }
  const size_t e = b + 16 < total ? b + 16 : total;
  for (size_t i = b; i < e; ++i) frame[i] = i < luma ? 0 : 128;
}

// One thread per star.
__global__ void k_display(int n, int width, int height, float size_scale, float4 lim, float unit_w, float unit_h,
                          const float2* __restrict__ pos, const float* __restrict__ weight,
                          uint8_t* __restrict__ luma) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 p = pos[i];
  const float fx = floorf((p.x - lim.x) / unit_w);
  const float fy = floorf((lim.w - p.y) / unit_h);
  // NaN and far-away stars fail these tests before any conversion to int
  if (!(fx >= 0.0f && fx < (float)width && fy >= 0.0f && fy < (float)height)) return;
  const int x = (int)fx, y = (int)fy;
  const float fs = floorf(sqrtf(weight[i]) / size_scale);
  const int s = fs >= 7.0f ? 7 : (fs >= 1.0f ? (int)fs : 0);  // NaN weight: the point only
  const int r2 = kStampR2[s];
  const int r = s >= 7 ? 4 : (s >= 5 ? 3 : (s >= 3 ? 2 : (s >= 1 ? 1 : 0)));
  const int y0 = max(y - r, 0), y1 = min(y + r, height - 1);
  const int x0 = max(x - r, 0), x1 = min(x + r, width - 1);
  for (int yy = y0; yy <= y1; ++yy) {
    const int dy = yy - y;
    uint8_t* row = luma + (size_t)yy * width;
    for (int xx = x0; xx <= x1; ++xx) {
      const int dx = xx - x;
      if (dx * dx + dy * dy <= r2) row[xx] = 255;
    }
  }
}

}  // namespace nbfmm_amd

using namespace nbfmm_amd;

struct nbfmm_y4m {
  FILE* fp = nullptr;
  size_t frame_bytes = 0;
  uint8_t* stage[2] = {nullptr, nullptr};  // pinned host staging, one frame each
  hipEvent_t done[2] = {nullptr, nullptr};
  int next = 0;        // staging slot of the next frame
  bool pending = false;  // slot next ^ 1 holds a frame not yet written to the file
  long frames = 0;
};

namespace {

int hip_error(hipError_t e, const char* what) {
  return report_error(NBFMM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define NB_DISPLAY_HIP(call)                          \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return hip_error(e_, #call); \
  } while (0)

// Waits for the staged frame in slot `k` and appends it to the file.
int flush_slot(nbfmm_y4m* y, int k) {
  NB_DISPLAY_HIP(hipEventSynchronize(y->done[k]));
  if (fputs("FRAME\n", y->fp) < 0 || fwrite(y->stage[k], 1, y->frame_bytes, y->fp) != y->frame_bytes)
    return report_error(NBFMM_ERR_IO, "nbfmm_y4m: writing a frame failed");
  ++y->frames;
  return NBFMM_OK;
}

void release(nbfmm_y4m* y) {
  for (int k = 0; k < 2; ++k) {
    if (y->done[k]) (void)hipEventDestroy(y->done[k]);
    if (y->stage[k]) (void)hipHostFree(y->stage[k]);
  }
  if (y->fp) fclose(y->fp);
  delete y;
}

}  // namespace

extern "C" {

int nbfmm_display(int num_star, int width, int height, float size_scale, const float display_limits[4],
                  const float* d_position, const float* d_weight, unsigned char* d_frame, void* stream) {
  if (width <= 0 || height <= 0 || (width & 1) || (height & 1))
    return report_error(NBFMM_ERR_INVALID_ARGUMENT, "nbfmm_display: width and height must be positive and even");
  if (num_star < 0 || !(size_scale > 0.0f) || !display_limits || !d_frame || (num_star > 0 && (!d_position || !d_weight)))
    return report_error(NBFMM_ERR_INVALID_ARGUMENT,
                        "nbfmm_display: num_star >= 0, size_scale > 0 and non-null limits and arrays");
  const float4 lim = make_float4(display_limits[0], display_limits[1], display_limits[2], display_limits[3]);
  if (!(lim.x < lim.z) || !(lim.y < lim.w))
    return report_error(NBFMM_ERR_INVALID_ARGUMENT, "nbfmm_display: display limits need x_min < x_max, y_min < y_max");
  hipStream_t st = (hipStream_t)stream;
  const size_t luma = (size_t)width * height;
  const size_t total = luma + luma / 2;
  const size_t chunks = (total + 15) / 16;
  hipLaunchKernelGGL(k_frame_clear, dim3((unsigned)((chunks + 255) / 256)), dim3(256), 0, st, luma, total,
                     (uint8_t*)d_frame);
  if (num_star > 0) {
    // display.cu:45-46
    const float unit_w = (lim.z - lim.x) / (float)width;
    const float unit_h = (lim.w - lim.y) / (float)height;
    hipLaunchKernelGGL(k_display, dim3((num_star + 255) / 256), dim3(256), 0, st, num_star, width, height, size_scale,
                       lim, unit_w, unit_h, (const float2*)d_position, d_weight, (uint8_t*)d_frame);
  }
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? NBFMM_OK : hip_error(e, "nbfmm_display launch");
}

int nbfmm_y4m_open(const char* path, int width, int height, int fps, nbfmm_y4m** writer) {
  if (!writer) return report_error(NBFMM_ERR_INVALID_ARGUMENT, "nbfmm_y4m_open: writer is null");
  *writer = nullptr;
  if (!path || width <= 0 || height <= 0 || (width & 1) || (height & 1) || fps <= 0)
    return report_error(NBFMM_ERR_INVALID_ARGUMENT,
                        "nbfmm_y4m_open: a path, positive even width and height, and fps > 0 are required");
  nbfmm_y4m* y = new nbfmm_y4m;
  y->frame_bytes = (size_t)width * height * 3 / 2;
  y->fp = fopen(path, "wb");
  if (!y->fp) {
    release(y);
    return report_error(NBFMM_ERR_IO, std::string("nbfmm_y4m_open: cannot open ") + path);
  }
  for (int k = 0; k < 2; ++k) {
    hipError_t e = hipHostMalloc((void**)&y->stage[k], y->frame_bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&y->done[k], hipEventDisableTiming);
    if (e != hipSuccess) {
      release(y);
      return hip_error(e, "nbfmm_y4m_open: staging buffers");
    }
  }
  // main.cu:235
  if (fprintf(y->fp, "YUV4MPEG2 W%d H%d F%d:%d Ip A1:1 C420\n", width, height, fps, 1) < 0) {
    release(y);
    return report_error(NBFMM_ERR_IO, "nbfmm_y4m_open: writing the header failed");
  }
  *writer = y;
  return NBFMM_OK;
}

int nbfmm_y4m_write_frame(nbfmm_y4m* y, const unsigned char* d_frame, void* stream) {
  if (!y || !d_frame) return report_error(NBFMM_ERR_INVALID_ARGUMENT, "nbfmm_y4m_write_frame: null argument");
  hipStream_t st = (hipStream_t)stream;
  const int k = y->next;
  NB_DISPLAY_HIP(hipMemcpyAsync(y->stage[k], d_frame, y->frame_bytes, hipMemcpyDeviceToHost, st));
  NB_DISPLAY_HIP(hipEventRecord(y->done[k], st));
  // the previous frame goes to the file while this one copies
  if (y->pending)
    if (int rc = flush_slot(y, k ^ 1)) return rc;
  y->pending = true;
  y->next = k ^ 1;
  return NBFMM_OK;
}

int nbfmm_y4m_close(nbfmm_y4m* y, long* num_frames) {
  if (!y) return report_error(NBFMM_ERR_INVALID_ARGUMENT, "nbfmm_y4m_close: null writer");
  int rc = NBFMM_OK;
  if (y->pending) rc = flush_slot(y, y->next ^ 1);
  if (num_frames) *num_frames = y->frames;
  if (rc == NBFMM_OK && fflush(y->fp) != 0) rc = report_error(NBFMM_ERR_IO, "nbfmm_y4m_close: flush failed");
  release(y);
  return rc;
}

}  // extern "C"
