// Initial conditions of the reference's star models (source/nbfmm/model/*.cu).
//
// Every model places particle i from its own counter-based random numbers,
// so a model is a pure function of (parameters, seed): the reference draws
// curand_uniform from curand_init(0, i, 0); here u = ((h >> 8) + 1) / 2^24,
// in (0, 1] like curand_uniform, with h = lowbias32(lowbias32(seed ^
// 0x9e3779b9) ^ (4 i + k)) for the k-th draw of particle i (lowbias32: C.
// Wellons, hash-prospector; the constant keeps seed 0 off the hash's fixed
// point at 0).  Circular models give each star its previous position on the circle of
// its radius, rotated back by the chord angle of one step,
//   angle_difference = acos(1 - a tick^2 / (2 r)),
// a being the acceleration towards the centre (disk.cu:49-55), and the
// double-disk models shift the second disk's previous positions to set the
// pair's orbit (double_disk.cu:61-86).  The kernel function's softening is a
// parameter (the reference's kernelFunction uses 0.1).
#include "launch.hpp"
#include "../../include/nbfmm_amd.h"

namespace nbfmm_amd {

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

__host__ __device__ inline uint32_t lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

// k-th uniform draw of particle i, in (0, 1]
__host__ __device__ inline float model_uniform(uint32_t seed, int i, int k) {
  const uint32_t h = lowbias32(lowbias32(seed ^ 0x9e3779b9u) ^ (uint32_t)(4 * i + k));
  return (float)((h >> 8) + 1u) * (1.0f / 16777216.0f);
}

// kernel_function.hpp:29-40 with correctly rounded sqrt and division: the
// x component of the effect at the origin of mass m at (r, 0)
__host__ __device__ inline float radial_effect(float r, float m, float eps) {
#pragma clang fp contract(off)
  const float dx = r;
  const float rr = sqrtf(dx * dx);
  const float tmp = m / (rr * rr * rr + eps);
  return dx * tmp;
}

// disk.cu:49-50: the chord angle of one step on a circle of radius r
__host__ __device__ inline float chord_angle(float a, float r, float tick) {
#pragma clang fp contract(off)
  const float q = a * tick * tick / r / 2.0f;
  return (float)acos(1.0 - (double)q);
}

__device__ inline void place(float2 c, float r, float ang, float2& p) {
#pragma clang fp contract(off)
  p.x = c.x + r * cosf(ang);
  p.y = c.y + r * sinf(ang);
}

// disk.cu:36-57 / disk_center.cu / disk_static.cu: radius (i+1)/n * R, angle
// uniform; enclosed mass weight * i (+ the centre star for the centred disk,
// which sits at i = 0 with center_weight); static: no rotation.
__global__ void k_model_disk(int n, float2 c, float R, float weight, float center_weight, int centered, int rotate,
                             float tick, float eps, uint32_t seed, float2* xcur, float2* xpre, float* w) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (centered && i == 0) {
    xcur[i] = c;
    xpre[i] = c;
    w[i] = center_weight;
    return;
  }
  const float r = ((float)(i + 1) / (float)n) * R;
  float dang = 0.0f;
  if (rotate) {
    const float m = centered ? weight * (float)(i - 1) + center_weight : weight * (float)i;
    dang = chord_angle(radial_effect(r, m, eps), r, tick);
  }
  const float ang = (float)(2.0 * 3.14159265358979323846 * (double)model_uniform(seed, i, 0));
  place(c, r, ang, xcur[i]);
  place(c, r, ang - dang, xpre[i]);
  w[i] = weight;
}

// circle_center.cu / circle_uniform_center.cu: a centre star of
// center_weight at i = 0 and the others on one circle, at random angles or
// at 2 pi i / n, all rotated back by the same chord angle.
__global__ void k_model_circle(int n, float2 c, float R, float weight, float center_weight, float dang, int uniform,
                               uint32_t seed, float2* xcur, float2* xpre, float* w) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i == 0) {
    xcur[i] = c;
    xpre[i] = c;
    w[i] = center_weight;
    return;
  }
  const float ang = uniform ? (float)((2.0 * 3.14159265358979323846 * (double)i) / (double)n)
                            : (float)(2.0 * 3.14159265358979323846 * (double)model_uniform(seed, i, 0));
  place(c, R, ang, xcur[i]);
  place(c, R, ang - dang, xpre[i]);
  w[i] = weight;
}

// rectangle.cu: uniform in the rectangle, at rest, weight max_weight * u
__global__ void k_model_rectangle(int n, float2 c, float width, float height, float max_weight, uint32_t seed,
                                  float2* xcur, float2* xpre, float* w) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float2 p;
  p.x = c.x + (model_uniform(seed, i, 0) - 0.5f) * width;
  p.y = c.y + (model_uniform(seed, i, 1) - 0.5f) * height;
  xcur[i] = p;
  xpre[i] = p;
  w[i] = max_weight * model_uniform(seed, i, 2);
}

__global__ void k_shift(int n, float2 d, float2* x) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i].x = x[i].x + d.x;
  x[i].y = x[i].y + d.y;
}

// double_disk.cu:61-86: the two disks' previous positions are shifted so
// that the pair starts on a (perturbed) circular orbit about the common
// centre of mass: each disk moves perpendicular to the other's pull by
// sqrt(r_k / a_k) * tick / 2^e and is pulled back by a_k * tick^2 * e.
// m1, m2: the disks' masses; q1, q2: the weights of the centre-of-mass split
// r_1 = r q2 / (q1 + q2) (particle counts for the plain pair, masses for the
// centred one, as the reference writes them).
static void orbit_offsets(float2 c1, float2 c2, float m1, float m2, float q1, float q2, float ecc, float tick,
                          float eps, float2& o1, float2& o2) {
#pragma clang fp contract(off)
  // kernelFunction(c1, c2, m2) and (c2, c1, m1)
  const float dx = c2.x - c1.x, dy = c2.y - c1.y;
  const float r = sqrtf(dx * dx + dy * dy);
  const float t1 = m2 / (r * r * r + eps), t2 = m1 / (r * r * r + eps);
  const float2 e1 = make_float2(dx * t1, dy * t1), e2 = make_float2(-dx * t2, -dy * t2);
  const float a1 = sqrtf(e1.x * e1.x + e1.y * e1.y);
  const float a2 = sqrtf(e2.x * e2.x + e2.y * e2.y);
  const float r1 = r * q2 / (q1 + q2);
  const float r2 = r * q1 / (q1 + q2);
  const float s1 = sqrtf(r1 / a1) * tick / exp2f(ecc);
  const float s2 = sqrtf(r2 / a2) * tick / exp2f(ecc);
  o1 = make_float2(-e1.y * s1 - e1.x * tick * tick * ecc, e1.x * s1 - e1.y * tick * tick * ecc);
  o2 = make_float2(-e2.y * s2 - e2.x * tick * tick * ecc, e2.x * s2 - e2.y * tick * tick * ecc);
}

}  // namespace nbfmm_amd

using namespace nbfmm_amd;

namespace {
int model_check(int n, const float* xcur, const float* xpre, const float* w) {
  if (n <= 0) return NBFMM_ERR_INVALID_ARGUMENT;
  if (!xcur || !xpre || !w) return NBFMM_ERR_INVALID_ARGUMENT;
  return NBFMM_OK;
}
int invalid(const char* model) {
  return report_error(NBFMM_ERR_INVALID_ARGUMENT,
                      std::string(model) + ": invalid argument (counts > 0, non-null arrays, positive sizes, "
                                           "weights and eps, eccentricity >= 0)");
}
int launched() {
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return NBFMM_OK;
  return report_error(NBFMM_ERR_HIP, std::string("model kernel launch: ") + hipGetErrorString(e));
}
}  // namespace

extern "C" {

int nbfmm_model_disk(int n, float cx, float cy, float radius, float weight, float tick, float eps, unsigned seed,
                     float* xcur, float* xpre, float* w, void* stream) {
  if (model_check(n, xcur, xpre, w) || !(radius > 0.0f) || !(weight > 0.0f) || !(eps > 0.0f))
    return invalid("nbfmm_model_disk");
  hipLaunchKernelGGL(k_model_disk, dim3(cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, n, make_float2(cx, cy),
                     radius, weight, 0.0f, 0, 1, tick, eps, seed, (float2*)xcur, (float2*)xpre, w);
  return launched();
}

int nbfmm_model_disk_static(int n, float cx, float cy, float radius, float weight, unsigned seed, float* xcur,
                            float* xpre, float* w, void* stream) {
  if (model_check(n, xcur, xpre, w) || !(radius > 0.0f) || !(weight > 0.0f)) return invalid("nbfmm_model_disk_static");
  hipLaunchKernelGGL(k_model_disk, dim3(cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, n, make_float2(cx, cy),
                     radius, weight, 0.0f, 0, 0, 0.0f, 1.0f, seed, (float2*)xcur, (float2*)xpre, w);
  return launched();
}

int nbfmm_model_disk_center(int n, float cx, float cy, float radius, float weight, float center_weight, float tick,
                            float eps, unsigned seed, float* xcur, float* xpre, float* w, void* stream) {
  if (model_check(n, xcur, xpre, w) || !(radius > 0.0f) || !(weight > 0.0f) || !(eps > 0.0f))
    return invalid("nbfmm_model_disk_center");
  hipLaunchKernelGGL(k_model_disk, dim3(cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, n, make_float2(cx, cy),
                     radius, weight, center_weight, 1, 1, tick, eps, seed, (float2*)xcur, (float2*)xpre, w);
  return launched();
}

int nbfmm_model_circle_center(int n, float cx, float cy, float radius, float weight, float center_weight, float tick,
                              float eps, int uniform_angles, unsigned seed, float* xcur, float* xpre, float* w,
                              void* stream) {
  if (model_check(n, xcur, xpre, w) || !(radius > 0.0f) || !(weight > 0.0f) || !(eps > 0.0f))
    return invalid("nbfmm_model_circle_center");
  // circle_center.cu:70-72: one chord angle for the whole circle, pulled by the centre star
  const float dang = chord_angle(radial_effect(radius, center_weight, eps), radius, tick);
  hipLaunchKernelGGL(k_model_circle, dim3(cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, n, make_float2(cx, cy),
                     radius, weight, center_weight, dang, uniform_angles ? 1 : 0, seed, (float2*)xcur,
                     (float2*)xpre, w);
  return launched();
}

int nbfmm_model_rectangle(int n, float cx, float cy, float width, float height, float max_weight, unsigned seed,
                          float* xcur, float* xpre, float* w, void* stream) {
  if (model_check(n, xcur, xpre, w) || !(width > 0.0f) || !(height > 0.0f) || !(max_weight > 0.0f))
    return invalid("nbfmm_model_rectangle");
  hipLaunchKernelGGL(k_model_rectangle, dim3(cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream, n,
                     make_float2(cx, cy), width, height, max_weight, seed, (float2*)xcur, (float2*)xpre, w);
  return launched();
}

int nbfmm_model_double_disk(int n1, int n2, const float c1[2], const float c2[2], float radius1, float radius2,
                            float weight, float center_weight1, float center_weight2, float eccentricity, float tick,
                            float eps, int centered, unsigned seed, float* xcur, float* xpre, float* w,
                            void* stream) {
  if (n1 <= 0 || n2 <= 0 || !c1 || !c2 || !xcur || !xpre || !w || !(eccentricity >= 0.0f) || !(eps > 0.0f))
    return invalid("nbfmm_model_double_disk");
  hipStream_t st = (hipStream_t)stream;
  int rc;
  if (centered) {
    if ((rc = nbfmm_model_disk_center(n1, c1[0], c1[1], radius1, weight, center_weight1, tick, eps, seed, xcur, xpre,
                                      w, stream)))
      return rc;
    if ((rc = nbfmm_model_disk_center(n2, c2[0], c2[1], radius2, weight, center_weight2, tick, eps, seed,
                                      xcur + 2 * n1, xpre + 2 * n1, w + n1, stream)))
      return rc;
  } else {
    if ((rc = nbfmm_model_disk(n1, c1[0], c1[1], radius1, weight, tick, eps, seed, xcur, xpre, w, stream))) return rc;
    if ((rc = nbfmm_model_disk(n2, c2[0], c2[1], radius2, weight, tick, eps, seed, xcur + 2 * n1, xpre + 2 * n1,
                               w + n1, stream)))
      return rc;
  }
  // masses that set the orbit: weight * n (double_disk.cu:64-65), or the
  // centred disks' weight * (n - 1) + center_weight (double_disk_center.cu).
  // Both disks draw the same uniform stream (the reference seeds each
  // generateDisk call alike), so their angles coincide index by index.
  const float m1 = centered ? weight * (float)(n1 - 1) + center_weight1 : weight * (float)n1;
  const float m2 = centered ? weight * (float)(n2 - 1) + center_weight2 : weight * (float)n2;
  const float q1 = centered ? m1 : (float)n1;
  const float q2 = centered ? m2 : (float)n2;
  float2 o1, o2;
  orbit_offsets(make_float2(c1[0], c1[1]), make_float2(c2[0], c2[1]), m1, m2, q1, q2, eccentricity, tick, eps, o1,
                o2);
  hipLaunchKernelGGL(k_shift, dim3(cdiv(n1, 256)), dim3(256), 0, st, n1, o1, (float2*)xpre);
  hipLaunchKernelGGL(k_shift, dim3(cdiv(n2, 256)), dim3(256), 0, st, n2, o2, (float2*)xpre + n1);
  return launched();
}

}  // extern "C"
