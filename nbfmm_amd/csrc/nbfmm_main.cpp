// nbfmm_main: the reference's command-line program (main/main.cu:25-262)
// over the nbfmm_amd C ABI, with the same positional arguments, so that the
// reference's launcher scripts (main/script/*.py) drive it unchanged:
//
//   nbfmm_main <output file> <width> <height> <fps> <time (second)> <FMM level> <number of stars>
//              <position width> <position height> <display scale> <grav_const> <size_scale>
//              <model> [model parameters] ...
//
// It simulates `time * fps` frames, one Verlet step per frame, and writes them
// as a YUV4MPEG2 file.  Frame 0 shows the initial stars; every later frame is
// Stars::update (solve, Verlet, prune; update.cu:26-55) followed by
// Stars::display (display.cu:219-229).  Display's own trailing prune is a
// no-op after update's prune (the positions have not moved), so it runs only
// after frame 0.
//
// Two departures from main.cu, both where it does not do what it says:
//  - double-disk-center gives disk 2 its own centre weight (main.cu:210
//    passes weight_center1 twice) and asks for all eight parameters
//    (main.cu:188 checks argc <= 20, then reads argv[21]);
//  - the double-disk models simulate n1 + n2 stars, n_k = (int)(num_star *
//    p_k); main.cu:172 leaves any remaining stars of num_star uninitialised.
// The random draws of the models come from nbfmm_amd's counter-based hash
// with seed 0 (the NBFMM_SEED environment variable overrides it), not from
// cuRAND, so the stars differ draw by draw from the reference's.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "nbfmm_amd.h"

namespace {

int g_rc = 0;

bool ok(int rc, const char* what) {
  if (rc == NBFMM_OK) return true;
  std::fprintf(stderr, "nbfmm_main: %s failed (status %d): %s\n", what, rc, nbfmm_last_error());
  g_rc = rc;
  return false;
}

bool hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "nbfmm_main: %s: %s\n", what, hipGetErrorString(e));
  g_rc = NBFMM_ERR_HIP;
  return false;
}

int usage(const char* argv0, const char* tail) {
  std::printf("Usage: %s %s\n\n", argv0, tail);
  return 1;
}

struct Float2 {
  float x, y;
};

// main.cu:164-167, in the float2 arithmetic of operator.hpp
void double_disk_centers(const float lim[4], float p1, float p2, Float2& c1, Float2& c2) {
#pragma clang fp contract(off)
  const float a1 = 1 * p1 + 0 * p2, b1 = 1 * p1 + 2 * p2;
  const float a2 = 1 * p2 + 0 * p1, b2 = 1 * p2 + 2 * p1;
  c1 = {(lim[0] * a1 + lim[2] * b1) / 2, (lim[1] * a1 + lim[3] * b1) / 2};
  c2 = {(lim[2] * a2 + lim[0] * b2) / 2, (lim[3] * a2 + lim[1] * b2) / 2};
}

}  // namespace

int main(int argc, char const* argv[]) {
  const int version = nbfmm_version();
  std::printf("NBFMM_AMD %d.%d.%d\n\n", version / 10000, version / 100 % 100, version % 100);
  if (argc <= 13)
    return usage(argv[0],
                 "<output file> <width> <height> <fps> <time (second)> <FMM level> <number of stars> "
                 "<position width> <position height> <display scale> <grav_const> <size_scale> "
                 "<model> [model parameters] ...");

  const char* result_y4m = argv[1];
  const int width = std::atoi(argv[2]);
  const int height = std::atoi(argv[3]);
  const int fps = std::atoi(argv[4]);
  const int num_second = std::atoi(argv[5]);
  const int num_frame = num_second * fps;
  const int fmm_level = std::atoi(argv[6]);
  const int num_star_arg = std::atoi(argv[7]);
  const float pos_width = (float)std::atof(argv[8]);
  const float pos_height = (float)std::atof(argv[9]);
  const float display_scale = (float)std::atof(argv[10]);
  const float tick = (float)(0.05 / fps);
  const float grav_const = (float)std::atof(argv[11]);
  const float size_scale = (float)std::atof(argv[12]);
  const char* model = argv[13];
  const float eps = 0.1f;  // the reference's kernel-function softening (kernel_function.hpp:35)
  const char* seed_env = std::getenv("NBFMM_SEED");
  const unsigned seed = seed_env ? (unsigned)std::strtoul(seed_env, nullptr, 10) : 0u;

  if (width <= 0 || height <= 0 || fps <= 0 || num_star_arg <= 0 || !(grav_const > 0.0f) || !(size_scale > 0.0f) ||
      !(pos_width > 0.0f) || !(pos_height > 0.0f) || !(display_scale > 0.0f)) {
    std::fprintf(stderr, "nbfmm_main: sizes, fps, star count, grav_const and size_scale must be positive\n");
    return 1;
  }

  const float lim[4] = {-pos_width / 2, -pos_height / 2, pos_width / 2, pos_height / 2};
  const float dlim[4] = {lim[0] * display_scale, lim[1] * display_scale, lim[2] * display_scale,
                         lim[3] * display_scale};
  // Stars::Stars (stars.cu:31): tick_ = tick * sqrt(G)
  const float tick_g = tick * std::sqrt(grav_const);

  std::printf("%d x %d pixel, %d fps, %d sec / %d frame\n", width, height, fps, num_second, num_frame);
  std::printf("%d FMM level, %d stars, tick = %.6f, grav const = %.2f, scale size = %.2f\n", fmm_level, num_star_arg,
              tick, grav_const, size_scale);
  std::printf("position size = %.2f x %.2f, display scale = %.2f\n", pos_width, pos_height, display_scale);
  std::printf("position range = [%.2f, %.2f] x [%.2f, %.2f], display range = [%.2f, %.2f] x [%.2f, %.2f]\n\n",
              lim[0], lim[2], lim[1], lim[3], dlim[0], dlim[2], dlim[1], dlim[3]);

  const int n_alloc = num_star_arg;
  float *x = nullptr, *xp = nullptr, *acc = nullptr, *w = nullptr;
  unsigned char* frame = nullptr;
  const size_t frame_bytes = (size_t)width * height * 3 / 2;
  if (!hip_ok(hipMalloc(&x, 8 * (size_t)n_alloc), "hipMalloc") ||
      !hip_ok(hipMalloc(&xp, 8 * (size_t)n_alloc), "hipMalloc") ||
      !hip_ok(hipMalloc(&acc, 8 * (size_t)n_alloc), "hipMalloc") ||
      !hip_ok(hipMalloc(&w, 4 * (size_t)n_alloc), "hipMalloc") ||
      !hip_ok(hipMalloc(&frame, frame_bytes), "hipMalloc"))
    return 1;
  hipStream_t stream;
  if (!hip_ok(hipStreamCreate(&stream), "hipStreamCreate")) return 1;
  void* st = (void*)stream;

  // ---- model (main.cu:78-222)
  int n = num_star_arg;
  const int nargs = argc - 14;
  auto arg = [&](int k) { return (float)std::atof(argv[14 + k]); };
  bool done = false;
  if (!std::strcmp(model, "circle") || !std::strcmp(model, "circle-uniform")) {
    const bool uniform = model[6] == '-';
    if (nargs < 3) return usage(argv[0], uniform ? "... circle-uniform <radius> <weight> <center weight>"
                                                 : "... circle <radius> <weight> <center weight>");
    std::printf("Model %sCircle: radius = %.2f, weight = %.2f, center weight = %.2f\n\n", uniform ? "Uniform " : "",
                arg(0), arg(1), arg(2));
    done = ok(nbfmm_model_circle_center(n, 0.0f, 0.0f, arg(0), arg(1), arg(2), tick_g, eps, uniform ? 1 : 0, seed, x,
                                        xp, w, st),
              "nbfmm_model_circle_center");
  } else if (!std::strcmp(model, "disk")) {
    if (nargs < 2) return usage(argv[0], "... disk <radius> <weight>");
    std::printf("Model Disk: radius = %.2f, weight = %.2f\n\n", arg(0), arg(1));
    done = ok(nbfmm_model_disk(n, 0.0f, 0.0f, arg(0), arg(1), tick_g, eps, seed, x, xp, w, st), "nbfmm_model_disk");
  } else if (!std::strcmp(model, "disk-center")) {
    if (nargs < 3) return usage(argv[0], "... disk-center <radius> <weight> <center weight>");
    std::printf("Model Centered Disk: radius = %.2f, weight = %.2f, center weight = %.2f\n\n", arg(0), arg(1),
                arg(2));
    done = ok(nbfmm_model_disk_center(n, 0.0f, 0.0f, arg(0), arg(1), arg(2), tick_g, eps, seed, x, xp, w, st),
              "nbfmm_model_disk_center");
  } else if (!std::strcmp(model, "disk-static")) {
    if (nargs < 2) return usage(argv[0], "... disk-static <radius> <weight>");
    std::printf("Model Static Disk: radius = %.2f, weight = %.2f\n\n", arg(0), arg(1));
    done = ok(nbfmm_model_disk_static(n, 0.0f, 0.0f, arg(0), arg(1), seed, x, xp, w, st), "nbfmm_model_disk_static");
  } else if (!std::strcmp(model, "double-disk") || !std::strcmp(model, "double-disk-center")) {
    const bool centered = std::strlen(model) > 11;
    if (nargs < (centered ? 8 : 6))
      return usage(argv[0], centered ? "... double-disk-center <proportion 1> <proportion 2> <radius 1> <radius 2> "
                                       "<weight> <center weight 1> <center weight 2> <eccentricity>"
                                     : "... double-disk <proportion 1> <proportion 2> <radius 1> <radius 2> "
                                       "<weight> <eccentricity>");
    const float q1 = arg(0), q2 = arg(1);
    if (!(q1 >= 0.0f && q2 >= 0.0f && q1 + q2 > 0.0f)) {
      std::fprintf(stderr, "nbfmm_main: proportions must be >= 0 and not both 0\n");
      return 1;
    }
    const float p1 = q1 / (q1 + q2), p2 = q2 / (q1 + q2);
    const int n1 = (int)(num_star_arg * p1), n2 = (int)(num_star_arg * p2);
    const float weight = arg(4);
    const float cw1 = centered ? arg(5) : 0.0f, cw2 = centered ? arg(6) : 0.0f;
    const float ecc = arg(centered ? 7 : 5);
    Float2 c1, c2;
    double_disk_centers(dlim, p1, p2, c1, c2);
    std::printf("Model Double Disk%s: proportion = %.2f / %.2f, radius = %.2f / %.2f, weight = %.2f, "
                "eccentricity = %.2f\n\n",
                centered ? " Center" : "", q1, q2, arg(2), arg(3), weight, ecc);
    const float cc1[2] = {c1.x, c1.y}, cc2[2] = {c2.x, c2.y};
    n = n1 + n2;
    done = ok(nbfmm_model_double_disk(n1, n2, cc1, cc2, arg(2), arg(3), weight, cw1, cw2, ecc, tick_g, eps,
                                      centered ? 1 : 0, seed, x, xp, w, st),
              "nbfmm_model_double_disk");
  } else if (!std::strcmp(model, "rectangle")) {
    if (nargs < 3) return usage(argv[0], "... rectangle <rectangle width> <rectangle height> <maximum weight>");
    std::printf("Model Rectangle: size = %.2f x %.2f, max weight = %.2f\n\n", arg(0), arg(1), arg(2));
    done = ok(nbfmm_model_rectangle(n, 0.0f, 0.0f, arg(0), arg(1), arg(2), seed, x, xp, w, st),
              "nbfmm_model_rectangle");
  } else {
    std::printf("Unknown model %s.\n\n", model);
    return 1;
  }
  if (!done) return 1;

  nbfmm_solver* solver = nullptr;
  if (!ok(nbfmm_solver_create(fmm_level, n_alloc, lim, eps, &solver), "nbfmm_solver_create")) return 1;
  // P2P configuration timed on the initial stars (every configuration gives
  // the same bits): the fixed default for small systems splits each leaf's
  // run over 4 waves, which loses to 8-wave blocks when a few leaves hold
  // thousands of stars, as a centred disk's do.
  if (n > 0 && !ok(nbfmm_autotune(solver, n, x, w, 3, nullptr, nullptr, nullptr, st), "nbfmm_autotune")) return 1;
  nbfmm_y4m* video = nullptr;
  if (!ok(nbfmm_y4m_open(result_y4m, width, height, fps, &video), "nbfmm_y4m_open")) return 1;

  // ---- frames (main.cu:227-252)
  for (int i = 0; i < 100; ++i) std::putchar('=');
  std::putchar('\r');
  std::fflush(stdout);
  int progress = 0;
  bool good = ok(nbfmm_display(n, width, height, size_scale, dlim, x, w, frame, st), "nbfmm_display") &&
              ok(nbfmm_y4m_write_frame(video, frame, st), "nbfmm_y4m_write_frame") &&
              ok(nbfmm_prune(solver, n, x, xp, w, &n, st), "nbfmm_prune");
  for (int j = 1; good && j < num_frame; ++j) {
    if (n > 0)
      good = ok(nbfmm_step(solver, n, tick_g, x, xp, w, acc, st), "nbfmm_step") &&
             ok(nbfmm_prune(solver, n, x, xp, w, &n, st), "nbfmm_prune");
    good = good && ok(nbfmm_display(n, width, height, size_scale, dlim, x, w, frame, st), "nbfmm_display") &&
           ok(nbfmm_y4m_write_frame(video, frame, st), "nbfmm_y4m_write_frame");
    if (100 * j > progress * num_frame) {
      ++progress;
      std::putchar('>');
      std::fflush(stdout);
    }
  }
  std::printf("\n\n");
  long frames = 0;
  good = ok(nbfmm_y4m_close(video, &frames), "nbfmm_y4m_close") && good;
  std::printf("%ld frames, %d stars left\n", frames, n);
  nbfmm_solver_destroy(solver);
  (void)hipStreamDestroy(stream);
  (void)hipFree(frame);
  (void)hipFree(w);
  (void)hipFree(acc);
  (void)hipFree(xp);
  (void)hipFree(x);
  return good ? 0 : (g_rc ? g_rc : 1);
}
