// C++ host-side mirror of nbfmm::Solver over the nbfmm_amd C ABI.
//
// Same class name, constructor arguments and methods as the reference
// (include/nbfmm/core/solver.hpp:22-209): Solver(num_level, max_num_particle,
// position_limits), solve(), the eight stage methods, and the solver's
// buffers (what NBFMM_CHECK makes public) through read()/write() in the
// reference layouts.  Two differences: the kernel softening is a constructor
// argument (the reference hard-codes 0.1, kernel_function.hpp:35), and a
// failing call throws nbfmm::Error instead of asserting.
//
// Header-only; link with libnbfmm_amd.so.  All pointers are device pointers.
#ifndef NBFMM_AMD_SOLVER_HPP_
#define NBFMM_AMD_SOLVER_HPP_

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "../nbfmm_amd.h"

namespace nbfmm {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc) {
  if (rc != NBFMM_OK) throw Error(rc, std::string("nbfmm: ") + nbfmm_last_error());
}

// Buffers of the reference Solver (include/nbfmm/core/solver.hpp:45-76).
enum class Buffer : int {
  position = NBFMM_BUF_POSITION,
  effect = NBFMM_BUF_EFFECT,
  weight = NBFMM_BUF_WEIGHT,
  index = NBFMM_BUF_INDEX,
  perm = NBFMM_BUF_PERM,
  head = NBFMM_BUF_HEAD,
  cell_position = NBFMM_BUF_CELL_POSITION,
  cell_effect = NBFMM_BUF_CELL_EFFECT,
  cell_weight = NBFMM_BUF_CELL_WEIGHT,
};

class Solver {
 public:
  Solver(int num_level, int max_num_particle, float4 position_limits, float epsilon = 0.1f,
         hipStream_t stream = nullptr)
      : num_level_(num_level), base_dim_(1 << (num_level + 1)), max_num_particle_(max_num_particle),
        stream_(stream) {
    const float lim[4] = {position_limits.x, position_limits.y, position_limits.z, position_limits.w};
    check(nbfmm_solver_create(num_level, max_num_particle, lim, epsilon, &h_));
  }
  ~Solver() {
    if (h_) nbfmm_solver_destroy(h_);
  }
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;

  int num_level() const { return num_level_; }
  int base_dim() const { return base_dim_; }
  int max_num_particle() const { return max_num_particle_; }
  nbfmm_solver* handle() const { return h_; }

  // solve.cu:11-27
  void solve(int num_particle, const float2* position_origin, const float* weight_origin, float2* effect_origin) {
    check(nbfmm_solve(h_, num_particle, (const float*)position_origin, weight_origin, (float*)effect_origin,
                      stream_));
  }

  // the eight stages (solver.hpp:111-207)
  void predo(int num_particle, const float2* position_origin, const float* weight_origin) {
    check(nbfmm_predo(h_, num_particle, (const float*)position_origin, weight_origin, stream_));
  }
  void postdo(int num_particle, float2* effect_origin) {
    check(nbfmm_postdo(h_, num_particle, (float*)effect_origin, stream_));
  }
  void p2p(int num_particle) { check(nbfmm_p2p(h_, num_particle, stream_)); }
  void p2m(int num_particle) { check(nbfmm_p2m(h_, num_particle, stream_)); }
  void m2m() { check(nbfmm_m2m(h_, stream_)); }
  void m2l() { check(nbfmm_m2l(h_, stream_)); }
  void l2l() { check(nbfmm_l2l(h_, stream_)); }
  void l2p(int num_particle) { check(nbfmm_l2p(h_, num_particle, stream_)); }

  // buffer access in the reference layouts; count = particles for the
  // per-particle buffers
  size_t bytes(Buffer b, int count = 0) const {
    size_t n = 0;
    check(nbfmm_buffer_bytes(h_, (int)b, count, &n));
    return n;
  }
  void read(Buffer b, void* device_dst, int count = 0) {
    check(nbfmm_buffer_read(h_, (int)b, count, device_dst, stream_));
  }
  void write(Buffer b, const void* device_src, int count = 0) {
    check(nbfmm_buffer_write(h_, (int)b, count, device_src, stream_));
  }

  // Stars::update without prune (update.cu:47-55)
  void step(int num_particle, float tick, float2* position_cur, float2* position_pre, const float* weight,
            float2* acceleration) {
    check(nbfmm_step(h_, num_particle, tick, (float*)position_cur, (float*)position_pre, weight,
                     (float*)acceleration, stream_));
  }

  // Stars::prune (prune.cu:28-62) with this solver's limits: returns the new star count
  int prune(int num_star, float2* position_cur, float2* position_pre, float* weight) {
    int kept = 0;
    check(nbfmm_prune(h_, num_star, (float*)position_cur, (float*)position_pre, weight, &kept, stream_));
    return kept;
  }

  // Stars::update (update.cu:47-55): solve, Verlet update, prune; returns the new star count
  int update_stars(int num_star, float tick, float2* position_cur, float2* position_pre, float* weight,
                   float2* acceleration) {
    step(num_star, tick, position_cur, position_pre, weight, acceleration);
    return prune(num_star, position_cur, position_pre, weight);
  }

  unsigned flags(bool clear = false) {
    unsigned f = 0;
    check(nbfmm_solver_flags(h_, &f, clear ? 1 : 0));
    return f;
  }

 private:
  int num_level_, base_dim_, max_num_particle_;
  hipStream_t stream_;
  nbfmm_solver* h_ = nullptr;
};

// updateDevice (update.cu:26-42)
inline void update(int num_star, float tick, float2* position_cur, float2* position_pre,
                   const float2* acceleration, hipStream_t stream = nullptr) {
  check(nbfmm_update(num_star, tick, (float*)position_cur, (float*)position_pre, (const float*)acceleration,
                     stream));
}

}  // namespace nbfmm

#endif  // NBFMM_AMD_SOLVER_HPP_
