// Stage tests of nbfmm::Solver written against the C++ mirror
// (include/nbfmm_amd/solver.hpp), in the shape of the reference's CppUnit
// suite (check/core/solver.hpp:18-68, check/core/solver/*.cu): one fixture
// with random particles, cell data and a random monotone head; each test
// writes one stage's inputs into the solver's buffers, runs that stage and
// compares with a host brute-force formula at absolute tolerance 1e-4.
//
// Build: see tests/cpp/Makefile.  Run: tests/cpp/_build/test_solver (needs a GPU).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <vector>

#include "nbfmm_amd/solver.hpp"

namespace {

int g_failures = 0;

#define EXPECT(cond)                                                      \
  do {                                                                    \
    if (!(cond)) {                                                        \
      std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);     \
      ++g_failures;                                                       \
      return;                                                             \
    }                                                                     \
  } while (0)

#define EXPECT_NEAR(a, b, tol) EXPECT(std::fabs((double)(a) - (double)(b)) <= (tol))

void hip_ok(hipError_t e) {
  if (e != hipSuccess) {
    std::printf("HIP error: %s\n", hipGetErrorString(e));
    std::exit(2);
  }
}

template <class T>
struct DeviceArray {
  T* p = nullptr;
  size_t n = 0;
  explicit DeviceArray(size_t count) : n(count) { hip_ok(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T))); }
  ~DeviceArray() { (void)hipFree(p); }
  void put(const std::vector<T>& h) { hip_ok(hipMemcpy(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice)); }
  std::vector<T> get() const {
    std::vector<T> h(n);
    hip_ok(hipDeviceSynchronize());
    hip_ok(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
    return h;
  }
};

// kernel_function.hpp:29-40 in double on the host, epsilon as the solver's
void kernel(float2 t, float2 s, float m, double eps, double& ax, double& ay) {
  const double dx = (double)s.x - t.x, dy = (double)s.y - t.y;
  const double r = std::sqrt(dx * dx + dy * dy);
  const double f = m / (r * r * r + eps);
  ax += dx * f;
  ay += dy * f;
}

struct Fixture {
  static constexpr int L = 3;
  static constexpr int B = 1 << (L + 1);
  static constexpr int C = B * B;
  static constexpr int N = 250;
  static constexpr float eps = 0.1f;
  const float4 limits = make_float4(0.0f, -1.0f, 8.0f, 2.0f);  // a non-square box
  std::vector<float2> position;
  std::vector<float> weight;
  std::vector<int2> index;
  std::vector<int> perm;
  std::vector<int> head;
  std::vector<float2> cell_position;  // cube [L][B][B]
  std::vector<float> cell_weight;

  static int cube(int l, int x, int y) { return l * C + y * B + x; }  // x, y in leaf units

  Fixture() : position(N), weight(N), index(N), perm(N), head(C + 1), cell_position(L * C), cell_weight(L * C) {
    std::mt19937 gen(20161205u);
    std::uniform_real_distribution<float> ux(limits.x, limits.z), uy(limits.y, limits.w);
    std::exponential_distribution<float> uw(1.0f);
    std::uniform_int_distribution<int> uh(0, 255);
    for (int i = 0; i < N; ++i) {
      position[i] = make_float2(ux(gen), uy(gen));
      weight[i] = uw(gen);
    }
    for (int l = 0; l < L; ++l) {
      const int s = 1 << l;
      for (int y = 0; y < B; y += s)
        for (int x = 0; x < B; x += s) {
          cell_position[cube(l, x, y)] = make_float2(ux(gen), uy(gen));
          cell_weight[cube(l, x, y)] = uw(gen);
        }
    }
    std::vector<long> csum(C + 1, 0);
    for (int c = 1; c <= C; ++c) csum[c] = csum[c - 1] + uh(gen);
    for (int c = 0; c <= C; ++c) head[c] = (int)(csum[c] * N / csum[C]);
    std::iota(perm.begin(), perm.end(), 0);
    std::shuffle(perm.begin(), perm.end(), gen);
    for (int c = 0; c < C; ++c)
      for (int i = head[c]; i < head[c + 1]; ++i) index[i] = make_int2(c % B, c / B);
  }
};

void test_predo(nbfmm::Solver& s, const Fixture& f) {
  const int N = Fixture::N, B = Fixture::B;
  DeviceArray<float2> d_pos(N);
  DeviceArray<float> d_w(N);
  d_pos.put(f.position);
  d_w.put(f.weight);
  s.predo(N, d_pos.p, d_w.p);
  DeviceArray<float2> pos(N);
  DeviceArray<float> w(N);
  DeviceArray<int2> idx(N);
  DeviceArray<int> perm(N), head(Fixture::C + 1);
  s.read(nbfmm::Buffer::position, pos.p, N);
  s.read(nbfmm::Buffer::weight, w.p, N);
  s.read(nbfmm::Buffer::index, idx.p, N);
  s.read(nbfmm::Buffer::perm, perm.p, N);
  s.read(nbfmm::Buffer::head, head.p);
  const auto P = pos.get();
  const auto W = w.get();
  const auto I = idx.get();
  const auto Q = perm.get();
  const auto H = head.get();
  const float cw = (f.limits.z - f.limits.x) / B, ch = (f.limits.w - f.limits.y) / B;
  for (int i = 0; i < N; ++i) {
    EXPECT(P[i].x == f.position[Q[i]].x && P[i].y == f.position[Q[i]].y);
    EXPECT(W[i] == f.weight[Q[i]]);
    EXPECT(P[i].x >= I[i].x * cw + f.limits.x && P[i].x < (I[i].x + 1) * cw + f.limits.x);
    EXPECT(P[i].y >= I[i].y * ch + f.limits.y && P[i].y < (I[i].y + 1) * ch + f.limits.y);
    // stable order: original index increases inside a cell
    if (i > 0 && I[i].x == I[i - 1].x && I[i].y == I[i - 1].y) EXPECT(Q[i] > Q[i - 1]);
  }
  for (int y = 0; y < B; ++y)
    for (int x = 0; x < B; ++x)
      for (int i = H[x + y * B]; i < H[x + y * B + 1]; ++i) EXPECT(I[i].x == x && I[i].y == y);
  EXPECT(H[0] == 0 && H[Fixture::C] == N);
}

void test_postdo(nbfmm::Solver& s, const Fixture& f) {
  const int N = Fixture::N;
  DeviceArray<float2> eff(N), out(N);
  DeviceArray<int> perm(N);
  eff.put(f.position);
  perm.put(f.perm);
  s.write(nbfmm::Buffer::effect, eff.p, N);
  s.write(nbfmm::Buffer::perm, perm.p, N);
  s.postdo(N, out.p);
  const auto O = out.get();
  for (int i = 0; i < N; ++i) EXPECT(O[f.perm[i]].x == f.position[i].x && O[f.perm[i]].y == f.position[i].y);
}

void load_particles(nbfmm::Solver& s, const Fixture& f) {
  const int N = Fixture::N;
  DeviceArray<float2> pos(N);
  DeviceArray<float> w(N);
  DeviceArray<int2> idx(N);
  DeviceArray<int> head(Fixture::C + 1);
  pos.put(f.position);
  w.put(f.weight);
  idx.put(f.index);
  head.put(f.head);
  s.write(nbfmm::Buffer::position, pos.p, N);
  s.write(nbfmm::Buffer::weight, w.p, N);
  s.write(nbfmm::Buffer::index, idx.p, N);
  s.write(nbfmm::Buffer::head, head.p);
  hip_ok(hipDeviceSynchronize());
}

void test_p2p(nbfmm::Solver& s, const Fixture& f) {
  const int N = Fixture::N;
  load_particles(s, f);
  s.p2p(N);
  DeviceArray<float2> eff(N);
  s.read(nbfmm::Buffer::effect, eff.p, N);
  const auto E = eff.get();
  for (int i = 0; i < N; ++i) {
    double ax = 0, ay = 0;
    for (int j = 0; j < N; ++j)
      if (i != j && std::abs(f.index[i].x - f.index[j].x) <= 1 && std::abs(f.index[i].y - f.index[j].y) <= 1)
        kernel(f.position[i], f.position[j], f.weight[j], Fixture::eps, ax, ay);
    EXPECT_NEAR(E[i].x, ax, 1e-4);
    EXPECT_NEAR(E[i].y, ay, 1e-4);
  }
}

void test_p2m(nbfmm::Solver& s, const Fixture& f) {
  const int N = Fixture::N, C = Fixture::C;
  load_particles(s, f);
  s.p2m(N);
  DeviceArray<float2> cp(Fixture::L * C);
  DeviceArray<float> cw(Fixture::L * C);
  s.read(nbfmm::Buffer::cell_position, cp.p);
  s.read(nbfmm::Buffer::cell_weight, cw.p);
  const auto P = cp.get();
  const auto W = cw.get();
  for (int c = 0; c < C; ++c) {
    double sx = 0, sy = 0, sw = 0;
    for (int i = f.head[c]; i < f.head[c + 1]; ++i) {
      sx += (double)f.position[i].x * f.weight[i];
      sy += (double)f.position[i].y * f.weight[i];
      sw += f.weight[i];
    }
    if (sw > 0) { sx /= sw; sy /= sw; }
    EXPECT_NEAR(P[c].x, sx, 1e-4);
    EXPECT_NEAR(P[c].y, sy, 1e-4);
    EXPECT_NEAR(W[c], sw, 1e-4);
  }
}

void test_m2m(nbfmm::Solver& s, const Fixture& f) {
  const int L = Fixture::L, B = Fixture::B, C = Fixture::C;
  std::vector<float2> p0(L * C, make_float2(0, 0));
  std::vector<float> w0(L * C, 0.0f);
  std::copy(f.cell_position.begin(), f.cell_position.begin() + C, p0.begin());
  std::copy(f.cell_weight.begin(), f.cell_weight.begin() + C, w0.begin());
  DeviceArray<float2> cp(L * C);
  DeviceArray<float> cw(L * C);
  cp.put(p0);
  cw.put(w0);
  s.write(nbfmm::Buffer::cell_position, cp.p);
  s.write(nbfmm::Buffer::cell_weight, cw.p);
  s.m2m();
  s.read(nbfmm::Buffer::cell_position, cp.p);
  s.read(nbfmm::Buffer::cell_weight, cw.p);
  const auto P = cp.get();
  const auto W = cw.get();
  // host recursion: parent = mass-weighted COM of its four children
  std::vector<double> hx(L * C), hy(L * C), hw(L * C);
  for (int c = 0; c < C; ++c) { hx[c] = p0[c].x; hy[c] = p0[c].y; hw[c] = w0[c]; }
  for (int l = 1; l < L; ++l) {
    const int sz = 1 << l, half = sz / 2;
    for (int y = 0; y < B; y += sz)
      for (int x = 0; x < B; x += sz) {
        double sx = 0, sy = 0, sw = 0;
        for (int dy = 0; dy <= half; dy += half)
          for (int dx = 0; dx <= half; dx += half) {
            const int k = Fixture::cube(l - 1, x + dx, y + dy);
            sx += hx[k] * hw[k];
            sy += hy[k] * hw[k];
            sw += hw[k];
          }
        const int k = Fixture::cube(l, x, y);
        hx[k] = sx / sw;
        hy[k] = sy / sw;
        hw[k] = sw;
        EXPECT_NEAR(P[k].x, hx[k], 1e-4);
        EXPECT_NEAR(P[k].y, hy[k], 1e-4);
        EXPECT_NEAR(W[k], hw[k], 1e-4);
      }
  }
}

void test_m2l(nbfmm::Solver& s, const Fixture& f) {
  const int L = Fixture::L, B = Fixture::B, C = Fixture::C;
  DeviceArray<float2> cp(L * C), ce(L * C);
  DeviceArray<float> cw(L * C);
  cp.put(f.cell_position);
  cw.put(f.cell_weight);
  s.write(nbfmm::Buffer::cell_position, cp.p);
  s.write(nbfmm::Buffer::cell_weight, cw.p);
  s.m2l();
  s.read(nbfmm::Buffer::cell_effect, ce.p);
  const auto E = ce.get();
  // O(cells^2): parents within one parent cell of each other, cells not adjacent
  for (int l = 0; l < L; ++l) {
    const int sz = 1 << l, psz = 2 * sz;
    for (int y = 0; y < B; y += sz)
      for (int x = 0; x < B; x += sz) {
        const int px = x / psz * psz, py = y / psz * psz;
        double ax = 0, ay = 0;
        for (int j = 0; j < B; j += sz)
          for (int i = 0; i < B; i += sz) {
            const int pi = i / psz * psz, pj = j / psz * psz;
            if (std::abs(pi - px) <= psz && std::abs(pj - py) <= psz && (std::abs(i - x) > sz || std::abs(j - y) > sz))
              kernel(f.cell_position[Fixture::cube(l, x, y)], f.cell_position[Fixture::cube(l, i, j)],
                     f.cell_weight[Fixture::cube(l, i, j)], Fixture::eps, ax, ay);
          }
        EXPECT_NEAR(E[Fixture::cube(l, x, y)].x, ax, 1e-4);
        EXPECT_NEAR(E[Fixture::cube(l, x, y)].y, ay, 1e-4);
      }
  }
}

void test_l2l(nbfmm::Solver& s, const Fixture& f) {
  const int L = Fixture::L, B = Fixture::B, C = Fixture::C;
  DeviceArray<float2> ce(L * C);
  ce.put(f.cell_position);  // random values used as effects
  s.write(nbfmm::Buffer::cell_effect, ce.p);
  s.l2l();
  s.read(nbfmm::Buffer::cell_effect, ce.p);
  const auto E = ce.get();
  // push every level's effect down to the leaves
  for (int y = 0; y < B; ++y)
    for (int x = 0; x < B; ++x) {
      double ax = 0, ay = 0;
      for (int l = 0; l < L; ++l) {
        const int sz = 1 << l;
        const float2 e = f.cell_position[Fixture::cube(l, x / sz * sz, y / sz * sz)];
        ax += e.x;
        ay += e.y;
      }
      EXPECT_NEAR(E[Fixture::cube(0, x, y)].x, ax, 1e-4);
      EXPECT_NEAR(E[Fixture::cube(0, x, y)].y, ay, 1e-4);
    }
}

void test_l2p(nbfmm::Solver& s, const Fixture& f) {
  const int N = Fixture::N, L = Fixture::L, C = Fixture::C;
  DeviceArray<float2> eff(N), ce(L * C);
  DeviceArray<int2> idx(N);
  eff.put(f.position);
  idx.put(f.index);
  ce.put(f.cell_position);
  s.write(nbfmm::Buffer::effect, eff.p, N);
  s.write(nbfmm::Buffer::index, idx.p, N);
  s.write(nbfmm::Buffer::cell_effect, ce.p);
  s.l2p(N);
  s.read(nbfmm::Buffer::effect, eff.p, N);
  const auto E = eff.get();
  for (int i = 0; i < N; ++i) {
    const float2 c = f.cell_position[Fixture::cube(0, f.index[i].x, f.index[i].y)];
    EXPECT_NEAR(E[i].x, (double)f.position[i].x + c.x, 1e-4);
    EXPECT_NEAR(E[i].y, (double)f.position[i].y + c.y, 1e-4);
  }
}

void test_update(nbfmm::Solver&, const Fixture& f) {
#pragma clang fp contract(off)
  const int N = Fixture::N;
  const float tick = 0.01f;
  std::vector<float2> pre(N), acc(N);
  for (int i = 0; i < N; ++i) {
    pre[i] = make_float2(f.position[i].x - 0.001f * f.weight[i], f.position[i].y + 0.002f);
    acc[i] = make_float2(f.weight[i] * 3.0f, -f.weight[i]);
  }
  DeviceArray<float2> d_cur(N), d_pre(N), d_acc(N);
  d_cur.put(f.position);
  d_pre.put(pre);
  d_acc.put(acc);
  nbfmm::update(N, tick, d_cur.p, d_pre.p, d_acc.p);
  const auto C = d_cur.get();
  const auto P = d_pre.get();
  for (int i = 0; i < N; ++i) {
    const float ex = 2.0f * f.position[i].x - pre[i].x + acc[i].x * tick * tick;
    const float ey = 2.0f * f.position[i].y - pre[i].y + acc[i].y * tick * tick;
    EXPECT(C[i].x == ex && C[i].y == ey);
    EXPECT(P[i].x == f.position[i].x && P[i].y == f.position[i].y);
  }
}

// Stars::prune (prune.cu:28-62): every third star is moved outside the box
// on one side or another (one lands exactly on x_max and stays); the kept
// ones come first, both groups in their previous order.
void test_prune(nbfmm::Solver& s, const Fixture& f) {
  const int N = Fixture::N;
  std::vector<float2> cur(f.position), pre(N);
  std::vector<float> w(f.weight);
  for (int i = 0; i < N; ++i) {
    pre[i] = make_float2(-f.position[i].y, f.position[i].x);
    if (i % 3 == 0) {
      const int side = (i / 3) % 5;
      if (side == 0) cur[i].x = f.limits.x - 0.5f;
      if (side == 1) cur[i].y = f.limits.y - 0.5f;
      if (side == 2) cur[i].x = f.limits.z + 0.5f;
      if (side == 3) cur[i].y = f.limits.w + 0.5f;
      if (side == 4) cur[i].x = f.limits.z;  // on the edge: kept
    }
  }
  std::vector<int> order;
  for (int pass = 0; pass < 2; ++pass)
    for (int i = 0; i < N; ++i) {
      const bool out = cur[i].x < f.limits.x || cur[i].y < f.limits.y || cur[i].x > f.limits.z || cur[i].y > f.limits.w;
      if (out == (pass == 1)) order.push_back(i);
    }
  int expect_kept = 0;
  for (int i = 0; i < N; ++i)
    expect_kept += !(cur[i].x < f.limits.x || cur[i].y < f.limits.y || cur[i].x > f.limits.z || cur[i].y > f.limits.w);
  DeviceArray<float2> d_cur(N), d_pre(N);
  DeviceArray<float> d_w(N);
  d_cur.put(cur);
  d_pre.put(pre);
  d_w.put(w);
  const int kept = s.prune(N, d_cur.p, d_pre.p, d_w.p);
  EXPECT(kept == expect_kept);
  const auto C = d_cur.get();
  const auto P = d_pre.get();
  const auto W = d_w.get();
  for (int j = 0; j < N; ++j) {
    const int i = order[j];
    EXPECT(C[j].x == cur[i].x && C[j].y == cur[i].y);
    EXPECT(P[j].x == pre[i].x && P[j].y == pre[i].y);
    EXPECT(W[j] == w[i]);
  }
}

}  // namespace

int main() {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    std::printf("no GPU\n");
    return 2;
  }
  Fixture f;
  nbfmm::Solver solver(Fixture::L, 256, f.limits, Fixture::eps);
  struct {
    const char* name;
    void (*fn)(nbfmm::Solver&, const Fixture&);
  } tests[] = {{"predo", test_predo}, {"postdo", test_postdo}, {"p2p", test_p2p}, {"p2m", test_p2m},
               {"m2m", test_m2m},     {"m2l", test_m2l},       {"l2l", test_l2l}, {"l2p", test_l2p},
               {"update", test_update}, {"prune", test_prune}};
  for (auto& t : tests) {
    const int before = g_failures;
    t.fn(solver, f);
    std::printf("%-8s %s\n", t.name, g_failures == before ? "ok" : "FAILED");
  }
  std::printf(g_failures ? "FAILURES: %d\n" : "ALL PASSED\n", g_failures);
  return g_failures ? 1 : 0;
}
