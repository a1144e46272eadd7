// nbfmm CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// A plain C++ restatement of the emfomy/nbfmm force-evaluation path
// (nbfmm::Solver::solve) and its Stormer-Verlet update, used by the tests as
// the parity checker and by bench.py as the timed CPU baseline.  Nothing in the
// product library links, loads or calls this file.
//
// Every stage follows the reference source expression by expression; each
// function names the reference file:line it restates.  Where the reference has
// a defect (head tail never written when N % 1024 == 0, the zero/scatter race in
// p2mAssigning, no clamp of the cell index) the oracle implements the intended
// semantics that the reference's own unit tests pin.
//
// Layouts are the reference's:
//   position float2[N] (x,y interleaved), weight float[N], index int2[N],
//   perm int[N], head int[B*B+1],
//   cell cubes [levels][B][B] where level-l cell (X,Y) lives at
//   l*B*B + (Y<<l)*B + (X<<l)                      (m2m.cu:35, m2l.cu:41).
//
// Arithmetic: compiled with -ffp-contract=off so every float operation rounds
// exactly as written (the reference is built without fast-math,
// CMakeLists.txt:18).  Templated on the real type; the float instantiation is
// the parity oracle, the double one serves finite-difference checks.
//
// Parity status of each function (which check pins it) is listed in the
// header comment of each function.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// include/nbfmm/core/kernel_function.hpp:29-40.  The reference hard-codes
// epsilon = 0.1 (:35); here it is a parameter.
//   d = s - t;  r = sqrt(d.d);  tmp = m / (r*r*r + eps);  a = d*tmp
// Pinned by: tests/test_oracle.py closed-form two-body values and the
// exact-regime checks against an independently written fp64 direct sum.
template <class T>
inline void kernel_accum(T tx, T ty, T sx, T sy, T m, T eps, T& ax, T& ay) {
  const T dx = sx - tx;
  const T dy = sy - ty;
  const T r = std::sqrt(dx * dx + dy * dy);
  const T tmp = m / (r * r * r + eps);
  ax += dx * tmp;
  ay += dy * tmp;
}

inline int cube_index(int B, int l, int X, int Y) { return l * B * B + (Y << l) * B + (X << l); }

void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

// source/nbfmm/core/solver/predo.cu:26-39 (computeParticleIndex) with the cell
// size of predo.cu:112-113.  The reference does not clamp (a particle at
// x_max gets index B); the oracle clamps into [0, B-1] and counts the clamps,
// the same contract the HIP library implements (out-of-box flag).
// Pinned by: bit-exact comparison with numpy float32 (x - lo) / cell, floor.
template <class T>
int compute_index(int n, int B, const T* pos, T xmin, T ymin, T cw, T ch, int* index) {
  int clamped = 0;
  for (int i = 0; i < n; ++i) {
    const T vx = std::floor((pos[2 * i] - xmin) / cw);
    const T vy = std::floor((pos[2 * i + 1] - ymin) / ch);
    int ix, iy;
    bool bad = false;
    if (!(vx >= T(0))) { ix = 0; bad = true; }
    else if (vx >= T(B)) { ix = B - 1; bad = true; }
    else ix = (int)vx;
    if (!(vy >= T(0))) { iy = 0; bad = true; }
    else if (vy >= T(B)) { iy = B - 1; bad = true; }
    else iy = (int)vy;
    index[2 * i] = ix;
    index[2 * i + 1] = iy;
    clamped += bad ? 1 : 0;
  }
  return clamped;
}

// predo.cu:105-135: index, perm = stable sort of (iy, ix) row-major
// (thrust::sort_by_key with operator< of include/nbfmm/utility/operator.hpp:107-109),
// head = first sorted slot of each cell with head[B*B] = N (intended
// semantics of extractHead, predo.cu:49-74), gather of position and weight
// (permuteInputVector, predo.cu:86-100).
// The stable sort is a counting sort on key = iy*B + ix: equal keys keep
// their input order.
// Pinned by: numpy argsort(kind="stable"), bincount/cumsum head and the
// predo checks of check/core/solver/predo.cu:50-66.
// given (finite differences with the cells held fixed, see orc_run_f64_cells):
// the cell index pairs to use instead of computing them; record: where to
// write the ones used.
template <class T>
int predo(int L, int n, const T* pos0, const T* w0, const T limits[4], T* pos, T* w, int* index,
          int* perm, int* head, const int* given = nullptr, int* record = nullptr) {
  const int B = 1 << (L + 1);
  const int C = B * B;
  const T cw = (limits[2] - limits[0]) / T(B);
  const T ch = (limits[3] - limits[1]) / T(B);
  std::vector<int> idx0(2 * (size_t)std::max(n, 1));
  int clamped = 0;
  if (given)
    std::copy(given, given + 2 * (size_t)n, idx0.begin());
  else
    clamped = compute_index<T>(n, B, pos0, limits[0], limits[1], cw, ch, idx0.data());
  if (record) std::copy(idx0.begin(), idx0.begin() + 2 * (size_t)n, record);
  std::vector<int> count(C + 1, 0);
  for (int i = 0; i < n; ++i) count[idx0[2 * i] + idx0[2 * i + 1] * B]++;
  head[0] = 0;
  for (int c = 0; c < C; ++c) head[c + 1] = head[c] + count[c];
  std::vector<int> fill(head, head + C);
  for (int i = 0; i < n; ++i) {
    const int key = idx0[2 * i] + idx0[2 * i + 1] * B;
    const int p = fill[key]++;
    perm[p] = i;
  }
  for (int p = 0; p < n; ++p) {
    const int i = perm[p];
    pos[2 * p] = pos0[2 * i];
    pos[2 * p + 1] = pos0[2 * i + 1];
    w[p] = w0[i];
    index[2 * p] = idx0[2 * i];
    index[2 * p + 1] = idx0[2 * i + 1];
  }
  return clamped;
}

// source/nbfmm/core/solver/p2p.cu:30-69 (p2pNaive): for each sorted particle,
// sum the kernel over the particles of the 3x3 neighbour cells, skipping
// itself; cells in row-major order, particles in sorted order.
// Pinned by: the all-pairs |dix|<=1 && |diy|<=1 formula of
// check/core/solver/p2p.cu:34-41 (restated in numpy, fp64), and L=0 == direct sum.
template <class T>
void p2p(int L, int n, const T* pos, const T* w, const int* index, const int* head, T eps, T* eff) {
  const int B = 1 << (L + 1);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    const int px = index[2 * i], py = index[2 * i + 1];
    const T tx = pos[2 * i], ty = pos[2 * i + 1];
    T ax = T(0), ay = T(0);
    for (int j = py - 1; j <= py + 1; ++j) {
      for (int k = px - 1; k <= px + 1; ++k) {
        if (k >= 0 && k < B && j >= 0 && j < B) {
          const int c = k + j * B;
          for (int s = head[c]; s < head[c + 1]; ++s) {
            if (s != i) kernel_accum<T>(tx, ty, pos[2 * s], pos[2 * s + 1], w[s], eps, ax, ay);
          }
        }
      }
    }
    eff[2 * i] = ax;
    eff[2 * i + 1] = ay;
  }
}

// source/nbfmm/core/solver/p2m.cu:24-110, intended semantics (every leaf is
// zeroed, then occupied leaves get centre of mass = sum(m*x)/sum(m) and mass
// sum(m); the reference's zero/scatter race p2m.cu:62-73 is not reproduced).
// The reference's summation order is Thrust's reduce_by_key, which nothing
// pins; the order here is the GPU's: a leaf's sorted range is cut at the
// multiples of kP2MSeg (its segments), each segment is summed in sorted
// order from 0, and the segment sums are added in order from 0.  A leaf that
// lies within one segment is summed plainly in sorted order.
// Pinned by: check/core/solver/p2m.cu:36-46 restated in numpy (fp64).
constexpr int kP2MSeg = 16;

template <class T>
void p2m(int L, int n, const T* pos, const T* w, const int* head, T* cell_pos, T* cell_w) {
  (void)n;
  if (L <= 0) return;
  const int B = 1 << (L + 1);
  for (int c = 0; c < B * B; ++c) {
    T sx = T(0), sy = T(0), sw = T(0);
    for (int s = head[c]; s < head[c + 1];) {
      const int e = std::min(head[c + 1], (s / kP2MSeg + 1) * kP2MSeg);
      T px = T(0), py = T(0), pw = T(0);
      for (int i = s; i < e; ++i) {
        px += pos[2 * i] * w[i];
        py += pos[2 * i + 1] * w[i];
        pw += w[i];
      }
      sx += px;
      sy += py;
      sw += pw;
      s = e;
    }
    if (head[c] != head[c + 1]) {
      cell_pos[2 * c] = sx / sw;
      cell_pos[2 * c + 1] = sy / sw;
    } else {
      cell_pos[2 * c] = T(0);
      cell_pos[2 * c + 1] = T(0);
    }
    cell_w[c] = sw;
  }
}

// source/nbfmm/core/solver/m2m.cu:22-74: levels 1..L-1, each parent from its
// four children (0,0),(1,0),(0,1),(1,1) in that order; mass 0 gives COM (0,0).
// Pinned by: check/core/solver/m2m.cu:32-48 restated in numpy, mass
// conservation and root COM == global COM.
template <class T>
void m2m(int L, T* cell_pos, T* cell_w) {
  if (L <= 1) return;
  const int B = 1 << (L + 1);
  for (int l = 1; l < L; ++l) {
    const int D = B >> l;
    for (int Y = 0; Y < D; ++Y) {
      for (int X = 0; X < D; ++X) {
        const int p = cube_index(B, l, X, Y);
        const int c00 = cube_index(B, l - 1, 2 * X, 2 * Y);
        const int c10 = cube_index(B, l - 1, 2 * X + 1, 2 * Y);
        const int c01 = cube_index(B, l - 1, 2 * X, 2 * Y + 1);
        const int c11 = cube_index(B, l - 1, 2 * X + 1, 2 * Y + 1);
        const T px = cell_pos[2 * c00] * cell_w[c00] + cell_pos[2 * c10] * cell_w[c10] +
                     cell_pos[2 * c01] * cell_w[c01] + cell_pos[2 * c11] * cell_w[c11];
        const T py = cell_pos[2 * c00 + 1] * cell_w[c00] + cell_pos[2 * c10 + 1] * cell_w[c10] +
                     cell_pos[2 * c01 + 1] * cell_w[c01] + cell_pos[2 * c11 + 1] * cell_w[c11];
        const T pw = cell_w[c00] + cell_w[c10] + cell_w[c01] + cell_w[c11];
        if (pw == T(0)) {
          cell_pos[2 * p] = T(0);
          cell_pos[2 * p + 1] = T(0);
        } else {
          cell_pos[2 * p] = px / pw;
          cell_pos[2 * p + 1] = py / pw;
        }
        cell_w[p] = pw;
      }
    }
  }
}

// source/nbfmm/core/solver/m2l.cu:24-81: at every level l = 0..L-1, each
// target cell sums the kernel, evaluated at its own COM, over the children of
// its parent's 3x3 neighbourhood (window [p-2, p+4) per axis, clipped) that
// are not adjacent to it (Chebyshev distance > 1); y outer, x inner.
// Pinned by: the O(cells^2) rule of check/core/solver/m2l.cu:34-50 restated
// in numpy, and partition exactness (each pair counted exactly once).
template <class T>
void m2l(int L, T eps, const T* cell_pos, const T* cell_w, T* cell_eff) {
  if (L <= 0) return;
  const int B = 1 << (L + 1);
  for (int l = 0; l < L; ++l) {
    const int D = B >> l;
#pragma omp parallel for schedule(static)
    for (int ty = 0; ty < D; ++ty) {
      for (int tx = 0; tx < D; ++tx) {
        const int px = tx & ~1, py = ty & ~1;
        const int t = cube_index(B, l, tx, ty);
        const T x0 = cell_pos[2 * t], y0 = cell_pos[2 * t + 1];
        T ax = T(0), ay = T(0);
        for (int y = py - 2; y < py + 4; ++y) {
          if (y < 0 || y >= D) continue;
          for (int x = px - 2; x < px + 4; ++x) {
            if (x < 0 || x >= D) continue;
            if (std::abs(x - tx) > 1 || std::abs(y - ty) > 1) {
              const int s = cube_index(B, l, x, y);
              kernel_accum<T>(x0, y0, cell_pos[2 * s], cell_pos[2 * s + 1], cell_w[s], eps, ax, ay);
            }
          }
        }
        cell_eff[2 * t] = ax;
        cell_eff[2 * t + 1] = ay;
      }
    }
  }
}

// source/nbfmm/core/solver/l2l.cu:21-56: for l = L-2 down to 0, every cell
// adds its parent's (already accumulated) effect.
// Pinned by: check/core/solver/l2l.cu:26-37 (push-down form) restated in numpy.
template <class T>
void l2l(int L, T* cell_eff) {
  if (L <= 1) return;
  const int B = 1 << (L + 1);
  for (int l = L - 2; l >= 0; --l) {
    const int D = B >> l;
    for (int Y = 0; Y < D; ++Y) {
      for (int X = 0; X < D; ++X) {
        const int c = cube_index(B, l, X, Y);
        const int p = cube_index(B, l + 1, X >> 1, Y >> 1);
        cell_eff[2 * c] = cell_eff[2 * c] + cell_eff[2 * p];
        cell_eff[2 * c + 1] = cell_eff[2 * c + 1] + cell_eff[2 * p + 1];
      }
    }
  }
}

// source/nbfmm/core/solver/l2p.cu:23-48: effect[i] += leaf cell effect.
// Pinned by: check/core/solver/l2p.cu:34-36 restated in numpy.
template <class T>
void l2p(int L, int n, const int* index, const T* cell_eff, T* eff) {
  if (L <= 0) return;
  const int B = 1 << (L + 1);
  for (int i = 0; i < n; ++i) {
    const int c = index[2 * i] + index[2 * i + 1] * B;
    eff[2 * i] = eff[2 * i] + cell_eff[2 * c];
    eff[2 * i + 1] = eff[2 * i + 1] + cell_eff[2 * c + 1];
  }
}

// source/nbfmm/core/solver/postdo.cu:22-45: eff0[perm[i]] = eff[i].
// Pinned by: check/core/solver/postdo.cu:35-37 (exact).
template <class T>
void postdo(int n, const int* perm, const T* eff, T* eff0) {
  for (int i = 0; i < n; ++i) {
    eff0[2 * perm[i]] = eff[2 * i];
    eff0[2 * perm[i] + 1] = eff[2 * i + 1];
  }
}

int cube_levels(int L) { return L < 1 ? 1 : L; }

// source/nbfmm/core/solver/solve.cu:11-27: predo, p2p, p2m, m2m, m2l, l2l,
// l2p, postdo.
// Pinned by: partition exactness, momentum conservation, the exact regime
// (<=1 particle per cell at every level == direct sum) and L=0 == direct sum.
template <class T>
int solve(int L, int n, const T limits[4], T eps, const T* pos0, const T* w0, T* eff0, const int* given = nullptr,
          int* record = nullptr) {
  const int B = 1 << (L + 1);
  const size_t C = (size_t)B * B;
  const size_t nn = (size_t)std::max(n, 1);
  std::vector<T> pos(2 * nn), w(nn), eff(2 * nn);
  std::vector<int> index(2 * nn), perm(nn), head(C + 1);
  const size_t cube = C * cube_levels(L);
  std::vector<T> cpos(2 * cube, T(0)), cw(cube, T(0)), ceff(2 * cube, T(0));
  const int clamped =
      predo<T>(L, n, pos0, w0, limits, pos.data(), w.data(), index.data(), perm.data(), head.data(), given, record);
  p2p<T>(L, n, pos.data(), w.data(), index.data(), head.data(), eps, eff.data());
  p2m<T>(L, n, pos.data(), w.data(), head.data(), cpos.data(), cw.data());
  m2m<T>(L, cpos.data(), cw.data());
  m2l<T>(L, eps, cpos.data(), cw.data(), ceff.data());
  l2l<T>(L, ceff.data());
  l2p<T>(L, n, index.data(), ceff.data(), eff.data());
  postdo<T>(n, perm.data(), eff.data(), eff0);
  return clamped;
}

// source/nbfmm/display/stars/update.cu:26-42 (updateDevice):
//   x_new = 2*x - x_prev + a*tick*tick;  x_prev = x
// evaluated left to right per component.  tick already includes sqrt(G)
// (source/nbfmm/display/stars.cu:31).
// Pinned by: the constant-acceleration closed form x_n = x_0 + n*v*dt + n^2*a*dt^2/2
// (fp64) and a bit-exact numpy float32 restatement.
template <class T>
void update(int n, T tick, T* xcur, T* xpre, const T* acc) {
  for (int i = 0; i < 2 * n; ++i) {
    const T x = xcur[i];
    xcur[i] = T(2) * x - xpre[i] + acc[i] * tick * tick;
    xpre[i] = x;
  }
}

// source/nbfmm/display/stars/prune.cu:28-62 (pruneDevice + Stars::prune):
// a star is eliminated when x < x_min || y < y_min || x > x_max || y > y_max
// (a NaN compares false and stays); thrust::sort_by_key on the bool flags
// moves (x_cur, x_pre, weight) so that the kept stars come first, both groups
// in their previous order, and the new count is the number kept.
// Pinned by: numpy argsort(flags, kind="stable") (the order a radix sort of
// bool keys gives) and hand-built cases on the box edges.
int prune(int n, const float limits[4], float* xcur, float* xpre, float* w) {
  std::vector<float> xc(xcur, xcur + 2 * n), xp(xpre, xpre + 2 * n), ww(w, w + n);
  std::vector<int> order;
  order.reserve(n);
  std::vector<char> gone(n);
  for (int i = 0; i < n; ++i) {
    const float x = xc[2 * i], y = xc[2 * i + 1];
    gone[i] = (x < limits[0] || y < limits[1] || x > limits[2] || y > limits[3]) ? 1 : 0;
    if (!gone[i]) order.push_back(i);
  }
  const int kept = (int)order.size();
  for (int i = 0; i < n; ++i)
    if (gone[i]) order.push_back(i);
  for (int j = 0; j < n; ++j) {
    const int i = order[j];
    xcur[2 * j] = xc[2 * i];
    xcur[2 * j + 1] = xc[2 * i + 1];
    xpre[2 * j] = xp[2 * i];
    xpre[2 * j + 1] = xp[2 * i + 1];
    w[j] = ww[i];
  }
  return kept;
}

// ---- initial-condition models (source/nbfmm/model/*.cu) ----
// Uniform draws: h = lowbias32(lowbias32(seed ^ 0x9e3779b9) ^ (4 i + k))
// (lowbias32: C. Wellons, hash-prospector) -> ((h >> 8) + 1) / 2^24 in
// (0, 1], the k-th draw of particle i (the reference uses curand_init(0, i, 0)
// streams).
// Pinned by: the closed-form orbit checks in tests/test_oracle.py (a star on
// its circle, one Verlet step under the central pull stays on it; the
// double-disk offsets are the circular-orbit step perpendicular to the pull).
uint32_t lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

float model_uniform(uint32_t seed, int i, int k) {
  const uint32_t h = lowbias32(lowbias32(seed ^ 0x9e3779b9u) ^ (uint32_t)(4 * i + k));
  return (float)((h >> 8) + 1u) * (1.0f / 16777216.0f);
}

// kernelFunction(origin, (r, 0), m).x with IEEE sqrt and division
float radial_effect(float r, float m, float eps) {
  const float dx = r;
  const float rr = std::sqrt(dx * dx);
  const float tmp = m / (rr * rr * rr + eps);
  return dx * tmp;
}

// disk.cu:49-50
float chord_angle(float a, float r, float tick) {
  const float q = a * tick * tick / r / 2.0f;
  return (float)std::acos(1.0 - (double)q);
}

void place(float cx, float cy, float r, float ang, float* p) {
  p[0] = cx + r * std::cos(ang);
  p[1] = cy + r * std::sin(ang);
}

constexpr double kTwoPi = 2.0 * 3.14159265358979323846;

void model_disk(int n, float cx, float cy, float R, float weight, float cw, int centered, int rotate, float tick,
                float eps, uint32_t seed, float* xcur, float* xpre, float* w) {
  for (int i = 0; i < n; ++i) {
    if (centered && i == 0) {
      xcur[0] = xpre[0] = cx;
      xcur[1] = xpre[1] = cy;
      w[0] = cw;
      continue;
    }
    const float r = ((float)(i + 1) / (float)n) * R;
    float dang = 0.0f;
    if (rotate) {
      const float m = centered ? weight * (float)(i - 1) + cw : weight * (float)i;
      dang = chord_angle(radial_effect(r, m, eps), r, tick);
    }
    const float ang = (float)(kTwoPi * (double)model_uniform(seed, i, 0));
    place(cx, cy, r, ang, xcur + 2 * i);
    place(cx, cy, r, ang - dang, xpre + 2 * i);
    w[i] = weight;
  }
}

}  // namespace

extern "C" {

int orc_version(void) { return 1; }

float orc_model_uniform(unsigned seed, int i, int k) { return model_uniform(seed, i, k); }

// model/disk.cu (centered 0, rotate 1), disk_center.cu (1, 1), disk_static.cu (0, 0)
void orc_model_disk(int n, float cx, float cy, float R, float weight, float cw, int centered, int rotate, float tick,
                    float eps, unsigned seed, float* xcur, float* xpre, float* w) {
  model_disk(n, cx, cy, R, weight, cw, centered, rotate, tick, eps, seed, xcur, xpre, w);
}

// model/circle_center.cu (uniform 0), circle_uniform_center.cu (uniform 1)
void orc_model_circle(int n, float cx, float cy, float R, float weight, float cw, float tick, float eps, int uniform,
                      unsigned seed, float* xcur, float* xpre, float* w) {
  const float dang = chord_angle(radial_effect(R, cw, eps), R, tick);
  for (int i = 0; i < n; ++i) {
    if (i == 0) {
      xcur[0] = xpre[0] = cx;
      xcur[1] = xpre[1] = cy;
      w[0] = cw;
      continue;
    }
    const float ang = uniform ? (float)((kTwoPi * (double)i) / (double)n) : (float)(kTwoPi * (double)model_uniform(seed, i, 0));
    place(cx, cy, R, ang, xcur + 2 * i);
    place(cx, cy, R, ang - dang, xpre + 2 * i);
    w[i] = weight;
  }
}

// model/rectangle.cu
void orc_model_rectangle(int n, float cx, float cy, float width, float height, float max_weight, unsigned seed,
                         float* xcur, float* xpre, float* w) {
  for (int i = 0; i < n; ++i) {
    xcur[2 * i] = xpre[2 * i] = cx + (model_uniform(seed, i, 0) - 0.5f) * width;
    xcur[2 * i + 1] = xpre[2 * i + 1] = cy + (model_uniform(seed, i, 1) - 0.5f) * height;
    w[i] = max_weight * model_uniform(seed, i, 2);
  }
}

// model/double_disk.cu:55-86 and double_disk_center.cu: the offsets added to
// the two disks' previous positions (o[0..1] disk 1, o[2..3] disk 2).
void orc_double_disk_offsets(float c1x, float c1y, float c2x, float c2y, float m1, float m2, float q1, float q2,
                             float ecc, float tick, float eps, float* o) {
  const float dx = c2x - c1x, dy = c2y - c1y;
  const float r = std::sqrt(dx * dx + dy * dy);
  const float t1 = m2 / (r * r * r + eps), t2 = m1 / (r * r * r + eps);
  const float e1x = dx * t1, e1y = dy * t1, e2x = -dx * t2, e2y = -dy * t2;
  const float a1 = std::sqrt(e1x * e1x + e1y * e1y), a2 = std::sqrt(e2x * e2x + e2y * e2y);
  const float r1 = r * q2 / (q1 + q2), r2 = r * q1 / (q1 + q2);
  const float s1 = std::sqrt(r1 / a1) * tick / std::exp2(ecc);
  const float s2 = std::sqrt(r2 / a2) * tick / std::exp2(ecc);
  o[0] = -e1y * s1 - e1x * tick * tick * ecc;
  o[1] = e1x * s1 - e1y * tick * tick * ecc;
  o[2] = -e2y * s2 - e2x * tick * tick * ecc;
  o[3] = e2x * s2 - e2y * tick * tick * ecc;
}

int orc_prune_f32(int n, const float limits[4], float* xcur, float* xpre, float* w) {
  return prune(n, limits, xcur, xpre, w);
}

// ---- video frames (source/nbfmm/display/stars/display.cu) ----
//
// Stars::display (display.cu:219-226): luma cleared to 0, chroma (the last
// width*height/2 bytes) to 128, then each star lights its stamp.  The stamp
// grows by one ring of offsets per size level (display.cu:54-211); each ring
// is listed by its generators (a, b), a >= b >= 0, and takes every sign and
// order of them: level 1 (1,0); 2 (1,1); 3 (2,0); 4 (2,1); 5 (3,0) (2,2);
// 6 (3,1); 7 (4,0) (3,2).  Intended semantics: every stamp pixel is clipped
// to the frame (the reference's y <= height at :53 and unguarded upward arms
// at :62, :90, :130, :182 write outside the luma plane).
// Pinned by hand-drawn stamps of every level, clipping at the edges and
// corners, and the pixel rule at cell boundaries (tests/test_oracle.py).
void orc_display(int n, int width, int height, float size_scale, const float lim[4], const float* pos,
                 const float* weight, uint8_t* frame) {
  const size_t luma = (size_t)width * height;
  std::fill(frame, frame + luma, (uint8_t)0);
  std::fill(frame + luma, frame + luma + luma / 2, (uint8_t)128);
  static const int gen[][3] = {{1, 1, 0}, {2, 1, 1}, {3, 2, 0}, {4, 2, 1}, {5, 3, 0},
                               {5, 2, 2}, {6, 3, 1}, {7, 4, 0}, {7, 3, 2}};  // {level, a, b}
  const float unit_w = (lim[2] - lim[0]) / (float)width;    // display.cu:45
  const float unit_h = (lim[3] - lim[1]) / (float)height;   // display.cu:46
  for (int i = 0; i < n; ++i) {
    const float fx = std::floor((pos[2 * i] - lim[0]) / unit_w);       // display.cu:47
    const float fy = std::floor((lim[3] - pos[2 * i + 1]) / unit_h);   // display.cu:48
    if (!(fx >= 0.0f && fx < (float)width && fy >= 0.0f && fy < (float)height)) continue;
    const int x = (int)fx, y = (int)fy;
    const float fs = std::floor(std::sqrt(weight[i]) / size_scale);    // display.cu:49
    const int level = fs >= 1.0f ? (int)std::min(fs, 7.0f) : 0;
    auto light = [&](int dx, int dy) {
      const int px = x + dx, py = y + dy;
      if (px >= 0 && px < width && py >= 0 && py < height) frame[(size_t)py * width + px] = 255;
    };
    light(0, 0);
    for (const auto& g : gen) {
      if (g[0] > level) continue;
      for (int sa : {-1, 1})
        for (int sb : {-1, 1}) {
          light(sa * g[1], sb * g[2]);
          light(sb * g[2], sa * g[1]);
        }
    }
  }
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int nthreads) { set_threads(nthreads); }

void orc_cell_size_f32(const float limits[4], int B, float out[2]) {
  out[0] = (limits[2] - limits[0]) / (float)B;
  out[1] = (limits[3] - limits[1]) / (float)B;
}

int orc_compute_index_f32(int L, int n, const float* pos, const float limits[4], int* index) {
  const int B = 1 << (L + 1);
  const float cw = (limits[2] - limits[0]) / (float)B;
  const float ch = (limits[3] - limits[1]) / (float)B;
  return compute_index<float>(n, B, pos, limits[0], limits[1], cw, ch, index);
}

int orc_predo_f32(int L, int n, const float* pos0, const float* w0, const float limits[4], float* pos, float* w,
                  int* index, int* perm, int* head) {
  return predo<float>(L, n, pos0, w0, limits, pos, w, index, perm, head);
}

void orc_p2p_f32(int L, int n, const float* pos, const float* w, const int* index, const int* head, float eps,
                 float* eff) {
  p2p<float>(L, n, pos, w, index, head, eps, eff);
}

void orc_p2m_f32(int L, int n, const float* pos, const float* w, const int* head, float* cell_pos, float* cell_w) {
  p2m<float>(L, n, pos, w, head, cell_pos, cell_w);
}

void orc_m2m_f32(int L, float* cell_pos, float* cell_w) { m2m<float>(L, cell_pos, cell_w); }

void orc_m2l_f32(int L, float eps, const float* cell_pos, const float* cell_w, float* cell_eff) {
  m2l<float>(L, eps, cell_pos, cell_w, cell_eff);
}

void orc_l2l_f32(int L, float* cell_eff) { l2l<float>(L, cell_eff); }

void orc_l2p_f32(int L, int n, const int* index, const float* cell_eff, float* eff) {
  l2p<float>(L, n, index, cell_eff, eff);
}

void orc_postdo_f32(int n, const int* perm, const float* eff, float* eff0) { postdo<float>(n, perm, eff, eff0); }

int orc_solve_f32(int L, int n, const float limits[4], float eps, const float* pos0, const float* w0, float* eff0) {
  return solve<float>(L, n, limits, eps, pos0, w0, eff0);
}

int orc_solve_f64(int L, int n, const double limits[4], double eps, const double* pos0, const double* w0,
                  double* eff0) {
  return solve<double>(L, n, limits, eps, pos0, w0, eff0);
}

void orc_update_f32(int n, float tick, float* xcur, float* xpre, const float* acc) {
  update<float>(n, tick, xcur, xpre, acc);
}

void orc_update_f64(int n, double tick, double* xcur, double* xpre, const double* acc) {
  update<double>(n, tick, xcur, xpre, acc);
}

// Stars::update without prune (source/nbfmm/display/stars/update.cu:47-55):
// solve at the current positions, then the Verlet update; `steps` times.
int orc_run_f32(int L, int n, const float limits[4], float eps, float tick, int steps, float* xcur, float* xpre,
                const float* w, float* acc) {
  int clamped = 0;
  for (int s = 0; s < steps; ++s) {
    clamped += solve<float>(L, n, limits, eps, xcur, w, acc);
    update<float>(n, tick, xcur, xpre, acc);
  }
  return clamped;
}

int orc_run_f64(int L, int n, const double limits[4], double eps, double tick, int steps, double* xcur,
                double* xpre, const double* w, double* acc) {
  int clamped = 0;
  for (int s = 0; s < steps; ++s) {
    clamped += solve<double>(L, n, limits, eps, xcur, w, acc);
    update<double>(n, tick, xcur, xpre, acc);
  }
  return clamped;
}

// orc_run_f64 with the cells of every step recorded (record = 1: cells gets
// steps x n index pairs) or replayed from cells (record = 0), for finite
// differences of the trajectory with the cells held fixed, the function a
// cells-fixed adjoint differentiates.  No reference counterpart.
// Pinned by: recording gives orc_run_f64's bits, replaying the recorded cells
// gives them again, and central differences of a replayed loss at h = 1e-3
// and h = 1e-6 agree (the frozen-cell loss is smooth).
int orc_run_f64_cells(int L, int n, const double limits[4], double eps, double tick, int steps, double* xcur,
                      double* xpre, const double* w, double* acc, int* cells, int record) {
  int clamped = 0;
  for (int s = 0; s < steps; ++s) {
    int* c = cells + (size_t)s * 2 * (size_t)n;
    clamped += solve<double>(L, n, limits, eps, xcur, w, acc, record ? nullptr : c, record ? c : nullptr);
    update<double>(n, tick, xcur, xpre, acc);
  }
  return clamped;
}

// fp64 O(N^2) direct sum with the same kernel form and epsilon
// (main/demo.cu:59-66 is the reference's only end-to-end comparison).
// Evaluates the targets listed in `targets` (all particles when targets is
// null), writing eff[2*t] for the t-th listed target.
void orc_direct_f64(int n, const float* pos, const float* w, double eps, int ntargets, const int* targets,
                    double* eff) {
  const int nt = targets ? ntargets : n;
#pragma omp parallel for schedule(dynamic, 16)
  for (int t = 0; t < nt; ++t) {
    const int i = targets ? targets[t] : t;
    const double tx = pos[2 * i], ty = pos[2 * i + 1];
    double ax = 0.0, ay = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      kernel_accum<double>(tx, ty, (double)pos[2 * j], (double)pos[2 * j + 1], (double)w[j], eps, ax, ay);
    }
    eff[2 * t] = ax;
    eff[2 * t + 1] = ay;
  }
}

}  // extern "C"
