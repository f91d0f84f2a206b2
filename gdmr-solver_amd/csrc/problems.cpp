// Synthetic operators and random vectors shared by the control program, tests and benchmarks.
//
// The reference builds its only test operator inside progDMR.c:75-104 (a dense 1-D Laplacian
// seeded by time()); the configurations this package is measured on are PDE stencils, so the
// generators below assemble them directly as CSR.  Random numbers come from SplitMix64
// (Steele, Lea, Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014), used as a
// counter-based generator: element i of stream (seed, stream) is mix(key + (i+1)*gamma), so any
// slice can be generated independently and every consumer sees identical values.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "gDMRC.h"

namespace {

constexpr uint64_t kGamma = 0x9e3779b97f4a7c15ULL;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return mix64(mix64(seed + kGamma) ^ (stream + 0x632be59bd9b4e019ULL));
}

// Uniform in [0, 1) with 53 random bits.
inline double u01(uint64_t bits) { return (double)(bits >> 11) * 0x1.0p-53; }

void uniform_fill(uint64_t seed, uint64_t stream, long long n, double lo, double hi, double *out) {
  const uint64_t key = stream_key(seed, stream);
  const double w = hi - lo;
  for (long long i = 0; i < n; ++i) out[i] = lo + w * u01(mix64(key + (uint64_t)(i + 1) * kGamma));
}

// Per-cell multiplicative jitter s_c = 1 + delta_c, delta ~ U(-jitter, jitter); exactly 1 when
// jitter == 0 so unjittered operators have constant coefficients.
std::vector<double> cell_scale(long long n, double jitter, uint64_t seed, uint64_t stream) {
  std::vector<double> s((size_t)n, 1.0);
  if (jitter != 0.0) {
    uniform_fill(seed, stream, n, -jitter, jitter, s.data());
    for (auto &v : s) v = 1.0 + v;
  }
  return s;
}

// 2-D 5-point: -div(a grad u) + b.grad u, h^2-scaled.  a[d] is the base diffusion per direction.
long long gen2d(int N, const double a[2], const double b[2], double jitter, uint64_t seed,
                uint64_t stream, int *rp, int *ci, double *va) {
  const long long n = (long long)N * N;
  const long long nnz = 5 * n - 4LL * N;
  if (!rp) return nnz;
  const double h = 1.0 / (N + 1);
  const std::vector<double> s = cell_scale(n, jitter, seed, stream);
  const long long stride[2] = {1, N};
  long long e = 0;
  rp[0] = 0;
  for (int j = 0; j < N; ++j) {
    for (int i = 0; i < N; ++i) {
      const long long r = i + (long long)N * j;
      const int idx[2] = {i, j};
      double cm[2], cp[2];  // face diffusion on the minus / plus side per direction
      bool hm[2], hp[2];
      for (int d = 0; d < 2; ++d) {
        hm[d] = idx[d] > 0;
        hp[d] = idx[d] < N - 1;
        cm[d] = a[d] * (hm[d] ? 0.5 * (s[r] + s[r - stride[d]]) : s[r]);
        cp[d] = a[d] * (hp[d] ? 0.5 * (s[r] + s[r + stride[d]]) : s[r]);
      }
      const double diag = ((cm[0] + cp[0]) + cm[1]) + cp[1];
      // Columns in ascending order: r-N, r-1, r, r+1, r+N.
      if (hm[1]) { ci[e] = (int)(r - N); va[e++] = -cm[1] - b[1] * h * 0.5; }
      if (hm[0]) { ci[e] = (int)(r - 1); va[e++] = -cm[0] - b[0] * h * 0.5; }
      ci[e] = (int)r; va[e++] = diag;
      if (hp[0]) { ci[e] = (int)(r + 1); va[e++] = -cp[0] + b[0] * h * 0.5; }
      if (hp[1]) { ci[e] = (int)(r + N); va[e++] = -cp[1] + b[1] * h * 0.5; }
      rp[r + 1] = (int)e;
    }
  }
  return e;
}

long long gen3d(int N, const double a[3], const double b[3], double jitter, uint64_t seed,
                uint64_t stream, int *rp, int *ci, double *va) {
  const long long n = (long long)N * N * N;
  const long long nnz = 7 * n - 6LL * N * N;
  if (!rp) return nnz;
  const double h = 1.0 / (N + 1);
  const std::vector<double> s = cell_scale(n, jitter, seed, stream);
  const long long stride[3] = {1, N, (long long)N * N};
  long long e = 0;
  rp[0] = 0;
  for (int l = 0; l < N; ++l) {
    for (int j = 0; j < N; ++j) {
      for (int i = 0; i < N; ++i) {
        const long long r = i + (long long)N * j + (long long)N * N * l;
        const int idx[3] = {i, j, l};
        double cm[3], cp[3];
        bool hm[3], hp[3];
        for (int d = 0; d < 3; ++d) {
          hm[d] = idx[d] > 0;
          hp[d] = idx[d] < N - 1;
          cm[d] = a[d] * (hm[d] ? 0.5 * (s[r] + s[r - stride[d]]) : s[r]);
          cp[d] = a[d] * (hp[d] ? 0.5 * (s[r] + s[r + stride[d]]) : s[r]);
        }
        const double diag = ((((cm[0] + cp[0]) + cm[1]) + cp[1]) + cm[2]) + cp[2];
        if (hm[2]) { ci[e] = (int)(r - stride[2]); va[e++] = -cm[2] - b[2] * h * 0.5; }
        if (hm[1]) { ci[e] = (int)(r - N); va[e++] = -cm[1] - b[1] * h * 0.5; }
        if (hm[0]) { ci[e] = (int)(r - 1); va[e++] = -cm[0] - b[0] * h * 0.5; }
        ci[e] = (int)r; va[e++] = diag;
        if (hp[0]) { ci[e] = (int)(r + 1); va[e++] = -cp[0] + b[0] * h * 0.5; }
        if (hp[1]) { ci[e] = (int)(r + N); va[e++] = -cp[1] + b[1] * h * 0.5; }
        if (hp[2]) { ci[e] = (int)(r + stride[2]); va[e++] = -cp[2] + b[2] * h * 0.5; }
        rp[r + 1] = (int)e;
      }
    }
  }
  return e;
}

}  // namespace

extern "C" {

long long dmrGenConvDiff2D(int N, double nu0, double bx, double by, double jitter,
                           unsigned long long seed, unsigned long long stream, int *rowPtr,
                           int *colInd, double *val) {
  if (N < 1) return -1;
  const double a[2] = {nu0, nu0}, b[2] = {bx, by};
  return gen2d(N, a, b, jitter, seed, stream, rowPtr, colInd, val);
}

long long dmrGenConvDiff3D(int N, double nu0, double bx, double by, double bz, double jitter,
                           unsigned long long seed, unsigned long long stream, int *rowPtr,
                           int *colInd, double *val) {
  if (N < 1) return -1;
  const double a[3] = {nu0, nu0, nu0}, b[3] = {bx, by, bz};
  return gen3d(N, a, b, jitter, seed, stream, rowPtr, colInd, val);
}

long long dmrGenAniso3D(int N, double ax, double ay, double az, double jitter,
                        unsigned long long seed, unsigned long long stream, int *rowPtr,
                        int *colInd, double *val) {
  if (N < 1) return -1;
  const double a[3] = {ax, ay, az}, b[3] = {0.0, 0.0, 0.0};
  return gen3d(N, a, b, jitter, seed, stream, rowPtr, colInd, val);
}

void dmrRandUniform(unsigned long long seed, unsigned long long stream, long long n, double lo,
                    double hi, double *out) {
  uniform_fill(seed, stream, n, lo, hi, out);
}

void dmrRandNormal(unsigned long long seed, unsigned long long stream, long long n, double *out) {
  // Box-Muller on consecutive pairs of uniforms; u1 in (0, 1] keeps the logarithm finite.
  const uint64_t key = stream_key(seed, stream);
  const double two_pi = 6.283185307179586476925286766559;
  for (long long i = 0; i < n; i += 2) {
    const double u1 = 1.0 - u01(mix64(key + (uint64_t)(i + 1) * kGamma));
    const double u2 = u01(mix64(key + (uint64_t)(i + 2) * kGamma));
    const double rad = std::sqrt(-2.0 * std::log(u1));
    out[i] = rad * std::cos(two_pi * u2);
    if (i + 1 < n) out[i + 1] = rad * std::sin(two_pi * u2);
  }
}

void dmrSplitMix64(unsigned long long seed, long long n, unsigned long long *out) {
  uint64_t x = seed;
  for (long long i = 0; i < n; ++i) {
    x += kGamma;
    out[i] = mix64(x);
  }
}

}  // extern "C"
