// FP64 host oracle of the deflated minimal residual (DMR) solve -- TEST INFRASTRUCTURE.
//
// This file is the checker, not the product: only tests, the smoke check and the CPU-baseline
// leg of the benchmark load it.  It restates, in real FP64 arithmetic on the host, the algorithm
// of seanMolesky/gDMR-solver (gDMR.cu) that the HIP solver implements, function by function:
//
//   oracle_solve / run_cycles  solverDMR        gDMR.cu:2638-2901 (setup, guards, restart loop)
//   start-vector choice        arnoldiSetupDMR  gDMR.cu:2586-2632
//   arnoldi()                  arnoldiDMR       gDMR.cu:1790-2332 (A Z QR, v0 deflation, CGS)
//                              deflatorDMR      gDMR.cu:1710-1787
//   update()                   updateTrgCoeffs  gDMR.cu:2334-2408 (t = P^T r, full projection)
//                              updateAppResDMR  gDMR.cu:2411-2582 (LSQ, r and x updates, flag)
//   lsq()                      qrHessArrDMR / qrSolveDMR / hessenbergMultDMR gDMR.cu:591-839
//   svd_smallest()             svDecompHessDMR  gDMR.cu:634-689 + genDeflatDMR gDMR.cu:1583-1704
//   spmv()                     the user operator contract, testLS.cu:97-128 (here CSR)
//
// Operation order.  Dot products are summed in the fixed order of the device reductions
// (256-thread blocks over 2-element units, xor butterflies per 64-lane wavefront, block partials
// in block order) with the block count G passed in by the caller, and every fused multiply-add
// of the device code is an explicit std::fma here, with contraction disabled elsewhere.  With the
// same inputs the oracle therefore reproduces the device results bit for bit where the device
// follows this order; the tests bound the difference either way.
//
// Mathematical pins live in tests/: CSR assembly against an independent dense assembly,
// LSQ against LAPACK (numpy.linalg.lstsq), SVD against numpy.linalg.svd, k = 0 against an
// independent textbook GMRES(m), Arnoldi relation and orthogonality, breakdown on a diagonal
// operator with few distinct eigenvalues, and FP64-checked true residuals.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int kB = 256;  // device block size
constexpr double kSvdTol = 1e-14;
constexpr int kSvdMaxSweeps = 128;
constexpr double kStagnation = 0.995;

// ---- device reduction order ----
// In-place xor butterfly over one 64-lane wavefront; every lane ends with the same sum.
double wave_sum(double *v) {
  double t[64];
  for (int off = 32; off >= 1; off >>= 1) {
    for (int l = 0; l < 64; ++l) t[l] = v[l] + v[l ^ off];
    std::memcpy(v, t, sizeof(t));
  }
  return v[0];
}
// block_sum of the device: four wavefront butterflies, then ((w0 + w1) + w2) + w3.
double block_sum(double *v /*256*/) {
  const double w0 = wave_sum(v), w1 = wave_sum(v + 64), w2 = wave_sum(v + 128),
               w3 = wave_sum(v + 192);
  return ((w0 + w1) + w2) + w3;
}
// Sum of G block partials as the device's last block does it: one wavefront, lane l adds
// partials l, l+64, ... in order, then the xor butterfly.
double sum_partials(const double *part, int G) {
  double v[64];
  for (int l = 0; l < 64; ++l) {
    double s = 0.0;
    for (int q = l; q < G; q += 64) s = s + part[q];
    v[l] = s;
  }
  return wave_sum(v);
}

// Generic fixed-order reduction: f(i, acc) folds element i into a thread accumulator, elements
// visited per thread in unit order (unit u = elements 2u, 2u+1), padded elements included.
template <typename F>
double reduce(long long n, int G, F f) {
  const long long nv = (n + 1) / 2;
  std::vector<double> part(G);
#pragma omp parallel for schedule(static)
  for (int b = 0; b < G; ++b) {
    double acc[kB];
    for (int t = 0; t < kB; ++t) acc[t] = 0.0;
    for (long long base = (long long)b * kB; base < nv; base += (long long)G * kB) {
      for (int t = 0; t < kB; ++t) {
        const long long u = base + t;
        if (u >= nv) break;
        acc[t] = f(2 * u, acc[t]);
        acc[t] = f(2 * u + 1, acc[t]);
      }
    }
    part[b] = block_sum(acc);
  }
  return sum_partials(part.data(), G);
}

struct Csr {
  int n = 0;
  std::vector<int> rp, ci;
  std::vector<double> va;
};

// y = A x: per row, products rounded, then added in column order (device LDS path).  Rows with
// more than 2048 nonzeros follow the device's block-strided path.
void spmv(const Csr &A, const double *x, double *y) {
  const int n = A.n;
#pragma omp parallel for schedule(static)
  for (int r = 0; r < n; ++r) {
    const int e0 = A.rp[r], e1 = A.rp[r + 1];
    if (e1 - e0 > 2048) {
      double v[kB];
      for (int t = 0; t < kB; ++t) {
        double s = 0.0;
        for (int e = e0 + t; e < e1; e += kB) {
          const double p = A.va[e] * x[A.ci[e]];
          s = s + p;
        }
        v[t] = s;
      }
      y[r] = block_sum(v);
    } else {
      double s = 0.0;
      for (int e = e0; e < e1; ++e) {
        const double p = A.va[e] * x[A.ci[e]];
        s = s + p;
      }
      y[r] = s;
    }
  }
}

// Vectors are stored padded to ld (multiple of 256) with zero padding, like the device.
struct Vecs {
  long long n = 0, ld = 0;
  int m = 0, k = 0;
  std::vector<double> P, Z, Zn, r, xa, g, u;
  double *col(std::vector<double> &v, long long j) { return v.data() + j * ld; }
};

inline double at(const double *v, long long i, long long n) { return i < n ? v[i] : 0.0; }

// out[l] = B_l . x (multidot).
void multidot(const double *B, long long ld, int p, const double *x, long long n, int G,
              double *out) {
  for (int l = 0; l < p; ++l) {
    const double *b = B + (long long)l * ld;
    out[l] = reduce(n, G, [&](long long i, double acc) { return std::fma(at(b, i, n), at(x, i, n), acc); });
  }
}

// out = in - sum c1 B1 - sum c2 B2 (COMB_SUB); out may alias in.  Returns (ss, sg).
void comb_sub(const double *in, double *out, const double *B1, int p1, const double *c1,
              const double *B2, int p2, const double *c2, long long ld, long long n) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < n; ++i) {
    double o = in[i];
    for (int l = 0; l < p1; ++l) o = std::fma(-c1[l], B1[(long long)l * ld + i], o);
    for (int l = 0; l < p2; ++l) o = std::fma(-c2[l], B2[(long long)l * ld + i], o);
    out[i] = o;
  }
}

// out = in + beta * (sum c1 B1 + sum c2 B2) (COMB_ACC).
void comb_acc(const double *in, double *out, const double *B1, int p1, const double *c1,
              const double *B2, int p2, const double *c2, long long ld, long long n, double beta) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < n; ++i) {
    double a = 0.0;
    for (int l = 0; l < p1; ++l) a = std::fma(c1[l], B1[(long long)l * ld + i], a);
    for (int l = 0; l < p2; ++l) a = std::fma(c2[l], B2[(long long)l * ld + i], a);
    const double t = beta * a;
    out[i] = in ? in[i] + t : t;
  }
}

double sumsq(const double *v, long long n, int G) {
  return reduce(n, G, [&](long long i, double acc) {
    const double x = at(v, i, n);
    return std::fma(x, x, acc);
  });
}
double dotp(const double *v, const double *g, long long n, int G) {
  return reduce(n, G, [&](long long i, double acc) { return std::fma(at(v, i, n), at(g, i, n), acc); });
}

void scale(double *v, long long n, double s) {
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < n; ++i) v[i] = v[i] * s;
}

}  // namespace

// =============================================================================================
// Hessenberg LSQ (Householder QR of [H | t], LAPACK dlarfg conventions) -- same operation order
// as the device kernel.  H is (m+1) x m column-major; y has m+1 entries (y[m] = 0).
// =============================================================================================
extern "C" void oracle_lsq(const double *H, int m, const double *t, double *y, double *z) {
  const int ldh = m + 1;
  std::vector<double> A(H, H + (size_t)ldh * m), b(t, t + ldh), yv(m);
  for (int j = 0; j < m; ++j) {
    double *vj = A.data() + (size_t)j * ldh;
    const double alpha = vj[j];
    double xn2 = 0.0;
    for (int i = j + 1; i <= m; ++i) xn2 = std::fma(vj[i], vj[i], xn2);
    double tau = 0.0;
    if (xn2 != 0.0) {
      const double beta = -std::copysign(std::sqrt(std::fma(alpha, alpha, xn2)), alpha);
      tau = (beta - alpha) / beta;
      const double sc = 1.0 / (alpha - beta);
      for (int i = j + 1; i <= m; ++i) vj[i] = vj[i] * sc;
      vj[j] = beta;
    }
    if (tau != 0.0) {
      for (int c = j + 1; c <= m; ++c) {
        double *col = (c < m) ? A.data() + (size_t)c * ldh : b.data();
        double w = col[j];
        for (int i = j + 1; i <= m; ++i) w = std::fma(vj[i], col[i], w);
        const double tw = tau * w;
        col[j] = col[j] - tw;
        for (int i = j + 1; i <= m; ++i) col[i] = std::fma(-tw, vj[i], col[i]);
      }
    }
  }
  for (int i = m - 1; i >= 0; --i) {
    yv[i] = b[i] / A[i + (size_t)i * ldh];
    for (int r = 0; r < i; ++r) b[r] = std::fma(-A[r + (size_t)i * ldh], yv[i], b[r]);
  }
  for (int i = 0; i < m; ++i) y[i] = yv[i];
  y[m] = 0.0;
  for (int r = 0; r <= m; ++r) {
    double s = 0.0;
    for (int c = 0; c < m; ++c) s = std::fma(H[r + (size_t)c * ldh], yv[c], s);
    z[r] = s;
  }
}

// =============================================================================================
// One-sided Jacobi SVD with round-robin ordering -- same operation order as the device kernel
// (one wavefront per pair: lane-strided partial sums, xor butterfly).  Returns the number of
// sweeps.  C (m x k): right singular vectors of the k smallest singular values, descending, each
// with its largest-magnitude entry positive.  sig (m): singular values descending.
// =============================================================================================
// Dot product in the order of a 32-lane half wavefront: lane l adds rows l, l+32, ... with fma,
// then an xor butterfly over the 32 lanes (offsets 16, 8, 4, 2, 1).
static double half_dot(const double *a, const double *b, int len) {
  double v[32], t[32];
  for (int l = 0; l < 32; ++l) {
    double s = 0.0;
    for (int i = l; i < len; i += 32) s = std::fma(a[i], b[i], s);
    v[l] = s;
  }
  for (int off = 16; off >= 1; off >>= 1) {
    for (int l = 0; l < 32; ++l) t[l] = v[l] + v[l ^ off];
    std::memcpy(v, t, sizeof(v));
  }
  return v[0];
}

extern "C" int oracle_svd(const double *H, int m, int k, double tol, int maxSweeps, double *C,
                          double *sig) {
  const int mm = m + (m & 1), ldh = m + 1;
  std::vector<double> A((size_t)ldh * mm, 0.0), V((size_t)m * mm, 0.0), sg(mm);
  for (int c = 0; c < m; ++c)
    for (int r = 0; r < ldh; ++r) A[r + (size_t)c * ldh] = H[r + (size_t)c * ldh];
  for (int c = 0; c < m; ++c) V[c + (size_t)c * m] = 1.0;
  int sweep = 0;
  while (sweep < maxSweeps) {
    // Exact squared column norms per sweep, updated by each rotation within the sweep.
    for (int c = 0; c < m; ++c) {
      const double *ac = A.data() + (size_t)c * ldh;
      sg[c] = half_dot(ac, ac, ldh);
    }
    int rot = 0;
    for (int r = 0; r < mm - 1; ++r) {
      for (int pi = 0; pi < mm / 2; ++pi) {
        int p, q;
        if (pi == 0) {
          p = r;
          q = mm - 1;
        } else {
          p = (r + pi) % (mm - 1);
          q = (r - pi + (mm - 1)) % (mm - 1);
        }
        if (p > q) std::swap(p, q);
        if (q >= m) continue;
        double *ap = A.data() + (size_t)p * ldh, *aq = A.data() + (size_t)q * ldh;
        const double ga = half_dot(ap, aq, ldh);
        const double al = sg[p], be = sg[q];
        if (ga != 0.0 && std::fabs(ga) > tol * std::sqrt(al * be)) {
          const double zeta = (be - al) / (2.0 * ga);
          const double tt =
              std::copysign(1.0, zeta) / (std::fabs(zeta) + std::sqrt(std::fma(zeta, zeta, 1.0)));
          const double cs = 1.0 / std::sqrt(std::fma(tt, tt, 1.0));
          const double sn = cs * tt;
          for (int i = 0; i < ldh; ++i) {
            const double x = ap[i], y = aq[i];
            ap[i] = cs * x - sn * y;
            aq[i] = sn * x + cs * y;
          }
          double *vp = V.data() + (size_t)p * m, *vq = V.data() + (size_t)q * m;
          for (int i = 0; i < m; ++i) {
            const double x = vp[i], y = vq[i];
            vp[i] = cs * x - sn * y;
            vq[i] = sn * x + cs * y;
          }
          const double tg = tt * ga;
          sg[p] = al - tg;
          sg[q] = be + tg;
          ++rot;
        }
      }
    }
    ++sweep;
    if (rot == 0) break;
  }
  for (int c = 0; c < m; ++c) {
    const double *ac = A.data() + (size_t)c * ldh;
    sg[c] = std::sqrt(half_dot(ac, ac, ldh));
  }
  for (int c = 0; c < m; ++c) {
    const double sc = sg[c];
    int rank = 0;
    for (int j = 0; j < m; ++j) rank += (sg[j] > sc || (sg[j] == sc && j < c)) ? 1 : 0;
    if (sig) sig[rank] = sc;
    if (rank >= m - k) {
      const int i = rank - (m - k);
      const double *vc = V.data() + (size_t)c * m;
      int imax = 0;
      double vmax = std::fabs(vc[0]);
      for (int r = 1; r < m; ++r) {
        const double a = std::fabs(vc[r]);
        if (a > vmax) {
          vmax = a;
          imax = r;
        }
      }
      const bool neg = vc[imax] < 0.0;
      for (int r = 0; r < m; ++r) C[r + (size_t)m * i] = neg ? -vc[r] : vc[r];
    }
  }
  return sweep;
}

// =============================================================================================
// Solver context
// =============================================================================================
struct OracleCtx {
  Csr A;
  Vecs v;
  int G = 1;
  bool hasZ = false;
  int kryLoopFlag = 0;
  int kryFlagCount = 0;
  std::vector<double> H, t, y, z, C, sig, cv;
  // statistics of the last solve
  std::vector<double> hist;
  int cycles = 0, matvecs = 0, breakdowns = 0, status = 0;
  double beta = 0.0, rho = 0.0;
  // optional copy of the last cycle (tests of the Arnoldi relation)
  bool keepLast = false;
  std::vector<double> lastP, lastH, lastZ, lastW;
  int lastD = 0;
  // With keepLast: the start choice of every cycle of the last solve (1 = from the residual,
  // 0 = from the old Krylov vector (m-k)/2), and, for the most recent cycle that restarted from
  // the old vector, the basis before its setup (stagPrevP) and after its Arnoldi process (stagP).
  std::vector<int> startLog;
  std::vector<double> stagPrevP, stagP;
  int stagD = -1;
  // The stagnation restart rule of arnoldiSetupDMR; switched off only by tests that show the
  // rule changes the residual history.
  bool stagRule = true;
};

namespace {

// One restart cycle's Arnoldi process (arnoldiSetupDMR + arnoldiDMR).  Returns the breakdown
// step (0 if none); *fromResidualOut: whether the start vector was the residual.
int arnoldi(OracleCtx &c, int d, int matRank, double tol, bool *fromResidualOut) {
  Vecs &v = c.v;
  const long long n = v.n, ld = v.ld;
  const int m = v.m, k = v.k, G = c.G;
  const int o = d > 0 ? k : 0;
  // arnoldiSetupDMR (gDMR.cu:2600-2621): restart from the old Krylov vector (m-k)/2 while the
  // stagnation flag is set, except every period-th flagged cycle.  A zero period (the reference
  // divides by it) means "from the residual".
  const int period = (m - k) > 0 ? (2 * matRank) / (m - k) : 0;
  const bool fromResidual = !c.stagRule || c.kryLoopFlag == 0 || period == 0 ||
                            (c.kryFlagCount % period) == 0;
  *fromResidualOut = fromResidual;
  double *Po = v.col(v.P, o);
  if (fromResidual)
    std::copy(v.r.begin(), v.r.end(), Po);
  else {
    const double *src = v.col(v.P, o + (m - k) / 2);
    if (src != Po) std::copy(src, src + ld, Po);
  }
  std::fill(c.H.begin(), c.H.end(), 0.0);
  double *H = c.H.data();
  // Deflation image A Z = W R.
  for (int i = 0; i < d; ++i) {
    spmv(c.A, v.col(v.Z, i), v.g.data());
    ++c.matvecs;
    double *hcol = H + (size_t)i * (m + 1);
    if (i > 0) multidot(v.P.data(), ld, i, v.g.data(), n, G, hcol);
    double *Pi = v.col(v.P, i);
    comb_sub(v.g.data(), Pi, v.P.data(), i, hcol, nullptr, 0, nullptr, ld, n);
    const double ss = sumsq(Pi, n, G), sg = dotp(Pi, v.g.data(), n, G);
    const double sc = 1.0 / std::sqrt(ss);
    hcol[i] = sg * sc;
    scale(Pi, n, sc);
  }
  if (d > 0) {
    multidot(v.P.data(), ld, k, Po, n, G, c.cv.data());
    comb_sub(Po, Po, v.P.data(), k, c.cv.data(), nullptr, 0, nullptr, ld, n);
    const double ss = sumsq(Po, n, G);
    scale(Po, n, 1.0 / std::sqrt(ss));
  }
  int brk = 0;
  for (int j = 1; j <= m - d; ++j) {
    spmv(c.A, v.col(v.P, o + j - 1), v.g.data());
    ++c.matvecs;
    double *hcol = H + (size_t)(d + j - 1) * (m + 1);
    const double *src = v.g.data();
    if (d > 0) {
      multidot(v.P.data(), ld, k, v.g.data(), n, G, hcol);
      comb_sub(v.g.data(), v.u.data(), v.P.data(), k, hcol, nullptr, 0, nullptr, ld, n);
      src = v.u.data();
    }
    multidot(Po, ld, j, src, n, G, hcol + d);
    double *vj = v.col(v.P, o + j);
    comb_sub(src, vj, Po, j, hcol + d, nullptr, 0, nullptr, ld, n);
    const double ss = sumsq(vj, n, G), sg = dotp(vj, v.g.data(), n, G);
    const double nrm = std::sqrt(ss);
    if (tol > 0.0 && nrm < tol) {
      brk = j;
      break;
    }
    const double sc = 1.0 / std::sqrt(ss);
    hcol[d + j] = sg * sc;
    scale(vj, n, sc);
  }
  if (brk) {
    // Zero slots o+j..m; ones on the diagonal from column d+j (the reference starts at column j,
    // gDMR.cu:2275, which overwrites filled columns when d > 0).
    std::fill(v.P.begin() + (long long)(o + brk) * ld, v.P.begin() + (long long)(m + 1) * ld, 0.0);
    for (int cc = d + brk; cc < m; ++cc) H[cc + (size_t)cc * (m + 1)] = 1.0;
    ++c.breakdowns;
  }
  return brk;
}

int run_cycles(OracleCtx &c, int startDef, double tol, int matRank, double beta, double &rho,
               int &grdItr) {
  Vecs &v = c.v;
  const long long n = v.n, ld = v.ld;
  const int m = v.m, k = v.k, G = c.G;
  while (rho > tol * beta) {
    if (grdItr > matRank) return 6;
    if (rho / beta > 1.0) return 7;
    const int d = (grdItr == 0 && startDef == 0) ? 0 : k;
    const int o = d > 0 ? k : 0;
    std::vector<double> Zused, Pbefore;
    if (c.keepLast) {
      Zused = v.Z;
      Pbefore = v.P;
    }
    bool fromResidual = true;
    const int brk = arnoldi(c, d, matRank, tol, &fromResidual);
    if (c.keepLast) {
      c.lastP = v.P;
      c.lastH = c.H;
      c.lastZ = Zused;
      c.lastD = d;
      c.startLog.push_back(fromResidual ? 1 : 0);
      if (!fromResidual) {
        c.stagPrevP = Pbefore;
        c.stagP = v.P;
        c.stagD = d;
      }
    }
    double *Po = v.col(v.P, o);
    // updateAppResDMR
    multidot(v.P.data(), ld, m + 1, v.r.data(), n, G, c.t.data());
    oracle_lsq(c.H.data(), m, c.t.data(), c.y.data(), c.z.data());
    comb_acc(v.r.data(), v.r.data(), Po, m + 1 - o, c.z.data() + d, v.P.data(), d, c.z.data(), ld,
             n, -1.0);
    const double rr = sumsq(v.r.data(), n, G);
    const double q = std::sqrt(rr);
    scale(v.r.data(), n, 1.0 / std::sqrt(rr));
    comb_acc(v.xa.data(), v.xa.data(), Po, m - d, c.y.data() + d, v.Z.data(), d, c.y.data(), ld, n,
             rho);
    // genDeflatDMR; skipped after a breakdown, whose identity-padded H would select combinations
    // of zeroed basis vectors (the reference normalises them by 1/sqrt(0)).
    if (k > 0 && brk == 0) {
      oracle_svd(c.H.data(), m, k, kSvdTol, kSvdMaxSweeps, c.C.data(), c.sig.data());
      std::vector<double> ss(k);
      for (int j = 0; j < k; ++j) {
        double *zj = v.col(v.Zn, j);
        comb_acc(nullptr, zj, Po, m - d, c.C.data() + d + (size_t)m * j, v.Z.data(), d,
                 c.C.data() + (size_t)m * j, ld, n, 1.0);
        ss[j] = sumsq(zj, n, G);
      }
      for (int j = 0; j < k; ++j) scale(v.col(v.Zn, j), n, 1.0 / std::sqrt(ss[j]));
      std::swap(v.Z, v.Zn);
      c.hasZ = true;
    }
    if (q > kStagnation) {
      c.kryLoopFlag = 1;
      ++c.kryFlagCount;
    } else {
      c.kryLoopFlag = 0;
    }
    rho = q * rho;
    c.hist.push_back(rho / beta);
    ++c.cycles;
    ++grdItr;
  }
  return 0;
}

}  // namespace

extern "C" {

OracleCtx *oracle_create(int n, int m, int k, int G) {
  if (n < 1 || m < 1 || k < 0 || k >= m || G < 1) return nullptr;
  OracleCtx *c = new OracleCtx();
  c->G = G;
  Vecs &v = c->v;
  v.n = n;
  v.ld = (n + 255) / 256 * 256;
  v.m = m;
  v.k = k;
  v.P.assign((size_t)v.ld * (m + 1), 0.0);
  v.Z.assign((size_t)v.ld * std::max(k, 1), 0.0);
  v.Zn.assign((size_t)v.ld * std::max(k, 1), 0.0);
  v.r.assign(v.ld, 0.0);
  v.xa.assign(v.ld, 0.0);
  v.g.assign(v.ld, 0.0);
  v.u.assign(v.ld, 0.0);
  c->H.assign((size_t)(m + 1) * m, 0.0);
  c->t.assign(m + 2, 0.0);
  c->y.assign(m + 2, 0.0);
  c->z.assign(m + 2, 0.0);
  c->cv.assign(m + 2, 0.0);
  c->C.assign((size_t)m * std::max(k, 1), 0.0);
  c->sig.assign(m + 1, 0.0);
  return c;
}

void oracle_destroy(OracleCtx *c) { delete c; }

void oracle_keep_last(OracleCtx *c, int on) { c->keepLast = on != 0; }

int oracle_set_csr(OracleCtx *c, const int *rp, const int *ci, const double *va) {
  const int n = (int)c->v.n;
  c->A.n = n;
  c->A.rp.assign(rp, rp + n + 1);
  const int nnz = rp[n];
  c->A.ci.assign(ci, ci + nnz);
  c->A.va.assign(va, va + nnz);
  return 0;
}

// New values for the operator's sparsity pattern (nnz of them), keeping the recycled subspace
// and the stagnation counter: the oracle side of dmrSetOperatorValues on a sequence.
void oracle_set_values(OracleCtx *c, const double *va) {
  std::copy(va, va + c->A.va.size(), c->A.va.begin());
}

// Same contract as dmrSolve / solverDMR (FP64).  x: in x0, out solution.  Returns the value the
// solver returns; *numIts as the solver sets it; status: 0 converged, 6 cycle limit, 7 runaway,
// -1 early exit.
double oracle_solve(OracleCtx *c, int mode, const double *b, double *x, double tol, int matRank,
                    int *numIts, int *status) {
  Vecs &v = c->v;
  const long long n = v.n;
  const int k = v.k, G = c->G;
  c->hist.clear();
  c->startLog.clear();
  c->stagD = -1;
  c->cycles = c->matvecs = c->breakdowns = 0;
  *numIts = matRank * matRank;
  // mode != 0 without a stored subspace runs the first cycle undeflated, as mode 0.
  const int startDef = (mode != 0 && c->hasZ) ? k : 0;
  c->kryLoopFlag = 0;
  std::fill(v.r.begin(), v.r.end(), 0.0);
  std::fill(v.xa.begin(), v.xa.end(), 0.0);
  std::copy(b, b + n, v.r.begin());
  std::copy(x, x + n, v.xa.begin());
  double bb = 0.0;
  {
    const double *r = v.r.data();
    bb = reduce(n, G, [&](long long i, double acc) { return std::fma(at(r, i, n), at(r, i, n), acc); });
  }
  spmv(c->A, x, v.g.data());
  ++c->matvecs;
  const double one = 1.0;
  comb_sub(v.r.data(), v.r.data(), v.g.data(), 1, &one, nullptr, 0, nullptr, v.ld, n);
  const double rr = sumsq(v.r.data(), n, G);
  const double beta = std::sqrt(bb);
  double rho = std::sqrt(rr);
  c->beta = beta;
  if (rho < beta * tol) {
    std::copy(v.xa.begin(), v.xa.begin() + n, x);
    *status = -1;
    c->rho = rho;
    return rho;
  }
  scale(v.r.data(), n, 1.0 / std::sqrt(rr));
  int grdItr = 0;
  const int rc = run_cycles(*c, startDef, tol, matRank, beta, rho, grdItr);
  c->rho = rho;
  *status = rc;
  if (rc) {
    // The cycle-limit and runaway exits return before the solution export (gDMR.cu:2816-2829
    // against 2885): x is left as passed in.
    *numIts = grdItr;
    return rho / beta;
  }
  *numIts = grdItr - 1;
  std::copy(v.xa.begin(), v.xa.begin() + n, x);
  return rho / beta;
}

int oracle_history(OracleCtx *c, double *hist, int cap) {
  const int cnt = (int)c->hist.size();
  for (int i = 0; i < std::min(cap, cnt); ++i) hist[i] = c->hist[i];
  return cnt;
}

void oracle_counts(OracleCtx *c, int *cycles, int *matvecs, int *breakdowns) {
  *cycles = c->cycles;
  *matvecs = c->matvecs;
  *breakdowns = c->breakdowns;
}

int oracle_has_z(OracleCtx *c) { return c->hasZ ? 1 : 0; }

void oracle_get_z(OracleCtx *c, double *Z) {
  const Vecs &v = c->v;
  for (int j = 0; j < v.k; ++j)
    std::copy(v.Z.begin() + (long long)j * v.ld, v.Z.begin() + (long long)j * v.ld + v.n,
              Z + (long long)j * v.n);
}

void oracle_set_z(OracleCtx *c, const double *Z) {
  Vecs &v = c->v;
  for (int j = 0; j < v.k; ++j)
    std::copy(Z + (long long)j * v.n, Z + (long long)(j + 1) * v.n, v.Z.begin() + (long long)j * v.ld);
  c->hasZ = true;
}

// Last cycle (when keep_last is on): P (n x (m+1)), H ((m+1) x m), Z used by the cycle (n x k),
// and d.  Returns d.
int oracle_last_cycle(OracleCtx *c, double *P, double *H, double *Z) {
  const Vecs &v = c->v;
  if (c->lastP.empty()) return -1;
  for (int j = 0; j <= v.m; ++j)
    std::copy(c->lastP.begin() + (long long)j * v.ld, c->lastP.begin() + (long long)j * v.ld + v.n,
              P + (long long)j * v.n);
  std::copy(c->lastH.begin(), c->lastH.end(), H);
  for (int j = 0; j < v.k; ++j)
    std::copy(c->lastZ.begin() + (long long)j * v.ld, c->lastZ.begin() + (long long)j * v.ld + v.n,
              Z + (long long)j * v.n);
  return c->lastD;
}

// Start choice of each cycle of the last solve (keep_last): 1 from the residual, 0 from the old
// Krylov vector.  Returns the cycle count.
int oracle_start_log(OracleCtx *c, int *log, int cap) {
  const int cnt = (int)c->startLog.size();
  for (int i = 0; i < std::min(cap, cnt); ++i) log[i] = c->startLog[i];
  return cnt;
}

// The most recent cycle of the last solve that restarted from the old Krylov vector (keep_last):
// the basis before its setup (prevP) and after its Arnoldi process (P), n x (m+1) each.  Returns
// its d, or -1 when no cycle restarted that way.
int oracle_stagnation_cycle(OracleCtx *c, double *prevP, double *P) {
  const Vecs &v = c->v;
  if (c->stagD < 0) return -1;
  for (int j = 0; j <= v.m; ++j) {
    std::copy(c->stagPrevP.begin() + (long long)j * v.ld,
              c->stagPrevP.begin() + (long long)j * v.ld + v.n, prevP + (long long)j * v.n);
    std::copy(c->stagP.begin() + (long long)j * v.ld, c->stagP.begin() + (long long)j * v.ld + v.n,
              P + (long long)j * v.n);
  }
  return c->stagD;
}

// Stagnation restart rule on (default) or off; off only to show in tests that the rule matters.
void oracle_set_stagnation_rule(OracleCtx *c, int on) { c->stagRule = on != 0; }

// Persistent stagnation counter (kryFlagCountDMR, gDMR.cu:890).
int oracle_get_stagnation_count(OracleCtx *c) { return c->kryFlagCount; }

void oracle_spmv(int n, const int *rp, const int *ci, const double *va, const double *x,
                 double *y) {
  Csr A;
  A.n = n;
  A.rp.assign(rp, rp + n + 1);
  A.ci.assign(ci, ci + rp[n]);
  A.va.assign(va, va + rp[n]);
  spmv(A, x, y);
}

// OpenMP threads of the oracle's parallel loops (the CPU baseline times it at 1 thread and at
// the host's core count).
void oracle_set_threads(int threads) {
#ifdef _OPENMP
  omp_set_num_threads(threads > 0 ? threads : 1);
#else
  (void)threads;
#endif
}

int oracle_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Host DRAM bandwidth: a[i] = b[i] + 3 c[i] over three arrays of n doubles with the current
// thread count, best of `reps` passes after a first-touch pass; returns bytes per second
// (24 n bytes moved per pass).
double oracle_stream_triad(long long n, int reps) {
  std::vector<double> a(n), b(n), c(n);
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < n; ++i) {
    a[i] = 0.0;
    b[i] = 1.0;
    c[i] = 2.0;
  }
  double best = 0.0;
  for (int r = 0; r < reps; ++r) {
    const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < n; ++i) a[i] = b[i] + 3.0 * c[i];
    const double dt =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (dt > 0.0) best = std::max(best, 24.0 * (double)n / dt);
  }
  volatile double sink = a[n / 2];
  (void)sink;
  return best;
}

// Dot product in the device reduction order with G blocks.
double oracle_dot(const double *a, const double *b, long long n, int G) {
  return reduce(n, G, [&](long long i, double acc) { return std::fma(at(a, i, n), at(b, i, n), acc); });
}

}  // extern "C"
