// One-workgroup FP64 kernels on the (m+1) x m Hessenberg matrix of a DMR cycle.
//
// * lsq: least-squares solve min ||t - H y|| by Householder QR (Golub & Van Loan, Matrix
//   Computations, Alg. 5.1.1 house + 5.2.1, LAPACK dlarfg conventions) applied to [H | t],
//   column-oriented back substitution, then z = H y.  Replaces cusolverDnCgeqrf + Cungqr +
//   the Q^H t GEMV + cublasCtrsv + cublasCgemv of qrHessArrDMR / qrSolveDMR / hessenbergMultDMR
//   (gDMR.cu:591-633, 690-839).  The implicit application of the reflectors to t gives the same
//   y as forming Q explicitly.
// * svd_smallest: right singular vectors by one-sided Jacobi (Hestenes 1958; Demmel & Veselic
//   1992) with the round-robin ("chess tournament") pair ordering, one 32-lane half wavefront
//   per column pair, m/2 disjoint pairs per round.  Replaces cusolverDnCgesvdj
//   (gDMR.cu:634-689).  Singular values are sorted descending; the k vectors of the smallest ones
//   are returned, each with its largest-magnitude component made positive.
// * image_qr / image_fill: the image A Z of a refreshed deflation subspace from the cycle's
//   Arnoldi relation (QR of H C by modified Gram-Schmidt), for the FP32 inner solves.
//
// Every operation order here (sequential per-thread sums, wavefront xor butterflies, explicit
// fma) is fixed, so the FP64 oracle can reproduce the results bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>

#include "dmr_internal.h"
#include "dmr_lanes.h"

namespace dmr {
namespace {

// Xor butterfly over the 32 lanes of a half wavefront.
__device__ inline double half_sum(double v) {
  v = xor_sum<32>(v);
  return v;
}
// Sum of squares of a[0..len): lane hl adds rows hl, hl+32, ... in order, then half_sum.
// Every lane of the half wavefront gets the result.
__device__ inline double half_sumsq(const double *a, int len, int hl) {
  double s = 0.0;
  for (int i = hl; i < len; i += 32) s = fma(a[i], a[i], s);
  return half_sum(s);
}

constexpr int kLsqThreads = 256;
constexpr int kSvdThreads = 1024;

__global__ __launch_bounds__(kLsqThreads) void k_lsq(const double *__restrict__ H, int m,
                                                     const double *__restrict__ t,
                                                     double *__restrict__ y,
                                                     double *__restrict__ z) {
  extern __shared__ double sm[];
  const int ldh = m + 1;
  double *A = sm;                       // (m+1) x m, column-major
  double *b = A + (long long)ldh * m;   // m+1
  double *yv = b + ldh;                 // m
  __shared__ double s_tau;
  const int tid = threadIdx.x, bd = blockDim.x;
  for (int i = tid; i < ldh * m; i += bd) A[i] = H[i];
  for (int i = tid; i < ldh; i += bd) b[i] = t[i];
  __syncthreads();
  for (int j = 0; j < m; ++j) {
    double *vj = A + (long long)j * ldh;
    if (tid == 0) {
      const double alpha = vj[j];
      double xn2 = 0.0;
      for (int i = j + 1; i <= m; ++i) xn2 = fma(vj[i], vj[i], xn2);
      double tau = 0.0;
      if (xn2 != 0.0) {
        const double beta = -copysign(sqrt(fma(alpha, alpha, xn2)), alpha);
        tau = (beta - alpha) / beta;
        const double sc = 1.0 / (alpha - beta);
        for (int i = j + 1; i <= m; ++i) vj[i] = vj[i] * sc;
        vj[j] = beta;
      }
      s_tau = tau;
    }
    __syncthreads();
    const double tau = s_tau;
    if (tau != 0.0) {
      for (int c = j + 1 + tid; c <= m; c += bd) {
        double *col = (c < m) ? A + (long long)c * ldh : b;
        double w = col[j];
        for (int i = j + 1; i <= m; ++i) w = fma(vj[i], col[i], w);
        const double tw = tau * w;
        col[j] = col[j] - tw;
        for (int i = j + 1; i <= m; ++i) col[i] = fma(-tw, vj[i], col[i]);
      }
    }
    __syncthreads();
  }
  for (int i = m - 1; i >= 0; --i) {
    if (tid == 0) yv[i] = b[i] / A[i + (long long)i * ldh];
    __syncthreads();
    const double yi = yv[i];
    for (int r = tid; r < i; r += bd) b[r] = fma(-A[r + (long long)i * ldh], yi, b[r]);
    __syncthreads();
  }
  for (int i = tid; i < m; i += bd) y[i] = yv[i];
  if (tid == 0) y[m] = 0.0;
  for (int r = tid; r <= m; r += bd) {
    double s = 0.0;
    for (int c = 0; c < m; ++c) s = fma(H[r + (long long)c * ldh], yv[c], s);
    z[r] = s;
  }
}

// H: rows x m (column-major, leading dimension ldH); the single-vector Hessenberg has
// rows = ldH = m + 1.
__global__ __launch_bounds__(kSvdThreads) void k_svd(const double *__restrict__ H, int rows,
                                                     int ldH, int m, int k, double tol,
                                                     int maxSweeps, int zeroLast,
                                                     double *__restrict__ C,
                                                     double *__restrict__ sig,
                                                     int *__restrict__ sweepsOut) {
  extern __shared__ double sm[];
  const int mm = m + (m & 1);  // even column count; column m is a zero dummy when m is odd
  const int ldh = rows;
  double *A = sm;                          // ldh x mm
  double *V = A + (long long)ldh * mm;     // m x mm (ld m)
  double *sg = V + (long long)m * mm;      // mm
  __shared__ int s_rot;
  const int tid = threadIdx.x, bd = blockDim.x;
  // One half wavefront (32 lanes) per column pair: a round's m/2 pairs all run at once.
  const int hl = tid & 31, hw = tid >> 5, nhw = bd >> 5;
  for (int idx = tid; idx < ldh * mm; idx += bd) {
    const int c = idx / ldh;
    A[idx] = c < m ? H[(long long)c * ldH + (idx - c * ldh)] : 0.0;
  }
  for (int idx = tid; idx < m * mm; idx += bd) {
    const int r = idx % m, c = idx / m;
    V[idx] = (r == c) ? 1.0 : 0.0;
  }
  __syncthreads();
  int sweep = 0;
  while (sweep < maxSweeps) {
    // Exact squared column norms at the start of each sweep; within the sweep a rotation
    // updates them as ||a_p||^2 - t g and ||a_q||^2 + t g.
    for (int c = hw; c < m; c += nhw) sg[c] = half_sumsq(A + (long long)c * ldh, ldh, hl);
    if (tid == 0) s_rot = 0;
    __syncthreads();
    for (int r = 0; r < mm - 1; ++r) {
      for (int pi = hw; pi < mm / 2; pi += nhw) {
        int p, q;
        if (pi == 0) {
          p = r;
          q = mm - 1;
        } else {
          p = (r + pi) % (mm - 1);
          q = (r - pi + (mm - 1)) % (mm - 1);
        }
        if (p > q) {
          const int tmp = p;
          p = q;
          q = tmp;
        }
        if (q >= m) continue;
        double *ap = A + (long long)p * ldh, *aq = A + (long long)q * ldh;
        double ga = 0.0;
        for (int i = hl; i < ldh; i += 32) ga = fma(ap[i], aq[i], ga);
        ga = half_sum(ga);
        const double al = sg[p], be = sg[q];
        if (ga != 0.0 && fabs(ga) > tol * sqrt(al * be)) {
          const double zeta = (be - al) / (2.0 * ga);
          const double tt = copysign(1.0, zeta) / (fabs(zeta) + sqrt(fma(zeta, zeta, 1.0)));
          const double cs = 1.0 / sqrt(fma(tt, tt, 1.0));
          const double sn = cs * tt;
          for (int i = hl; i < ldh; i += 32) {
            const double x = ap[i], y = aq[i];
            ap[i] = cs * x - sn * y;
            aq[i] = sn * x + cs * y;
          }
          double *vp = V + (long long)p * m, *vq = V + (long long)q * m;
          for (int i = hl; i < m; i += 32) {
            const double x = vp[i], y = vq[i];
            vp[i] = cs * x - sn * y;
            vq[i] = sn * x + cs * y;
          }
          if (hl == 0) {
            const double tg = tt * ga;
            sg[p] = al - tg;
            sg[q] = be + tg;
            atomicAdd(&s_rot, 1);
          }
        }
      }
      __syncthreads();
    }
    const int rot = s_rot;
    __syncthreads();
    ++sweep;
    if (rot == 0) break;
  }
  // Column norms = singular values.
  for (int c = hw; c < m; c += nhw) {
    const double s = half_sumsq(A + (long long)c * ldh, ldh, hl);
    // zeroLast: an exactly zero column (a zero basis vector of a block cycle) is never selected.
    if (hl == 0) sg[c] = (zeroLast && s == 0.0) ? INFINITY : sqrt(s);
  }
  __syncthreads();
  for (int c = tid; c < m; c += bd) {
    const double sc = sg[c];
    int rank = 0;
    for (int j = 0; j < m; ++j) rank += (sg[j] > sc || (sg[j] == sc && j < c)) ? 1 : 0;
    sig[rank] = sc;
    if (rank >= m - k) {
      const int i = rank - (m - k);
      const double *vc = V + (long long)c * m;
      int imax = 0;
      double vmax = fabs(vc[0]);
      for (int r = 1; r < m; ++r) {
        const double a = fabs(vc[r]);
        if (a > vmax) {
          vmax = a;
          imax = r;
        }
      }
      const bool neg = vc[imax] < 0.0;
      for (int r = 0; r < m; ++r) C[r + (long long)m * i] = neg ? -vc[r] : vc[r];
    }
  }
  if (tid == 0) *sweepsOut = sweep;
}

// Sum over a 64-lane wavefront (xor butterfly); every lane gets the result.
__device__ inline double wave_sum64(double v) {
  v = xor_sum<64>(v);
  return v;
}

// The image of a refreshed deflation subspace from the Arnoldi relation of the cycle that
// produced it.  The cycle satisfies A [Z V] = P H with P = [W V_0 .. V_{m-d}] (m + 1 orthonormal
// basis columns) and H its (m+1) x m Hessenberg matrix, and the refresh takes Z~ = [Z V] C
// (C: m x k, ld m; H with leading dimension ldH).  So A Z~ = P (H C), and the QR factorisation
// H C = Q R gives
// A Z = (P Q)(R D^-1), D = diag(||z~_i||): an orthonormal W = P Q and an upper triangular factor,
// in exact arithmetic the same pair as the QR of the recomputed image A Z (k operator
// applications and a Gram-Schmidt QR, gDMR.cu:1823-2031).  Modified Gram-Schmidt in FP64, one
// wavefront per remaining column.  Q: (m+1) x k (ld m+1); R: k x k (ld k), zero below the
// diagonal.  *ok = 0 when a column is numerically dependent on the previous ones.
__global__ __launch_bounds__(256) void k_image_qr(const double *__restrict__ H, int ldH, int m,
                                                  const double *__restrict__ C, int k,
                                                  double *__restrict__ Q, double *__restrict__ R,
                                                  int *__restrict__ ok) {
  extern __shared__ double sm[];
  const int rows = m + 1;
  double *M = sm;  // rows x k, ld rows
  __shared__ double s_nrm;
  __shared__ int s_ok;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  for (int idx = tid; idx < rows * k; idx += blockDim.x) {
    const int i = idx % rows, j = idx / rows;
    double a = 0.0;
    for (int l = 0; l < m; ++l) a = fma(H[i + (long long)ldH * l], C[l + (long long)m * j], a);
    M[idx] = a;
  }
  for (int idx = tid; idx < k * k; idx += blockDim.x) R[idx] = 0.0;
  if (tid == 0) s_ok = 1;
  __syncthreads();
  double first = 0.0;
  for (int j = 0; j < k; ++j) {
    double *mj = M + (long long)rows * j;
    if (wv == 0) {
      double s = 0.0;
      for (int r = lane; r < rows; r += 64) s = fma(mj[r], mj[r], s);
      s = wave_sum64(s);
      if (lane == 0) s_nrm = sqrt(s);
    }
    __syncthreads();
    double nrm = s_nrm;
    if (j == 0) first = nrm;
    if (!(nrm > 1e-12 * first)) {
      if (tid == 0) s_ok = 0;
      nrm = 1.0;
    }
    for (int r = tid; r < rows; r += blockDim.x) {
      const double q = mj[r] / nrm;
      mj[r] = q;
      Q[r + (long long)rows * j] = q;
    }
    if (tid == 0) R[j + (long long)k * j] = nrm;
    __syncthreads();
    for (int l = j + 1 + wv; l < k; l += nw) {
      double *ml = M + (long long)rows * l;
      double s = 0.0;
      for (int r = lane; r < rows; r += 64) s = fma(mj[r], ml[r], s);
      s = wave_sum64(s);
      for (int r = lane; r < rows; r += 64) ml[r] = fma(-s, mj[r], ml[r]);
      if (lane == 0) R[j + (long long)k * l] = s;
    }
    __syncthreads();
  }
  if (tid == 0) *ok = s_ok;
}

// H[r, i] = R[r, i] / ||z~_i|| (scalZ[i] = 1 / ||z~_i||) for r <= i < k: the deflation columns
// of a cycle whose image comes from k_image_qr (H zeroed before).
__global__ void k_image_fill(double *__restrict__ H, int m, int k, const double *__restrict__ R,
                             const double *__restrict__ scalZ) {
  for (int idx = threadIdx.x; idx < k * k; idx += blockDim.x) {
    const int r = idx % k, i = idx / k;
    if (r <= i) H[r + (long long)(m + 1) * i] = R[r + (long long)k * i] * scalZ[i];
  }
}

// Check of an implicit image on its last column (the smallest singular value, where the
// relation's error weighs most).  Prep: col[r] = R[r, k-1] / ||z~_{k-1}||, the coefficients of
// the expected image W col of the normalised z_{k-1}.  Final (after e = A z_{k-1} - W col with
// ss = ||e||^2 in chk[0]): *ok = 0 when ||e|| > relTol ||col|| (W orthonormal: ||W col|| = ||col||)
// or the check is not finite.
__global__ void k_image_check_prep(const double *__restrict__ R, int k,
                                   const double *__restrict__ scalZ, double *__restrict__ col) {
  for (int r = threadIdx.x; r < k; r += blockDim.x)
    col[r] = R[r + (long long)k * (k - 1)] * scalZ[k - 1];
}
__global__ void k_image_check_final(const double *__restrict__ col, int k,
                                    const double *__restrict__ chk, double relTol, int *ok) {
  if (threadIdx.x != 0) return;
  double cc = 0.0;
  for (int r = 0; r < k; ++r) cc = fma(col[r], col[r], cc);
  const double e = sqrt(chk[0]);
  if (!(e <= relTol * sqrt(cc))) *ok = 0;
}

}  // namespace

size_t image_qr_lds(int m, int k) { return sizeof(double) * (size_t)(m + 1) * (size_t)k; }

void image_check_prep(const double *R, int k, const double *scalZ, double *col, hipStream_t s) {
  hipLaunchKernelGGL(k_image_check_prep, dim3(1), dim3(64), 0, s, R, k, scalZ, col);
}
void image_check_final(const double *col, int k, const double *chk, double relTol, int *ok,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_image_check_final, dim3(1), dim3(64), 0, s, col, k, chk, relTol, ok);
}

void image_qr(const double *H, int ldH, int m, const double *C, int k, double *Q, double *R,
              int *ok, hipStream_t s) {
  const size_t lds = image_qr_lds(m, k);
  if (lds > 64 * 1024) prepare_dense();
  hipLaunchKernelGGL(k_image_qr, dim3(1), dim3(256), lds, s, H, ldH, m, C, k, Q, R, ok);
}

void image_fill(double *H, int m, int k, const double *R, const double *scalZ, hipStream_t s) {
  hipLaunchKernelGGL(k_image_fill, dim3(1), dim3(256), 0, s, H, m, k, R, scalZ);
}

void lsq(const double *H, int m, const double *t, double *y, double *z, hipStream_t s) {
  const size_t lds = sizeof(double) * ((size_t)(m + 1) * m + (m + 1) + m);
  if (lds > 64 * 1024) prepare_dense();
  hipLaunchKernelGGL(k_lsq, dim3(1), dim3(kLsqThreads), lds, s, H, m, t, y, z);
}

size_t svd_lds(int rows, int m) {
  const int mm = m + (m & 1);
  return sizeof(double) * ((size_t)rows * mm + (size_t)m * mm + mm);
}

void allow_max_lds(const void *fn) {
  hipFuncAttributes fa{};
  if (hipFuncGetAttributes(&fa, fn) == hipSuccess)
    hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)(160 * 1024 - fa.sharedSizeBytes));
  // Leave no error behind for the caller's next error check.
  (void)hipGetLastError();
}

void prepare_dense() {
  // Once per process: the largest dynamic LDS (160 KiB less the static LDS) of the
  // one-workgroup kernels.  Solver contexts call it at creation, so no launch needs it while
  // its stream is being captured.
  static const bool done = [] {
    allow_max_lds(reinterpret_cast<const void *>(k_svd));
    allow_max_lds(reinterpret_cast<const void *>(k_lsq));
    allow_max_lds(reinterpret_cast<const void *>(k_image_qr));
    return true;
  }();
  (void)done;
}

void svd_smallest_rect(const double *H, int rows, int ldH, int m, int k, double tol, int maxSweeps,
                       int zeroLast, double *C, double *sig, int *sweeps, hipStream_t s) {
  const size_t lds = svd_lds(rows, m);
  if (lds > 64 * 1024) prepare_dense();
  hipLaunchKernelGGL(k_svd, dim3(1), dim3(kSvdThreads), lds, s, H, rows, ldH, m, k, tol, maxSweeps,
                     zeroLast, C, sig, sweeps);
}

void svd_smallest(const double *H, int m, int k, double tol, int maxSweeps, double *C, double *sig,
                  int *sweeps, hipStream_t s) {
  svd_smallest_rect(H, m + 1, m + 1, m, k, tol, maxSweeps, 0, C, sig, sweeps, s);
}

}  // namespace dmr
