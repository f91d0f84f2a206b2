// Internal interface between the host driver (dmr_solver.cpp, dmr_ls.cpp) and the HIP kernels
// (dmr_kernels.hip, dmr_dense.hip).  Nothing here is part of the public C ABI.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace dmr {

// All streaming kernels run 256-thread blocks (four 64-wide wavefronts).
constexpr int kBlock = 256;
// Cap on the number of row blocks of a fixed-order reduction.
constexpr int kMaxRedBlocks = 1024;
constexpr int kDefaultRedBlocks = 512;
// Widest column group of the multi-dot (GEMV-T) kernel.
constexpr int kDotCols = 64;
// Largest single-output coefficient count the combination kernels stage in LDS.
constexpr int kMaxCoef = 256;
// Row blocks of the CSR SpMV hold at most this many nonzeros (staged in LDS) and 256 rows.
constexpr int kSpmvNnzCap = 2048;

// bfloat16 storage (upper 16 bits of an IEEE binary32), round-to-nearest-even on store.
struct bf16 {
  uint16_t bits;
};

// Elements per 16-byte vector unit of a storage type.
template <typename T> struct Unit;
template <> struct Unit<double> { static constexpr int N = 2; };
template <> struct Unit<float> { static constexpr int N = 4; };
template <> struct Unit<bf16> { static constexpr int N = 8; };

template <typename T> inline long long units(long long n) {
  return (n + Unit<T>::N - 1) / Unit<T>::N;
}

// Number of row blocks of the fixed-order reduction over vectors of n elements of type T.
// Every reduction kernel over such vectors uses exactly this geometry, so results do not depend
// on which kernel computed them, on co-scheduled work, or on the run.
// The cap is kMaxRedBlocks unless the environment variable DMR_RED_BLOCKS (read once per
// process, 1..kMaxRedBlocks) lowers it; dmrReductionBlocks reports the value in use.
int red_blocks_cap();
// The same cap applies to every storage type (measured: halving it for FP32 vectors slowed the
// FP32 multi-dot from 3.1 to 2.85 TB/s at n = 2^21).
template <typename T> inline int red_blocks(long long n) {
  long long g = (units<T>(n) + kBlock - 1) / kBlock;
  const int cap = red_blocks_cap();
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

// The thread-local error channel behind dmrGetLastError / dmrGetLastErrorString.
int set_error(int code, const char *msg);
void clear_error();

// Leading dimension (elements) of the basis columns for vectors of n elements: n rounded up to
// 256, plus a pad so that columns streamed together do not start at addresses a power of two
// apart (which maps them onto the same HBM channels).  DMR_LD_PAD (elements, multiple of 256)
// overrides the pad.
long long basis_ld(long long n);

// Workspace of the fixed-order reductions: block partials and arrival counters.  Counters start
// at zero and every kernel returns its counter to zero, so one workspace serves any number of
// stream-ordered launches.  Kernels running concurrently need distinct workspaces.
struct RedWork {
  void *partials;        // kMaxPartials accumulators (double or float)
  unsigned *counters;    // kCounters counters
};
constexpr int kMaxPartials = 128 * kMaxRedBlocks;
// Arrival counters of a grid reduction: one top counter and kArrivalFan sub-counters per column
// group (blocks arrive on sub-counter b % kArrivalFan, the last of each sub-group on the top
// counter), each on its own 128-byte line, for up to kCounterGroups column groups per launch.
constexpr int kArrivalFan = 16;
constexpr int kCounterStride = 32;
constexpr int kCounterGroups = 8;
constexpr int kCounters = kCounterGroups * (1 + kArrivalFan) * kCounterStride;

// How the last block of a combination kernel finalises its reductions.
enum FinMode : int {
  FIN_NONE = 0,
  FIN_NORM = 1,     // scal[0] = ss, scal[1] = 1/sqrt(ss), scal[2] = sqrt(ss)
  FIN_ARNOLDI = 2,  // breakdown test sqrt(ss) < tol -> *flag = step; else scal[1] = 1/sqrt(ss),
                    // *hdst = sg * scal[1]
  FIN_STORE = 3     // scal[0] = ss, scal[1] = sg
};
struct Fin {
  int mode;
  double *scal;
  double *hdst;
  double tol;
  int step;
  int *flag;
};

// Device CSR operator with its SpMV row-block partition.
struct CsrDev {
  int n = 0;
  int nnz = 0;
  const int *rowPtr = nullptr;
  const int *colInd = nullptr;
  const double *val = nullptr;
  const float *val32 = nullptr;
  int *blockRows = nullptr;   // nBlocks + 1 row boundaries (owned by the caller)
  int nBlocks = 0;
};

// Host-side partition of rows into SpMV blocks (<= 256 rows, <= kSpmvNnzCap nonzeros each; a
// longer row gets a block of its own).  Returns the number of blocks; fills br[0..nb].
int spmv_partition(const int *rowPtrHost, int n, int *br);

// ---- launchers (stream-ordered, no host synchronisation) ----
// dst = A src (FP64 values, FP64 vectors).
void spmv(const CsrDev &A, const double *x, double *y, hipStream_t s);
// FP32 values and vectors, FP32 accumulation.
void spmv32(const CsrDev &A, const float *x, float *y, hipStream_t s);
// FP32 values, bf16 vectors (FP32 products and sums, rounded to bf16 on store).
void spmv16(const CsrDev &A, const bf16 *x, bf16 *y, hipStream_t s);
// r = b - A x in FP64 with ss = ||r||^2 finalised by FIN_NORM into scal.
void residual64(const CsrDev &A, const double *x, const double *b, double *r, RedWork w,
                double *scal, hipStream_t s);

template <typename TS>
void fill(TS *p, long long count, double v, hipStream_t s);
void fill_int(int *p, long long count, int v, hipStream_t s);

// out[l * ostride] = sum_i B[i + l*ld] * x[i], l < p (one fixed-order reduction per column).
template <typename TS>
void multidot(const TS *B, long long ld, int p, const TS *x, long long n, double *out,
              int ostride, RedWork w, const int *flag, hipStream_t s);

// Single-output combination.
//   mode SUB: out = in - sum_l c1[l] B1_l - sum_l c2[l] B2_l   (sequential, one column at a time)
//   mode ACC: out = in + beta * (sum_l c1[l] B1_l + sum_l c2[l] B2_l)   (in may equal out)
// red: bit 0 accumulates ss = out.out, bit 1 sg = out.g; fin says what the last block does.
// betaDev (optional): beta read from device memory at run time instead of the argument (a
// value that changes between replays of a captured launch).
enum CombMode : int { COMB_SUB = 0, COMB_ACC = 1 };
template <typename TS>
void combo(int mode, const TS *in, TS *out, const TS *B1, int p1, const double *c1, const TS *B2,
           int p2, const double *c2, long long ld, long long n, double beta, int red,
           const TS *g, const Fin &fin, RedWork w, const int *flag, hipStream_t s,
           const double *betaDev = nullptr);

// out_j = sum_l C1[l + ldc*j] B1_l + sum_l C2[l + ldc*j] B2_l, j < J, with ss_j; scal[j] =
// 1/sqrt(ss_j).  Columns of out are at stride ld.
template <typename TS>
void combo_multi(TS *out, int J, const TS *B1, int p1, const double *C1, const TS *B2, int p2,
                 const double *C2, int ldc, long long ld, long long n, double *scal, RedWork w,
                 hipStream_t s);

// Fused orthogonalisation of one Arnoldi vector in one cooperative launch:
//   u = g; if pa: hA = A^T u, u -= A hA; if pb: hB = B^T u, u -= B hB; norms (red 1: ss, 3: ss
//   and sg); finalize (fin); dst = u / ||u|| (u unscaled if the launch set the breakdown flag).
// hA / hB receive the coefficients.  dst may equal g.  bar: two grid-barrier words reset at the
// start of every cycle; epoch0: phases already completed in the cycle.  Returns the number of
// barrier phases used (>= 1), or 0 when the grid cannot be co-resident or the shape is not
// supported (nothing enqueued: the caller runs the unfused kernels, which give the same bits).
template <typename TS>
int orth(const TS *g, TS *dst, const TS *A, int pa, double *ha, const TS *B, int pb, double *hb,
         long long ld, long long n, int red, const Fin &fin, RedWork w, unsigned *bar,
         unsigned epoch0, int *err, const int *flag, hipStream_t s);

// v *= scal[idx] (scal on device), skipped when *flag != 0.
template <typename TS>
void scale(TS *v, long long n, const double *scal, const int *flag, hipStream_t s);
// Column j of v (stride ld) *= scal[j], j < J.
template <typename TS>
void scale_cols(TS *v, int J, long long ld, long long n, const double *scal, hipStream_t s);
// dst = src converted (FP64 <-> FP32/BF16), n elements.
template <typename TD, typename TSrc>
void convert(TD *dst, const TSrc *src, long long n, hipStream_t s);
// x += d (FP64 x, d of type TS).
template <typename TS>
void add_to(double *x, const TS *d, long long n, hipStream_t s);

// Cycle start: H[(m+1) x m] = 0, *flag = 0 and the two grid-barrier words bar[0..1] = 0.
void cycle_begin(double *H, int m, int *flag, unsigned *bar, hipStream_t s);
// After the Arnoldi loop: if *flag = j != 0 zero basis columns o+j..m and set H[c][c] = 1 for
// columns c >= d+j.
template <typename TS>
void breakdown_fixup(TS *P, long long ld, long long n, int o, int d, int m, double *H,
                     const int *flag, hipStream_t s);

// Hessenberg least squares: y = argmin ||t - H y|| by Householder QR of [H t], back
// substitution; y[m] = 0; z = H y.  One workgroup.
void lsq(const double *H, int m, const double *t, double *y, double *z, hipStream_t s);
// Right singular vectors of H ((m+1) x m) by one-sided Jacobi with round-robin ordering.
// Writes the k vectors of the smallest singular values, in descending order of sigma, as
// columns of C (m x k, ld m), each with its largest-magnitude entry positive; sig[0..m-1] the
// singular values in descending order; *sweeps the number of sweeps.  One workgroup.
void svd_smallest(const double *H, int m, int k, double tol, int maxSweeps, double *C,
                  double *sig, int *sweeps, hipStream_t s);
// The same for a rows x m matrix with leading dimension ldH (rows >= m).  The matrix and the
// right vectors live in LDS: svd_lds(rows, m) bytes, at most 160 KiB.
size_t svd_lds(int rows, int m);
// Sets the dynamic LDS limits of the one-workgroup dense kernels (once per process).
void prepare_dense();
// Allows a kernel the dynamic LDS left by its static LDS within 160 KiB; clears the error slot.
void allow_max_lds(const void *fn);
// zeroLast != 0: exactly zero columns are ranked largest (never among the k selected).
void svd_smallest_rect(const double *H, int rows, int ldH, int m, int k, double tol, int maxSweeps,
                       int zeroLast, double *C, double *sig, int *sweeps, hipStream_t s);

// Streaming read of units16 16-byte units of buf (the HBM read ceiling probe); sink is written
// only in the practically impossible case that the folded words equal a fixed constant.
void probe_read(const void *buf, long long units16, unsigned *sink, hipStream_t s);

// The image A Z = W R_Z of the subspace refreshed from a cycle, from that cycle's Arnoldi
// relation instead of k operator applications: Q R = H C (H: the leading (m+1) x m block of the
// cycle's Hessenberg matrix, leading dimension ldH; C the m x k refresh coefficients, ld m; Q
// (m+1) x k, R k x k upper triangular, *ok = 0 for a dependent column), W = P Q by combo_multi,
// and image_fill puts R diag(scalZ) into the first k columns of the next cycle's H.  One
// workgroup each; image_qr needs image_qr_lds(m, k) <= 160 KiB of LDS.
size_t image_qr_lds(int m, int k);
void image_qr(const double *H, int ldH, int m, const double *C, int k, double *Q, double *R,
              int *ok, hipStream_t s);
void image_fill(double *H, int m, int k, const double *R, const double *scalZ, hipStream_t s);
// Check of an implicit image on column k - 1: image_check_prep writes the coefficients
// col = R[:, k-1] / ||z~_{k-1}|| of its expected image W col; after the caller forms
// e = A z_{k-1} - W col with ||e||^2 in chk[0], image_check_final clears *ok when
// ||e|| > relTol ||col||.
void image_check_prep(const double *R, int k, const double *scalZ, double *col, hipStream_t s);
void image_check_final(const double *col, int k, const double *chk, double relTol, int *ok,
                       hipStream_t s);

}  // namespace dmr
