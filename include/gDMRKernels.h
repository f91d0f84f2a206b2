/*
 * Kernel-level entry points of the block (multi-right-hand-side) DMR: the tall-skinny products,
 * the block SpMM, the Cholesky QR factor and the banded least squares that dmrBlockSolve chains.
 * No counterpart in the reference (its solver has no block form); they exist for tests of each
 * kernel against an independent reference and for callers that compose their own block
 * methods.  Every pointer is device memory, every call is enqueued on `stream` and returns 0 or a
 * DMR_ERR_* code (DMR_ERR_ARG for a shape the kernels do not take, checked before any launch).
 * Blocks of vectors are column-major with the given leading dimension; a bf16 array holds the
 * upper 16 bits of each binary32 (round to nearest even), as uint16_t.
 */
#ifndef GDMR_KERNELS_H
#define GDMR_KERNELS_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Basis storage of P in the products below. */
enum { DMR_KBASIS_FP32 = 0, DMR_KBASIS_BF16 = 1 };

/* Floats of partial-sum workspace a projection C = P^T U over n rows with p <= 384 columns of
 * P needs at rowsPerBlock rows per workgroup (a multiple of 64). */
size_t dmrKernelTnWorkspace(long long n, int p, int rowsPerBlock);
/* C = P[:, 0:p]^T U[:, 0:s] (s <= 32) on the matrix cores with per-workgroup partials summed in a
 * fixed order in FP64: C64 (optional, column-major p x s, ld ldc) and C32 (optional, row-major
 * p x 32, zero beyond s).  variant: 0 (FP32 32x32x2), 1 (FP32, LDS-staged), 2 (FP32 16x16x4; FP32
 * or bf16 P), 3 (bf16 16x16x16 with U split into hi + lo bf16 parts; bf16 P only). */
int dmrKernelGemmTn(int basis, const void *P, long long ldp, int p, const float *U,
                    long long ldu, int s, long long n, int rowsPerBlock, int variant, float *part,
                    double *C64, int ldc, float *C32, hipStream_t stream);
/* Two blocks U and U2 of s <= 16 columns against a bf16 P in one pass: P^T U to C64a and C32,
 * P^T U2 to C64b. */
int dmrKernelGemmTnPair(const uint16_t *P, long long ldp, int p, const float *U, const float *U2,
                        long long ldu, int s, long long n, int rowsPerBlock, float *part,
                        double *C64a, int ldca, float *C32, double *C64b, int ldcb,
                        hipStream_t stream);
/* Out = a In + b P[:, 0:p] C for s <= 32 output columns (C row-major p x 32; In unused when
 * a == 0).  Out may alias In.  unroll: loads in flight (2, 4, 8, 16) on the FP32 matrix cores, or
 * 100 / 101 (bf16 P only) for the bf16 matrix cores with 128 / 256 rows per wavefront.  gram
 * (bf16 P, unroll >= 100, s <= 16, optional): per-workgroup partial Gram matrices Out^T Out,
 * dmrKernelGramBlocks(n, unroll) of them in the projection's partial layout, which
 * dmrKernelGramReduce sums. */
int dmrKernelGemmNn(int basis, const void *P, long long ldp, int p, const float *C,
                    const float *In, long long ldi, float a, float b, float *Out, long long ldo,
                    int s, long long n, int unroll, float *gram, hipStream_t stream);
int dmrKernelGramBlocks(long long n, int unroll);
/* G64 (s x s, ld s) = the fixed-order FP64 sum of G partial Gram matrices. */
int dmrKernelGramReduce(const float *part, int G, int s, double *G64, hipStream_t stream);
/* One pass over a bf16 P for two blocks of s <= 16 columns (ld ld): X1 -= P C[:, 0:16] and
 * X2 -= P C[:, 16:32] (C row-major p x 32).  unroll 100 or 101. */
int dmrKernelGemmNnPair(const uint16_t *P, long long ldp, int p, const float *C, float *X1,
                        float *X2, long long ld, int s, long long n, int unroll,
                        hipStream_t stream);
/* Y[:, 0:s] = A X[:, 0:s] (s <= 32) for a CSR operator with FP32 values val32 (rowPtr n + 1,
 * colInd nnz), X FP32 (xBasis DMR_KBASIS_FP32) or bf16, Y FP32, FP32 accumulation over each
 * row's entries in order.  variant: cg (4, 8, 16 columns per thread), 100 + cg (indices loaded up
 * front), 200 + cg (cg 8 or 16, the workgroup's entries staged in LDS); bf16 X takes the last two
 * only.  All variants give the same bits. */
int dmrKernelSpmm(int xBasis, int n, int nnz, const int *rowPtr, const int *colInd,
                  const float *val32, const void *X, long long ldx, float *Y, long long ldy,
                  int s, int variant, hipStream_t stream);
/* Cholesky QR factor of the Gram matrix G (s x s, ld s, FP64, s <= 32): G = S^T S with S upper
 * triangular.  A pivot below relTol x the largest diagonal marks the column dependent: its row of
 * S and column of S^-1 are zero and flags[0] gets 1.  Writes Sinv32 (FP32 row-major 32 x 32:
 * S^-1 in its leading s x s, zero elsewhere) and the accumulated factor Sacc <- S Sacc (S itself
 * when first != 0; upper triangular s x s, ld s), copied to Hdst (ld ldh) when given. */
int dmrKernelCholQr(const double *G, int s, double relTol, float *Sinv32, double *Sacc, int first,
                    double *Hdst, int ldh, int *flags, hipStream_t stream);
/* dmrKernelGramReduce of G partials (s <= 16) followed by dmrKernelCholQr on the sum (first = 1),
 * in one launch: workgroups sum segments of the partials and the last one to finish adds the
 * segments in a fixed order and factorises.  The sum is the FP64 sum of the same partials in
 * another fixed order, so it may differ from dmrKernelGramReduce's in the last bits.  work:
 * DMR_KGRAMCHOL_WORK_BYTES bytes of device memory, zero before the first call; each call leaves
 * its counter at zero again, and calls that share a work buffer must not overlap. */
#define DMR_KGRAMCHOL_WORK_BYTES (256 + 64 * 256 * 8)
int dmrKernelGramChol(const float *part, int G, int s, double relTol, float *Sinv32, double *Sacc,
                      double *Hdst, int ldh, int *flags, void *work, hipStream_t stream);
/* Out[:, 0:s] = U[:, 0:s] C (C row-major s x 32, s <= 16, ldu and ldo even) into FP32 or bf16
 * (outBasis). */
int dmrKernelScaleBlock(const float *U, long long ldu, const float *C, void *Out, long long ldo,
                        int outBasis, int s, long long n, hipStream_t stream);
/* Least squares min ||T[:, j] - H Y[:, j]|| for each of s <= 32 right-hand sides on a block
 * Hessenberg H (rows = cols + bw, lower bandwidth bw <= 32, column-major ld rows, FP64): Y (cols x
 * 32 row-major FP64), Yr32 = Y diag(rho) (FP32 cols x 32), ZH32 = H Y (FP32 rows x 32).
 * Workspaces R (cols x cols doubles) and Qt (cols x 32 doubles).  Returns DMR_ERR_ARG when the
 * shape's working set does not fit one workgroup's LDS (dmrKernelLsqFits). */
int dmrKernelLsqBlock(const double *H, int rows, int cols, int bw, const double *T, int s,
                      const double *rho, double *R, double *Qt, double *Y, float *Yr32,
                      float *ZH32, hipStream_t stream);
int dmrKernelLsqFits(int cols, int bw, int s);

/* Measured HBM read bandwidth of the device: `bytes` (a multiple of 16, at least 64 MiB) of
 * `buf` streamed by 16-byte loads, eight in flight per thread, 2048 workgroups; two warm-up
 * launches, then `reps` timed launches (HIP events on `stream`), *gbps = the best.  A buffer
 * several times the 256 MiB memory-side cache measures HBM.  The stream is synchronised.  No
 * counterpart in the reference: it gives the practical ceiling that kernel bandwidths are
 * compared with beside the nominal peak. */
int dmrProbeReadBandwidth(const void *buf, size_t bytes, int reps, hipStream_t stream,
                          double *gbps);

#ifdef __cplusplus
}
#endif

#endif /* GDMR_KERNELS_H */
