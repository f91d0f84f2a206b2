// Internal interface of the block (multi-right-hand-side) DMR kernels (dmr_block.hip) used by the
// block driver (dmr_block.cpp).  Blocks of vectors are column-major with leading dimension ld.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>

#include "dmr_internal.h"

namespace dmr {
namespace blk {

// Widest block: one 32-column MFMA tile.
constexpr int kMaxRhs = 32;
// gemm_tn holds up to kMaxTiles 32-column tiles of P (4 per wavefront): p <= 32 * kMaxTiles.
constexpr int kMaxTiles = 12;
constexpr int kMaxCols = 32 * kMaxTiles;
// Cap on the row ranges (workgroups) of gemm_tn.
constexpr int kMaxTnBlocks = 1024;

// Y[:, 0:s] = A X[:, 0:s] with FP32 values and vectors, FP32 accumulation; cg (4, 8 or 16)
// columns per thread, or 100 + cg for the variant that loads each row's indices up front, or
// 200 + cg (cg 8 or 16) for the one that stages the workgroup's indices in LDS.  Every variant
// sums each row's entries in order: the results are identical.
void spmm32(const CsrDev &A, const float *X, long long ldx, float *Y, long long ldy, int s,
            int cg, hipStream_t st);
// bf16 X (basis storage): the up-front variants only (cg or 100 + cg; 200 + cg).
void spmm32(const CsrDev &A, const bf16 *X, long long ldx, float *Y, long long ldy, int s, int cg,
            hipStream_t st);
// R = B - A X in FP64 (s columns).  With val, column c uses the values val + c * ldv (same
// sparsity as A); without, A's FP64 values.
void resid_block(const CsrDev &A, const double *val, long long ldv, const double *X,
                 long long ldx, const double *B, long long ldb, double *R, long long ldr, int s,
                 hipStream_t st);
// out = FP32 entrywise mean of count value arrays (val + c * ldv, nnz each), summed in order.
void mean_values(const double *val, long long ldv, int count, long long nnz, float *out,
                 hipStream_t st);
// out[c] = ||X[:, c]||^2 (sqrtMode 0) or ||X[:, c]|| (1), FP64 accumulation in a fixed order.
// partial: colnorm_blocks(n) * s doubles.
int colnorm_blocks(long long n);
template <typename T>
void colnorms(const T *X, long long ld, int s, long long n, double *partial, double *out,
              int sqrtMode, hipStream_t st);
// C = P[:, 0:p]^T U[:, 0:s] (p <= kMaxCols, s <= 32) over G row ranges (gemm_tn_blocks), MFMA in
// FP32, partials summed in order in FP64.  C64 (optional): column-major, ld ldc.  C32 (optional):
// row-major p x 32, zero beyond s.  part: G * roundup32(p) * 32 floats.
int gemm_tn_blocks(long long n, int cap, int rowsTarget);
size_t gemm_tn_lds(int pPad);
// variant 0: register-direct 32x32x2 MFMA; 1: LDS-staged; 2: register-direct 16x16x4 MFMA.
// reduce = false launches the partials only; gemm_tn_reduce then sums them (the same two launches
// as reduce = true, for callers that time them apart).
void gemm_tn(const float *P, long long ldp, int p, const float *U, long long ldu, int s,
             long long n, int G, int variant, float *part, double *C64, int ldc, float *C32,
             hipStream_t st, bool reduce = true);
// bf16 P (basis storage), FP32 U: variant 3 uses the bf16 matrix cores (U split into hi + lo
// bf16 parts), any other the FP32 ones on unpacked values.
void gemm_tn(const bf16 *P, long long ldp, int p, const float *U, long long ldu, int s,
             long long n, int G, int variant, float *part, double *C64, int ldc, float *C32,
             hipStream_t st, bool reduce = true);
// The fixed-order sum of gemm_tn's G partials into C64 / C32 (as gemm_tn describes them).  pair:
// the two-block layout of gemm_tn_pair (s columns of each block, the second block to C64b).
void gemm_tn_reduce(const float *part, int G, int p, int s, double *C64, int ldc, float *C32,
                    hipStream_t st, bool pair = false, double *C64b = nullptr, int ldcb = 0);
// Two blocks U and U2 of s <= 16 columns each (pitch ldu) against a bf16 P in one pass (the
// matrix-core kernel): P^T U to C64a (ld ldca) and C32 (row-major p x 32, columns s .. 31 zero),
// P^T U2 to C64b (ld ldcb).  part as for gemm_tn (pPad x 32 floats per row range).
void gemm_tn_pair(const bf16 *P, long long ldp, int p, const float *U, const float *U2,
                  long long ldu, int s, long long n, int G, float *part, double *C64a, int ldca,
                  float *C32, double *C64b, int ldcb, hipStream_t st, bool reduce = true);
// Out = a In + b P[:, 0:p] C (C row-major p x 32; In unused when a == 0).  Out may alias In or P.
// un (2, 4, 8 or 16): column pairs of P loaded ahead of their MFMAs.
void gemm_nn(const float *P, long long ldp, int p, const float *C, const float *In, long long ldi,
             float a, float b, float *Out, long long ldo, int s, long long n, int un,
             hipStream_t st);
// bf16 basis P (un >= 100: the bf16 matrix cores with C split into hi + lo parts, else the FP32
// ones on unpacked values); or FP32 P written to a bf16 Out (rounded to nearest).
// gram (un >= 100, s <= 16): each workgroup also writes the partial Gram matrix Out^T Out of its
// rows in gemm_tn's partial layout (pPad = 32); gemm_nn_gram_blocks partials; gram_reduce sums
// them in a fixed order into G64 (s x s, ld s).
void gemm_nn(const bf16 *P, long long ldp, int p, const float *C, const float *In, long long ldi,
             float a, float b, float *Out, long long ldo, int s, long long n, int un,
             hipStream_t st, float *gram = nullptr);
int gemm_nn_gram_blocks(long long n, int un);
// One pass over a bf16 basis for two blocks of s <= 16 columns (ld ld): X1 -= P[:, 0:p] C[:, 0:16]
// and X2 -= P[:, 0:p] C[:, 16:32] (C row-major p x 32; un >= 100 as for gemm_nn).
void gemm_nn_pair(const bf16 *P, long long ldp, int p, const float *C, float *X1, float *X2,
                  long long ld, int s, long long n, int un, hipStream_t st);
// Cp (rows x 32) = [ZH32[:, 0:16] | -Yr32 on rows p0 .. cols - 1, zero elsewhere] for
// gemm_nn_pair: the residual and solution updates of a cycle in one pass.
void pack_pair(const float *ZH32, const float *Yr32, int p0, int cols, int rows, int s, float *Cp,
               hipStream_t st);
void gram_reduce(const float *part, int G, int s, double *G64, hipStream_t st);
void gemm_nn(const float *P, long long ldp, int p, const float *C, const float *In, long long ldi,
             float a, float b, bf16 *Out, long long ldo, int s, long long n, int un,
             hipStream_t st);
// Out[:, 0:s] = U[:, 0:s] C (C row-major s x 32), s <= 16, even ldu and ldo (else returns false
// and launches nothing).  Out may alias U.
bool scale_block(const float *U, long long ldu, const float *C, float *Out, long long ldo, int s,
                 long long n, hipStream_t st);
bool scale_block(const float *U, long long ldu, const float *C, bf16 *Out, long long ldo, int s,
                 long long n, hipStream_t st);
// The Gram reduce of G partials (s <= 16, the layout gram_reduce reads) and chol_qr's step on
// the sum, in one launch of gram_chol_segments(G) workgroups.  seg: 256 doubles per segment;
// ctr: one counter, zero on entry, left at zero.
int gram_chol_segments(int G);
constexpr int kGramCholSegs = 64;  // most segments
void gram_chol(const float *part, int G, int s, double relTol, float *Sinv32, double *Sacc,
               int first, double *Hdst, int ldh, int *brk, double *seg, unsigned *ctr,
               hipStream_t st);
// Cholesky QR step on the Gram matrix G (s x s, ld s): see k_chol_qr.
void chol_qr(const double *G, int s, double relTol, float *Sinv32, double *Sacc, int first,
             double *Hdst, int ldh, int *brk, hipStream_t st);
// Block Hessenberg least squares (rows = cols + bw rows, lower bandwidth bw <= 32, ld rows) with
// s right-hand sides T (rows x s, ld rows): Y (cols x 32 row-major FP64), Yr32 = Y diag(rho),
// ZH32 = H Y.  Workspaces R (cols x cols) and Qt (cols x 32).  LDS: lsq_block_lds bytes.
size_t lsq_block_lds(int cols, int bw, int s);
void lsq_block(const double *H, int rows, int cols, int bw, const double *T, int s,
               const double *rho, double *R, double *Qt, double *Y, float *Yr32, float *ZH32,
               hipStream_t st);
void pack_coef(const double *C, int cols, int k, float *C32, hipStream_t st);
void track(double *rho, const double *q, double *scal, int s, hipStream_t st);
void add_scaled(double *X, long long ldx, const float *D, long long ldd, const double *f, int s,
                long long n, hipStream_t st);
void scale_to32(const double *R, long long ldr, const double *f, float *D, long long ldd, int s,
                long long n, hipStream_t st);
// out[j] = 1 / q[j] (0 where q[j] = 0); rho[j] = 1 when rho is given (s <= 64).
void recip(const double *q, double *out, int s, double *rho, hipStream_t st);
// Sets the LDS limits of the one-workgroup kernels; call before any stream capture.
void prepare();
// p[0..cnt) = 0 (cnt <= 64), a kernel rather than a memset node.
void zero_ints(int *p, int cnt, hipStream_t st);

}  // namespace blk
}  // namespace dmr
