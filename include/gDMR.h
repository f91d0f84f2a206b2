/*
 * gDMR.h -- C ABI of libgDMR.so, the MI355X-native deflated minimal residual (DMR) solver.
 *
 * DMR is restarted GMRES(m) in which k of the m search directions are recycled approximate
 * smallest right-singular vectors of A (Al Daas, Grigori, Henon, Ricoux, Inria RR-9206, 2018).
 *
 * The first block keeps the reference solver interface (names, arity, argument order, ownership,
 * the void(void) operator callback and its in-place contract).  Element type is real double
 * instead of complex float and the stream is a hipStream_t; everything else means what it means
 * in the reference.  The second block adds what the reference lacks: an error channel, an
 * out-of-place operator, statistics, handle-based contexts and the mixed-precision solve.
 */
#ifndef GDMR_H
#define GDMR_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------------------
 * Reference interface (replaces gDMR.h:13-34 of seanMolesky/gDMR-solver)
 * ------------------------------------------------------------------------------------------- */

/* Initialise the solver (replaces gDMR.h:13, gDMR.cu:840-1208).
 * blocks, threadsPerBlock: accepted and ignored; launch shapes are chosen per kernel.
 * devList, numDevs: must name exactly one device (numDevs == 1).
 * elements: vector length n.  basisDim: total basis m (Krylov + deflation).
 * deflatDim: deflation dimension k, 0 <= k < m.
 * linComp: operator callback.  It must overwrite the source vector most recently handed to
 * solverDMR (extSrcVec) with A * extSrcVec, enqueued on extStream (in place, as testLS.cu:97-128). */
void initDMR(int blocks, int threadsPerBlock, int *devList, int numDevs, int elements,
             int basisDim, int deflatDim, void (*linComp)(void));

/* Free all solver memory (replaces gDMR.h:15, gDMR.cu:1210-1475). */
void freeDMR(void);

/* Solve A x = b (replaces gDMR.h:29, gDMR.cu:2638-2901).
 * mode != 0 starts from the stored deflation subspace (previous solve or impSVDMR); with none
 *   stored the first cycle runs undeflated, as for mode 0.
 * extSrcVec: device, n doubles; read as x0, overwritten by A x0 and the operator inputs, and on
 *   return holds the solution.  extImgVec: device, n doubles, the right-hand side b (read only).
 * solTol: relative residual target; also the absolute Krylov breakdown threshold.
 * matRank: maximum number of restart cycles; also sets the stagnation restart period.
 * Returns ||b - A x|| / ||b|| tracked by the recursion, the absolute ||b - A x0|| when x0 already
 * meets the target (then *numIts = matRank*matRank), and *numIts = cycles - 1 on convergence. */
double solverDMR(int mode, double *extSrcVec, double *extImgVec, hipStream_t extStream,
                 int extDevNum, double solTol, int matRank, int *numIts);

/* Import / export the source-side deflation vectors Z: a host buffer of deflatDim contiguous
 * vectors of `elements` doubles (replaces gDMR.h:33-34, gDMR.cu:431-477). */
void impSVDMR(double *cMem);
void expSVDMR(double *cMem);

/* ---------------------------------------------------------------------------------------------
 * Extensions (no counterpart in the reference)
 * ------------------------------------------------------------------------------------------- */

/* Error codes.  The reference reports errors only on stderr and by resetting the device. */
enum {
  DMR_OK = 0,
  DMR_ERR_ARG = 1,        /* invalid argument (sizes, m/k, device list)            */
  DMR_ERR_STATE = 2,      /* call out of order (solve before init, mode!=0 w/o Z)  */
  DMR_ERR_HIP = 3,        /* a HIP runtime call failed                             */
  DMR_ERR_ALLOC = 4,      /* device allocation failed / not enough free HBM        */
  DMR_ERR_OPERATOR = 5,   /* the operator callback reported failure                */
  DMR_ERR_NOCONV = 6,     /* cycle limit reached (matRank)                         */
  DMR_ERR_RUNAWAY = 7,    /* residual grew above ||b|| ("runaway")                 */
  DMR_ERR_NOTBUILT = 8    /* requested feature is not available in this build      */
};
int dmrGetLastError(void);
const char *dmrGetLastErrorString(void);
/* An operator can raise an error from inside a callback; the solver returns at the next check. */
void dmrRaiseOperatorError(const char *msg);

/* Out-of-place operator: dst = A * src, both device vectors of n doubles, enqueued on stream. */
typedef void (*dmrOperatorFn)(const double *src, double *dst, hipStream_t stream, void *ctx);
void initDMRx(int device, int elements, int basisDim, int deflatDim, dmrOperatorFn op, void *ctx);

/* Statistics of the last solve of the global solver: relative residual rho/||b|| after each
 * cycle (up to cap entries), number of cycles and of operator applications, Krylov breakdowns. */
typedef struct {
  int cycles;
  int matvecs;
  int breakdowns;
  int status;          /* 0 converged, DMR_ERR_NOCONV, DMR_ERR_RUNAWAY, or -1 early exit */
  double beta;         /* ||b||                                                           */
  double rho;          /* final tracked ||r||                                             */
  double seconds;      /* wall time of the solve                                          */
  int svdSweeps;       /* Jacobi sweeps summed over the deflation refreshes read back     */
  int svdCalls;        /* deflation refreshes whose sweep count is included               */
} dmrStats;
int dmrGetStats(dmrStats *st, double *residHist, int cap);

/* Device geometry of the fixed-order reductions: every dot product over n elements is summed
 * by this many 256-thread blocks (2 elements per thread step), then block partials in order. */
int dmrReductionBlocks(int elements);

/* ---- handle-based contexts (serving, several operator families at once) ---- */
typedef struct dmrContext *dmrHandle;

/* Precision of the Krylov basis and SpMV.  FP64 is the plain solver; FP32 / BF16 run the inner
 * DMR in that storage (FP32 arithmetic, FP32-accumulated reductions) inside FP64 iterative
 * refinement; Hessenberg LSQ, QR and SVD always run in FP64. */
enum { DMR_PREC_FP64 = 0, DMR_PREC_FP32 = 1, DMR_PREC_BF16 = 2 };

int dmrCreate(dmrHandle *h, int device, int elements, int basisDim, int deflatDim, int precision);
int dmrDestroy(dmrHandle h);
/* Built-in CSR operator (device arrays, int32 indices, sorted columns).  val32 may be NULL for an
 * FP64-only handle; mixed-precision handles use val32 for the inner SpMV and val for the FP64
 * refinement residual. */
int dmrSetOperatorCsr(dmrHandle h, int nnz, const int *rowPtr, const int *colInd,
                      const double *val, const float *val32);
/* Replace the values of the CSR operator (same sparsity pattern), e.g. the next system of a
 * sequence whose coefficients change; no allocation when val32 is given or the handle is FP64. */
int dmrSetOperatorValues(dmrHandle h, const double *val, const float *val32);
int dmrSetOperatorFn(dmrHandle h, dmrOperatorFn op, void *ctx);
int dmrSetOperatorLegacy(dmrHandle h, void (*linComp)(void));
/* Same contract as solverDMR on a handle.  x: device, x0 in / solution out; b: device. */
double dmrSolve(dmrHandle h, int mode, double *x, const double *b, hipStream_t stream,
                double solTol, int matRank, int *numIts);
int dmrImportZ(dmrHandle h, const double *hostZ);
int dmrExportZ(dmrHandle h, double *hostZ);
int dmrHasZ(dmrHandle h);
/* Device-side copies of the recycled subspace in the handle's own layout (k columns of
 * dmrSubspaceBytes(h)/k bytes, the handle's storage precision), enqueued on stream: lets a
 * caller keep one resident subspace per operator family and hand it from handle to handle. */
size_t dmrSubspaceBytes(dmrHandle h);
int dmrSubspaceToDevice(dmrHandle h, void *dst, hipStream_t stream);
int dmrSubspaceFromDevice(dmrHandle h, const void *src, hipStream_t stream);
/* The stagnation counter that persists across the solves of one context (kryFlagCountDMR,
 * gDMR.cu:890, reset only by initDMR); with the subspace it is a recycled family's state. */
int dmrGetStagnationCount(dmrHandle h);
int dmrSetStagnationCount(dmrHandle h, int count);
int dmrGetStatsH(dmrHandle h, dmrStats *st, double *residHist, int cap);
/* Mixed-precision refinement: relative tolerance of each inner FP32 solve (default 1e-5) and the
 * maximum number of refinement steps (default 60). */
int dmrSetRefinement(dmrHandle h, double innerTol, int maxOuter);
/* The handle behind the global-state interface (NULL before initDMR / initLS). */
dmrHandle dmrGlobalHandle(void);
/* Per-phase profile of the restart cycles, measured with HIP events on the solve stream around
 * each launch (adds two event records per launch while on).  Phases: */
enum {
  DMR_PH_SPMV = 0,    /* operator applications                                  */
  DMR_PH_GEMVT = 1,   /* multi-dot projections (Gram-Schmidt, t = P^T r)        */
  DMR_PH_UPDATE = 2,  /* Gram-Schmidt updates with fused norms                  */
  DMR_PH_SCALE = 3,   /* normalisations                                         */
  DMR_PH_COMBO = 4,   /* residual / solution basis combinations                 */
  DMR_PH_ZGEMM = 5,   /* deflation-subspace refresh [V Z] C and normalisation   */
  DMR_PH_LSQ = 6,     /* Hessenberg least squares                               */
  DMR_PH_SVD = 7,     /* Hessenberg Jacobi SVD                                  */
  DMR_PH_ORTH = 8,    /* fused Gram-Schmidt step (projections, updates, norms)  */
  DMR_PH_REDUCE = 9,  /* fixed-order sums of block projection partials          */
  DMR_PH_RESID = 10,  /* FP64 refinement residuals b - A x                      */
  DMR_NPHASES = 11
};
int dmrSetProfiling(dmrHandle h, int on);
/* Accumulated milliseconds, algorithmic bytes and launch counts per phase since profiling was
 * switched on; each array has DMR_NPHASES entries (any may be NULL).  Returns DMR_NPHASES. */
int dmrGetProfile(dmrHandle h, double *ms, double *bytes, long long *launches);
/* x <- A x in place, allocating nothing: the contract of linOptLS (x <- A x on the device source
 * vector, testLS.cu:97-128) for a CSR operator and vector passed in as device pointers (rowPtr
 * n+1 and colInd nnz int32, val nnz doubles, x n doubles), enqueued on stream.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The product is staged in a
 * static device array of 4,194,304 doubles (32 MiB, held on every GPU the library runs on until
 * the process ends) and copied over x; n above that is rejected with DMR_ERR_ARG.  Every call on
 * one GPU stages through the same array, so calls that overlap on one GPU race.  Each row sums its
 * products in the order linOptLS does, so the results are the same bits.  Returns 0 or a
 * DMR_ERR_* code (also through dmrGetLastError). */
int dmrCsrApplyStatic(int n, int nnz, const int *rowPtr, const int *colInd, const double *val,
                      double *x, hipStream_t stream);
/* Bytes of HBM a handle of this shape allocates (for admission by free HBM). */
size_t dmrFootprint(int elements, int basisDim, int deflatDim, int precision);

/* ---- block (multi-right-hand-side) DMR ----
 * Up to 32 right-hand sides of one operator share one block Krylov space and one recycled
 * deflation subspace Z (n x k, k <= 32).  Per restart cycle: the deflation image A Z = W R_Z,
 * blockSteps block Arnoldi steps with one-stage block classical Gram-Schmidt (tall-skinny GEMMs
 * on the matrix cores) and Cholesky-QR normalisation, the full-projection least squares on the
 * block Hessenberg, and the refresh of Z from the smallest right singular vectors of its leading
 * block.  Storage and matrix-core arithmetic are FP32 inside FP64 iterative refinement per
 * column; least squares, Cholesky factors and SVD are FP64.  Basis columns k + (blockSteps+1) *
 * maxRhs must not exceed 384. */
typedef struct dmrBlockContext *dmrBlockHandle;
int dmrBlockCreate(dmrBlockHandle *h, int device, int elements, int blockSteps, int deflatDim,
                   int maxRhs);
int dmrBlockDestroy(dmrBlockHandle h);
/* Same contracts as dmrSetOperatorCsr / dmrSetOperatorValues; without val32 the FP32 values are
 * converted on the next solve's stream. */
int dmrBlockSetOperatorCsr(dmrBlockHandle h, int nnz, const int *rowPtr, const int *colInd,
                           const double *val, const float *val32);
int dmrBlockSetOperatorValues(dmrBlockHandle h, const double *val, const float *val32);
/* Per-column operators (no counterpart in the reference): column c of the following solves is
 * the system A_c X_c = B_c with A_c = the CSR pattern of dmrBlockSetOperatorCsr and the FP64
 * values val + c * ldv (device, nnz each, ldv >= nnz, count <= maxRhs columns; the caller keeps
 * them alive).  The FP64 refinement residual of column c uses A_c; the inner FP32 block solve
 * shares one operator for all columns: the entrywise mean of the count operators when innerMean
 * is nonzero (recomputed at the start of every solve), else the values of
 * dmrBlockSetOperatorCsr / dmrBlockSetOperatorValues.  This solves a sequence of systems that
 * differ in their values (a jittered or time-stepped family) as one block: the refinement
 * converges while the inner operator is close to every A_c.  val == NULL clears the setting;
 * dmrBlockSetOperatorCsr clears it and dmrBlockSetOperatorValues turns innerMean off.  A solve
 * with more columns than count is rejected. */
int dmrBlockSetColumnOperators(dmrBlockHandle h, const double *val, long long ldv, int count,
                               int innerMean);
/* Storage of the block basis [W V_0 V_1 ...] (no counterpart in the reference): DMR_PREC_FP32
 * (default) or DMR_PREC_BF16.  With bf16 the GEMMs and the SpMM read half the bytes and compute
 * in FP32 (MFMA FP32 accumulation); the newest block, the residual and solution blocks and the
 * recycled subspace stay FP32, and the refinement, least squares and SVD stay FP64.  Each inner
 * solve then reaches about 1e-3 relative, so the inner tolerance should be at least that; the
 * FP64 refinement still delivers the requested tolerance. */
int dmrBlockSetPrecision(dmrBlockHandle h, int precision);
/* Relative tolerance of the inner FP32 block solves (default 1e-5) and the maximum number of
 * refinement steps (default 30). */
int dmrBlockSetRefinement(dmrBlockHandle h, double innerTol, int maxOuter);
/* Solve A X = B for nrhs <= maxRhs columns.  X: device, column-major n x nrhs (leading dimension
 * ldx), x0 in / solution out.  B: device, ldb.  mode != 0 starts from the stored subspace (else
 * the first cycle is undeflated).  maxCycles caps the restart cycles of the whole solve, summed
 * over its refinement steps: once they are spent the solve returns DMR_ERR_NOCONV with X holding
 * the last refined iterate.  relRes (host, nrhs): the FP64-checked ||b - A x|| / ||b|| of each
 * column.  Returns 0 when every column meets solTol, else a DMR_ERR_* code. */
int dmrBlockSolve(dmrBlockHandle h, int mode, int nrhs, double *X, long long ldx, const double *B,
                  long long ldb, hipStream_t stream, double solTol, int maxCycles,
                  double *relRes);
/* cycles: inner restart cycles; matvecs: operator applications to single columns; breakdowns:
 * cycles with a dependent block column; rho: largest final relative residual; residHist: per
 * cycle, the largest estimated relative residual. */
int dmrBlockGetStats(dmrBlockHandle h, dmrStats *st, double *residHist, int cap);
int dmrBlockHasZ(dmrBlockHandle h);
/* graphs: replay each restart cycle as a captured HIP graph (default on; off while profiling).
 * autotune: time candidate launch shapes on the first solve of each block width and keep the
 * fastest (default on; cached per problem shape for the process).  A negative value leaves the
 * setting unchanged.  DMR_BLOCK_GRAPHS=0 / DMR_BLOCK_TUNE=0 change the defaults. */
int dmrBlockSetOptions(dmrBlockHandle h, int graphs, int autotune);
/* out[7]: gemm_tn rows per workgroup and variant for projections, the same for Gram matrices,
 * gemm_nn load-ahead depth, SpMM columns per thread, number of captured cycle graphs. */
int dmrBlockGetTuning(dmrBlockHandle h, int *out);
/* File of persisted launch shapes (no counterpart in the reference): one text line per tuned
 * problem shape, keyed by the device architecture and (n, nnz, block width, deflatDim, block
 * steps, basis storage).  Read before the process's first tuning, appended to after each new
 * one; NULL or "" turns persistence off (the default unless DMR_TUNE_FILE names a file).
 * Processes that share the file launch the same kernels for a shape, so their block solves give
 * the same bits. */
int dmrBlockTuningFile(const char *path);
int dmrBlockImportZ(dmrBlockHandle h, const double *hostZ);
int dmrBlockExportZ(dmrBlockHandle h, double *hostZ);
/* Phase profile as dmrSetProfiling / dmrGetProfile: SPMV = SpMM and FP64 residuals, GEMVT =
 * projections onto the basis (gemm_tn), UPDATE = block updates U -= P C (gemm_nn over the basis),
 * COMBO = residual and solution updates over the basis (gemm_nn), ZGEMM = products with the
 * subspace Z (solution update and refresh), LSQ = least squares, SVD = Jacobi SVD, SCALE =
 * column norms and scaling, ORTH = block normalisation (Cholesky QR: Gram matrices, factors,
 * scaling).  flops: the matrix-core and SpMM floating-point operations. */
int dmrBlockSetProfiling(dmrBlockHandle h, int on);
int dmrBlockGetProfile(dmrBlockHandle h, double *ms, double *bytes, double *flops,
                       long long *launches);
size_t dmrBlockFootprint(int elements, int blockSteps, int deflatDim, int maxRhs);

#ifdef __cplusplus
}
#endif

#endif /* GDMR_H */
