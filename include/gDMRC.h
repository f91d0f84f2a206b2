/*
 * gDMRC.h -- linear-system (user operator) C ABI bundled into libgDMR.so.
 *
 * Mirrors gDMRC.h:6-29 of seanMolesky/gDMR-solver: a control program drives a sequence of solves
 * through these calls only, and the linear-system translation unit registers its operator with
 * the solver at initLS.  The reference operator is a dense n x n complex matrix applied with a
 * library GEMV (testLS.cu:97-128); here the operator is a CSR matrix applied by the package's own
 * HIP SpMV, and vectors are real double.  impOptLS / expOptLS keep the dense interface
 * (column-major n x n, converted to / from CSR); impCsrLS is the native way in.
 */
#ifndef GDMRC_H
#define GDMRC_H

#ifdef __cplusplus
extern "C" {
#endif

/* Initialise the linear system and the solver (replaces gDMRC.h:6, testLS.cu:130-193).
 * devNumLO: device of the operator.  cells: vector length n.  devListMDR / numDevsMDR: the
 * solver device list (exactly one device, equal to devNumLO).  basisSizeMDR: m.
 * deflatSizeMDR: k.  blocksMDR / threadsPerBlockMDR: accepted and ignored. */
void initLS(int blocksMDR, int threadsPerBlockMDR, int devNumLO, int cells, int *devListMDR,
            int numDevsMDR, int basisSizeMDR, int deflatSizeMDR);
/* Free the linear system and the solver (replaces gDMRC.h:8, testLS.cu:195-244). */
void freeLS(void);
/* Import / export the image vector b, n doubles (replaces gDMRC.h:10-11). */
void impImgLS(double *cImg);
void expImgLS(double *cImg);
/* Import / export the source vector x, n doubles (replaces gDMRC.h:13-14). */
void impSrcLS(double *cSrc);
void expSrcLS(double *cSrc);
/* Import / export the operator as a dense column-major n x n matrix (replaces gDMRC.h:16-17).
 * Import keeps the nonzeros as CSR; export expands CSR back to dense. */
void impOptLS(double *cOpt);
void expOptLS(double *cOpt);
/* Solve A x = b with the imported x as x0 (replaces gDMRC.h:22, testLS.cu:386-391). */
double solveLS(int mode, double solTol, int matRank, int *numIts);
/* x <- A x in place on the device source vector (replaces gDMRC.h:24, testLS.cu:97-128). */
void linOptLS(void);
/* Re-declared from gDMR.h as in gDMRC.h:28-29. */
void impSVDMR(double *cMem);
void expSVDMR(double *cMem);

/* ---- extensions (no counterpart in the reference) ---- */
/* Import a CSR operator from host arrays: rowPtr[n+1], colInd[nnz] (0-based, sorted per row),
 * val[nnz].  Returns 0 or a DMR_ERR_* code. */
int impCsrLS(const int *rowPtr, const int *colInd, const double *val, int nnz);
/* Number of stored nonzeros of the current operator. */
int nnzLS(void);

/* Matrix Market files (the NIST exchange format) as on-disk operators.
 * Reads `matrix coordinate real|integer|pattern general|symmetric|skew-symmetric` and
 * `matrix array real|integer general`, 1-based, into 0-based CSR with columns sorted per row and
 * duplicate entries summed (symmetric files are expanded to both triangles).  Complex / Hermitian
 * files return DMR_ERR_NOTBUILT.  Call with rowPtr, colInd and val NULL to get the sizes, then
 * again with arrays of rows+1, nnz and nnz entries.  Returns 0 or a DMR_ERR_* code. */
int dmrReadMatrixMarket(const char *path, int *rows, int *cols, long long *nnz, int *rowPtr,
                        int *colInd, double *val);
/* Writes a square CSR matrix as `matrix coordinate real general` with 17 significant digits. */
int dmrWriteMatrixMarket(const char *path, int n, const int *rowPtr, const int *colInd,
                         const double *val);
/* Reads a Matrix Market file and imports it as the operator (it must be cells x cells). */
int impMtxLS(const char *path);

/* Synthetic problems shared by the control program, tests and benchmarks.
 * All produce CSR with int32 indices, columns sorted, rows in natural order
 * (x fastest: row = i + N*j (+ N*N*l)).  The matrix is h^2 times the discrete operator
 * -div(nu grad u) + beta . grad u on the unit square / cube with homogeneous Dirichlet
 * boundaries, central differences, h = 1/(N+1).  nu is per cell nu0*(1+delta), delta ~
 * U(-jitter, jitter) drawn from (seed, stream); face coefficients are the mean of the two cells
 * (the cell's own value on boundary faces).  Pass NULL arrays to query sizes. */
long long dmrGenConvDiff2D(int N, double nu0, double bx, double by, double jitter,
                           unsigned long long seed, unsigned long long stream,
                           int *rowPtr, int *colInd, double *val);
long long dmrGenConvDiff3D(int N, double nu0, double bx, double by, double bz, double jitter,
                           unsigned long long seed, unsigned long long stream,
                           int *rowPtr, int *colInd, double *val);
/* Anisotropic diffusion -(ax u_xx + ay u_yy + az u_zz), 7-point, h^2-scaled, no convection. */
long long dmrGenAniso3D(int N, double ax, double ay, double az, double jitter,
                        unsigned long long seed, unsigned long long stream,
                        int *rowPtr, int *colInd, double *val);
/* n i.i.d. N(0,1) doubles from the counter-based generator keyed by (seed, stream). */
void dmrRandNormal(unsigned long long seed, unsigned long long stream, long long n, double *out);
/* n i.i.d. U(lo, hi) doubles keyed by (seed, stream). */
void dmrRandUniform(unsigned long long seed, unsigned long long stream, long long n, double lo,
                    double hi, double *out);
/* Raw 64-bit output of the SplitMix64 generator (Steele, Lea, Flood 2014) from state `seed`. */
void dmrSplitMix64(unsigned long long seed, long long n, unsigned long long *out);

#ifdef __cplusplus
}
#endif

#endif /* GDMRC_H */
