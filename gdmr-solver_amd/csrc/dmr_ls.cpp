// Linear-system translation unit bundled into libgDMR.so (the gDMRC.h interface).
//
// Mirrors testLS.cu of seanMolesky/gDMR-solver: it owns the device source / image vectors and
// the operator, creates the operator stream, registers the operator with the solver at initLS
// (testLS.cu:130-193) and forwards solveLS to the solver (testLS.cu:386-391).  The operator is a
// CSR matrix applied by the package's HIP SpMV instead of a dense cuBLAS GEMV; it is registered
// out of place (y = A x into the solver's buffer), which avoids the temporary-and-copy that the
// in-place callback contract needs.  linOptLS keeps the in-place contract for control programs.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dmr_internal.h"
#include "gDMR.h"
#include "gDMRC.h"

extern "C" dmrHandle dmrGlobalHandle(void);

namespace {

struct LinearSystem {
  int dev = 0;
  int n = 0;
  hipStream_t stream = nullptr;
  double *src = nullptr;
  double *img = nullptr;
  double *tmp = nullptr;
  int *rowPtr = nullptr;
  int *colInd = nullptr;
  double *val = nullptr;
  int *blockRows = nullptr;
  dmr::CsrDev csr;
  bool live = false;
};

LinearSystem g_ls;

void report(const char *where, hipError_t e) {
  std::fprintf(stderr, "%s HIP error: %s\n", where, hipGetErrorString(e));
}

void release_csr() {
  if (g_ls.rowPtr) hipFree(g_ls.rowPtr);
  if (g_ls.colInd) hipFree(g_ls.colInd);
  if (g_ls.val) hipFree(g_ls.val);
  if (g_ls.blockRows) hipFree(g_ls.blockRows);
  g_ls.rowPtr = nullptr;
  g_ls.colInd = nullptr;
  g_ls.val = nullptr;
  g_ls.blockRows = nullptr;
  g_ls.csr = dmr::CsrDev{};
}

// Out-of-place operator registered with the solver.
void ls_apply(const double *x, double *y, hipStream_t s, void *) { dmr::spmv(g_ls.csr, x, y, s); }

int upload_csr(const int *rp, const int *ci, const double *va, int nnz) {
  const int n = g_ls.n;
  if (!rp || rp[0] != 0 || rp[n] != nnz || nnz < 0) return DMR_ERR_ARG;
  for (int r = 0; r < n; ++r) {
    if (rp[r + 1] < rp[r]) return DMR_ERR_ARG;
    for (int e = rp[r]; e < rp[r + 1]; ++e)
      if (ci[e] < 0 || ci[e] >= n) return DMR_ERR_ARG;
  }
  hipSetDevice(g_ls.dev);
  hipStreamSynchronize(g_ls.stream);
  release_csr();
  std::vector<int> br(n + 1);
  const int nb = dmr::spmv_partition(rp, n, br.data());
  hipError_t e;
  if ((e = hipMalloc(&g_ls.rowPtr, sizeof(int) * (n + 1))) != hipSuccess ||
      (e = hipMalloc(&g_ls.colInd, sizeof(int) * (size_t)std::max(nnz, 1))) != hipSuccess ||
      (e = hipMalloc(&g_ls.val, sizeof(double) * (size_t)std::max(nnz, 1))) != hipSuccess ||
      (e = hipMalloc(&g_ls.blockRows, sizeof(int) * (nb + 1))) != hipSuccess) {
    report("impCsrLS", e);
    release_csr();
    return DMR_ERR_ALLOC;
  }
  if ((e = hipMemcpy(g_ls.rowPtr, rp, sizeof(int) * (n + 1), hipMemcpyHostToDevice)) !=
          hipSuccess ||
      (nnz > 0 && (e = hipMemcpy(g_ls.colInd, ci, sizeof(int) * nnz, hipMemcpyHostToDevice)) !=
                      hipSuccess) ||
      (nnz > 0 && (e = hipMemcpy(g_ls.val, va, sizeof(double) * nnz, hipMemcpyHostToDevice)) !=
                      hipSuccess) ||
      (e = hipMemcpy(g_ls.blockRows, br.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice)) !=
          hipSuccess) {
    report("impCsrLS", e);
    release_csr();
    return DMR_ERR_HIP;
  }
  g_ls.csr.n = n;
  g_ls.csr.nnz = nnz;
  g_ls.csr.rowPtr = g_ls.rowPtr;
  g_ls.csr.colInd = g_ls.colInd;
  g_ls.csr.val = g_ls.val;
  g_ls.csr.blockRows = g_ls.blockRows;
  g_ls.csr.nBlocks = nb;
  return DMR_OK;
}

}  // namespace

extern "C" {

void initLS(int blocksMDR, int threadsPerBlockMDR, int devNumLO, int cells, int *devListMDR,
            int numDevsMDR, int basisSizeMDR, int deflatSizeMDR) {
  (void)blocksMDR;
  (void)threadsPerBlockMDR;
  if (g_ls.live) freeLS();
  if (!devListMDR || numDevsMDR != 1 || devListMDR[0] != devNumLO || cells < 1) {
    std::fprintf(stderr, "initLS: exactly one device, equal to the operator device, is supported\n");
    return;
  }
  hipError_t e;
  if ((e = hipSetDevice(devNumLO)) != hipSuccess) {
    report("initLS", e);
    return;
  }
  g_ls.dev = devNumLO;
  g_ls.n = cells;
  if ((e = hipMalloc(&g_ls.src, sizeof(double) * cells)) != hipSuccess ||
      (e = hipMalloc(&g_ls.img, sizeof(double) * cells)) != hipSuccess ||
      (e = hipMalloc(&g_ls.tmp, sizeof(double) * cells)) != hipSuccess ||
      (e = hipMemset(g_ls.src, 0, sizeof(double) * cells)) != hipSuccess ||
      (e = hipMemset(g_ls.img, 0, sizeof(double) * cells)) != hipSuccess ||
      (e = hipStreamCreate(&g_ls.stream)) != hipSuccess) {
    report("initLS", e);
    return;
  }
  // Identity operator until one is imported (the reference initialises the identity too,
  // testLS.cu:21-53, 187).
  std::vector<int> rp(cells + 1), ci(cells);
  std::vector<double> va(cells, 1.0);
  for (int i = 0; i <= cells; ++i) rp[i] = i;
  for (int i = 0; i < cells; ++i) ci[i] = i;
  g_ls.live = true;
  upload_csr(rp.data(), ci.data(), va.data(), cells);
  initDMRx(devNumLO, cells, basisSizeMDR, deflatSizeMDR, ls_apply, nullptr);
}

void freeLS(void) {
  if (!g_ls.live) return;
  hipSetDevice(g_ls.dev);
  if (g_ls.stream) hipStreamSynchronize(g_ls.stream);
  release_csr();
  if (g_ls.src) hipFree(g_ls.src);
  if (g_ls.img) hipFree(g_ls.img);
  if (g_ls.tmp) hipFree(g_ls.tmp);
  if (g_ls.stream) hipStreamDestroy(g_ls.stream);
  g_ls = LinearSystem{};
  freeDMR();
}

static void copy_in(double *dst, const double *host, const char *where) {
  if (!g_ls.live || !host) {
    std::fprintf(stderr, "%s: linear system not initialised\n", where);
    return;
  }
  hipSetDevice(g_ls.dev);
  hipStreamSynchronize(g_ls.stream);
  hipError_t e = hipMemcpy(dst, host, sizeof(double) * g_ls.n, hipMemcpyHostToDevice);
  if (e != hipSuccess) report(where, e);
}

static void copy_out(double *host, const double *src, const char *where) {
  if (!g_ls.live || !host) {
    std::fprintf(stderr, "%s: linear system not initialised\n", where);
    return;
  }
  hipSetDevice(g_ls.dev);
  hipStreamSynchronize(g_ls.stream);
  hipError_t e = hipMemcpy(host, src, sizeof(double) * g_ls.n, hipMemcpyDeviceToHost);
  if (e != hipSuccess) report(where, e);
}

void impImgLS(double *cImg) { copy_in(g_ls.img, cImg, "impImgLS"); }
void expImgLS(double *cImg) { copy_out(cImg, g_ls.img, "expImgLS"); }
void impSrcLS(double *cSrc) { copy_in(g_ls.src, cSrc, "impSrcLS"); }
void expSrcLS(double *cSrc) { copy_out(cSrc, g_ls.src, "expSrcLS"); }

void impOptLS(double *cOpt) {
  if (!g_ls.live || !cOpt) return;
  const int n = g_ls.n;
  // Column-major dense n x n (the reference's cublasCgemv(OP_N, lda = n) layout) to CSR.
  std::vector<int> rp(n + 1, 0);
  for (int c = 0; c < n; ++c)
    for (int r = 0; r < n; ++r)
      if (cOpt[(size_t)c * n + r] != 0.0) ++rp[r + 1];
  for (int r = 0; r < n; ++r) rp[r + 1] += rp[r];
  const int nnz = rp[n];
  std::vector<int> ci(std::max(nnz, 1));
  std::vector<double> va(std::max(nnz, 1));
  std::vector<int> fillp(rp.begin(), rp.end() - 1);
  for (int c = 0; c < n; ++c)
    for (int r = 0; r < n; ++r) {
      const double v = cOpt[(size_t)c * n + r];
      if (v != 0.0) {
        ci[fillp[r]] = c;
        va[fillp[r]++] = v;
      }
    }
  if (upload_csr(rp.data(), ci.data(), va.data(), nnz) != DMR_OK)
    std::fprintf(stderr, "impOptLS: invalid operator\n");
}

void expOptLS(double *cOpt) {
  if (!g_ls.live || !cOpt) return;
  const int n = g_ls.n, nnz = g_ls.csr.nnz;
  hipSetDevice(g_ls.dev);
  hipStreamSynchronize(g_ls.stream);
  std::vector<int> rp(n + 1), ci(std::max(nnz, 1));
  std::vector<double> va(std::max(nnz, 1));
  hipMemcpy(rp.data(), g_ls.rowPtr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost);
  if (nnz > 0) {
    hipMemcpy(ci.data(), g_ls.colInd, sizeof(int) * nnz, hipMemcpyDeviceToHost);
    hipMemcpy(va.data(), g_ls.val, sizeof(double) * nnz, hipMemcpyDeviceToHost);
  }
  std::memset(cOpt, 0, sizeof(double) * (size_t)n * n);
  for (int r = 0; r < n; ++r)
    for (int e = rp[r]; e < rp[r + 1]; ++e) cOpt[(size_t)ci[e] * n + r] = va[e];
}

int impCsrLS(const int *rowPtr, const int *colInd, const double *val, int nnz) {
  if (!g_ls.live) return DMR_ERR_STATE;
  return upload_csr(rowPtr, colInd, val, nnz);
}

int nnzLS(void) { return g_ls.live ? g_ls.csr.nnz : 0; }

int impMtxLS(const char *path) {
  if (!g_ls.live) return DMR_ERR_STATE;
  int rows = 0, cols = 0;
  long long nnz = 0;
  int rc = dmrReadMatrixMarket(path, &rows, &cols, &nnz, nullptr, nullptr, nullptr);
  if (rc != DMR_OK) return rc;
  if (rows != g_ls.n || cols != g_ls.n) return DMR_ERR_ARG;
  std::vector<int> rp(rows + 1), ci(std::max<long long>(nnz, 1));
  std::vector<double> va(std::max<long long>(nnz, 1));
  rc = dmrReadMatrixMarket(path, &rows, &cols, &nnz, rp.data(), ci.data(), va.data());
  if (rc != DMR_OK) return rc;
  return upload_csr(rp.data(), ci.data(), va.data(), (int)nnz);
}

double solveLS(int mode, double solTol, int matRank, int *numIts) {
  if (!g_ls.live) {
    std::fprintf(stderr, "solveLS: linear system not initialised\n");
    return 1.0;
  }
  return solverDMR(mode, g_ls.src, g_ls.img, g_ls.stream, g_ls.dev, solTol, matRank, numIts);
}

void linOptLS(void) {
  if (!g_ls.live) return;
  hipSetDevice(g_ls.dev);
  dmr::spmv(g_ls.csr, g_ls.src, g_ls.tmp, g_ls.stream);
  hipMemcpyAsync(g_ls.src, g_ls.tmp, sizeof(double) * g_ls.n, hipMemcpyDeviceToDevice,
                 g_ls.stream);
}

}  // extern "C"
