// C ABI of the block DMR kernels (include/gDMRKernels.h): shape checks on the host, then the same
// launch wrappers the block driver uses (dmr_block.h).
#include <hip/hip_runtime_api.h>

#include <string>

#include "dmr_block.h"
#include "gDMR.h"
#include "gDMRKernels.h"

namespace {

using dmr::bf16;
namespace blk = dmr::blk;

int bad(const char *what) { return dmr::set_error(DMR_ERR_ARG, what); }

int launched(const char *what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return dmr::set_error(DMR_ERR_HIP, (std::string(what) + ": " + hipGetErrorString(e)).c_str());
  return DMR_OK;
}

bool rows_ok(int rowsPerBlock) { return rowsPerBlock >= 64 && rowsPerBlock % 64 == 0; }

}  // namespace

extern "C" {

int dmrProbeReadBandwidth(const void *buf, size_t bytes, int reps, hipStream_t stream,
                          double *gbps) {
  dmr::clear_error();
  if (!buf || !gbps || bytes < ((size_t)64 << 20) || bytes % 16 != 0 || reps < 1)
    return bad("dmrProbeReadBandwidth: invalid arguments");
  unsigned *sink = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipMalloc(&sink, sizeof(unsigned)) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    if (sink) hipFree(sink);
    if (e0) hipEventDestroy(e0);
    return dmr::set_error(DMR_ERR_HIP, "dmrProbeReadBandwidth: setup failed");
  }
  const long long units = (long long)(bytes / 16);
  float best = 1e30f;
  int rc = DMR_OK;
  for (int r = 0; r < reps + 2 && rc == DMR_OK; ++r) {
    hipEventRecord(e0, stream);
    dmr::probe_read(buf, units, sink, stream);
    hipEventRecord(e1, stream);
    if (hipEventSynchronize(e1) != hipSuccess) {
      rc = dmr::set_error(DMR_ERR_HIP, "dmrProbeReadBandwidth: launch failed");
      break;
    }
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    if (r >= 2 && ms > 0.f && ms < best) best = ms;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  hipFree(sink);
  if (rc == DMR_OK) *gbps = (double)bytes / (best * 1e-3) / 1e9;
  return rc;
}

size_t dmrKernelTnWorkspace(long long n, int p, int rowsPerBlock) {
  if (n < 1 || p < 1 || !rows_ok(rowsPerBlock)) return 0;
  const int G = blk::gemm_tn_blocks(n, blk::kMaxTnBlocks, rowsPerBlock);
  return (size_t)G * (size_t)((p + 31) & ~31) * 32;
}

int dmrKernelGemmTn(int basis, const void *P, long long ldp, int p, const float *U,
                    long long ldu, int s, long long n, int rowsPerBlock, int variant, float *part,
                    double *C64, int ldc, float *C32, hipStream_t stream) {
  dmr::clear_error();
  if (!P || !U || !part || (!C64 && !C32) || n < 1 || p < 1 || p > blk::kMaxCols || s < 1 ||
      s > blk::kMaxRhs || ldp < n || ldu < n || (C64 && ldc < p) || !rows_ok(rowsPerBlock))
    return bad("dmrKernelGemmTn: invalid shape");
  const int G = blk::gemm_tn_blocks(n, blk::kMaxTnBlocks, rowsPerBlock);
  if (basis == DMR_KBASIS_FP32) {
    if (variant < 0 || variant > 2) return bad("dmrKernelGemmTn: FP32 P takes variants 0-2");
    blk::gemm_tn(static_cast<const float *>(P), ldp, p, U, ldu, s, n, G, variant, part, C64, ldc,
                 C32, stream);
  } else if (basis == DMR_KBASIS_BF16) {
    if (variant != 2 && variant != 3) return bad("dmrKernelGemmTn: bf16 P takes variants 2, 3");
    blk::gemm_tn(static_cast<const bf16 *>(P), ldp, p, U, ldu, s, n, G, variant, part, C64, ldc,
                 C32, stream);
  } else {
    return bad("dmrKernelGemmTn: unknown basis storage");
  }
  return launched("dmrKernelGemmTn");
}

int dmrKernelGemmTnPair(const uint16_t *P, long long ldp, int p, const float *U, const float *U2,
                        long long ldu, int s, long long n, int rowsPerBlock, float *part,
                        double *C64a, int ldca, float *C32, double *C64b, int ldcb,
                        hipStream_t stream) {
  dmr::clear_error();
  if (!P || !U || !U2 || !part || !C64a || !C64b || n < 1 || p < 1 || p > blk::kMaxCols ||
      s < 1 || s > 16 || ldp < n || ldu < n || ldca < p || ldcb < p || !rows_ok(rowsPerBlock))
    return bad("dmrKernelGemmTnPair: invalid shape");
  const int G = blk::gemm_tn_blocks(n, blk::kMaxTnBlocks, rowsPerBlock);
  blk::gemm_tn_pair(reinterpret_cast<const bf16 *>(P), ldp, p, U, U2, ldu, s, n, G, part, C64a,
                    ldca, C32, C64b, ldcb, stream);
  return launched("dmrKernelGemmTnPair");
}

int dmrKernelGemmNn(int basis, const void *P, long long ldp, int p, const float *C,
                    const float *In, long long ldi, float a, float b, float *Out, long long ldo,
                    int s, long long n, int unroll, float *gram, hipStream_t stream) {
  dmr::clear_error();
  if (!P || !C || !Out || (a != 0.0f && !In) || n < 1 || p < 1 || p > blk::kMaxCols || s < 1 ||
      s > blk::kMaxRhs || ldp < n || ldo < n || (a != 0.0f && ldi < n))
    return bad("dmrKernelGemmNn: invalid shape");
  const bool core = unroll == 100 || unroll == 101;
  if (!core && unroll != 2 && unroll != 4 && unroll != 8 && unroll != 16)
    return bad("dmrKernelGemmNn: unroll is 2, 4, 8, 16, 100 or 101");
  if (gram && !(basis == DMR_KBASIS_BF16 && core && s <= 16))
    return bad("dmrKernelGemmNn: Gram partials need a bf16 basis, unroll 100/101 and s <= 16");
  if (basis == DMR_KBASIS_FP32) {
    if (core) return bad("dmrKernelGemmNn: unroll 100/101 needs a bf16 basis");
    blk::gemm_nn(static_cast<const float *>(P), ldp, p, C, In, ldi, a, b, Out, ldo, s, n, unroll,
                 stream);
  } else if (basis == DMR_KBASIS_BF16) {
    blk::gemm_nn(static_cast<const bf16 *>(P), ldp, p, C, In, ldi, a, b, Out, ldo, s, n, unroll,
                 stream, gram);
  } else {
    return bad("dmrKernelGemmNn: unknown basis storage");
  }
  return launched("dmrKernelGemmNn");
}

int dmrKernelGramBlocks(long long n, int unroll) {
  if (n < 1 || (unroll != 100 && unroll != 101)) return 0;
  return blk::gemm_nn_gram_blocks(n, unroll);
}

int dmrKernelGramReduce(const float *part, int G, int s, double *G64, hipStream_t stream) {
  dmr::clear_error();
  if (!part || !G64 || G < 1 || s < 1 || s > 16) return bad("dmrKernelGramReduce: invalid shape");
  blk::gram_reduce(part, G, s, G64, stream);
  return launched("dmrKernelGramReduce");
}

int dmrKernelGemmNnPair(const uint16_t *P, long long ldp, int p, const float *C, float *X1,
                        float *X2, long long ld, int s, long long n, int unroll,
                        hipStream_t stream) {
  dmr::clear_error();
  if (!P || !C || !X1 || !X2 || n < 1 || p < 1 || p > blk::kMaxCols || s < 1 || s > 16 ||
      ldp < n || ld < n || (unroll != 100 && unroll != 101))
    return bad("dmrKernelGemmNnPair: invalid shape");
  blk::gemm_nn_pair(reinterpret_cast<const bf16 *>(P), ldp, p, C, X1, X2, ld, s, n, unroll,
                    stream);
  return launched("dmrKernelGemmNnPair");
}

int dmrKernelSpmm(int xBasis, int n, int nnz, const int *rowPtr, const int *colInd,
                  const float *val32, const void *X, long long ldx, float *Y, long long ldy,
                  int s, int variant, hipStream_t stream) {
  dmr::clear_error();
  if (!rowPtr || !colInd || !val32 || !X || !Y || n < 1 || nnz < 0 || s < 1 ||
      s > blk::kMaxRhs || ldx < n || ldy < n)
    return bad("dmrKernelSpmm: invalid shape");
  const int cg = variant % 100, kind = variant / 100;
  if ((cg != 4 && cg != 8 && cg != 16) || kind > 2 || (kind == 2 && cg == 4))
    return bad("dmrKernelSpmm: variant is 4, 8, 16, 104, 108, 116, 208 or 216");
  dmr::CsrDev A;
  A.n = n;
  A.nnz = nnz;
  A.rowPtr = rowPtr;
  A.colInd = colInd;
  A.val32 = val32;
  if (xBasis == DMR_KBASIS_FP32) {
    blk::spmm32(A, static_cast<const float *>(X), ldx, Y, ldy, s, variant, stream);
  } else if (xBasis == DMR_KBASIS_BF16) {
    if (kind == 0) return bad("dmrKernelSpmm: bf16 X takes variants 104-216");
    blk::spmm32(A, static_cast<const bf16 *>(X), ldx, Y, ldy, s, variant, stream);
  } else {
    return bad("dmrKernelSpmm: unknown vector storage");
  }
  return launched("dmrKernelSpmm");
}

int dmrKernelCholQr(const double *G, int s, double relTol, float *Sinv32, double *Sacc, int first,
                    double *Hdst, int ldh, int *flags, hipStream_t stream) {
  dmr::clear_error();
  if (!G || !Sinv32 || !Sacc || !flags || s < 1 || s > 32 || !(relTol >= 0.0) ||
      (Hdst && ldh < s))
    return bad("dmrKernelCholQr: invalid shape");
  blk::chol_qr(G, s, relTol, Sinv32, Sacc, first, Hdst, ldh, flags, stream);
  return launched("dmrKernelCholQr");
}

int dmrKernelGramChol(const float *part, int G, int s, double relTol, float *Sinv32, double *Sacc,
                      double *Hdst, int ldh, int *flags, void *work, hipStream_t stream) {
  dmr::clear_error();
  static_assert(DMR_KGRAMCHOL_WORK_BYTES == 256 + blk::kGramCholSegs * 256 * sizeof(double),
                "work layout");
  if (!part || !Sinv32 || !Sacc || !flags || !work || G < 1 || s < 1 || s > 16 ||
      !(relTol >= 0.0) || (Hdst && ldh < s))
    return bad("dmrKernelGramChol: invalid shape");
  char *w = static_cast<char *>(work);
  blk::gram_chol(part, G, s, relTol, Sinv32, Sacc, 1, Hdst, ldh, flags,
                 reinterpret_cast<double *>(w + 256), reinterpret_cast<unsigned *>(w), stream);
  return launched("dmrKernelGramChol");
}

int dmrKernelScaleBlock(const float *U, long long ldu, const float *C, void *Out, long long ldo,
                        int outBasis, int s, long long n, hipStream_t stream) {
  dmr::clear_error();
  if (!U || !C || !Out || n < 1 || s < 1 || s > 16 || ldu < n || ldo < n || (ldu & 1) ||
      (ldo & 1))
    return bad("dmrKernelScaleBlock: invalid shape");
  bool ok = false;
  if (outBasis == DMR_KBASIS_FP32)
    ok = blk::scale_block(U, ldu, C, static_cast<float *>(Out), ldo, s, n, stream);
  else if (outBasis == DMR_KBASIS_BF16)
    ok = blk::scale_block(U, ldu, C, static_cast<bf16 *>(Out), ldo, s, n, stream);
  if (!ok) return bad("dmrKernelScaleBlock: unknown output storage");
  return launched("dmrKernelScaleBlock");
}

int dmrKernelLsqFits(int cols, int bw, int s) {
  if (cols < 1 || bw < 1 || bw > 32 || s < 1 || s > 32) return 0;
  return blk::lsq_block_lds(cols, bw, s) + 1024 <= 160 * 1024 ? 1 : 0;
}

int dmrKernelLsqBlock(const double *H, int rows, int cols, int bw, const double *T, int s,
                      const double *rho, double *R, double *Qt, double *Y, float *Yr32,
                      float *ZH32, hipStream_t stream) {
  dmr::clear_error();
  if (!H || !T || !rho || !R || !Qt || !Y || !Yr32 || !ZH32 || rows != cols + bw ||
      !dmrKernelLsqFits(cols, bw, s))
    return bad("dmrKernelLsqBlock: invalid shape");
  blk::lsq_block(H, rows, cols, bw, T, s, rho, R, Qt, Y, Yr32, ZH32, stream);
  return launched("dmrKernelLsqBlock");
}

}  // extern "C"
