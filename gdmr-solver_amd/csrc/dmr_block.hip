// Kernels of the block (multi-right-hand-side) DMR: s right-hand sides share one block Krylov
// space and one deflation subspace.  Blocks of vectors are column-major (element (r, c) at
// c * ld + r) with the padded leading dimension of the single-vector basis.
//
// * spmm / resid_block: CSR times a block of vectors, one thread per row, the row's nonzeros
//   reused across the block's columns.
// * gemm_tn: C = P^T U (p x s, reduction over the n rows), the block Gram-Schmidt projection and
//   the Gram matrix of the Cholesky QR.  Row ranges per workgroup; each 32-row chunk of P and U is
//   staged through LDS (transposed for the MFMA operand layout) and contracted with
//   v_mfma_f32_32x32x2f32; per-workgroup partials are summed in a fixed order by gemm_tn_reduce
//   (deterministic, FP64 accumulation of the partials).
// * gemm_nn: Out = a In + b P C (n x s), the block update, the basis normalisation U S^-1, the
//   residual and solution updates and the subspace refresh.  A wave owns 128 rows as four
//   interleaved 32-row MFMA tiles, so every P load is a 16-byte vector per lane and a
//   half-wavefront reads 512 contiguous bytes of one column.
// * chol_qr: Cholesky factor of the s x s Gram matrix and its inverse (one wavefront, FP64).
// * lsq_block: least squares on the block Hessenberg (lower bandwidth s) by Householder QR with
//   reflectors of length <= s + 1, back substitution, Z_H = H Y (one workgroup, FP64).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>

#include "dmr_block.h"
#include "dmr_lanes.h"

namespace dmr {
namespace blk {
namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kThreads = 256;
constexpr int kRC = 32;        // rows per LDS chunk of gemm_tn
constexpr int kLdsStride = kRC + 1;

__device__ inline f32x16 mfma32(float a, float b, f32x16 c) {
  return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
}

// Row i of the 32 x 32 MFMA result held in accumulator v by lane l (column l % 32).
__device__ inline int mfma_row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }

// Per-workgroup partial products P^T U are stored row by row: element (row, col) of workgroup g
// of a G-workgroup launch at (row * G + g) * 32 + col, so the G partials of one row are one
// contiguous run that the fixed-order reduction streams with 16-byte loads.
__device__ inline long long part_index(int row, int col) {
  return ((long long)row * gridDim.x + blockIdx.x) * 32 + col;
}

// bf16 storage: the upper 16 bits of a binary32, rounded to nearest even on store.
__device__ inline float bf_to_f(uint32_t bits16) { return __uint_as_float(bits16 << 16); }
__device__ inline uint32_t f_to_bf(float f) {  // finite inputs
  const uint32_t x = __float_as_uint(f);
  return (x + 0x7fffu + ((x >> 16) & 1u)) >> 16;
}
__device__ inline float ldf(const float *p, long long i) { return p[i]; }
__device__ inline float ldf(const bf16 *p, long long i) { return bf_to_f(p[i].bits); }
// Unsigned 32-bit offsets from a wavefront-uniform base: one global load with a scalar base and a
// 32-bit vector offset, no 64-bit address arithmetic per element.  The byte offset must fit 32
// bits: i < 2^30 for FP32 (a block basis of that many rows would not fit in device memory).
__device__ inline float ldf(const float *p, unsigned i) {
  return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(p) + (i << 2));
}
__device__ inline float ldf(const bf16 *p, unsigned i) {
  return bf_to_f(
      *reinterpret_cast<const uint16_t *>(reinterpret_cast<const char *>(p) + (i << 1)));
}

// Four consecutive outputs (row .. row + 3 of one column; full when all four are below n).
__device__ inline void store_rows4(float *dst, const float (&w)[4], int cnt) {
  if (cnt == 4) {
    *reinterpret_cast<float4 *>(dst) = make_float4(w[0], w[1], w[2], w[3]);
  } else {
    for (int q = 0; q < cnt; ++q) dst[q] = w[q];
  }
}
__device__ inline void store_rows4(bf16 *dst, const float (&w)[4], int cnt) {
  if (cnt == 4) {
    uint2 t;
    t.x = f_to_bf(w[0]) | (f_to_bf(w[1]) << 16);
    t.y = f_to_bf(w[2]) | (f_to_bf(w[3]) << 16);
    *reinterpret_cast<uint2 *>(dst) = t;
  } else {
    for (int q = 0; q < cnt; ++q) dst[q].bits = (uint16_t)f_to_bf(w[q]);
  }
}

// ---------------------------------------------------------------------------------------------
// Thread (row, column group): blockIdx.y selects a group of CG columns, so one launch has
// ceil(s / CG) threads per row and each thread reuses its row's nonzeros across CG columns.
template <typename TV, typename TX, typename TY, typename TA, int CG>
__global__ __launch_bounds__(kThreads) void k_spmm(int n, const int *__restrict__ rowPtr,
                                                   const int *__restrict__ colInd,
                                                   const TV *__restrict__ val,
                                                   const TX *__restrict__ X, long long ldx,
                                                   TY *__restrict__ Y, long long ldy, int s) {
  const int row = blockIdx.x * kThreads + threadIdx.x;
  const int c0 = blockIdx.y * CG;
  if (row >= n) return;
  const int b = rowPtr[row], e = rowPtr[row + 1];
  TA acc[CG];
#pragma unroll
  for (int c = 0; c < CG; ++c) acc[c] = (TA)0;
  for (int q = b; q < e; ++q) {
    const long long col = colInd[q];
    const TA v = (TA)val[q];
#pragma unroll
    for (int c = 0; c < CG; ++c)
      if (c0 + c < s) acc[c] = fma(v, (TA)X[col + (c0 + c) * ldx], acc[c]);
  }
#pragma unroll
  for (int c = 0; c < CG; ++c)
    if (c0 + c < s) Y[row + (c0 + c) * ldy] = (TY)acc[c];
}

// Same product with every row's first W column indices and values loaded up front (one memory
// latency), then all W x CG gathers of the row issued independently; rows with more than W entries
// take the loop of k_spmm.  Each column's sum runs over the row's entries in order, as in k_spmm.
template <typename TV, typename TX, typename TY, typename TA, int CG, int W>
__global__ __launch_bounds__(kThreads) void k_spmm_w(int n, const int *__restrict__ rowPtr,
                                                     const int *__restrict__ colInd,
                                                     const TV *__restrict__ val,
                                                     const TX *__restrict__ X, long long ldx,
                                                     TY *__restrict__ Y, long long ldy, int s) {
  const int row = blockIdx.x * kThreads + threadIdx.x;
  const int c0 = blockIdx.y * CG;
  if (row >= n) return;
  const int b = rowPtr[row], e = rowPtr[row + 1];
  TA acc[CG];
#pragma unroll
  for (int c = 0; c < CG; ++c) acc[c] = (TA)0;
  if (e - b <= W) {
    unsigned col[W];
    TA v[W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
      const bool in = b + q < e;
      col[q] = (unsigned)(in ? colInd[b + q] : row);
      v[q] = in ? (TA)val[b + q] : (TA)0;
    }
#pragma unroll
    for (int c = 0; c < CG; ++c) {
      if (c0 + c < s) {
        const TX *xc = X + (long long)(c0 + c) * ldx;
        TA xv[W];
#pragma unroll
        for (int q = 0; q < W; ++q) xv[q] = (TA)ldf(xc, col[q]);
#pragma unroll
        for (int q = 0; q < W; ++q)
          if (b + q < e) acc[c] = fma(v[q], xv[q], acc[c]);
      }
    }
  } else {
    for (int q = b; q < e; ++q) {
      const long long col = colInd[q];
      const TA v = (TA)val[q];
#pragma unroll
      for (int c = 0; c < CG; ++c)
        if (c0 + c < s) acc[c] = fma(v, (TA)ldf(X, col + (c0 + c) * ldx), acc[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < CG; ++c)
    if (c0 + c < s) Y[row + (c0 + c) * ldy] = (TY)acc[c];
}

// Same product with the workgroup's nonzeros (its rows' index range of colInd and val, when it
// holds at most W entries per row on average) staged in LDS by coalesced loads first; each row of
// up to W entries then reads its indices from LDS and issues its W x CG gathers at once.  Rows
// with more entries, and workgroups whose range does not fit, read colInd and val directly.  The
// sums run over each row's entries in order, as in k_spmm.  Y is
// written with the non-temporal hint: this kernel is bound by its gathers, so the write-back of Y
// happens during it instead of during the bandwidth-bound projection that reads Y next (which went
// from 4.26 to 4.67 TB/s in the block cycle).
template <typename TV, typename TX, typename TY, typename TA, int CG, int W>
__global__ __launch_bounds__(kThreads) void k_spmm_lds(int n, const int *__restrict__ rowPtr,
                                                       const int *__restrict__ colInd,
                                                       const TV *__restrict__ val,
                                                       const TX *__restrict__ X, long long ldx,
                                                       TY *__restrict__ Y, long long ldy, int s) {
  __shared__ int sCol[kThreads * W];
  __shared__ TV sVal[kThreads * W];
  const int r0 = blockIdx.x * kThreads, row = r0 + threadIdx.x;
  const int c0 = blockIdx.y * CG;
  const int gb = rowPtr[r0], ge = rowPtr[r0 + kThreads < n ? r0 + kThreads : n];
  const bool staged = ge - gb <= kThreads * W;  // uniform over the workgroup
  if (staged) {
    for (int i = threadIdx.x; i < ge - gb; i += kThreads) {
      sCol[i] = colInd[gb + i];
      sVal[i] = val[gb + i];
    }
  }
  __syncthreads();
  if (row >= n) return;
  const int b = rowPtr[row], e = rowPtr[row + 1];
  TA acc[CG];
#pragma unroll
  for (int c = 0; c < CG; ++c) acc[c] = (TA)0;
  if (staged && e - b <= W) {
    unsigned col[W];
    TA v[W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
      const bool in = b + q < e;
      col[q] = (unsigned)(in ? sCol[b - gb + q] : row);
      v[q] = in ? (TA)sVal[b - gb + q] : (TA)0;
    }
#pragma unroll
    for (int c = 0; c < CG; ++c) {
      if (c0 + c < s) {
        const TX *xc = X + (long long)(c0 + c) * ldx;
        TA xv[W];
#pragma unroll
        for (int q = 0; q < W; ++q) xv[q] = (TA)ldf(xc, col[q]);
#pragma unroll
        for (int q = 0; q < W; ++q)
          if (b + q < e) acc[c] = fma(v[q], xv[q], acc[c]);
      }
    }
  } else {
    for (int q = b; q < e; ++q) {
      const long long col = colInd[q];
      const TA v = (TA)val[q];
#pragma unroll
      for (int c = 0; c < CG; ++c)
        if (c0 + c < s) acc[c] = fma(v, (TA)ldf(X, col + (c0 + c) * ldx), acc[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < CG; ++c)
    if (c0 + c < s) __builtin_nontemporal_store((TY)acc[c], Y + row + (long long)(c0 + c) * ldy);
}

// R = B - A X in FP64.
// Column c uses the values val + c * ldv (ldv = 0: one operator for every column).  blockIdx.y
// selects a group of CG columns.  A row of up to W entries loads its column indices once, then
// issues every value load and gather of the group at once; longer rows loop.  Each column's sum
// runs over the row's entries in order from zero, then R = B - sum.
template <int CG, int W>
__global__ __launch_bounds__(kThreads) void k_resid_block(int n, const int *__restrict__ rowPtr,
                                                          const int *__restrict__ colInd,
                                                          const double *__restrict__ val,
                                                          long long ldv,
                                                          const double *__restrict__ X,
                                                          long long ldx,
                                                          const double *__restrict__ B,
                                                          long long ldb, double *__restrict__ R,
                                                          long long ldr, int s) {
  const int row = blockIdx.x * kThreads + threadIdx.x;
  if (row >= n) return;
  const int b = rowPtr[row], e = rowPtr[row + 1];
  const int c0 = blockIdx.y * CG;
  double acc[CG];
#pragma unroll
  for (int c = 0; c < CG; ++c) acc[c] = 0.0;
  if (e - b <= W) {
    int col[W];
#pragma unroll
    for (int q = 0; q < W; ++q) col[q] = b + q < e ? colInd[b + q] : row;
    double v[CG][W], x[CG][W];
#pragma unroll
    for (int c = 0; c < CG; ++c)
#pragma unroll
      for (int q = 0; q < W; ++q) {
        const bool in = b + q < e && c0 + c < s;
        v[c][q] = in ? val[b + q + (c0 + c) * ldv] : 0.0;
        x[c][q] = in ? X[col[q] + (c0 + c) * ldx] : 0.0;
      }
#pragma unroll
    for (int c = 0; c < CG; ++c)
#pragma unroll
      for (int q = 0; q < W; ++q)
        if (b + q < e) acc[c] = fma(v[c][q], x[c][q], acc[c]);
  } else {
    for (int q = b; q < e; ++q) {
      const long long col = colInd[q];
#pragma unroll
      for (int c = 0; c < CG; ++c)
        if (c0 + c < s) acc[c] = fma(val[q + (c0 + c) * ldv], X[col + (c0 + c) * ldx], acc[c]);
    }
  }
#pragma unroll
  for (int c = 0; c < CG; ++c)
    if (c0 + c < s) R[row + (c0 + c) * ldr] = B[row + (c0 + c) * ldb] - acc[c];
}

// out[q] = (float)(sum_c val[q + c * ldv] / count), the sum in column order in FP64: the entrywise
// mean of count operators with one sparsity pattern.
__global__ __launch_bounds__(kThreads) void k_mean_values(const double *__restrict__ val,
                                                          long long ldv, int count, long long nnz,
                                                          float *__restrict__ out) {
  for (long long q = (long long)blockIdx.x * kThreads + threadIdx.x; q < nnz;
       q += (long long)gridDim.x * kThreads) {
    double acc = 0.0;
    for (int c = 0; c < count; ++c) acc += val[q + c * ldv];
    out[q] = (float)(acc / count);
  }
}

// ---------------------------------------------------------------------------------------------
// Column sums of squares, fixed order: block g sums rows [g*rows, (g+1)*rows) of every column
// (each thread in row order, then the block's threads in a fixed tree), partial[g * s + c].
// Block (g, c) = (blockIdx.x, blockIdx.y) sums rows [g*rows, (g+1)*rows) of column c: each thread
// in row order, then a wavefront xor butterfly, then the four wavefronts in order.
template <typename T>
__global__ __launch_bounds__(kThreads) void k_colsumsq(const T *__restrict__ X, long long ld,
                                                       int s, long long n, long long rows,
                                                       double *__restrict__ partial) {
  __shared__ double red[kThreads / 64];
  const int c = blockIdx.y;
  const long long r0 = (long long)blockIdx.x * rows;
  const long long r1 = r0 + rows < n ? r0 + rows : n;
  double acc = 0.0;
  const T *col = X + c * ld;
  for (long long r = r0 + threadIdx.x; r < r1; r += kThreads) {
    const double x = (double)col[r];
    acc = fma(x, x, acc);
  }
  acc = xor_sum<64>(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    partial[(long long)blockIdx.x * s + c] = ((red[0] + red[1]) + red[2]) + red[3];
}

// out[c] = sum_g partial[g * s + c]: thread (c, q) = (t % 32, t / 32) sums g = q, q + 8, ... in
// order, then the eight sums in order of q.  mode 0: out = ss; mode 1: out = sqrt(ss).
__global__ __launch_bounds__(kThreads) void k_colsumsq_final(const double *__restrict__ partial,
                                                             int G, int s,
                                                             double *__restrict__ out, int mode) {
  __shared__ double red[kThreads];
  const int t = threadIdx.x, c = t & 31, q = t >> 5;
  double a = 0.0;
  if (c < s)
    for (int g = q; g < G; g += 8) a = a + partial[(long long)g * s + c];
  red[t] = a;
  __syncthreads();
  if (q != 0 || c >= s) return;
  double sum = 0.0;
#pragma unroll
  for (int w = 0; w < 8; ++w) sum = sum + red[w * 32 + c];
  out[c] = mode ? sqrt(sum) : sum;
}

// ---------------------------------------------------------------------------------------------
// Eight 16-byte loads of rows [rr, rr + 32) of one column (zeros past rEnd or for a null column).
__device__ inline void load_seg(const float *col, long long rr, long long rEnd, float4 *out) {
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const long long r = rr + 4 * q;
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (col) {
      if (r + 3 < rEnd) {
        v = *reinterpret_cast<const float4 *>(col + r);
      } else {
        if (r < rEnd) v.x = col[r];
        if (r + 1 < rEnd) v.y = col[r + 1];
        if (r + 2 < rEnd) v.z = col[r + 2];
      }
    }
    out[q] = v;
  }
}

// Partial C_g = P[r in range g, 0:p]^T U[r in range g, 0:s], as pPad x 32 FP32 per workgroup.
// Register-direct: no LDS staging.  The pPad / 32 tiles form nTG = ceil(tiles / kTPW) groups of
// at most kTPW tiles; the four wavefronts are nTG tile groups x nRG = 4 / nTG row groups.  A
// wavefront walks its row group's 64-row rounds; in a round lane (i, h) streams rows
// [r0 + 32h, r0 + 32h + 32) of column i of its tiles and of U as eight 16-byte loads, and MFMA
// k-step 4q + e takes row r0 + 32h + 4q + e of both operands (any row order is a valid reduction
// order).  The row groups' sums are added through LDS in a fixed order.
constexpr int kTPW = 3;
__global__ __launch_bounds__(kThreads) void k_gemm_tn(const float *__restrict__ P, long long ldp,
                                                      int p, int pPad,
                                                      const float *__restrict__ U, long long ldu,
                                                      int s, long long n, long long rowsPerBlock,
                                                      float *__restrict__ part) {
  __shared__ float red[4 * kTPW * 1024];
  const int tiles = pPad / 32;
  const int nTG = (tiles + kTPW - 1) / kTPW;
  const int nRG = 4 / nTG;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tg = wave % nTG, rg = wave / nTG;
  const int i = lane & 31, h = lane >> 5;
  const long long rBeg = (long long)blockIdx.x * rowsPerBlock;
  const long long rEnd = rBeg + rowsPerBlock < n ? rBeg + rowsPerBlock : n;
  f32x16 acc[kTPW];
#pragma unroll
  for (int t = 0; t < kTPW; ++t)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[t][v] = 0.0f;
  const float *pc[kTPW];
#pragma unroll
  for (int t = 0; t < kTPW; ++t) {
    const int tile = tg * kTPW + t, c = tile * 32 + i;
    pc[t] = (tile < tiles && c < p) ? P + c * ldp : nullptr;
  }
  const float *uc = i < s ? U + i * ldu : nullptr;
  if (rg < nRG) {
    for (long long r0 = rBeg + 64 * rg; r0 < rEnd; r0 += 64 * nRG) {
      const long long rr = r0 + 32 * h;
      float4 ub[8], pb[kTPW][8];
      load_seg(uc, rr, rEnd, ub);
#pragma unroll
      for (int t = 0; t < kTPW; ++t)
        if (tg * kTPW + t < tiles) load_seg(pc[t], rr, rEnd, pb[t]);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float bq[4] = {ub[q].x, ub[q].y, ub[q].z, ub[q].w};
#pragma unroll
        for (int t = 0; t < kTPW; ++t) {
          if (tg * kTPW + t < tiles) {
            const float aq[4] = {pb[t][q].x, pb[t][q].y, pb[t][q].z, pb[t][q].w};
#pragma unroll
            for (int e = 0; e < 4; ++e) acc[t] = mfma32(aq[e], bq[e], acc[t]);
          }
        }
      }
    }
#pragma unroll
    for (int t = 0; t < kTPW; ++t)
#pragma unroll
      for (int v = 0; v < 16; ++v) red[(wave * kTPW + t) * 1024 + v * 64 + lane] = acc[t][v];
  }
  __syncthreads();
  if (rg != 0) return;
#pragma unroll
  for (int t = 0; t < kTPW; ++t) {
    const int tile = tg * kTPW + t;
    if (tile < tiles) {
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        float a = 0.0f;
        for (int g = 0; g < nRG; ++g) a = a + red[((g * nTG + tg) * kTPW + t) * 1024 + v * 64 + lane];
        part[part_index(tile * 32 + mfma_row(v, lane), i)] = a;
      }
    }
  }
}

// 16 x 16 x 4 MFMA variant (no half-empty tiles for s <= 16).  P's columns form nT = ceil(p/16)
// tiles of 16 and U's columns NU = ceil(s/16) tiles.  The four wavefronts are nTG tile groups (of
// at most TPW tiles) x nRG = 4 / nTG row groups.  A row group walks 32-row rounds; in a round lane
// (m, kk) = (l % 16, l / 16) loads rows r0 + 8kk .. r0 + 8kk + 7 of column m of each of its tiles
// and of U as two 16-byte loads, and MFMA k-step e takes row r0 + 8kk + e of both operands.  The
// next round's loads are issued before the current round's MFMAs.  The row groups' sums are
// added through LDS in a fixed order.  Output layout of v_mfma_f32_16x16x4f32: lane l, register v
// holds D[4 (l / 16) + v][l % 16].
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline f32x4 mfma16(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Eight rows of one column as loaded (raw storage, unpacked only where consumed, so the loads of
// a prefetch ring stay in flight): two float4 for FP32, one 16-byte word for bf16.
template <typename TP> struct Raw8;
template <> struct Raw8<float> { float4 a, b; };
template <> struct Raw8<bf16> { uint4 w; };

__device__ inline void load_raw8(const float *col, long long r, long long rEnd, Raw8<float> &o) {
  o.a = make_float4(0.f, 0.f, 0.f, 0.f);
  o.b = o.a;
  if (!col) return;
  if (r + 7 < rEnd) {
    o.a = *reinterpret_cast<const float4 *>(col + r);
    o.b = *reinterpret_cast<const float4 *>(col + r + 4);
  } else {
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = r + e < rEnd ? col[r + e] : 0.0f;
    o.a = make_float4(v[0], v[1], v[2], v[3]);
    o.b = make_float4(v[4], v[5], v[6], v[7]);
  }
}
// NT: the 16-byte load carries the non-temporal hint.  The block update reads its basis that way
// (each column once per launch), which keeps the operator and the small blocks that the SpMM, the
// projection and the normalisation reuse in the memory-side cache: the update went from 4.96 to
// 5.50 TB/s and the whole block cycle took 4% less time.  The same hint on the projection's basis
// loads made it slower (4.30 -> 4.14 TB/s), so those stay ordinary loads.
typedef unsigned u32x4v __attribute__((ext_vector_type(4)));
template <bool NT>
__device__ inline void load_raw8b(const bf16 *col, long long r, long long rEnd, Raw8<bf16> &o) {
  o.w = make_uint4(0u, 0u, 0u, 0u);
  if (!col) return;
  if (r + 7 < rEnd) {
    if (NT) {
      const u32x4v v = __builtin_nontemporal_load(reinterpret_cast<const u32x4v *>(col + r));
      o.w = make_uint4(v.x, v.y, v.z, v.w);
    } else {
      o.w = *reinterpret_cast<const uint4 *>(col + r);
    }
  } else {
    uint32_t h[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) h[e] = r + e < rEnd ? (uint32_t)col[r + e].bits : 0u;
    o.w = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16),
                     h[6] | (h[7] << 16));
  }
}
__device__ inline void load_raw8(const bf16 *col, long long r, long long rEnd, Raw8<bf16> &o) {
  load_raw8b<false>(col, r, rEnd, o);
}
__device__ inline void unpack8(const Raw8<float> &o, float (&v)[8]) {
  v[0] = o.a.x; v[1] = o.a.y; v[2] = o.a.z; v[3] = o.a.w;
  v[4] = o.b.x; v[5] = o.b.y; v[6] = o.b.z; v[7] = o.b.w;
}
__device__ inline void unpack8(const Raw8<bf16> &o, float (&v)[8]) {
  const uint32_t w[4] = {o.w.x, o.w.y, o.w.z, o.w.w};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    v[2 * i] = __uint_as_float(w[i] << 16);
    v[2 * i + 1] = __uint_as_float(w[i] & 0xffff0000u);
  }
}

template <typename TP, int TPW, int NU>
__global__ __launch_bounds__(kThreads) void k_gemm_tn16(const TP *__restrict__ P, long long ldp,
                                                        int p, int pPad,
                                                        const float *__restrict__ U,
                                                        long long ldu, int s, long long n,
                                                        long long rowsPerBlock,
                                                        float *__restrict__ part) {
  __shared__ float red[4 * TPW * NU * 256];
  const int nT = (p + 15) / 16;
  const int nTG = (nT + TPW - 1) / TPW;
  const int nRG = 4 / nTG;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tg = wave % nTG, rg = wave / nTG;
  const int m = lane & 15, kk = lane >> 4;
  const long long rBeg = (long long)blockIdx.x * rowsPerBlock;
  const long long rEnd = rBeg + rowsPerBlock < n ? rBeg + rowsPerBlock : n;
  f32x4 acc[TPW][NU];
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int u = 0; u < NU; ++u) acc[t][u] = f32x4{0.f, 0.f, 0.f, 0.f};
  const TP *pc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tile = tg * TPW + t, c = tile * 16 + m;
    pc[t] = (tile < nT && c < p) ? P + c * ldp : nullptr;
  }
  const float *uc[NU];
#pragma unroll
  for (int u = 0; u < NU; ++u) uc[u] = u * 16 + m < s ? U + (u * 16 + m) * ldu : nullptr;
  if (rg < nRG) {
    const long long step = 32LL * nRG;
    long long r0 = rBeg + 32 * rg;
    Raw8<TP> pr[TPW];
    Raw8<float> ur[NU];
#pragma unroll
    for (int t = 0; t < TPW; ++t) load_raw8(pc[t], r0 + 8 * kk, rEnd, pr[t]);
#pragma unroll
    for (int u = 0; u < NU; ++u) load_raw8(uc[u], r0 + 8 * kk, rEnd, ur[u]);
    while (r0 < rEnd) {
      const long long rn = r0 + step;
      Raw8<TP> qr[TPW];
      Raw8<float> vr[NU];
#pragma unroll
      for (int t = 0; t < TPW; ++t) load_raw8(rn < rEnd ? pc[t] : nullptr, rn + 8 * kk, rEnd, qr[t]);
#pragma unroll
      for (int u = 0; u < NU; ++u) load_raw8(rn < rEnd ? uc[u] : nullptr, rn + 8 * kk, rEnd, vr[u]);
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        float bu[8];
        unpack8(ur[u], bu);
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
          float at[8];
          unpack8(pr[t], at);
#pragma unroll
          for (int e = 0; e < 8; ++e) acc[t][u] = mfma16(at[e], bu[e], acc[t][u]);
        }
      }
#pragma unroll
      for (int t = 0; t < TPW; ++t) pr[t] = qr[t];
#pragma unroll
      for (int u = 0; u < NU; ++u) ur[u] = vr[u];
      r0 = rn;
    }
#pragma unroll
    for (int t = 0; t < TPW; ++t)
#pragma unroll
      for (int u = 0; u < NU; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) red[((wave * TPW + t) * NU + u) * 256 + v * 64 + lane] = acc[t][u][v];
  }
  __syncthreads();
  if (rg != 0) return;
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tile = tg * TPW + t;
    if (tile < nT) {
#pragma unroll
      for (int u = 0; u < NU; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          float a = 0.0f;
          for (int g = 0; g < nRG; ++g)
            a = a + red[(((g * nTG + tg) * TPW + t) * NU + u) * 256 + v * 64 + lane];
          part[part_index(tile * 16 + 4 * kk + v, u * 16 + m)] = a;
        }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// bf16 basis on the bf16 matrix cores (v_mfma_f32_16x16x16bf16_1k, FP32 accumulation).  The FP32
// operand (U, or the coefficients C) is split into hi + lo bf16 parts (hi = RNE(x), lo =
// RNE(x - hi), |x - hi - lo| <= 2^-17 |x|), and both parts are multiplied, so the products keep
// about FP32 accuracy while the basis is read as stored.  Operand layouts: A (16 x 16): lane l
// holds A[l % 16][4 (l / 16) + 0..3]; B (16 x 16): lane l holds B[4 (l / 16) + 0..3][l % 16];
// D: lane l, register v holds D[4 (l / 16) + v][l % 16].
typedef short s16x4 __attribute__((ext_vector_type(4)));

__device__ inline f32x4 mfma16b(s16x4 a, s16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ inline s16x4 pack_s16x4(uint32_t lo, uint32_t hi) {
  s16x4 r;
  r[0] = (short)(lo & 0xffffu);
  r[1] = (short)(lo >> 16);
  r[2] = (short)(hi & 0xffffu);
  r[3] = (short)(hi >> 16);
  return r;
}
// hi and lo bf16 parts of four floats.
__device__ inline void split4(float x0, float x1, float x2, float x3, s16x4 &hi, s16x4 &lo) {
  const uint32_t h0 = f_to_bf(x0), h1 = f_to_bf(x1), h2 = f_to_bf(x2), h3 = f_to_bf(x3);
  const uint32_t l0 = f_to_bf(x0 - bf_to_f(h0)), l1 = f_to_bf(x1 - bf_to_f(h1)),
                 l2 = f_to_bf(x2 - bf_to_f(h2)), l3 = f_to_bf(x3 - bf_to_f(h3));
  hi = pack_s16x4(h0 | (h1 << 16), h2 | (h3 << 16));
  lo = pack_s16x4(l0 | (l1 << 16), l2 | (l3 << 16));
}

// C = P^T U with bf16 P: the tile-group x row-group structure of k_gemm_tn16 with 64-row rounds:
// lane (m, kk) holds rows r0 + 16kk .. r0 + 16kk + 15 of its columns (two 16-byte loads per bf16
// column, so a column's four lanes read 128 contiguous bytes), and MFMA group g (g = 0..3) takes
// rows r0 + 16kk + 4g .. + 3 as K indices 4kk .. 4kk + 3, once with the hi and once with the lo
// part of U.
// U2 (NU = 2): columns 16 .. 15 + s2 of the product are columns 0 .. s2 - 1 of U2 (pitch ldu),
// and U has s <= 16 columns.
template <int TPW, int NU>
__global__ __launch_bounds__(kThreads) void k_gemm_tn16b(const bf16 *__restrict__ P, long long ldp,
                                                         int p, int pPad,
                                                         const float *__restrict__ U,
                                                         long long ldu, int s, long long n,
                                                         long long rowsPerBlock,
                                                         float *__restrict__ part,
                                                         const float *__restrict__ U2, int s2) {
  __shared__ float red[4 * TPW * NU * 256];
  const int nT = (p + 15) / 16;
  const int nTG = (nT + TPW - 1) / TPW;
  const int nRG = 4 / nTG;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tg = wave % nTG, rg = wave / nTG;
  const int m = lane & 15, kk = lane >> 4;
  const long long rBeg = (long long)blockIdx.x * rowsPerBlock;
  const long long rEnd = rBeg + rowsPerBlock < n ? rBeg + rowsPerBlock : n;
  f32x4 acc[TPW][NU];
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int u = 0; u < NU; ++u) acc[t][u] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16 *pc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tile = tg * TPW + t, c = tile * 16 + m;
    pc[t] = (tile < nT && c < p) ? P + c * ldp : nullptr;
  }
  const float *uc[NU];
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    if (u == 1 && U2)
      uc[u] = m < s2 ? U2 + m * ldu : nullptr;
    else
      uc[u] = u * 16 + m < s ? U + (u * 16 + m) * ldu : nullptr;
  }
  if (rg < nRG) {
    const long long step = 64LL * nRG;
    long long r0 = rBeg + 64 * rg;
    Raw8<bf16> pr[TPW][2];
    Raw8<float> ur[NU][2];
#pragma unroll
    for (int t = 0; t < TPW; ++t)
#pragma unroll
      for (int h = 0; h < 2; ++h) load_raw8(pc[t], r0 + 16 * kk + 8 * h, rEnd, pr[t][h]);
#pragma unroll
    for (int u = 0; u < NU; ++u)
#pragma unroll
      for (int h = 0; h < 2; ++h) load_raw8(uc[u], r0 + 16 * kk + 8 * h, rEnd, ur[u][h]);
    while (r0 < rEnd) {
      const long long rn = r0 + step;
      Raw8<bf16> qr[TPW][2];
      Raw8<float> vr[NU][2];
#pragma unroll
      for (int t = 0; t < TPW; ++t)
#pragma unroll
        for (int h = 0; h < 2; ++h)
          load_raw8(rn < rEnd ? pc[t] : nullptr, rn + 16 * kk + 8 * h, rEnd, qr[t][h]);
#pragma unroll
      for (int u = 0; u < NU; ++u)
#pragma unroll
        for (int h = 0; h < 2; ++h)
          load_raw8(rn < rEnd ? uc[u] : nullptr, rn + 16 * kk + 8 * h, rEnd, vr[u][h]);
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        s16x4 bh[4], bl[4];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          split4(ur[u][h].a.x, ur[u][h].a.y, ur[u][h].a.z, ur[u][h].a.w, bh[2 * h], bl[2 * h]);
          split4(ur[u][h].b.x, ur[u][h].b.y, ur[u][h].b.z, ur[u][h].b.w, bh[2 * h + 1],
                 bl[2 * h + 1]);
        }
#pragma unroll
        for (int t = 0; t < TPW; ++t) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const s16x4 a0 = pack_s16x4(pr[t][h].w.x, pr[t][h].w.y);
            const s16x4 a1 = pack_s16x4(pr[t][h].w.z, pr[t][h].w.w);
            acc[t][u] = mfma16b(a0, bh[2 * h], acc[t][u]);
            acc[t][u] = mfma16b(a0, bl[2 * h], acc[t][u]);
            acc[t][u] = mfma16b(a1, bh[2 * h + 1], acc[t][u]);
            acc[t][u] = mfma16b(a1, bl[2 * h + 1], acc[t][u]);
          }
        }
      }
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
        pr[t][0] = qr[t][0];
        pr[t][1] = qr[t][1];
      }
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        ur[u][0] = vr[u][0];
        ur[u][1] = vr[u][1];
      }
      r0 = rn;
    }
#pragma unroll
    for (int t = 0; t < TPW; ++t)
#pragma unroll
      for (int u = 0; u < NU; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) red[((wave * TPW + t) * NU + u) * 256 + v * 64 + lane] = acc[t][u][v];
  }
  __syncthreads();
  if (rg != 0) return;
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tile = tg * TPW + t;
    if (tile < nT) {
#pragma unroll
      for (int u = 0; u < NU; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          float a = 0.0f;
          for (int g = 0; g < nRG; ++g)
            a = a + red[(((g * nTG + tg) * TPW + t) * NU + u) * 256 + v * 64 + lane];
          part[part_index(tile * 16 + 4 * kk + v, u * 16 + m)] = a;
        }
    }
  }
}

// Out[r, j] = a In[r, j] + b sum_k P[r, k] C[k][j] with bf16 P (C: p x 32 row-major FP32, staged
// in LDS as hi and lo bf16 parts, column-major by output column).  A wavefront owns 128 rows:
// lane (m, kg) loads rows r0 + 8m .. r0 + 8m + 7 of the four columns c0 + 4kg .. c0 + 4kg + 3 of
// each 16-column chunk (16-byte loads, a ring of two chunks in flight); MFMA i (i = 0..7) covers
// rows r0 + 8m + i, m = 0..15, so its A operand is element i of the four loaded words.  Output
// row r0 + 8 (4kg + v) + i, column 16t + m of N-tile t comes out in lane (m, kg), register v of
// accumulator (i, t); each N-tile is transposed through LDS so that In and Out move as whole
// columns (512 contiguous bytes per 128 rows).  Every wavefront runs the whole loop (rows past n
// load zeros and store nothing), so the workgroup barriers are uniform.
constexpr int kNnbTile = 128 * 16 + 16 * 4;  // floats per wavefront: 16 columns of 128 (+4 pad)

// With gram (NT = 1 only), the workgroup also writes the partial Gram matrix Out^T Out of its
// rows, element (i, j) at gram[part_index(i, j)] for i, j < 16 (FP32 16x16x4 MFMA over the output
// tile in LDS, the wavefronts added in order): the layout of gemm_tn's partials.
template <int NT, typename TO, int RB>
__global__ __launch_bounds__(kThreads) void k_gemm_nn_b(const bf16 *__restrict__ P, long long ldp,
                                                        int p, const float *__restrict__ C,
                                                        const float *In, long long ldi, float a,
                                                        float b, TO *Out, long long ldo, int s,
                                                        long long n, float *__restrict__ gram,
                                                        const float *In2, TO *Out2) {
  extern __shared__ uint16_t Cl[];  // [hi | lo][16 NT output columns][KP]; then the epilogue
  const int KP = (p + 15) & ~15;
  uint16_t *Ch = Cl, *Clo = Cl + (size_t)16 * NT * KP;
  // Eight coefficient loads in flight per thread, then their hi/lo parts.
  for (int i0 = threadIdx.x; i0 < 16 * NT * KP; i0 += 8 * kThreads) {
    float x[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int idx = i0 + q * kThreads;
      const int col = idx / KP, k = idx % KP;
      x[q] = (idx < 16 * NT * KP && k < p && col < s) ? C[k * 32 + col] : 0.0f;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int idx = i0 + q * kThreads;
      if (idx < 16 * NT * KP) {
        const uint32_t h = f_to_bf(x[q]);
        Ch[idx] = (uint16_t)h;
        Clo[idx] = (uint16_t)f_to_bf(x[q] - bf_to_f(h));
      }
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // RB consecutive 128-row blocks per wavefront: each column is read in RB * 256 contiguous bytes.
  const long long rw = ((long long)blockIdx.x * 4 + wave) * 128 * RB;
  const int m = lane & 15, kg = lane >> 4;
  f32x4 acc[RB][8][NT];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[rb][i][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  auto load_chunk = [&](int c0, Raw8<bf16> (&w)[RB][4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int cc = c0 + 4 * kg + j;
#pragma unroll
      for (int rb = 0; rb < RB; ++rb)
        load_raw8b<true>(cc < p ? P + cc * ldp : nullptr, rw + 128 * rb + 8 * m, n, w[rb][j]);
    }
  };
  Raw8<bf16> cur[RB][4], nxt[RB][4];
  load_chunk(0, cur);
  for (int c0 = 0; c0 < p; c0 += 16) {
    load_chunk(c0 + 16, nxt);  // past p: zero words, no loads
    s16x4 bh[NT], bl[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int base = (16 * t + m) * KP + c0 + 4 * kg;
      const uint2 h = *reinterpret_cast<const uint2 *>(Ch + base);
      const uint2 l = *reinterpret_cast<const uint2 *>(Clo + base);
      bh[t] = pack_s16x4(h.x, h.y);
      bl[t] = pack_s16x4(l.x, l.y);
    }
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      const uint32_t w[4][4] = {{cur[rb][0].w.x, cur[rb][0].w.y, cur[rb][0].w.z, cur[rb][0].w.w},
                                {cur[rb][1].w.x, cur[rb][1].w.y, cur[rb][1].w.z, cur[rb][1].w.w},
                                {cur[rb][2].w.x, cur[rb][2].w.y, cur[rb][2].w.z, cur[rb][2].w.w},
                                {cur[rb][3].w.x, cur[rb][3].w.y, cur[rb][3].w.z, cur[rb][3].w.w}};
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int q = i >> 1, sh = (i & 1) * 16;
        const uint32_t e0 = (w[0][q] >> sh) & 0xffffu, e1 = (w[1][q] >> sh) & 0xffffu;
        const uint32_t e2 = (w[2][q] >> sh) & 0xffffu, e3 = (w[3][q] >> sh) & 0xffffu;
        const s16x4 av = pack_s16x4(e0 | (e1 << 16), e2 | (e3 << 16));
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          acc[rb][i][t] = mfma16b(av, bh[t], acc[rb][i][t]);
          acc[rb][i][t] = mfma16b(av, bl[t], acc[rb][i][t]);
        }
      }
    }
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int j = 0; j < 4; ++j) cur[rb][j] = nxt[rb][j];
  }
  // Epilogue: through LDS (the coefficient area is free once every wavefront is done with it).
  f32x4 gacc = f32x4{0.f, 0.f, 0.f, 0.f};
  float *tile = reinterpret_cast<float *>(Cl) + wave * kNnbTile;
  const int stride = 128 + 4;
#pragma unroll
  for (int rb = 0; rb < RB; ++rb)
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const long long r0 = rw + 128 * rb;
    __syncthreads();
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      float *dst = tile + m * stride + 8 * (4 * kg + v);
      *reinterpret_cast<float4 *>(dst) = make_float4(acc[rb][0][t][v], acc[rb][1][t][v],
                                                     acc[rb][2][t][v], acc[rb][3][t][v]);
      *reinterpret_cast<float4 *>(dst + 4) = make_float4(acc[rb][4][t][v], acc[rb][5][t][v],
                                                         acc[rb][6][t][v], acc[rb][7][t][v]);
    }
    __syncthreads();
    // Lane (h, q) = (l / 32, l % 32) moves rows r0 + 4q .. r0 + 4q + 3 of column 2cc + h.
    const int h = lane >> 5, q4 = (lane & 31) * 4;
    const long long row = r0 + q4;
    if (row < n) {
      const int cnt = row + 3 < n ? 4 : (int)(n - row);
#pragma unroll
      for (int cc = 0; cc < 8; ++cc) {
        const int jl = 2 * cc + h, j = 16 * t + jl;
        // With Out2 (NT = 2), output columns 16 .. 31 are columns 0 .. 15 of In2 / Out2, and
        // each of the two blocks has s - 16 columns.
        if (NT == 2 && Out2 ? jl >= s - 16 : j >= s) continue;
        const float *inc = (NT == 2 && t == 1 && Out2) ? In2 + jl * ldi : In + j * ldi;
        TO *outc = (NT == 2 && t == 1 && Out2) ? Out2 + jl * ldo : Out + j * ldo;
        const float4 pv = *reinterpret_cast<const float4 *>(tile + jl * stride + q4);
        const float pr[4] = {pv.x, pv.y, pv.z, pv.w};
        float in[4] = {0.f, 0.f, 0.f, 0.f};
        if (a != 0.0f) {
          if (cnt == 4) {
            const float4 iv = *reinterpret_cast<const float4 *>(inc + row);
            in[0] = iv.x;
            in[1] = iv.y;
            in[2] = iv.z;
            in[3] = iv.w;
          } else {
            for (int e = 0; e < cnt; ++e) in[e] = inc[row + e];
          }
        }
        float o[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) o[e] = a != 0.0f ? a * in[e] + b * pr[e] : b * pr[e];
        store_rows4(outc + row, o, cnt);
        if (NT == 1 && gram)  // the output as stored (rows past n stay zero) for the Gram matrix
          *reinterpret_cast<float4 *>(tile + jl * stride + q4) = make_float4(o[0], o[1], o[2], o[3]);
      }
    }
    if (NT == 1 && gram) {
      __syncthreads();
      // Lane l contributes rows 4q + l / 16 of column l % 16 to k-step q: A = B = this value.
#pragma unroll 8
      for (int q = 0; q < 32; ++q) {
        const float x = tile[(lane & 15) * stride + 4 * q + (lane >> 4)];
        gacc = mfma16(x, x, gacc);
      }
    }
  }
  if (NT == 1 && gram) {
    __syncthreads();
    float *red = reinterpret_cast<float *>(Cl) + wave * kNnbTile;
#pragma unroll
    for (int v = 0; v < 4; ++v) red[v * 64 + lane] = gacc[v];
    __syncthreads();
    if (wave == 0) {

#pragma unroll
      for (int v = 0; v < 4; ++v) {
        float g = 0.0f;
#pragma unroll
        for (int w = 0; w < 4; ++w)
          g = g + reinterpret_cast<const float *>(Cl)[w * kNnbTile + v * 64 + lane];
        gram[part_index(4 * kg + v, m)] = g;
      }
    }
  }
}

// Out[r, j] = sum_{k < s} U[r, k] C[k][j] for j < s <= 16 (C row-major, 32 wide): the product of
// a block by a small square factor (Cholesky QR's normalisation).  Thread: rows 2t and 2t + 1,
// one 8-byte load and one store of two rows per column; C is uniform (scalar loads).  Each sum
// runs over k in order with fused multiply-adds.  ldu and ldo are even (the basis pitch).  Out
// may alias U: a thread reads all of its rows before it writes them.
__device__ inline void store_rows2(float *dst, float v0, float v1, bool two) {
  if (two)
    *reinterpret_cast<float2 *>(dst) = make_float2(v0, v1);
  else
    dst[0] = v0;
}
__device__ inline void store_rows2(bf16 *dst, float v0, float v1, bool two) {
  if (two)
    *reinterpret_cast<uint32_t *>(dst) = f_to_bf(v0) | (f_to_bf(v1) << 16);
  else
    dst[0].bits = (uint16_t)f_to_bf(v0);
}

template <typename TO>
__global__ __launch_bounds__(kThreads) void k_scale_block(const float *U, long long ldu,
                                                          const float *__restrict__ C, TO *Out,
                                                          long long ldo, int s, long long n) {
  const long long r = 2 * ((long long)blockIdx.x * kThreads + threadIdx.x);
  if (r >= n) return;
  const bool two = r + 1 < n;
  float u0[16], u1[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    u0[k] = u1[k] = 0.0f;
    if (k < s) {
      if (two) {
        const float2 v = *reinterpret_cast<const float2 *>(U + k * ldu + r);
        u0[k] = v.x;
        u1[k] = v.y;
      } else {
        u0[k] = U[k * ldu + r];
      }
    }
  }
  float o0[16], o1[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) o0[j] = o1[j] = 0.0f;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    if (k < s) {
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        const float cv = C[k * 32 + j];
        o0[j] = fmaf(u0[k], cv, o0[j]);
        o1[j] = fmaf(u1[k], cv, o1[j]);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 16; ++j)
    if (j < s) store_rows2(Out + j * ldo + r, o0[j], o1[j], two);
}

// LDS-staged variant: each 32-row chunk of P and U is transposed through LDS.
__global__ __launch_bounds__(kThreads) void k_gemm_tn_lds(const float *__restrict__ P, long long ldp,
                                                      int p, int pPad,
                                                      const float *__restrict__ U, long long ldu,
                                                      int s, long long n, long long rowsPerBlock,
                                                      float *__restrict__ part) {
  extern __shared__ float lds[];
  float *Ps = lds;                               // pPad columns x kLdsStride
  float *Us = lds + (long long)pPad * kLdsStride;  // 32 columns x kLdsStride
  const int tiles = pPad / 32;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long rBeg = (long long)blockIdx.x * rowsPerBlock;
  const long long rEnd = rBeg + rowsPerBlock < n ? rBeg + rowsPerBlock : n;
  f32x16 acc[kMaxTiles / 4];
#pragma unroll
  for (int t = 0; t < kMaxTiles / 4; ++t)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[t][v] = 0.0f;
  const int ncol = pPad + 32;
  for (long long r0 = rBeg; r0 < rEnd; r0 += kRC) {
    for (int idx = threadIdx.x; idx < ncol * (kRC / 4); idx += kThreads) {
      const int col = idx / (kRC / 4), rq = (idx % (kRC / 4)) * 4;
      const long long r = r0 + rq;
      const bool isU = col >= pPad;
      const int c = isU ? col - pPad : col;
      float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f;
      if (isU ? c < s : c < p) {
        const float *src = (isU ? U + c * ldu : P + c * ldp) + r;
        if (r + 3 < rEnd) {
          const float4 v = *reinterpret_cast<const float4 *>(src);
          v0 = v.x;
          v1 = v.y;
          v2 = v.z;
          v3 = v.w;
        } else {
          if (r < rEnd) v0 = src[0];
          if (r + 1 < rEnd) v1 = src[1];
          if (r + 2 < rEnd) v2 = src[2];
        }
      }
      float *dst = (isU ? Us + c * kLdsStride : Ps + c * kLdsStride) + rq;
      dst[0] = v0;
      dst[1] = v1;
      dst[2] = v2;
      dst[3] = v3;
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < kMaxTiles / 4; ++t) {
      const int tile = wave + 4 * t;
      if (tile < tiles) {
        const float *pa = Ps + (tile * 32 + (lane & 31)) * kLdsStride + (lane >> 5);
        const float *pb = Us + (lane & 31) * kLdsStride + (lane >> 5);
#pragma unroll 4
        for (int k0 = 0; k0 < kRC; k0 += 2) acc[t] = mfma32(pa[k0], pb[k0], acc[t]);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int t = 0; t < kMaxTiles / 4; ++t) {
    const int tile = wave + 4 * t;
    if (tile < tiles) {
#pragma unroll
      for (int v = 0; v < 16; ++v)
        part[part_index(tile * 32 + mfma_row(v, lane), lane & 31)] = acc[t][v];
    }
  }
}

// C[i][j] = sum_g part_g[i][j] (FP64), for i < p, j < s: to C64 (column-major, element (i, j) at
// j * ldc + i) and to C32 (row-major p x 32, zero beyond s).  One block of kRedThreads threads
// per row i, whose G partials are contiguous (part_index); JW = 16 (s <= 16, only the columns
// that hold data are read) or 32 columns, read four at a time: thread (c4, q) = (t % (JW / 4),
// t / (JW / 4)) adds partials g = q, q + NQ, ... (NQ = kRedThreads / (JW / 4)) in order, all of a
// batch's loads issued before its adds; then per column the NQ sums are added in order of q, in
// kRedSeg consecutive segments and the segments in order (a fixed order, independent of timing).
// split > 0 (JW = 32): columns 0 .. s - 1 (s <= split) go to C64 and C32, columns
// split .. split + s2 - 1 to C64b (its columns 0 .. s2 - 1).
constexpr int kRedThreads = 512;
constexpr int kRedLoads = 8;  // partials loaded per batch and thread
constexpr int kRedSeg = 8;

template <int JW>
__global__ __launch_bounds__(kRedThreads) void k_gemm_tn_reduce(const float *__restrict__ part,
                                                                int G, int p, int s,
                                                                double *__restrict__ C64, int ldc,
                                                                float *__restrict__ C32,
                                                                double *__restrict__ C64b,
                                                                int ldcb, int split, int s2) {
  constexpr int TPG = JW / 4;          // threads per partial (four columns each)
  constexpr int NQ = kRedThreads / TPG;
  constexpr int QPS = NQ / kRedSeg;
  __shared__ double red[NQ * JW];
  __shared__ double seg[kRedSeg * JW];
  const int i = blockIdx.x, t = threadIdx.x, c4 = t % TPG, q = t / TPG;
  const float4 *src = reinterpret_cast<const float4 *>(part + (long long)i * G * 32) + c4;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int g0 = q; g0 < G; g0 += NQ * kRedLoads) {
    float4 v[kRedLoads];
#pragma unroll
    for (int u = 0; u < kRedLoads; ++u) {
      const int g = g0 + u * NQ;
      v[u] = g < G ? src[(long long)g * 8] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < kRedLoads; ++u) {
      if (g0 + u * NQ < G) {
        a0 = a0 + (double)v[u].x;
        a1 = a1 + (double)v[u].y;
        a2 = a2 + (double)v[u].z;
        a3 = a3 + (double)v[u].w;
      }
    }
  }
  red[q * JW + 4 * c4 + 0] = a0;
  red[q * JW + 4 * c4 + 1] = a1;
  red[q * JW + 4 * c4 + 2] = a2;
  red[q * JW + 4 * c4 + 3] = a3;
  __syncthreads();
  if (t < kRedSeg * JW) {
    const int c = t % JW, z = t / JW;
    double b = 0.0;
    for (int qq = z * QPS; qq < (z + 1) * QPS; ++qq) b = b + red[qq * JW + c];
    seg[z * JW + c] = b;
  }
  __syncthreads();
  if (t >= 32) return;
  // Thread t < 32 writes column t (columns JW .. 31 are zero).
  double sum = 0.0;
  if (t < JW) {
#pragma unroll
    for (int z = 0; z < kRedSeg; ++z) sum = sum + seg[z * JW + t];
  }
  if (split > 0 && t >= split) {
    // Columns split .. split + s2 - 1 go to C64b (its columns 0 ..), and are zero in C32.
    if (C64b && t - split < s2) C64b[(long long)(t - split) * ldcb + i] = sum;
    if (C32) C32[i * 32 + t] = 0.0f;
    return;
  }
  if (t >= s) sum = 0.0;
  if (C64 && t < s) C64[(long long)t * ldc + i] = sum;
  if (C32) C32[i * 32 + t] = (float)sum;
}

void tn_reduce(const float *part, int G, int pPad, int p, int s, double *C64, int ldc, float *C32,
               hipStream_t st, double *C64b = nullptr, int ldcb = 0, int split = 0, int s2 = 0) {
  (void)pPad;
  if (s <= 16 && split == 0)
    hipLaunchKernelGGL(k_gemm_tn_reduce<16>, dim3(p), dim3(kRedThreads), 0, st, part, G, p, s,
                       C64, ldc, C32, C64b, ldcb, split, s2);
  else
    hipLaunchKernelGGL(k_gemm_tn_reduce<32>, dim3(p), dim3(kRedThreads), 0, st, part, G, p, s,
                       C64, ldc, C32, C64b, ldcb, split, s2);
}

// ---------------------------------------------------------------------------------------------
// Out[r, j] = a In[r, j] + b sum_k P[r, k] C[k][j] for j < s, r < n.  C: p x 32 row-major.
// Out may alias In or P (each row depends on its own row only).  A wavefront owns 128 rows as
// four interleaved 32-row MFMA tiles; lane (i, h) loads rows r0 + 4i .. r0 + 4i + 3 of column
// c + h as one 16-byte vector, and the loads run D column pairs ahead of their MFMAs (a register
// ring), so each wavefront keeps D KiB in flight.
__device__ inline float4 load_rows4(const float *P, long long ldp, int cc, int p, long long rl,
                                    long long n) {
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  if (cc < p) {
    const float *src = P + cc * ldp + rl;
    if (rl + 3 < n) {
      v = *reinterpret_cast<const float4 *>(src);
    } else {
      if (rl < n) v.x = src[0];
      if (rl + 1 < n) v.y = src[1];
      if (rl + 2 < n) v.z = src[2];
    }
  }
  return v;
}

// Four rows of one column as loaded (raw, unpacked where consumed): float4 or two bf16 words.
template <typename TP> struct Raw4;
template <> struct Raw4<float> { float4 v; };
template <> struct Raw4<bf16> { uint2 w; };
__device__ inline Raw4<float> load_raw4(const float *P, long long ldp, int cc, int p,
                                        long long rl, long long n) {
  return Raw4<float>{load_rows4(P, ldp, cc, p, rl, n)};
}
__device__ inline Raw4<bf16> load_raw4(const bf16 *P, long long ldp, int cc, int p, long long rl,
                                       long long n) {
  Raw4<bf16> o{make_uint2(0u, 0u)};
  if (cc < p) {
    const bf16 *src = P + cc * ldp + rl;
    if (rl + 3 < n) {
      o.w = *reinterpret_cast<const uint2 *>(src);
    } else {
      uint32_t h[4] = {0u, 0u, 0u, 0u};
      for (int q = 0; q < 3; ++q)
        if (rl + q < n) h[q] = src[q].bits;
      o.w = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
    }
  }
  return o;
}
// load_raw4 with the non-temporal hint on the full four-row loads: the basis reads of k_gemm_nn,
// each column read once per launch (FP32-basis block update 14.47 -> 14.79 solves/s on the 96^3
// batched benchmark, same bits).
typedef float f32x4v_nt __attribute__((ext_vector_type(4)));
typedef unsigned u32x2v_nt __attribute__((ext_vector_type(2)));
__device__ inline Raw4<float> load_raw4_nt(const float *P, long long ldp, int cc, int p,
                                           long long rl, long long n) {
  if (cc < p && rl + 3 < n) {
    const f32x4v_nt v = __builtin_nontemporal_load(reinterpret_cast<const f32x4v_nt *>(P + cc * ldp + rl));
    return Raw4<float>{make_float4(v.x, v.y, v.z, v.w)};
  }
  return load_raw4(P, ldp, cc, p, rl, n);
}
__device__ inline Raw4<bf16> load_raw4_nt(const bf16 *P, long long ldp, int cc, int p,
                                          long long rl, long long n) {
  if (cc < p && rl + 3 < n) {
    const u32x2v_nt v = __builtin_nontemporal_load(reinterpret_cast<const u32x2v_nt *>(P + cc * ldp + rl));
    return Raw4<bf16>{make_uint2(v.x, v.y)};
  }
  return load_raw4(P, ldp, cc, p, rl, n);
}
template <bool NT, typename TP>
__device__ inline Raw4<TP> load_raw4_h(const TP *P, long long ldp, int cc, int p, long long rl,
                                       long long n) {
  if constexpr (NT) return load_raw4_nt(P, ldp, cc, p, rl, n);
  else return load_raw4(P, ldp, cc, p, rl, n);
}
__device__ inline float4 unpack4(const Raw4<float> &o) { return o.v; }
__device__ inline float4 unpack4(const Raw4<bf16> &o) {
  return make_float4(__uint_as_float(o.w.x << 16), __uint_as_float(o.w.x & 0xffff0000u),
                     __uint_as_float(o.w.y << 16), __uint_as_float(o.w.y & 0xffff0000u));
}

template <int D, typename TP, typename TO>
__global__ __launch_bounds__(kThreads) void k_gemm_nn(const TP *P, long long ldp, int p,
                                                      const float *__restrict__ C,
                                                      const float *In, long long ldi, float a,
                                                      float b, TO *Out, long long ldo, int s,
                                                      long long n) {
  extern __shared__ float Cs[];
  const int pPad = (p + 2 * D - 1) / (2 * D) * (2 * D) + 2 * D;  // ring reads past p are zero
  for (int idx = threadIdx.x; idx < pPad * 32; idx += kThreads)
    Cs[idx] = (idx >> 5) < p ? C[idx] : 0.0f;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long r0 = ((long long)blockIdx.x * 4 + wave) * 128;
  if (r0 >= n) return;
  f32x16 acc0, acc1, acc2, acc3;
#pragma unroll
  for (int v = 0; v < 16; ++v) acc0[v] = acc1[v] = acc2[v] = acc3[v] = 0.0f;
  const long long rl = r0 + 4 * (lane & 31);
  const int kh = lane >> 5, jl = lane & 31;
  Raw4<TP> ring[D];
#pragma unroll
  for (int i = 0; i < D; ++i) ring[i] = load_raw4_h<true>(P, ldp, 2 * i + kh, p, rl, n);
  for (int c = 0; c < p; c += 2 * D) {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      const float4 x = unpack4(ring[i]);
      ring[i] = load_raw4_h<true>(P, ldp, c + 2 * (i + D) + kh, p, rl, n);
      const float bb = Cs[(c + 2 * i + kh) * 32 + jl];
      acc0 = mfma32(x.x, bb, acc0);
      acc1 = mfma32(x.y, bb, acc1);
      acc2 = mfma32(x.z, bb, acc2);
      acc3 = mfma32(x.w, bb, acc3);
    }
  }
  const int j = jl;
  if (j >= s) return;
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    const long long row = r0 + 4 * mfma_row(v, lane);
    const float o[4] = {acc0[v], acc1[v], acc2[v], acc3[v]};
    if (row >= n) continue;
    const int cnt = row + 3 < n ? 4 : (int)(n - row);
    float w[4];
    if (a != 0.0f) {
      float in[4] = {0.f, 0.f, 0.f, 0.f};
      if (cnt == 4) {
        const float4 iv = *reinterpret_cast<const float4 *>(In + j * ldi + row);
        in[0] = iv.x;
        in[1] = iv.y;
        in[2] = iv.z;
        in[3] = iv.w;
      } else {
        for (int q = 0; q < cnt; ++q) in[q] = In[j * ldi + row + q];
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) w[q] = a * in[q] + b * o[q];
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) w[q] = b * o[q];
    }
    store_rows4(Out + j * ldo + row, w, cnt);
  }
}

// ---------------------------------------------------------------------------------------------
// Broadcast of lane `l`'s value to the wavefront (l uniform).
__device__ inline double bcast(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// Cholesky QR factor of the Gram matrix G (s x s, column-major ld s, FP64): G = S^T S with S
// upper triangular.  A pivot below relTol * max diag marks the column dependent: its row of S
// and its column of S^-1 are zero (the normalised block gets a zero column) and *brk is set.
// Writes Sinv (FP32 row-major s x 32 for gemm_nn) and the combined factor Sacc <- S * Sacc
// (Sacc upper triangular s x s, ld s; S itself on the first pass), optionally copied into the
// block Hessenberg at Hdst (ld ldh).  One wavefront: lane j keeps column j of G (then of S, of
// S^-1 and of the product) in registers; every cross-lane operand is a broadcast of one lane's
// register, so the factorisation needs no LDS and no barrier.
__global__ __launch_bounds__(64) void k_chol_qr(const double *__restrict__ Gin, int s,
                                                double relTol, float *__restrict__ Sinv32,
                                                double *__restrict__ Sacc, int first,
                                                double *__restrict__ Hdst, int ldh,
                                                int *__restrict__ brk) {
  const int j = threadIdx.x;
  const bool own = j < s;
  double g[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) g[i] = (own && i < s) ? Gin[j * s + i] : 0.0;
  double dmax = 0.0;
#pragma unroll
  for (int i = 0; i < 32; ++i)
    if (i < s) dmax = fmax(dmax, bcast(g[i], i));
  unsigned dep = 0u;
  // Right-looking factorisation: after step k lane j holds S(k, j) in g[k].
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    if (k < s) {
      const double d = bcast(g[k], k);
      const bool bad = !(d > relTol * dmax);
      const double skk = bad ? 0.0 : sqrt(d);
      dep |= bad ? (1u << k) : 0u;
      if (j == k) g[k] = skk;
      if (j > k) g[k] = bad ? 0.0 : g[k] / skk;
      const double skj = g[k];
#pragma unroll
      for (int i = k + 1; i < 32; ++i) {
        if (i < s) {
          const double ski = bcast(g[k], i);
          if (j >= i) g[i] = g[i] - ski * skj;
        }
      }
    }
  }
  // Column j of S^-1 by back substitution, columns of S broadcast from their lanes.
  const bool depj = own ? ((dep >> j) & 1u) != 0u : true;
  double x[32], acc[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    x[i] = 0.0;
    acc[i] = (i == j) ? 1.0 : 0.0;
  }
#pragma unroll
  for (int l = 31; l >= 0; --l) {
    if (l < s) {
      const double sll = bcast(g[l], l);
      const bool skip = depj || ((dep >> l) & 1u) != 0u || l > j;
      const double xl = skip ? 0.0 : acc[l] / sll;
      x[l] = xl;
#pragma unroll
      for (int r = 0; r < l; ++r) acc[r] = fma(-bcast(g[r], l), xl, acc[r]);
    }
  }
  if (j < 32) {
#pragma unroll
    for (int i = 0; i < 32; ++i) Sinv32[i * 32 + j] = own && i < s ? (float)x[i] : 0.0f;
  }
  // Sacc <- S * Sacc: column j of the product, S(i, l) broadcast from lane l.
#pragma unroll
  for (int i = 0; i < 32; ++i) acc[i] = (!first && own && i < s) ? Sacc[j * s + i] : 0.0;
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    if (i < s) {
      double a = 0.0;
      if (first) {
        a = i <= j ? g[i] : 0.0;
      } else {
#pragma unroll
        for (int l = i; l < 32; ++l)
          if (l < s) a = fma(bcast(g[i], l), l <= j ? acc[l] : 0.0, a);
      }
      x[i] = a;
    }
  }
  if (own) {
#pragma unroll
    for (int i = 0; i < 32; ++i) {
      if (i < s) {
        Sacc[j * s + i] = x[i];
        if (Hdst) Hdst[(long long)j * ldh + i] = x[i];
      }
    }
  }
  if (j == 0 && dep != 0u) atomicOr(brk, 1);
}

// The same factorisation with one thread per matrix element (D x D threads, D = 16 or 32 >= s):
// the operands of every step come from LDS, and the operations and their order per element are
// those of k_chol_qr, so both give the same bits.
// Thread t = D i + j of a D x D workgroup brings element (i, j) of the Gram matrix in g (zero
// outside s x s).
template <int D>
__device__ inline void chol_el(double g, int s, double relTol, float *__restrict__ Sinv32,
                               double *__restrict__ Sacc, int first, double *__restrict__ Hdst,
                               int ldh, int *__restrict__ brk) {
  constexpr int LD = D + 1;
  __shared__ double A[D * LD];  // G, then S in the upper triangle
  __shared__ double X[D * LD];  // unit columns, then S^-1 (column j of S^-1 in column j)
  __shared__ double Y[D * LD];  // the previous accumulated factor
  __shared__ double sh_dmax, sh_skk;
  __shared__ unsigned sh_dep;
  __shared__ int sh_bad;
  const int t = threadIdx.x, i = t / D, j = t % D;
  const bool own = i < s && j < s;
  A[i * LD + j] = own ? g : 0.0;
  X[i * LD + j] = i == j ? 1.0 : 0.0;
  Y[i * LD + j] = (!first && own) ? Sacc[j * s + i] : 0.0;
  if (t == 0) sh_dep = 0u;
  __syncthreads();
  if (t < 64) {
    double d = t < s ? A[t * LD + t] : 0.0;
    d = fmax(d, 0.0);
    d = xor_max64(d);
    if (t == 0) sh_dmax = d;
  }
  __syncthreads();
  const double dmax = sh_dmax;
  // Right-looking factorisation: after step k row k of A holds S(k, :).
  for (int k = 0; k < s; ++k) {
    if (t == 0) {
      const double d = A[k * LD + k];
      const bool bad = !(d > relTol * dmax);
      const double skk = bad ? 0.0 : sqrt(d);
      A[k * LD + k] = skk;
      sh_skk = skk;
      sh_bad = bad ? 1 : 0;
      if (bad) sh_dep |= 1u << k;
    }
    __syncthreads();
    if (i == k && j > k && j < s) A[k * LD + j] = sh_bad ? 0.0 : A[k * LD + j] / sh_skk;
    __syncthreads();
    if (i > k && j >= i && j < s) A[i * LD + j] = A[i * LD + j] - A[k * LD + i] * A[k * LD + j];
    __syncthreads();
  }
  const unsigned dep = sh_dep;
  // Column j of S^-1 by back substitution (row l of X becomes x_l once step l is done).
  for (int l = s - 1; l >= 0; --l) {
    if (i == l) {
      const bool depj = j < s ? ((dep >> j) & 1u) != 0u : true;
      const bool skip = depj || ((dep >> l) & 1u) != 0u || l > j;
      X[l * LD + j] = skip ? 0.0 : X[l * LD + j] / A[l * LD + l];
    }
    __syncthreads();
    if (i < l) X[i * LD + j] = fma(-A[i * LD + l], X[l * LD + j], X[i * LD + j]);
    __syncthreads();
  }
  if (i < 32 && j < 32)
    Sinv32[i * 32 + j] = own ? (float)X[i * LD + j] : 0.0f;
  for (int q = t; q < 32 * 32; q += D * D) {  // the rest of the 32-wide FP32 factor is zero
    const int qi = q >> 5, qj = q & 31;
    if (qi >= D || qj >= D) Sinv32[q] = 0.0f;
  }
  // Sacc <- S * Sacc (S alone on the first pass), into Hdst as well.
  if (own) {
    double a = 0.0;
    if (first) {
      a = i <= j ? A[i * LD + j] : 0.0;
    } else {
      for (int l = i; l < s; ++l) a = fma(A[i * LD + l], l <= j ? Y[l * LD + j] : 0.0, a);
    }
    Sacc[j * s + i] = a;
    if (Hdst) Hdst[(long long)j * ldh + i] = a;
  }
  if (t == 0 && dep != 0u) atomicOr(brk, 1);
}

template <int D>
__global__ __launch_bounds__(D * D) void k_chol_qr_el(const double *__restrict__ Gin, int s,
                                                      double relTol, float *__restrict__ Sinv32,
                                                      double *__restrict__ Sacc, int first,
                                                      double *__restrict__ Hdst, int ldh,
                                                      int *__restrict__ brk) {
  const int i = threadIdx.x / D, j = threadIdx.x % D;
  chol_el<D>(i < s && j < s ? Gin[j * s + i] : 0.0, s, relTol, Sinv32, Sacc, first, Hdst, ldh,
             brk);
}

// The Gram reduce of G partials (rows i < s <= 16 of the layout part[(i G + g) 32 + j]) and the
// Cholesky factorisation of k_chol_qr_el<16> in one launch.  Workgroup b sums partials
// g0 .. g0 + L - 1 (L = ceil(G / gridDim.x), g0 = b L) in FP64.  Phase 1: the s rows times
// nc4 = ceil(s / 4) four-column groups are the ng = s nc4 element groups; thread t < H ng
// (H = min(8, 256 / ng)) takes group t % ng and adds partials g0 + h, g0 + h + H, ... in order
// (h = t / ng, kGcLoads loads in flight); the H sums of each element are added in order of h and
// the workgroup's s x 16 segment sum goes to seg[b].  Phase 2, in the last workgroup to arrive:
// T2 = min(4, 256 / s^2) threads per element of the s x s sum, thread part c adding segments
// c, c + T2, ... in order, then the T2 parts in order of c; then the factorisation.  The order is
// fixed by G, s and the grid, independent of timing.  ctr is zero on entry and on exit.
constexpr int kGcSegs = kGramCholSegs;
constexpr int kGcLoads = 16;
constexpr int kGcMaxH = 8;

__global__ __launch_bounds__(256) void k_gram_chol(const float *__restrict__ part, int G, int L,
                                                   int s, double relTol,
                                                   float *__restrict__ Sinv32,
                                                   double *__restrict__ Sacc, int first,
                                                   double *__restrict__ Hdst, int ldh,
                                                   int *__restrict__ brk,
                                                   double *__restrict__ seg, unsigned *ctr) {
  __shared__ double red[kGcMaxH * 256];
  __shared__ int s_last;
  const int t = threadIdx.x;
  const int nc4 = (s + 3) >> 2, ng = s * nc4;
  const int H = 256 / ng < kGcMaxH ? 256 / ng : kGcMaxH;
  const int g0 = blockIdx.x * L, g1 = g0 + L < G ? g0 + L : G;
  if (t < H * ng) {
    const int grp = t % ng, h = t / ng, r = grp / nc4, c4 = grp % nc4;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    const float4 *src = reinterpret_cast<const float4 *>(part + (long long)r * G * 32) + c4;
    for (int gb = g0 + h; gb < g1; gb += H * kGcLoads) {
      float4 v[kGcLoads];
#pragma unroll
      for (int u = 0; u < kGcLoads; ++u) {
        const int g = gb + H * u;
        v[u] = g < g1 ? src[(long long)g * 8] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
      for (int u = 0; u < kGcLoads; ++u) {
        if (gb + H * u < g1) {
          a0 = a0 + (double)v[u].x;
          a1 = a1 + (double)v[u].y;
          a2 = a2 + (double)v[u].z;
          a3 = a3 + (double)v[u].w;
        }
      }
    }
    double *rd = red + h * 256 + r * 16 + 4 * c4;
    rd[0] = a0;
    rd[1] = a1;
    rd[2] = a2;
    rd[3] = a3;
  }
  __syncthreads();
  // Thread t = 16 i + j: element (i, j) of this segment (i < s, j < 4 nc4).
  if ((t >> 4) < s && (t & 15) < 4 * nc4) {
    double a = red[t];
    for (int h = 1; h < H; ++h) a = a + red[h * 256 + t];
    seg[(long long)blockIdx.x * 256 + t] = a;
  }
  __syncthreads();
  if (t == 0) {
    const unsigned total = gridDim.x;
    const bool last =
        __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == total - 1u;
    if (last) __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_last = last ? 1 : 0;
  }
  __syncthreads();
  if (!s_last) return;
  const int ne = s * s, nb = gridDim.x;
  const int T2 = 256 / ne < 4 ? 256 / ne : 4;
  if (t < T2 * ne) {
    const int e = t % ne, c = t / ne, i = e / s, j = e % s;
    const double *src = seg + i * 16 + j;
    double a = 0.0;
    for (int b0 = c; b0 < nb; b0 += T2 * kGcLoads) {
      double v[kGcLoads];
#pragma unroll
      for (int u = 0; u < kGcLoads; ++u) {
        const int b = b0 + T2 * u;
        v[u] = b < nb ? __hip_atomic_load(src + (long long)b * 256, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT)
                      : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kGcLoads; ++u)
        if (b0 + T2 * u < nb) a = a + v[u];
    }
    red[c * 256 + e] = a;
  }
  __syncthreads();
  const int i = t >> 4, j = t & 15;
  double g = 0.0;
  if (i < s && j < s) {
    const int e = i * s + j;
    g = red[e];
    for (int c = 1; c < T2; ++c) g = g + red[c * 256 + e];
  }
  __syncthreads();  // red is free again before chol_el's own shared arrays are written
  chol_el<16>(g, s, relTol, Sinv32, Sacc, first, Hdst, ldh, brk);
}


// ---------------------------------------------------------------------------------------------
// Least squares min ||T - H Y|| (per column) on the block Hessenberg H (rows x cols, ld rows,
// lower bandwidth bw <= 32, rows = cols + bw): Householder QR with reflectors over rows c..c+bw.
// Only rows c..c+bw change at step c, so they live in an LDS window of bw + 1 rows of [H | T]
// (a ring: row r in slot r % (bw + 1)); after step c row c is final (written to R, column-major
// ld cols, and Qt = Q^T T) and its slot takes row c + bw + 1 from H, which no reflector has
// touched yet (fetched into registers while the reflector is applied).  Back substitution runs in
// 32-row blocks: one wavefront solves the diagonal block from LDS, then all threads update the
// rows above with the block's columns of R.  Then Yr32 = Y diag(rho) (FP32 row-major cols x 32)
// and ZH32 = H Y (FP32 row-major rows x 32).  One workgroup of 1024 threads.
// BWT: the bandwidth as a compile-time constant (16) for the unrolled reflector application, or
// 0 for a runtime bandwidth.  Both compute the same bits: rows of the window past the matrix are
// zero, so a reflector padded to the full bandwidth adds exact zeros.
// STG (BWT = 16 or 32): the register path described in the kernel, with the rows entering the
// window fetched kLsqChunk steps at a time, one chunk ahead, through an LDS stage.
constexpr int kLsqChunk = 8;

template <int BWT, bool STG>
__global__ __launch_bounds__(BWT == 32 ? 512 : 1024) void k_lsq_block(const double *__restrict__ H,
                                                                      int rows,
                                                    int cols, int bw,
                                                    const double *__restrict__ T, int s,
                                                    const double *__restrict__ rho,
                                                    double *__restrict__ R,
                                                    double *__restrict__ Qt,
                                                    double *__restrict__ Y,
                                                    float *__restrict__ Yr32,
                                                    float *__restrict__ ZH32) {
  // Right-hand-side columns that can be nonzero: s, which equals the bandwidth (16 when BWT = 16).
  constexpr int SW = BWT > 0 && BWT <= 16 ? 16 : 32;
  extern __shared__ double sm[];
  __shared__ double s_tau, s_rmax;
  const int tid = threadIdx.x, bd = blockDim.x;
  const int W = bw + 1, nc = cols + s;
  double *win = sm;  // W x nc
  auto src = [&](int r, int col) -> double {
    return col < cols ? H[(long long)col * rows + r] : T[(long long)(col - cols) * rows + r];
  };
  if constexpr (STG) {
    // Register path (BWT = 16 or 32, bandwidth bw <= BWT, cols + s <= blockDim.x): thread tid
    // owns column tid and holds its window rows c .. c + BWT in registers (x[0] = row c).  Per
    // step: the owner of column c forms the reflector from its registers and posts it (v and tau,
    // double-buffered by step parity) in LDS; one barrier; every later column takes the
    // reflector, writes its final row c to R or Qt and shifts its window up by one row, the new
    // row c + BWT + 1 coming from the stage.  A bandwidth below BWT gives reflectors whose tail
    // past row c + bw is zero: those entries add exact zeros, and the window rows they would
    // touch keep their values.
    constexpr int WS = BWT + 1;  // rows held per column
    double x[BWT + 1];
    const int col = tid;
    const bool own = col < nc;
#pragma unroll
    for (int l = 0; l <= BWT; ++l) x[l] = (own && l < rows) ? src(l, col) : 0.0;
    double *stage = sm;                       // kLsqChunk x nc
    double *vbuf = sm + kLsqChunk * nc;       // 2 x 64: v[0 .. BWT - 1], tau at [BWT]
    double pre[kLsqChunk];
#pragma unroll
    for (int q = 0; q < kLsqChunk; ++q) pre[q] = (own && WS + q < rows) ? src(WS + q, col) : 0.0;
    for (int cb = 0; cb < cols; cb += kLsqChunk) {
      // Rows cb + WS + q for the steps cb + q (each thread reads back only its own column, so no
      // barrier guards the stage); then the next chunk's rows are fetched.
#pragma unroll
      for (int q = 0; q < kLsqChunk; ++q)
        if (own) stage[q * nc + col] = pre[q];
#pragma unroll
      for (int q = 0; q < kLsqChunk; ++q) {
        const int r = cb + kLsqChunk + WS + q;
        pre[q] = (own && r < rows) ? src(r, col) : 0.0;
      }
      const int ce = cb + kLsqChunk < cols ? cb + kLsqChunk : cols;
      for (int c = cb; c < ce; ++c) {
        double *vb = vbuf + (c & 1) * 64;
        if (col == c) {
          const int L = (c + bw < rows - 1 ? c + bw : rows - 1) - c;  // reflector tail length
          double xn2 = 0.0;
#pragma unroll
          for (int l = 1; l <= BWT; ++l)
            if (l <= L) xn2 = fma(x[l], x[l], xn2);
          const double alpha = x[0];
          double tau = 0.0, sc = 0.0, beta = alpha;
          if (xn2 != 0.0) {
            beta = -copysign(sqrt(fma(alpha, alpha, xn2)), alpha);
            tau = (beta - alpha) / beta;
            sc = 1.0 / (alpha - beta);
          }
#pragma unroll
          for (int l = 1; l <= BWT; ++l) vb[l - 1] = (xn2 != 0.0 && l <= L) ? x[l] * sc : 0.0;
          vb[BWT] = tau;
          R[(long long)c * cols + c] = beta;
        }
        __syncthreads();
        if (own && col > c) {
          const double tau = vb[BWT];
          if (tau != 0.0) {
            double v[BWT];
#pragma unroll
            for (int l = 0; l < BWT; ++l) v[l] = vb[l];
            double w = x[0];
#pragma unroll
            for (int l = 0; l < BWT; ++l) w = fma(v[l], x[l + 1], w);
            const double tw = tau * w;
            x[0] = x[0] - tw;
#pragma unroll
            for (int l = 0; l < BWT; ++l) x[l + 1] = fma(-tw, v[l], x[l + 1]);
          }
          if (col < cols)
            R[(long long)col * cols + c] = x[0];
          else
            Qt[(long long)c * 32 + (col - cols)] = x[0];
#pragma unroll
          for (int l = 0; l < BWT; ++l) x[l] = x[l + 1];
          x[BWT] = stage[(c - cb) * nc + col];
        }
      }
    }
    __syncthreads();
  } else {
    for (int e = tid; e < W * nc; e += bd) {
      const int r = e / nc, col = e % nc;
      win[e] = r >= rows ? 0.0 : src(r, col);
    }
    __syncthreads();
    for (int c = 0; c < cols; ++c) {
      const int L = (c + bw < rows - 1 ? c + bw : rows - 1) - c;  // reflector tail length
      double *rowc = win + (c % W) * nc;
      // The row entering the window after this step, fetched now.
      const int rn = c + W;
      const double incoming = (tid < nc && rn < rows) ? src(rn, tid) : 0.0;
      if (tid < 64) {
        double x = 0.0;
        if (tid >= 1 && tid <= L) x = win[((c + tid) % W) * nc + c];
        double xn2 = x * x;
        xn2 = xor_sum<64>(xn2);
        const double alpha = rowc[c];
        double tau = 0.0, sc = 0.0, beta = alpha;
        if (xn2 != 0.0) {
          beta = -copysign(sqrt(fma(alpha, alpha, xn2)), alpha);
          tau = (beta - alpha) / beta;
          sc = 1.0 / (alpha - beta);
        }
        if (xn2 != 0.0 && tid >= 1 && tid <= L) win[((c + tid) % W) * nc + c] = x * sc;
        if (tid == 0) {
          rowc[c] = beta;
          s_tau = tau;
        }
      }
      __syncthreads();
      const double tau = s_tau;
      if (tau != 0.0) {
        const int c0 = c % W;
        for (int col = c + 1 + tid; col < nc; col += bd) {
          if constexpr (BWT > 0) {
            // All loads of the column's window entries issued before the sums.
            double vr[BWT], xr[BWT];
  #pragma unroll
            for (int l = 1; l <= BWT; ++l) {
              const int sl = c0 + l < W ? c0 + l : c0 + l - W;
              vr[l - 1] = win[sl * nc + c];
              xr[l - 1] = win[sl * nc + col];
            }
            double w = rowc[col];
  #pragma unroll
            for (int l = 0; l < BWT; ++l) w = fma(vr[l], xr[l], w);
            const double tw = tau * w;
            rowc[col] = rowc[col] - tw;
  #pragma unroll
            for (int l = 1; l <= BWT; ++l) {
              const int sl = c0 + l < W ? c0 + l : c0 + l - W;
              win[sl * nc + col] = fma(-tw, vr[l - 1], xr[l - 1]);
            }
          } else {
            double w = rowc[col];
            for (int l = 1; l <= L; ++l) {
              const int sl = c0 + l < W ? c0 + l : c0 + l - W;
              const double *rl = win + sl * nc;
              w = fma(rl[c], rl[col], w);
            }
            const double tw = tau * w;
            rowc[col] = rowc[col] - tw;
            for (int l = 1; l <= L; ++l) {
              const int sl = c0 + l < W ? c0 + l : c0 + l - W;
              double *rl = win + sl * nc;
              rl[col] = fma(-tw, rl[c], rl[col]);
            }
          }
        }
      }
      __syncthreads();
      for (int col = tid; col < nc; col += bd) {
        const double v = rowc[col];
        if (col >= c && col < cols) R[(long long)col * cols + c] = v;
        if (col >= cols) Qt[(long long)c * 32 + (col - cols)] = v;
        rowc[col] = col == tid ? incoming : (rn < rows ? src(rn, col) : 0.0);
      }
      __syncthreads();
    }
  }
  // Back substitution R Y = Qt in 32-row blocks; a negligible pivot leaves its unknown at zero.
  double *Tl = sm;              // cols x 32: right-hand sides, updated in place
  double *Dg = sm + cols * 32;  // 32 x 33: diagonal block of R (row-major, padded)
  __shared__ double Yb[32 * 32];
  __shared__ double Rinv[32];
  for (int e = tid; e < cols * 32; e += bd) Tl[e] = (e & 31) < s ? Qt[e] : 0.0;
  if (tid < 64) {
    double m = 0.0;
    for (int c = tid; c < cols; c += 64) m = fmax(m, fabs(R[(long long)c * cols + c]));
    m = xor_max64(m);
    if (tid == 0) s_rmax = m;
  }
  __syncthreads();
  const double piv = 1e-14 * s_rmax;
  for (int i1 = cols; i1 > 0; i1 -= 32) {
    const int i0 = i1 - 32 > 0 ? i1 - 32 : 0, b = i1 - i0;
    for (int e = tid; e < b * b; e += bd) {
      const int r = e % b, l = e / b;  // R(i0 + r, i0 + l)
      const double v = l >= r ? R[(long long)(i0 + l) * cols + i0 + r] : 0.0;
      Dg[r * 33 + l] = v;
      // The pivots' reciprocals (used only where the pivot is not negligible), here so that the
      // solve below multiplies instead of dividing on its dependency chain.
      if (l == r) Rinv[r] = 1.0 / v;
    }
    __syncthreads();
    if (tid < 32) {
      // Lane jr solves column jr of the diagonal block, bottom up, its unknowns in registers
      // (the rows r >= b of a partial block do not exist).
      const int jr = tid;
      double yv[32];
#pragma unroll
      for (int r = 31; r >= 0; --r) {
        yv[r] = 0.0;
        if (r < b) {
          double a = Tl[(i0 + r) * 32 + jr];
          // The row's entries in two halves, each loaded at once before its multiply-adds.
#pragma unroll
          for (int h0 = r + 1; h0 < 32; h0 += 16) {
            double dr[16];
#pragma unroll
            for (int l = h0; l < h0 + 16 && l < 32; ++l) dr[l - h0] = Dg[r * 33 + l];
#pragma unroll
            for (int l = h0; l < h0 + 16 && l < 32; ++l)
              if (l < b) a = fma(-dr[l - h0], yv[l], a);
          }
          const double d = Dg[r * 33 + r];
          yv[r] = (jr < s && fabs(d) > piv) ? a * Rinv[r] : 0.0;
          Yb[r * 32 + jr] = yv[r];
          Y[(long long)(i0 + r) * 32 + jr] = yv[r];
        }
      }
    }
    __syncthreads();
    // Rows above the block: Tl[r][:] -= sum_l R(r, i0 + l) Y(i0 + l, :) over the SW columns that
    // can be nonzero, a thread per row (the loads of a column of R are contiguous across threads;
    // Y is read by broadcast).
    for (int r = tid; r < i0; r += bd) {
      double a[SW];
#pragma unroll
      for (int jr = 0; jr < SW; ++jr) a[jr] = Tl[r * 32 + jr];
#pragma unroll 1
      for (int l0 = 0; l0 < 32; l0 += 16) {
        double rv[16];  // sixteen loads of R in flight
#pragma unroll
        for (int q = 0; q < 16; ++q)
          rv[q] = l0 + q < b ? R[(long long)(i0 + l0 + q) * cols + r] : 0.0;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          if (l0 + q < b) {
#pragma unroll
            for (int jr = 0; jr < SW; ++jr) a[jr] = fma(-rv[q], Yb[(l0 + q) * 32 + jr], a[jr]);
          }
        }
      }
#pragma unroll
      for (int jr = 0; jr < SW; ++jr) Tl[r * 32 + jr] = a[jr];
    }
    __syncthreads();
  }
  for (int e = tid; e < cols * 32; e += bd) {
    const int j = e & 31;
    Yr32[e] = j < s ? (float)(Y[e] * rho[j]) : 0.0f;
  }
  // ZH = H Y with Y staged in LDS: thread (r, g) forms columns 8g .. 8g + 7 of row r (SW / 8
  // groups), sixteen loads of H in flight.
  double *Yl = sm;
  for (int e = tid; e < cols * 32; e += bd) Yl[e] = Y[e];
  __syncthreads();
  constexpr int NG = SW / 8;
  for (int e = tid; e < rows * NG; e += bd) {
    const int r = e % rows, g = e / rows;
    double z[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) z[j] = 0.0;
    for (int c0 = 0; c0 < cols; c0 += 16) {
      double hv[16];
#pragma unroll
      for (int q = 0; q < 16; ++q)
        hv[q] = c0 + q < cols ? H[(long long)(c0 + q) * rows + r] : 0.0;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        if (c0 + q < cols) {
#pragma unroll
          for (int j = 0; j < 8; ++j) z[j] = fma(hv[q], Yl[(c0 + q) * 32 + 8 * g + j], z[j]);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) ZH32[r * 32 + 8 * g + j] = 8 * g + j < s ? (float)z[j] : 0.0f;
  }
  if constexpr (SW < 32) {
    for (int e = tid; e < rows * (32 - SW); e += bd)  // columns no group forms
      ZH32[(e / (32 - SW)) * 32 + SW + e % (32 - SW)] = 0.0f;
  }
}

// Coefficients of the paired update (gemm_nn_pair), rows x 32 row-major: columns 0 .. 15 are
// ZH32's, columns 16 .. 16 + s - 1 are -Yr32's on rows p0 .. cols - 1 and zero elsewhere.
__global__ void k_pack_pair(const float *__restrict__ ZH32, const float *__restrict__ Yr32, int p0,
                            int cols, int rows, int s, float *__restrict__ Cp) {
  const int e = blockIdx.x * kThreads + threadIdx.x;
  if (e >= rows * 32) return;
  const int r = e >> 5, j = e & 31;
  float v = 0.0f;
  if (j < 16)
    v = ZH32[e];
  else if (j - 16 < s && r >= p0 && r < cols)
    v = -Yr32[r * 32 + (j - 16)];
  Cp[e] = v;
}

// Converts the k selected right singular vectors C (cols x k, ld cols, FP64) to FP32 row-major
// cols x 32.
__global__ void k_pack_coef(const double *__restrict__ C, int cols, int k, float *__restrict__ C32) {
  const int e = blockIdx.x * kThreads + threadIdx.x;
  if (e >= cols * 32) return;
  const int i = e >> 5, j = e & 31;
  C32[e] = j < k ? (float)C[(long long)j * cols + i] : 0.0f;
}

// rho[j] *= q[j]; scal[j] = 1 / q[j] (0 for q = 0): residual tracking per column.
__global__ void k_track(double *__restrict__ rho, const double *__restrict__ q,
                        double *__restrict__ scal, int s) {
  const int j = threadIdx.x;
  if (j >= s) return;
  rho[j] = rho[j] * q[j];
  scal[j] = q[j] > 0.0 ? 1.0 / q[j] : 0.0;
}

// x[:, j] += f[j] * d[:, j]  (FP64 x, FP32 d).
__global__ __launch_bounds__(kThreads) void k_add_scaled(double *__restrict__ X, long long ldx,
                                                         const float *__restrict__ D,
                                                         long long ldd,
                                                         const double *__restrict__ f, int s,
                                                         long long n) {
  const long long r = (long long)blockIdx.x * kThreads + threadIdx.x;
  const int j = blockIdx.y;
  if (r >= n || j >= s) return;
  X[j * ldx + r] = fma(f[j], (double)D[j * ldd + r], X[j * ldx + r]);
}

// D[:, j] = (FP32) (R[:, j] * f[j])
__global__ __launch_bounds__(kThreads) void k_scale_to32(const double *__restrict__ R,
                                                         long long ldr, const double *__restrict__ f,
                                                         float *__restrict__ D, long long ldd,
                                                         int s, long long n) {
  const long long r = (long long)blockIdx.x * kThreads + threadIdx.x;
  const int j = blockIdx.y;
  if (r >= n || j >= s) return;
  D[j * ldd + r] = (float)(R[j * ldr + r] * f[j]);
}

// out[j] = 1 / q[j] (0 for q = 0), and optionally rho[j] = 1: per-column scale factors.
__global__ void k_recip(const double *__restrict__ q, double *__restrict__ out, int s,
                        double *__restrict__ rho) {
  const int j = threadIdx.x;
  if (j >= s) return;
  out[j] = q[j] > 0.0 ? 1.0 / q[j] : 0.0;
  if (rho) rho[j] = 1.0;
}

__global__ void k_zero_ints(int *__restrict__ p, int cnt) {
  const int i = threadIdx.x;
  if (i < cnt) p[i] = 0;
}

inline unsigned grid1(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

}  // namespace

void prepare() {
  // Once per process, before any stream capture: the LDS limits of the one-workgroup kernels.
  static const bool done = [] {
    dmr::allow_max_lds(reinterpret_cast<const void *>(k_lsq_block<0, false>));
    dmr::allow_max_lds(reinterpret_cast<const void *>(k_lsq_block<16, false>));
    dmr::allow_max_lds(reinterpret_cast<const void *>(k_lsq_block<16, true>));
    dmr::allow_max_lds(reinterpret_cast<const void *>(k_lsq_block<32, true>));
    dmr::prepare_dense();
    return true;
  }();
  (void)done;
}

void recip(const double *q, double *out, int s, double *rho, hipStream_t st) {
  hipLaunchKernelGGL(k_recip, dim3(1), dim3(64), 0, st, q, out, s, rho);
}

void zero_ints(int *p, int cnt, hipStream_t st) {
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, st, p, cnt);
}

template <typename TX>
void spmm_upfront(const CsrDev &A, const TX *X, long long ldx, float *Y, long long ldy, int s,
                  int g, hipStream_t st) {
  if (g >= 100) {
    // LDS-staged nonzeros, rows of up to 7 entries unrolled.
    g -= 100;
    const dim3 grid(grid1(A.n), (unsigned)((s + g - 1) / g));
    if (g == 16)
      hipLaunchKernelGGL((k_spmm_lds<float, TX, float, float, 16, 7>), grid, dim3(kThreads), 0, st,
                         A.n, A.rowPtr, A.colInd, A.val32, X, ldx, Y, ldy, s);
    else
      hipLaunchKernelGGL((k_spmm_lds<float, TX, float, float, 8, 7>), grid, dim3(kThreads), 0, st,
                         A.n, A.rowPtr, A.colInd, A.val32, X, ldx, Y, ldy, s);
    return;
  }
  const dim3 grid(grid1(A.n), (unsigned)((s + g - 1) / g));
  if (g == 4)
    hipLaunchKernelGGL((k_spmm_w<float, TX, float, float, 4, 8>), grid, dim3(kThreads), 0, st, A.n,
                       A.rowPtr, A.colInd, A.val32, X, ldx, Y, ldy, s);
  else if (g == 16)
    hipLaunchKernelGGL((k_spmm_w<float, TX, float, float, 16, 8>), grid, dim3(kThreads), 0, st,
                       A.n, A.rowPtr, A.colInd, A.val32, X, ldx, Y, ldy, s);
  else
    hipLaunchKernelGGL((k_spmm_w<float, TX, float, float, 8, 8>), grid, dim3(kThreads), 0, st, A.n,
                       A.rowPtr, A.colInd, A.val32, X, ldx, Y, ldy, s);
}

void spmm32(const CsrDev &A, const bf16 *X, long long ldx, float *Y, long long ldy, int s, int cg,
            hipStream_t st) {
  spmm_upfront<bf16>(A, X, ldx, Y, ldy, s, cg >= 100 ? cg - 100 : cg, st);
}

void spmm32(const CsrDev &A, const float *X, long long ldx, float *Y, long long ldy, int s,
            int cg, hipStream_t st) {
  if (cg >= 100) {
    // Up-front index variant: cg - 100 columns per thread, rows of up to 8 entries unrolled
    // (cg >= 200: the LDS-staged one, up to 7).
    spmm_upfront<float>(A, X, ldx, Y, ldy, s, cg - 100, st);
    return;
  }
  const dim3 grid(grid1(A.n), (unsigned)((s + cg - 1) / cg));
  if (cg == 4)
    hipLaunchKernelGGL((k_spmm<float, float, float, float, 4>), grid, dim3(kThreads), 0, st, A.n,
                       A.rowPtr, A.colInd, A.val32, X, ldx, Y, ldy, s);
  else if (cg == 16)
    hipLaunchKernelGGL((k_spmm<float, float, float, float, 16>), grid, dim3(kThreads), 0, st, A.n,
                       A.rowPtr, A.colInd, A.val32, X, ldx, Y, ldy, s);
  else
    hipLaunchKernelGGL((k_spmm<float, float, float, float, 8>), grid, dim3(kThreads), 0, st, A.n,
                       A.rowPtr, A.colInd, A.val32, X, ldx, Y, ldy, s);
}

void resid_block(const CsrDev &A, const double *val, long long ldv, const double *X,
                 long long ldx, const double *B, long long ldb, double *R, long long ldr, int s,
                 hipStream_t st) {
  hipLaunchKernelGGL((k_resid_block<4, 7>), dim3(grid1(A.n), (s + 3) / 4), dim3(kThreads), 0, st,
                     A.n, A.rowPtr, A.colInd, val ? val : A.val, val ? ldv : 0LL, X, ldx, B, ldb,
                     R, ldr, s);
}

void mean_values(const double *val, long long ldv, int count, long long nnz, float *out,
                 hipStream_t st) {
  long long g = (nnz + kThreads - 1) / kThreads;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  hipLaunchKernelGGL(k_mean_values, dim3((unsigned)g), dim3(kThreads), 0, st, val, ldv, count,
                     nnz, out);
}

int colnorm_blocks(long long n) {
  long long g = (n + 4095) / 4096;
  if (g < 1) g = 1;
  if (g > 256) g = 256;
  return (int)g;
}

template <typename T>
void colnorms(const T *X, long long ld, int s, long long n, double *partial, double *out,
              int sqrtMode, hipStream_t st) {
  const int G = colnorm_blocks(n);
  const long long rows = (n + G - 1) / G;
  hipLaunchKernelGGL((k_colsumsq<T>), dim3(G, s), dim3(kThreads), 0, st, X, ld, s, n, rows,
                     partial);
  hipLaunchKernelGGL(k_colsumsq_final, dim3(1), dim3(kThreads), 0, st, partial, G, s, out,
                     sqrtMode);
}
template void colnorms<float>(const float *, long long, int, long long, double *, double *, int,
                              hipStream_t);
template void colnorms<double>(const double *, long long, int, long long, double *, double *, int,
                               hipStream_t);

int gemm_tn_blocks(long long n, int cap, int rowsTarget) {
  long long g = (n + rowsTarget - 1) / rowsTarget;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

size_t gemm_tn_lds(int pPad) { return sizeof(float) * (size_t)(pPad + 32) * kLdsStride; }

template <typename TP>
void gemm_tn16(const TP *P, long long ldp, int p, int pPad, const float *U, long long ldu, int s,
               long long n, int G, long long rows, float *part, hipStream_t st) {
  // Tile groups: one group of up to 6 tiles (four row groups), else two, else four.
  const int nT = (p + 15) / 16;
  const int nTG = nT <= 6 ? 1 : (nT <= 12 ? 2 : 4);
  const int tpw = (nT + nTG - 1) / nTG;
  const int nu = s <= 16 ? 1 : 2;
#define DMR_TN16(T)                                                                              \
  if (tpw == T) {                                                                                \
    if (nu == 1)                                                                                 \
      hipLaunchKernelGGL((k_gemm_tn16<TP, T, 1>), dim3(G), dim3(kThreads), 0, st, P, ldp, p,     \
                         pPad, U, ldu, s, n, rows, part);                                        \
    else                                                                                         \
      hipLaunchKernelGGL((k_gemm_tn16<TP, T, 2>), dim3(G), dim3(kThreads), 0, st, P, ldp, p,     \
                         pPad, U, ldu, s, n, rows, part);                                        \
  }
  DMR_TN16(1)
  DMR_TN16(2)
  DMR_TN16(3)
  DMR_TN16(4)
  DMR_TN16(5)
  DMR_TN16(6)
#undef DMR_TN16
}

void gram_reduce(const float *part, int G, int s, double *G64, hipStream_t st) {
  tn_reduce(part, G, 32, s, s, G64, s, nullptr, st);
}

long long tn_rows(long long n, int G) {
  long long rows = (n + G - 1) / G;
  return (rows + 63) / 64 * 64;
}

// bf16 P with the matrix-core kernel (variant 3) or the FP32 one (else); U2 as for k_gemm_tn16b
// (variant 3, s <= 16 columns of U and s2 of U2), the product of U2 then going to C64b.
void gemm_tn_b(const bf16 *P, long long ldp, int p, const float *U, const float *U2,
               long long ldu, int s, long long n, int G, int variant, float *part, double *C64,
               int ldc, float *C32, double *C64b, int ldcb, int s2, hipStream_t st, bool reduce) {
  const int pPad = (p + 31) & ~31;
  const long long rows = tn_rows(n, G);
  if (variant == 3) {
    const int nT = (p + 15) / 16;
    const int nTG = nT <= 6 ? 1 : (nT <= 12 ? 2 : 4);
    const int tpw = (nT + nTG - 1) / nTG;
    const int nu = (s <= 16 && !U2) ? 1 : 2;
#define DMR_TN16B(T)                                                                             \
  if (tpw == T) {                                                                                \
    if (nu == 1)                                                                                 \
      hipLaunchKernelGGL((k_gemm_tn16b<T, 1>), dim3(G), dim3(kThreads), 0, st, P, ldp, p, pPad,  \
                         U, ldu, s, n, rows, part, U2, s2);                                      \
    else                                                                                         \
      hipLaunchKernelGGL((k_gemm_tn16b<T, 2>), dim3(G), dim3(kThreads), 0, st, P, ldp, p, pPad,  \
                         U, ldu, s, n, rows, part, U2, s2);                                      \
  }
    DMR_TN16B(1)
    DMR_TN16B(2)
    DMR_TN16B(3)
    DMR_TN16B(4)
    DMR_TN16B(5)
    DMR_TN16B(6)
#undef DMR_TN16B
  } else {
    gemm_tn16<bf16>(P, ldp, p, pPad, U, ldu, s, n, G, rows, part, st);
  }
  if (reduce) tn_reduce(part, G, pPad, p, s, C64, ldc, C32, st, C64b, ldcb, U2 ? 16 : 0, s2);
}

void gemm_tn(const bf16 *P, long long ldp, int p, const float *U, long long ldu, int s,
             long long n, int G, int variant, float *part, double *C64, int ldc, float *C32,
             hipStream_t st, bool reduce) {
  gemm_tn_b(P, ldp, p, U, nullptr, ldu, s, n, G, variant, part, C64, ldc, C32, nullptr, 0, 0, st,
            reduce);
}

void gemm_tn_pair(const bf16 *P, long long ldp, int p, const float *U, const float *U2,
                  long long ldu, int s, long long n, int G, float *part, double *C64a, int ldca,
                  float *C32, double *C64b, int ldcb, hipStream_t st, bool reduce) {
  gemm_tn_b(P, ldp, p, U, U2, ldu, s, n, G, 3, part, C64a, ldca, C32, C64b, ldcb, s, st, reduce);
}

void gemm_tn_reduce(const float *part, int G, int p, int s, double *C64, int ldc, float *C32,
                    hipStream_t st, bool pair, double *C64b, int ldcb) {
  tn_reduce(part, G, (p + 31) & ~31, p, s, C64, ldc, C32, st, C64b, ldcb, pair ? 16 : 0,
            pair ? s : 0);
}

void gemm_tn(const float *P, long long ldp, int p, const float *U, long long ldu, int s,
             long long n, int G, int variant, float *part, double *C64, int ldc, float *C32,
             hipStream_t st, bool reduce) {
  const int pPad = (p + 31) & ~31;
  const long long rows = tn_rows(n, G);
  if (variant == 2) {
    gemm_tn16<float>(P, ldp, p, pPad, U, ldu, s, n, G, rows, part, st);
  } else if (variant == 1) {
    const size_t lds = gemm_tn_lds(pPad);
    hipLaunchKernelGGL(k_gemm_tn_lds, dim3(G), dim3(kThreads), lds, st, P, ldp, p, pPad, U, ldu,
                       s, n, rows, part);
  } else {
    hipLaunchKernelGGL(k_gemm_tn, dim3(G), dim3(kThreads), 0, st, P, ldp, p, pPad, U, ldu, s, n,
                       rows, part);
  }
  if (reduce) tn_reduce(part, G, pPad, p, s, C64, ldc, C32, st);
}

template <typename TP, typename TO>
void gemm_nn_t(const TP *P, long long ldp, int p, const float *C, const float *In, long long ldi,
               float a, float b, TO *Out, long long ldo, int s, long long n, int un,
               hipStream_t st) {
  const long long waves = (n + 127) / 128;
  const unsigned grid = (unsigned)((waves + 3) / 4);
  const int D = un >= 16 ? 16 : (un >= 8 ? 8 : (un >= 4 ? 4 : 2));
  const int pPad = (p + 2 * D - 1) / (2 * D) * (2 * D) + 2 * D;
  const size_t lds = sizeof(float) * (size_t)pPad * 32;
  if (D == 16)
    hipLaunchKernelGGL((k_gemm_nn<16, TP, TO>), dim3(grid), dim3(kThreads), lds, st, P, ldp, p, C,
                       In, ldi, a, b, Out, ldo, s, n);
  else if (D == 8)
    hipLaunchKernelGGL((k_gemm_nn<8, TP, TO>), dim3(grid), dim3(kThreads), lds, st, P, ldp, p, C,
                       In, ldi, a, b, Out, ldo, s, n);
  else if (D == 4)
    hipLaunchKernelGGL((k_gemm_nn<4, TP, TO>), dim3(grid), dim3(kThreads), lds, st, P, ldp, p, C,
                       In, ldi, a, b, Out, ldo, s, n);
  else
    hipLaunchKernelGGL((k_gemm_nn<2, TP, TO>), dim3(grid), dim3(kThreads), lds, st, P, ldp, p, C,
                       In, ldi, a, b, Out, ldo, s, n);
}

void gemm_nn(const float *P, long long ldp, int p, const float *C, const float *In, long long ldi,
             float a, float b, float *Out, long long ldo, int s, long long n, int un,
             hipStream_t st) {
  gemm_nn_t<float, float>(P, ldp, p, C, In, ldi, a, b, Out, ldo, s, n, un, st);
}
void gemm_nn_pair(const bf16 *P, long long ldp, int p, const float *C, float *X1, float *X2,
                  long long ld, int s, long long n, int un, hipStream_t st) {
  const int RB = un >= 101 ? 2 : 1;
  const long long waves = (n + 128 * RB - 1) / (128 * RB);
  const unsigned grid = (unsigned)((waves + 3) / 4);
  const int KP = (p + 15) & ~15;
  const size_t epi = sizeof(float) * 4 * kNnbTile;
  const size_t l2 = sizeof(uint16_t) * 2 * 2 * 16 * KP;
  const size_t lds = l2 > epi ? l2 : epi;
  const int sp = 16 + s;  // output columns: X1's s, then X2's from column 16
  if (RB == 2)
    hipLaunchKernelGGL((k_gemm_nn_b<2, float, 2>), dim3(grid), dim3(kThreads), lds, st, P, ldp, p,
                       C, X1, ld, 1.0f, -1.0f, X1, ld, sp, n, nullptr, X2, X2);
  else
    hipLaunchKernelGGL((k_gemm_nn_b<2, float, 1>), dim3(grid), dim3(kThreads), lds, st, P, ldp, p,
                       C, X1, ld, 1.0f, -1.0f, X1, ld, sp, n, nullptr, X2, X2);
}

void pack_pair(const float *ZH32, const float *Yr32, int p0, int cols, int rows, int s, float *Cp,
               hipStream_t st) {
  hipLaunchKernelGGL(k_pack_pair, dim3(grid1((long long)rows * 32)), dim3(kThreads), 0, st, ZH32,
                     Yr32, p0, cols, rows, s, Cp);
}

int gemm_nn_gram_blocks(long long n, int un) {
  const int RB = un >= 101 ? 2 : 1;
  const long long waves = (n + 128 * RB - 1) / (128 * RB);
  return (int)((waves + 3) / 4);
}

void gemm_nn(const bf16 *P, long long ldp, int p, const float *C, const float *In, long long ldi,
             float a, float b, float *Out, long long ldo, int s, long long n, int un,
             hipStream_t st, float *gram) {
  if (un >= 100) {
    // bf16 matrix cores: 128 (un = 100) or 256 (un = 101) rows per wavefront, four per workgroup.
    const int RB = un >= 101 ? 2 : 1;
    const long long waves = (n + 128 * RB - 1) / (128 * RB);
    const unsigned grid = (unsigned)((waves + 3) / 4);
    const int KP = (p + 15) & ~15;
    const size_t epi = sizeof(float) * 4 * kNnbTile;
    const size_t l1 = sizeof(uint16_t) * 2 * 16 * KP, l2 = 2 * l1;
    const size_t lds1 = l1 > epi ? l1 : epi, lds2 = l2 > epi ? l2 : epi;
    if (s <= 16 && RB == 2)
      hipLaunchKernelGGL((k_gemm_nn_b<1, float, 2>), dim3(grid), dim3(kThreads), lds1, st, P, ldp,
                         p, C, In, ldi, a, b, Out, ldo, s, n, gram, nullptr,
                         (float *)nullptr);
    else if (s <= 16)
      hipLaunchKernelGGL((k_gemm_nn_b<1, float, 1>), dim3(grid), dim3(kThreads), lds1, st, P, ldp,
                         p, C, In, ldi, a, b, Out, ldo, s, n, gram, nullptr,
                         (float *)nullptr);
    else if (RB == 2)
      hipLaunchKernelGGL((k_gemm_nn_b<2, float, 2>), dim3(grid), dim3(kThreads), lds2, st, P, ldp,
                         p, C, In, ldi, a, b, Out, ldo, s, n, nullptr, nullptr,
                         (float *)nullptr);
    else
      hipLaunchKernelGGL((k_gemm_nn_b<2, float, 1>), dim3(grid), dim3(kThreads), lds2, st, P, ldp,
                         p, C, In, ldi, a, b, Out, ldo, s, n, nullptr, nullptr,
                         (float *)nullptr);
    return;
  }
  gemm_nn_t<bf16, float>(P, ldp, p, C, In, ldi, a, b, Out, ldo, s, n, un, st);
}
template <typename TO>
void scale_block_t(const float *U, long long ldu, const float *C, TO *Out, long long ldo, int s,
                   long long n, hipStream_t st) {
  const long long pairs = (n + 1) / 2;
  hipLaunchKernelGGL(k_scale_block<TO>, dim3((unsigned)((pairs + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, st, U, ldu, C, Out, ldo, s, n);
}
bool scale_block(const float *U, long long ldu, const float *C, float *Out, long long ldo, int s,
                 long long n, hipStream_t st) {
  if (s > 16 || (ldu & 1) || (ldo & 1)) return false;
  scale_block_t<float>(U, ldu, C, Out, ldo, s, n, st);
  return true;
}
bool scale_block(const float *U, long long ldu, const float *C, bf16 *Out, long long ldo, int s,
                 long long n, hipStream_t st) {
  if (s > 16 || (ldu & 1) || (ldo & 1)) return false;
  scale_block_t<bf16>(U, ldu, C, Out, ldo, s, n, st);
  return true;
}

void gemm_nn(const float *P, long long ldp, int p, const float *C, const float *In, long long ldi,
             float a, float b, bf16 *Out, long long ldo, int s, long long n, int un,
             hipStream_t st) {
  gemm_nn_t<float, bf16>(P, ldp, p, C, In, ldi, a, b, Out, ldo, s, n, un, st);
}

int gram_chol_segments(int G) {
  const int nb = (G + 63) / 64;
  return nb < kGcSegs ? nb : kGcSegs;
}

void gram_chol(const float *part, int G, int s, double relTol, float *Sinv32, double *Sacc,
               int first, double *Hdst, int ldh, int *brk, double *seg, unsigned *ctr,
               hipStream_t st) {
  const int nb = gram_chol_segments(G);
  const int L = (G + nb - 1) / nb;
  hipLaunchKernelGGL(k_gram_chol, dim3(nb), dim3(256), 0, st, part, G, L, s, relTol, Sinv32,
                     Sacc, first, Hdst, ldh, brk, seg, ctr);
}

void chol_qr(const double *G, int s, double relTol, float *Sinv32, double *Sacc, int first,
             double *Hdst, int ldh, int *brk, hipStream_t st) {
  if (s <= 16)
    hipLaunchKernelGGL(k_chol_qr_el<16>, dim3(1), dim3(256), 0, st, G, s, relTol, Sinv32, Sacc,
                       first, Hdst, ldh, brk);
  else
    hipLaunchKernelGGL(k_chol_qr_el<32>, dim3(1), dim3(1024), 0, st, G, s, relTol, Sinv32, Sacc,
                       first, Hdst, ldh, brk);
}

// LDS of the least squares on the register path (reg) or the window path.
size_t lsq_lds(int cols, int bw, int s, bool reg) {
  const size_t win = reg ? (size_t)kLsqChunk * (cols + s) + 128 : (size_t)(bw + 1) * (cols + s);
  const size_t back = (size_t)cols * 32 + 32 * 33;
  return sizeof(double) * (win > back ? win : back);
}
// The register path: one thread per column of [H | T], its LDS fits; bandwidth up to 16 on the
// 16-row window (1024 threads), up to 32 on the 32-row window (512 threads: the window rows of
// a column take 66 registers).  DMR_LSQ_WINDOW=1 keeps every shape on the window path.
bool lsq_staged(int cols, int bw, int s) {
  if (const char *e = std::getenv("DMR_LSQ_WINDOW"))
    if (e[0] == '1') return false;
  const int threads = bw > 16 ? 512 : 1024;
  return ((bw <= 16 && s <= 16) || (bw <= 32 && s <= 32)) && bw >= 1 && cols + s <= threads &&
         lsq_lds(cols, bw, s, true) <= 140 * 1024;
}

size_t lsq_block_lds(int cols, int bw, int s) {
  return lsq_lds(cols, bw, s, lsq_staged(cols, bw, s));
}

void lsq_block(const double *H, int rows, int cols, int bw, const double *T, int s,
               const double *rho, double *R, double *Qt, double *Y, float *Yr32, float *ZH32,
               hipStream_t st) {
  const size_t lds = lsq_block_lds(cols, bw, s);
  const bool stg = lsq_staged(cols, bw, s);
  prepare();
#define DMR_LSQ(B, S)                                                                            \
  hipLaunchKernelGGL((k_lsq_block<B, S>), dim3(1), dim3(B == 32 ? 512 : 1024), lds, st, H, rows, \
                     cols, bw, T, s, rho, R, Qt, Y, Yr32, ZH32)
  if (stg && bw > 16)
    DMR_LSQ(32, true);
  else if (stg)
    DMR_LSQ(16, true);
  else if (bw == 16 && s <= 16)
    DMR_LSQ(16, false);
  else
    DMR_LSQ(0, false);
#undef DMR_LSQ
}

void pack_coef(const double *C, int cols, int k, float *C32, hipStream_t st) {
  hipLaunchKernelGGL(k_pack_coef, dim3(grid1((long long)cols * 32)), dim3(kThreads), 0, st, C,
                     cols, k, C32);
}

void track(double *rho, const double *q, double *scal, int s, hipStream_t st) {
  hipLaunchKernelGGL(k_track, dim3(1), dim3(64), 0, st, rho, q, scal, s);
}

void add_scaled(double *X, long long ldx, const float *D, long long ldd, const double *f, int s,
                long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_add_scaled, dim3(grid1(n), s), dim3(kThreads), 0, st, X, ldx, D, ldd, f,
                     s, n);
}

void scale_to32(const double *R, long long ldr, const double *f, float *D, long long ldd, int s,
                long long n, hipStream_t st) {
  hipLaunchKernelGGL(k_scale_to32, dim3(grid1(n), s), dim3(kThreads), 0, st, R, ldr, f, D, ldd,
                     s, n);
}

}  // namespace blk
}  // namespace dmr
