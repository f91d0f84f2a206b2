// Streaming HIP kernels of the DMR solve for CDNA4 (gfx950): CSR SpMV, multi-dot (GEMV-T),
// single- and multi-output basis combinations with fused reductions, scaling and fills.
//
// Design rules shared by every kernel here:
//  * 256-thread blocks (four 64-wide wavefronts); vectors are walked in 16-byte units
//    (2 doubles / 4 floats / 8 bf16) so every load is a dwordx4 and every wavefront touches
//    1 KiB contiguous per column per step.
//  * Vectors live in padded columns (ld a multiple of 256 elements, padding kept at zero), so a
//    unit never straddles the end of a column and needs no tail handling.
//  * Reductions are fixed-order and single-pass: each thread accumulates its grid-stride units
//    in order, a wavefront reduces by an xor butterfly, the four wavefront sums are added in
//    order, block partials go to a workspace, and the last block to arrive (counter with
//    agent-scope fences) adds the partials in block order.  The result is bitwise independent of
//    scheduling and of other work on the device, and is the order the FP64 oracle reproduces.
//    This replaces the reference's innerProdKerDMR (gDMR.cu:268-326), whose cross-block sum is
//    read without a grid-wide barrier.
//  * Floating-point contraction is off for this file (-ffp-contract=off); every fused
//    multiply-add is an explicit fma().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>

#include "dmr_internal.h"
#include "gDMR.h"

namespace dmr {
namespace {

template <typename TS> struct Acc { using type = float; };
template <> struct Acc<double> { using type = double; };

__device__ inline double fmac(double a, double b, double c) { return fma(a, b, c); }
__device__ inline float fmac(float a, float b, float c) { return fmaf(a, b, c); }

// ---- 16-byte unit loads / stores with conversion to the accumulation type ----
template <typename TA>
__device__ inline void ldu(const double *p, long long u, TA (&v)[2]) {
  const double2 t = reinterpret_cast<const double2 *>(p)[u];
  v[0] = (TA)t.x;
  v[1] = (TA)t.y;
}
template <typename TA>
__device__ inline void ldu(const float *p, long long u, TA (&v)[4]) {
  const float4 t = reinterpret_cast<const float4 *>(p)[u];
  v[0] = (TA)t.x;
  v[1] = (TA)t.y;
  v[2] = (TA)t.z;
  v[3] = (TA)t.w;
}
__device__ inline float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ inline float bf_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
template <typename TA>
__device__ inline void ldu(const bf16 *p, long long u, TA (&v)[8]) {
  const uint4 t = reinterpret_cast<const uint4 *>(p)[u];
  v[0] = (TA)bf_lo(t.x); v[1] = (TA)bf_hi(t.x);
  v[2] = (TA)bf_lo(t.y); v[3] = (TA)bf_hi(t.y);
  v[4] = (TA)bf_lo(t.z); v[5] = (TA)bf_hi(t.z);
  v[6] = (TA)bf_lo(t.w); v[7] = (TA)bf_hi(t.w);
}

// ldu with the non-temporal hint (NT).  The linear combinations (k_combo, k_combo_multi) read
// their basis columns that way: each is read once per launch, and the hint keeps the operator and
// the working vectors that the SpMV and the multi-dot reuse in the memory-side cache (FP64 leg on
// the 128^3 sequence: update 4.52 -> 5.21 TB/s, 4.09 -> 4.32 solves/s; same bits).  The same hint
// on the multi-dot's basis loads cost the mixed leg 2.5%, so those stay ordinary loads.
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef unsigned u4v __attribute__((ext_vector_type(4)));
template <bool NT, typename TA>
__device__ inline void ldu_t(const double *p, long long u, TA (&v)[2]) {
  if (!NT) return ldu(p, u, v);
  const d2v t = __builtin_nontemporal_load(reinterpret_cast<const d2v *>(p) + u);
  v[0] = (TA)t.x;
  v[1] = (TA)t.y;
}
template <bool NT, typename TA>
__device__ inline void ldu_t(const float *p, long long u, TA (&v)[4]) {
  if (!NT) return ldu(p, u, v);
  const f4v t = __builtin_nontemporal_load(reinterpret_cast<const f4v *>(p) + u);
  v[0] = (TA)t.x;
  v[1] = (TA)t.y;
  v[2] = (TA)t.z;
  v[3] = (TA)t.w;
}
template <bool NT, typename TA>
__device__ inline void ldu_t(const bf16 *p, long long u, TA (&v)[8]) {
  if (!NT) return ldu(p, u, v);
  const u4v t = __builtin_nontemporal_load(reinterpret_cast<const u4v *>(p) + u);
  v[0] = (TA)bf_lo(t.x); v[1] = (TA)bf_hi(t.x);
  v[2] = (TA)bf_lo(t.y); v[3] = (TA)bf_hi(t.y);
  v[4] = (TA)bf_lo(t.z); v[5] = (TA)bf_hi(t.z);
  v[6] = (TA)bf_lo(t.w); v[7] = (TA)bf_hi(t.w);
}

template <typename TA>
__device__ inline void stu(double *p, long long u, const TA (&v)[2]) {
  reinterpret_cast<double2 *>(p)[u] = make_double2((double)v[0], (double)v[1]);
}
template <typename TA>
__device__ inline void stu(float *p, long long u, const TA (&v)[4]) {
  reinterpret_cast<float4 *>(p)[u] = make_float4((float)v[0], (float)v[1], (float)v[2], (float)v[3]);
}
// Round-to-nearest-even float -> bf16 bits (finite inputs).
__device__ inline uint32_t bf_bits(float f) {
  const uint32_t x = __float_as_uint(f);
  return (x + 0x7fffu + ((x >> 16) & 1u)) >> 16;
}
template <typename TA>
__device__ inline void stu(bf16 *p, long long u, const TA (&v)[8]) {
  uint4 t;
  t.x = bf_bits((float)v[0]) | (bf_bits((float)v[1]) << 16);
  t.y = bf_bits((float)v[2]) | (bf_bits((float)v[3]) << 16);
  t.z = bf_bits((float)v[4]) | (bf_bits((float)v[5]) << 16);
  t.w = bf_bits((float)v[6]) | (bf_bits((float)v[7]) << 16);
  reinterpret_cast<uint4 *>(p)[u] = t;
}
// Value as it reads back after a store to storage type TS.
template <typename TS, typename TA> __device__ inline TA as_stored(TA v) { return v; }
template <> __device__ inline float as_stored<bf16, float>(float v) {
  return __uint_as_float(bf_bits(v) << 16);
}

// ---- fixed-order reductions ----
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
  return v;
}
__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v = v + __shfl_xor(v, off, 64);
  return v;
}

// Sum over the 256 threads of a block: wavefront butterflies, then ((w0 + w1) + w2) + w3.
// Every thread gets the result.
template <typename TA> __device__ inline TA block_sum(TA v) {
  __shared__ TA part[kBlock / 64];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  const TA s = ((part[0] + part[1]) + part[2]) + part[3];
  __syncthreads();
  return s;
}

// Partials are written and read with agent-scope atomics so the last block sees every other
// block's value whichever XCD (and L2) it ran on.
__device__ inline void put_partial(double *p, double v) {
  __hip_atomic_store(reinterpret_cast<unsigned long long *>(p), __double_as_longlong(v),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void put_partial(float *p, float v) {
  __hip_atomic_store(reinterpret_cast<unsigned *>(p), __float_as_uint(v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline double get_partial(const double *p) {
  return __longlong_as_double((long long)__hip_atomic_load(
      reinterpret_cast<const unsigned long long *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
__device__ inline float get_partial(const float *p) {
  return __uint_as_float(__hip_atomic_load(reinterpret_cast<const unsigned *>(p), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT));
}

// Arrival of block b (of total) at the end of a grid-wide reduction.  True in exactly one block,
// the last to arrive.  Every block writes its partials; thread 0 then counts the arrival with
// agent-scope acquire-release atomics, which order those stores before the count (one L2
// write-back per block, not one per wavefront as a block-wide __threadfence would cost) and, in
// the last block, order the partial loads after it.  The count runs in two levels: block b
// arrives on sub-counter b % kArrivalFan of its group, the last arrival of a sub-group on the
// group's top counter.  Atomics on one address are serialised by the memory system (measured:
// about 20 ns each, so 512 blocks on one counter cost 10 us per launch); here at most
// total / kArrivalFan + 1 of them meet on any one counter.  The chain of acquire-release pairs
// makes every block's partials visible to the last block.  Each counter is returned to zero by
// the block that completes it.
__device__ inline unsigned *group_counters(unsigned *base, int group) {
  return base + (long long)group * (1 + kArrivalFan) * kCounterStride;
}
__device__ inline bool arrive_last(unsigned *ctr, unsigned b, unsigned total) {
  __shared__ int s_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned fan = total < (unsigned)kArrivalFan ? total : (unsigned)kArrivalFan;
    const unsigned sub = b % fan;
    const unsigned subTotal = total / fan + (sub < total % fan ? 1u : 0u);
    unsigned *sc = ctr + (1 + sub) * kCounterStride;
    bool last = false;
    if (__hip_atomic_fetch_add(sc, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
        subTotal - 1u) {
      __hip_atomic_store(sc, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      last = __hip_atomic_fetch_add(ctr, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
             fan - 1u;
      if (last) __hip_atomic_store(ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    s_last = last ? 1 : 0;
  }
  __syncthreads();
  return s_last != 0;
}

// Sum of G block partials in block order, one wavefront per column: lane l adds partials l,
// l+64, ... in order, then the xor butterfly.  Returns the sum in every lane.  The loads are
// issued sixteen per lane at a time, ahead of the adds, so the last block waits one memory
// round trip per 1024 partials instead of one per 64.
template <typename TA> __device__ inline TA wave_sum_partials(const TA *part, int G) {
  const int lane = threadIdx.x & 63;
  TA v = 0;
  for (int q0 = 0; q0 < G; q0 += 64 * 16) {
    TA t[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int q = q0 + lane + 64 * i;
      t[i] = q < G ? get_partial(part + q) : TA(0);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i)
      if (q0 + lane + 64 * i < G) v = v + t[i];
  }
  return wave_sum(v);
}

// Column sums of partial columns c0..c0+pc-1 (column-major, stride G <= 64 * NQ), each in the
// order of wave_sum_partials.  Wavefront w takes CPB = min(8, 16 / NQ) adjacent columns at a
// time (columns w*CPB.., then +4*CPB), all their loads in flight together.  Column c's sum lands
// in res[c].
template <typename TA, int NQ, typename TR>
__device__ inline void sum_cols_nq(const TA *part, int G, int c0, int pc, TR *res) {
  constexpr int CPB = NQ >= 2 ? 16 / NQ : 8;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int cb = wv * CPB; cb < pc; cb += (kBlock / 64) * CPB) {
    TA t[CPB][NQ];
#pragma unroll
    for (int cc = 0; cc < CPB; ++cc)
#pragma unroll
      for (int i = 0; i < NQ; ++i) {
        const int q = lane + 64 * i;
        t[cc][i] = (cb + cc < pc && q < G)
                       ? get_partial(part + (long long)(c0 + cb + cc) * G + q)
                       : TA(0);
      }
#pragma unroll
    for (int cc = 0; cc < CPB; ++cc) {
      TA v = 0;
#pragma unroll
      for (int i = 0; i < NQ; ++i)
        if (lane + 64 * i < G) v = v + t[cc][i];
      v = wave_sum(v);
      if (lane == 0 && cb + cc < pc) res[cb + cc] = (TR)v;
    }
  }
}

__device__ inline void finalize(const Fin &fin, double ss, double sg) {
  switch (fin.mode) {
    case FIN_NORM:
      fin.scal[0] = ss;
      fin.scal[1] = 1.0 / sqrt(ss);
      fin.scal[2] = sqrt(ss);
      break;
    case FIN_ARNOLDI: {
      const double nrm = sqrt(ss);
      if (fin.tol > 0.0 && nrm < fin.tol) {
        *fin.flag = fin.step;
      } else {
        const double sc = 1.0 / sqrt(ss);
        fin.scal[1] = sc;
        *fin.hdst = sg * sc;
      }
      break;
    }
    case FIN_STORE:
      fin.scal[0] = ss;
      fin.scal[1] = sg;
      break;
    default:
      break;
  }
}

// o += sum_l cs[l] * B_l[u] for the NC columns of one chunk: all NC loads issued before the
// fused multiply-adds, which run in column order.
template <int NC, typename TS, typename TA, int N>
__device__ inline void fma_chunk(const TS *__restrict__ B, long long ld, long long u, const TA *cs,
                                 TA (&o)[N]) {
  TA b[NC][N];
#pragma unroll
  for (int i = 0; i < NC; ++i) ldu_t<true>(B + (long long)i * ld, u, b[i]);
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    const TA c = cs[i];
#pragma unroll
    for (int e = 0; e < N; ++e) o[e] = fmac(c, b[i][e], o[e]);
  }
}
// o += sum_l cs[l] * B_l[u] for l < p, in column order: chunks of 16 columns' loads in flight,
// then one chunk of 8, 4 and 1 to 3 for the rest, so no load reads a column past p.
template <typename TS, typename TA, int N>
__device__ inline void fma_cols(const TS *__restrict__ B, int p, long long ld, long long u,
                                const TA *cs, TA (&o)[N]) {
  int l0 = 0;
  for (; l0 + 16 <= p; l0 += 16) fma_chunk<16>(B + (long long)l0 * ld, ld, u, cs + l0, o);
  if (p - l0 >= 8) {
    fma_chunk<8>(B + (long long)l0 * ld, ld, u, cs + l0, o);
    l0 += 8;
  }
  if (p - l0 >= 4) {
    fma_chunk<4>(B + (long long)l0 * ld, ld, u, cs + l0, o);
    l0 += 4;
  }
  switch (p - l0) {
    case 3: fma_chunk<3>(B + (long long)l0 * ld, ld, u, cs + l0, o); break;
    case 2: fma_chunk<2>(B + (long long)l0 * ld, ld, u, cs + l0, o); break;
    case 1: fma_chunk<1>(B + (long long)l0 * ld, ld, u, cs + l0, o); break;
    default: break;
  }
}

// =============================================================================================
// Multi-dot: out[l] = B_l . x for the p columns of B (GEMV-T).  Replaces basisProjDMR's
// cublasCgemv(OP_C) (gDMR.cu:548-590).  grid = (G row blocks, ceil(p / kDotCols) column groups);
// x is re-read once per column group.
// =============================================================================================
// Block sums of CB values per thread at once: wavefront butterflies on all CB values, then the
// four wavefront sums in order.  Column c's result lands in res[c] (LDS) for c < CB.
template <typename TA, int CB>
__device__ inline void block_sum_cols(TA (&v)[CB], TA *res) {
  __shared__ TA red[kBlock / 64][CB];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < CB; ++c) v[c] = wave_sum(v[c]);
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < CB; ++c) red[wv][c] = v[c];
  }
  __syncthreads();
  if (threadIdx.x < CB) {
    const int c = threadIdx.x;
    res[c] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
  }
  __syncthreads();
}

// Last-block sum of the G partials (G <= kMaxRedBlocks = 1024) of columns c0..c0+pc-1
// (column-major, stride G), each in the order of wave_sum_partials.  Column c's sum lands in
// res[c].
template <typename TA, typename TR>
__device__ inline void sum_partials_cols(const TA *part, int G, int c0, int pc, TR *res) {
  static_assert(kMaxRedBlocks <= 64 * 16, "sixteen partials per lane cover every row block");
  const int nq = (G + 63) / 64;
  if (nq <= 1)
    sum_cols_nq<TA, 1>(part, G, c0, pc, res);
  else if (nq <= 2)
    sum_cols_nq<TA, 2>(part, G, c0, pc, res);
  else if (nq <= 4)
    sum_cols_nq<TA, 4>(part, G, c0, pc, res);
  else if (nq <= 8)
    sum_cols_nq<TA, 8>(part, G, c0, pc, res);
  else
    sum_cols_nq<TA, 16>(part, G, c0, pc, res);
  __syncthreads();
}

template <typename TS, int CB>
__global__ __launch_bounds__(kBlock) void k_multidot(const TS *__restrict__ B, long long ld, int p,
                                                     const TS *__restrict__ x, long long nv, int G,
                                                     typename Acc<TS>::type *partials,
                                                     unsigned *counters, double *out, int ostride,
                                                     const int *flag) {
  using TA = typename Acc<TS>::type;
  constexpr int N = Unit<TS>::N;
  if (flag && *flag) return;
  // Group of up to CB columns (CB a multiple of 8); all of p when p <= CB, so x is read once.
  const int c0 = blockIdx.y * CB;
  const int pc = min(CB, p - c0);
  TA acc[CB];
#pragma unroll
  for (int c = 0; c < CB; ++c) acc[c] = 0;
  const long long stride = (long long)G * kBlock;
  for (long long u = (long long)blockIdx.x * kBlock + threadIdx.x; u < nv; u += stride) {
    TA xv[N];
    ldu(x, u, xv);
    // The loads of CH columns of the unit issued together: all of a group of up to 32 columns
    // (one memory round trip per unit), else chunks of 16 columns (measured faster than 32 for
    // FP32 groups of 40-56: fewer registers, more wavefronts), the last chunk cut at CB.
    // Indices past the group alias its last column (a cache hit) and their sums are discarded.
    constexpr int CH = CB <= 32 && N <= 4 ? CB : CB < 16 ? CB : 16;
#pragma unroll
    for (int cc = 0; cc < CB; cc += CH) {
      if (cc < pc) {
        TA bv[CH][N];
#pragma unroll
        for (int i = 0; i < CH; ++i)
          if (cc + i < CB) ldu(B + (long long)(c0 + min(cc + i, pc - 1)) * ld, u, bv[i]);
#pragma unroll
        for (int i = 0; i < CH; ++i)
          if (cc + i < CB) {
#pragma unroll
            for (int e = 0; e < N; ++e) acc[cc + i] = fmac(bv[i][e], xv[e], acc[cc + i]);
          }
      }
    }
  }
  __shared__ TA res[CB];
  block_sum_cols<TA, CB>(acc, res);
  if (threadIdx.x < pc) put_partial(partials + (long long)(c0 + threadIdx.x) * G + blockIdx.x,
                                    res[threadIdx.x]);
  if (!arrive_last(group_counters(counters, blockIdx.y), blockIdx.x, (unsigned)G)) return;
  sum_partials_cols<TA>(partials, G, c0, pc, res);
  if (threadIdx.x < pc) out[(long long)(c0 + threadIdx.x) * ostride] = (double)res[threadIdx.x];
}

// =============================================================================================
// Single-output combination with fused reductions.  Replaces vecScaleDiffDMR (gDMR.cu:229-248,
// mode SUB), addBasisVecsDMR + vecAddDMR + normalizeVecDMR(mode 2) (gDMR.cu:185-209, 250-266,
// 328-355, mode ACC) and the innerProdDMR calls that follow them.
// =============================================================================================
template <typename TS, int MODE, int RED>
__global__ __launch_bounds__(kBlock) void k_combo(const TS *in, TS *out, const TS *__restrict__ B1,
                                                  int p1, const double *__restrict__ c1,
                                                  const TS *__restrict__ B2, int p2,
                                                  const double *__restrict__ c2, long long ld,
                                                  long long nv, double beta,
                                                  const TS *__restrict__ g, int G,
                                                  typename Acc<TS>::type *partials,
                                                  unsigned *counter, Fin fin, const int *flag,
                                                  const double *betaDev) {
  using TA = typename Acc<TS>::type;
  constexpr int N = Unit<TS>::N;
  if (flag && *flag) return;
  __shared__ TA cs[kMaxCoef];
  for (int l = threadIdx.x; l < p1 + p2; l += kBlock) {
    const double c = l < p1 ? c1[l] : c2[l - p1];
    cs[l] = (TA)(MODE == COMB_SUB ? -c : c);
  }
  __syncthreads();
  const TA tb = (TA)(betaDev ? *betaDev : beta);
  TA ss = 0, sg = 0;
  const long long stride = (long long)G * kBlock;
  for (long long u = (long long)blockIdx.x * kBlock + threadIdx.x; u < nv; u += stride) {
    TA o[N];
    if (MODE == COMB_SUB) {
      ldu(in, u, o);
      fma_cols(B1, p1, ld, u, cs, o);
      fma_cols(B2, p2, ld, u, cs + p1, o);
    } else {
      TA a[N];
#pragma unroll
      for (int e = 0; e < N; ++e) a[e] = 0;
      fma_cols(B1, p1, ld, u, cs, a);
      fma_cols(B2, p2, ld, u, cs + p1, a);
      if (in) {
        TA xi[N];
        ldu(in, u, xi);
#pragma unroll
        for (int e = 0; e < N; ++e) {
          const TA t = tb * a[e];
          o[e] = xi[e] + t;
        }
      } else {
#pragma unroll
        for (int e = 0; e < N; ++e) o[e] = tb * a[e];
      }
    }
    stu(out, u, o);
    if (RED & 1) {
#pragma unroll
      for (int e = 0; e < N; ++e) {
        const TA v = as_stored<TS>(o[e]);
        ss = fmac(v, v, ss);
      }
    }
    if (RED & 2) {
      TA gv[N];
      ldu(g, u, gv);
#pragma unroll
      for (int e = 0; e < N; ++e) sg = fmac(as_stored<TS>(o[e]), gv[e], sg);
    }
  }
  if (RED == 0) return;
  constexpr int NR = (RED & 2) ? 2 : 1;  // partial columns: ss, then sg
  __shared__ TA res[2];
  TA v[NR];
  v[0] = ss;
  if (NR == 2) v[NR - 1] = sg;
  block_sum_cols<TA, NR>(v, res);
  if (threadIdx.x < NR) put_partial(partials + (long long)threadIdx.x * G + blockIdx.x,
                                    res[threadIdx.x]);
  if (!arrive_last(counter, blockIdx.x, (unsigned)G)) return;
  sum_partials_cols<TA>(partials, G, 0, NR, res);
  if (threadIdx.x == 0) finalize(fin, (double)res[0], NR == 2 ? (double)res[1] : 0.0);
}

// =============================================================================================
// Multi-output combination out_j = [B1 B2] c_j with per-output sum of squares, j < J; the
// deflation refresh Z_new = [V Z] C (genDeflatDMR, gDMR.cu:1583-1704) as one skinny GEMM instead
// of k separate combinations.  grid = (G, ceil(J / JB)).
// =============================================================================================
template <typename TS, int JB>
__global__ __launch_bounds__(kBlock) void k_combo_multi(
    TS *__restrict__ out, int J, const TS *__restrict__ B1, int p1, const double *__restrict__ C1,
    const TS *__restrict__ B2, int p2, const double *__restrict__ C2, int ldc, long long ld,
    long long nv, int G, typename Acc<TS>::type *partials, unsigned *counters, double *scal) {
  using TA = typename Acc<TS>::type;
  constexpr int N = Unit<TS>::N;
  extern __shared__ unsigned char smem_raw[];
  TA *cs = reinterpret_cast<TA *>(smem_raw);  // (p1 + p2) x JB
  const int j0 = blockIdx.y * JB;
  const int Jg = min(JB, J - j0);
  for (int idx = threadIdx.x; idx < (p1 + p2) * JB; idx += kBlock) {
    const int l = idx / JB, j = idx % JB;
    double c = 0.0;
    if (j < Jg) c = l < p1 ? C1[l + (long long)ldc * (j0 + j)] : C2[(l - p1) + (long long)ldc * (j0 + j)];
    cs[idx] = (TA)c;
  }
  __syncthreads();
  TA ss[JB];
#pragma unroll
  for (int j = 0; j < JB; ++j) ss[j] = 0;
  const long long stride = (long long)G * kBlock;
  for (long long u = (long long)blockIdx.x * kBlock + threadIdx.x; u < nv; u += stride) {
    TA a[JB][N];
#pragma unroll
    for (int j = 0; j < JB; ++j)
#pragma unroll
      for (int e = 0; e < N; ++e) a[j][e] = 0;
    // Eight columns' loads in flight, then their multiply-adds in column order; the last
    // columns one at a time.
    const int pt = p1 + p2;
    int l = 0;
    for (; l + 8 <= pt; l += 8) {
      TA b[8][N];
#pragma unroll
      for (int i = 0; i < 8; ++i)
        ldu(l + i < p1 ? B1 + (long long)(l + i) * ld : B2 + (long long)(l + i - p1) * ld, u, b[i]);
#pragma unroll
      for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) {
          if (j < Jg) {
            const TA c = cs[(l + i) * JB + j];
#pragma unroll
            for (int e = 0; e < N; ++e) a[j][e] = fmac(c, b[i][e], a[j][e]);
          }
        }
    }
    for (; l < pt; ++l) {
      TA b[N];
      ldu(l < p1 ? B1 + (long long)l * ld : B2 + (long long)(l - p1) * ld, u, b);
#pragma unroll
      for (int j = 0; j < JB; ++j) {
        if (j < Jg) {
          const TA c = cs[l * JB + j];
#pragma unroll
          for (int e = 0; e < N; ++e) a[j][e] = fmac(c, b[e], a[j][e]);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < JB; ++j) {
      if (j < Jg) {
        stu(out + (long long)(j0 + j) * ld, u, a[j]);
#pragma unroll
        for (int e = 0; e < N; ++e) {
          const TA v = as_stored<TS>(a[j][e]);
          ss[j] = fmac(v, v, ss[j]);
        }
      }
    }
  }
  __shared__ TA res[JB];
  block_sum_cols<TA, JB>(ss, res);
  if (threadIdx.x < Jg)
    put_partial(partials + (long long)(j0 + threadIdx.x) * G + blockIdx.x, res[threadIdx.x]);
  if (!arrive_last(group_counters(counters, blockIdx.y), blockIdx.x, (unsigned)G)) return;
  sum_partials_cols<TA>(partials, G, j0, Jg, res);
  if (threadIdx.x < Jg) scal[j0 + threadIdx.x] = 1.0 / sqrt((double)res[threadIdx.x]);
}

// =============================================================================================
// Fused orthogonalisation of one Arnoldi vector (one cooperative launch per step).
//
// u = g;  [stage A: cA = A^T u -> hA;  u -= A cA];  [stage B: cB = B^T u -> hB;  u -= B cB];
// ss = u.u (and sg = u.g);  finalize;  dst = u * (1/sqrt(ss)).
//
// This is the sequence multidot -> combo(SUB) -> multidot -> combo(SUB, norms) -> scale of the
// unfused path in one kernel: every thread keeps its UPT units of g and u in registers across
// the phases, and the grid synchronises through an arrival counter and a release flag (all
// blocks are co-resident: the kernel is only launched cooperatively).  Per-thread unit mapping,
// block sums and the last block's partial sums are those of the unfused kernels, so the results
// are bitwise identical to them.  A phase's last block to arrive finishes its reduction, thread
// 0 publishes the results, then releases the others.
// =============================================================================================
struct OrthArgs {
  const void *g;
  void *dst;
  const void *A;
  int pa;
  double *ha;
  const void *B;
  int pb;
  double *hb;
  long long ld;
  long long nv;
  int G;
  void *partials;
  unsigned *bar;     // bar[0]: arrivals, bar[1]: released epoch
  unsigned epoch0;   // phases already completed in this cycle
  int red;           // 1: ss, 3: ss and sg
  Fin fin;
  const int *flag;
  int *err;          // set when a barrier wait exceeds its bound
};

constexpr unsigned kSpinLimit = 1u << 23;

// Phase barrier: true in the block that arrives last (it must call phase_release).
__device__ inline bool phase_arrive(unsigned *bar, unsigned target) {
  __shared__ int s_last;
  __syncthreads();
  if (threadIdx.x == 0)
    s_last = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
             target - 1u;
  __syncthreads();
  return s_last != 0;
}
__device__ inline void phase_release(unsigned *bar, unsigned epoch) {
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(bar + 1, epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
// Bounded wait for the release of `epoch`; on timeout the error word is set and the kernel goes
// on (its results are then discarded by the host), so every wave always exits.
__device__ inline void phase_wait(unsigned *bar, unsigned epoch, int *err) {
  if (threadIdx.x == 0) {
    unsigned spins = 0;
    while (__hip_atomic_load(bar + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
      // Back off (~500 cycles) so hundreds of waiting blocks do not flood one address.
      __builtin_amdgcn_s_sleep(8);
      if (++spins > kSpinLimit) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
  }
  __syncthreads();
}
__device__ inline double load_published(const double *p) {
  return __longlong_as_double((long long)__hip_atomic_load(
      reinterpret_cast<const unsigned long long *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

// Block partials of p dot products B_l . x (x in registers), in chunks of 16 columns; per column
// the same unit order and block sum as k_multidot.
template <typename TS, typename TA, int UPT, int N>
__device__ inline void orth_dots(const TS *__restrict__ Bm, int p, long long ld,
                                 const TA (&x)[UPT][N], long long u0, long long stride,
                                 long long nv, int G, TA *partials) {
  for (int c0 = 0; c0 < p; c0 += 16) {
    const int pc = min(16, p - c0);
    TA acc[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc[c] = 0;
#pragma unroll
    for (int s = 0; s < UPT; ++s) {
      const long long u = u0 + s * stride;
      if (u < nv) {
#pragma unroll
        for (int cc = 0; cc < 16; cc += 8) {
          if (cc < pc) {
            TA bv[8][N];
#pragma unroll
            for (int i = 0; i < 8; ++i)
              ldu(Bm + (long long)(c0 + min(cc + i, pc - 1)) * ld, u, bv[i]);
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
              for (int e = 0; e < N; ++e) acc[cc + i] = fmac(bv[i][e], x[s][e], acc[cc + i]);
          }
        }
      }
    }
    __shared__ TA r16[16];
    block_sum_cols<TA, 16>(acc, r16);
    if ((int)threadIdx.x < pc)
      put_partial(partials + (long long)(c0 + threadIdx.x) * G + blockIdx.x, r16[threadIdx.x]);
  }
}

// Last block: column sums of p partial columns into out[0..p) (published by thread 0).
template <typename TA>
__device__ inline void orth_finish_cols(const TA *partials, int G, int p, double *out) {
  __shared__ double fin[kMaxCoef];
  sum_partials_cols<TA, double>(partials, G, 0, p, fin);
  if (threadIdx.x == 0)
    for (int c = 0; c < p; ++c) out[c] = fin[c];
}

// x -= sum_l c_l B_l (coefficients in LDS, already negated), columns in order.
template <typename TS, typename TA, int UPT, int N>
__device__ inline void orth_update(const TS *__restrict__ Bm, int p, long long ld, const TA *cs,
                                   TA (&x)[UPT][N], long long u0, long long stride, long long nv) {
#pragma unroll
  for (int s = 0; s < UPT; ++s) {
    const long long u = u0 + s * stride;
    if (u < nv) fma_cols(Bm, p, ld, u, cs, x[s]);
  }
}

template <typename TS, int UPT>
__global__ __launch_bounds__(kBlock) void k_orth(OrthArgs a) {
  using TA = typename Acc<TS>::type;
  constexpr int N = Unit<TS>::N;
  if (a.flag && *a.flag) return;
  const TS *g = (const TS *)a.g;
  const int G = a.G;
  const long long stride = (long long)G * kBlock;
  const long long u0 = (long long)blockIdx.x * kBlock + threadIdx.x;
  TA *partials = (TA *)a.partials;
  TA gv[UPT][N], uv[UPT][N];
#pragma unroll
  for (int s = 0; s < UPT; ++s) {
    const long long u = u0 + s * stride;
    if (u < a.nv) {
      ldu(g, u, gv[s]);
    } else {
#pragma unroll
      for (int e = 0; e < N; ++e) gv[s][e] = 0;
    }
#pragma unroll
    for (int e = 0; e < N; ++e) uv[s][e] = gv[s][e];
  }
  __shared__ TA cs[kMaxCoef];
  unsigned epoch = a.epoch0;
  const TS *stB[2] = {(const TS *)a.A, (const TS *)a.B};
  const int stP[2] = {a.pa, a.pb};
  double *stH[2] = {a.ha, a.hb};
#pragma unroll
  for (int st = 0; st < 2; ++st) {
    const int p = stP[st];
    if (p <= 0) continue;
    orth_dots<TS, TA, UPT, N>(stB[st], p, a.ld, uv, u0, stride, a.nv, G, partials);
    ++epoch;
    if (phase_arrive(a.bar, epoch * (unsigned)G)) {
      orth_finish_cols<TA>(partials, G, p, stH[st]);
      phase_release(a.bar, epoch);
    } else {
      phase_wait(a.bar, epoch, a.err);
    }
    for (int l = threadIdx.x; l < p; l += kBlock) cs[l] = (TA)(-load_published(stH[st] + l));
    __syncthreads();
    orth_update<TS, TA, UPT, N>(stB[st], p, a.ld, cs, uv, u0, stride, a.nv);
    __syncthreads();
  }
  // Norms (same per-thread order as k_combo's fused reductions).
  TA v[2] = {0, 0};
#pragma unroll
  for (int s = 0; s < UPT; ++s) {
    if (u0 + s * stride < a.nv) {
#pragma unroll
      for (int e = 0; e < N; ++e) v[0] = fmac(uv[s][e], uv[s][e], v[0]);
      if (a.red & 2) {
#pragma unroll
        for (int e = 0; e < N; ++e) v[1] = fmac(uv[s][e], gv[s][e], v[1]);
      }
    }
  }
  const int nr = (a.red & 2) ? 2 : 1;
  __shared__ TA r2[2];
  block_sum_cols<TA, 2>(v, r2);
  if ((int)threadIdx.x < nr)
    put_partial(partials + (long long)threadIdx.x * G + blockIdx.x, r2[threadIdx.x]);
  ++epoch;
  if (phase_arrive(a.bar, epoch * (unsigned)G)) {
    __shared__ TA fin2[2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (wv < nr) {
      const TA s = wave_sum_partials(partials + (long long)wv * G, G);
      if (lane == 0) fin2[wv] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) finalize(a.fin, (double)fin2[0], nr == 2 ? (double)fin2[1] : 0.0);
    phase_release(a.bar, epoch);
  } else {
    phase_wait(a.bar, epoch, a.err);
  }
  // After a breakdown the column is zeroed by breakdown_fixup; it is left unscaled here.
  const bool broke = a.fin.flag && __hip_atomic_load(a.fin.flag, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT) != 0;
  const TA sc = broke ? (TA)1 : (TA)load_published(a.fin.scal + 1);
  TS *dst = (TS *)a.dst;
#pragma unroll
  for (int s = 0; s < UPT; ++s) {
    const long long u = u0 + s * stride;
    if (u < a.nv) {
      TA o[N];
#pragma unroll
      for (int e = 0; e < N; ++e) o[e] = uv[s][e] * sc;
      stu(dst, u, o);
    }
  }
}

// v *= scal (normalizeVecDMR, gDMR.cu:328-355), skipped after a breakdown.
template <typename TS>
__global__ __launch_bounds__(kBlock) void k_scale(TS *v, long long nv, const double *scal,
                                                  const int *flag) {
  using TA = typename Acc<TS>::type;
  constexpr int N = Unit<TS>::N;
  if (flag && *flag) return;
  const TA s = (TA)*scal;
  for (long long u = (long long)blockIdx.x * kBlock + threadIdx.x; u < nv;
       u += (long long)gridDim.x * kBlock) {
    TA x[N];
    ldu(v, u, x);
#pragma unroll
    for (int e = 0; e < N; ++e) x[e] = x[e] * s;
    stu(v, u, x);
  }
}

template <typename TS>
__global__ __launch_bounds__(kBlock) void k_scale_cols(TS *v, long long ld, long long nv,
                                                       const double *scal) {
  using TA = typename Acc<TS>::type;
  constexpr int N = Unit<TS>::N;
  TS *col = v + (long long)blockIdx.y * ld;
  const TA s = (TA)scal[blockIdx.y];
  for (long long u = (long long)blockIdx.x * kBlock + threadIdx.x; u < nv;
       u += (long long)gridDim.x * kBlock) {
    TA x[N];
    ldu(col, u, x);
#pragma unroll
    for (int e = 0; e < N; ++e) x[e] = x[e] * s;
    stu(col, u, x);
  }
}

template <typename TS>
__global__ __launch_bounds__(kBlock) void k_fill(TS *p, long long nv, double val) {
  using TA = typename Acc<TS>::type;
  constexpr int N = Unit<TS>::N;
  TA x[N];
#pragma unroll
  for (int e = 0; e < N; ++e) x[e] = (TA)val;
  for (long long u = (long long)blockIdx.x * kBlock + threadIdx.x; u < nv;
       u += (long long)gridDim.x * kBlock)
    stu(p, u, x);
}

__global__ __launch_bounds__(kBlock) void k_fill_int(int *p, long long count, int v) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < count;
       i += (long long)gridDim.x * kBlock)
    p[i] = v;
}

// Element-wise conversion; n need not be a multiple of a unit (external vectors).
template <typename TD, typename TSrc>
__global__ __launch_bounds__(kBlock) void k_convert(TD *dst, const TSrc *src, long long n) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    dst[i] = (TD)src[i];
}
template <>
__global__ __launch_bounds__(kBlock) void k_convert<bf16, double>(bf16 *dst, const double *src,
                                                                   long long n) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    dst[i].bits = (uint16_t)bf_bits((float)src[i]);
}
template <>
__global__ __launch_bounds__(kBlock) void k_convert<double, bf16>(double *dst, const bf16 *src,
                                                                   long long n) {
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (long long)gridDim.x * kBlock)
    dst[i] = (double)__uint_as_float((uint32_t)src[i].bits << 16);
}

template <typename TS>
__global__ __launch_bounds__(kBlock) void k_add_to(double *x, const TS *d, long long nv) {
  using TA = typename Acc<TS>::type;
  constexpr int N = Unit<TS>::N;
  // x is FP64 (2 per 16-byte unit); walk d in its own units and x element-wise.
  for (long long u = (long long)blockIdx.x * kBlock + threadIdx.x; u < nv;
       u += (long long)gridDim.x * kBlock) {
    TA dv[N];
    ldu(d, u, dv);
#pragma unroll
    for (int e = 0; e < N; ++e) x[u * N + e] = x[u * N + e] + (double)dv[e];
  }
}

__global__ void k_cycle_begin(double *H, int count, int *flag, unsigned *bar) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < count; i += gridDim.x * kBlock) H[i] = 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *flag = 0;
    bar[0] = 0u;
    bar[1] = 0u;
  }
}

// Breakdown at Krylov step j (gDMR.cu:2256-2290): zero basis slots o+j..m and put ones on the
// diagonal of the Hessenberg columns that were never filled (from column d+j; the reference
// starts at column j, which overwrites filled columns when d > 0).
template <typename TS>
__global__ __launch_bounds__(kBlock) void k_breakdown_fixup(TS *P, long long ld, int o, int d,
                                                            int m, double *H, const int *flag) {
  const int j = *flag;
  if (j == 0) return;
  using TA = typename Acc<TS>::type;
  constexpr int N = Unit<TS>::N;
  const long long nvcol = ld / N;
  const long long total = nvcol * (long long)(m + 1 - (o + j));
  TS *base = P + (long long)(o + j) * ld;
  TA z[N];
#pragma unroll
  for (int e = 0; e < N; ++e) z[e] = 0;
  for (long long u = (long long)blockIdx.x * kBlock + threadIdx.x; u < total;
       u += (long long)gridDim.x * kBlock)
    stu(base, u, z);
  if (blockIdx.x == 0)
    for (int c = d + j + threadIdx.x; c < m; c += kBlock) H[c + (long long)c * (m + 1)] = 1.0;
}

// =============================================================================================
// CSR SpMV, y = A x (replaces the user operator's dense cublasCgemv, testLS.cu:97-128).
// Row blocks of <= 256 rows and <= kSpmvNnzCap nonzeros: the block streams its contiguous slice
// of val / colInd with coalesced loads, forms the products in LDS, then each thread adds its row's
// products in column order.  A row with more than kSpmvNnzCap nonzeros has a block of its own
// and is reduced by the block.
// =============================================================================================
// Element loads / stores of vectors in their storage type (bf16: round to nearest even).
__device__ inline double ldv(const double *p, long long i) { return p[i]; }
__device__ inline float ldv(const float *p, long long i) { return p[i]; }
__device__ inline float ldv(const bf16 *p, long long i) {
  return __uint_as_float((uint32_t)p[i].bits << 16);
}
__device__ inline void stv(double *p, long long i, double v) { p[i] = v; }
__device__ inline void stv(float *p, long long i, float v) { p[i] = v; }
__device__ inline void stv(bf16 *p, long long i, float v) { p[i].bits = (uint16_t)bf_bits(v); }
// stv with the non-temporal hint (NT): the SpMV writes y that way, so its write-back overlaps the
// SpMV's gathers instead of the bandwidth-bound multi-dot that reads y next.
template <bool NT> __device__ inline void stv_t(double *p, long long i, double v) {
  if (NT) __builtin_nontemporal_store(v, p + i); else p[i] = v;
}
template <bool NT> __device__ inline void stv_t(float *p, long long i, float v) {
  if (NT) __builtin_nontemporal_store(v, p + i); else p[i] = v;
}
template <bool NT> __device__ inline void stv_t(bf16 *p, long long i, float v) { stv(p, i, v); }

// prod[e] = va[e] * x[ci[e]] for e < ne <= kSpmvNnzCap, entry e on thread e % kBlock: all index
// and value loads of the thread first, then all of its gathers, then the products.  Slots past
// ne load entry ne - 1 again (a cache hit), so no load waits on a branch.
constexpr int kSpmvEpt = kSpmvNnzCap / kBlock;
template <typename TV, typename TX, typename TA>
__device__ inline void spmv_products(const int *__restrict__ ci, const TV *__restrict__ va, int ne,
                                     const TX *__restrict__ x, TA *prod) {
  if (ne <= 0) return;
  int idx[kSpmvEpt];
  TV v[kSpmvEpt];
#pragma unroll
  for (int i = 0; i < kSpmvEpt; ++i) {
    const int e = min((int)threadIdx.x + kBlock * i, ne - 1);
    idx[i] = ci[e];
    v[i] = va[e];
  }
  TA xv[kSpmvEpt];
#pragma unroll
  for (int i = 0; i < kSpmvEpt; ++i) xv[i] = (TA)ldv(x, idx[i]);
#pragma unroll
  for (int i = 0; i < kSpmvEpt; ++i) {
    const int e = (int)threadIdx.x + kBlock * i;
    if (e < ne) prod[e] = (TA)v[i] * xv[i];
  }
}

template <typename TV, typename TX, typename TA>
__global__ __launch_bounds__(kBlock) void k_spmv(const int *__restrict__ rp,
                                                 const int *__restrict__ ci,
                                                 const TV *__restrict__ va,
                                                 const int *__restrict__ br,
                                                 const TX *__restrict__ x, TX *__restrict__ y) {
  __shared__ TA prod[kSpmvNnzCap];
  const int rs = br[blockIdx.x], re = br[blockIdx.x + 1];
  const int e0 = rp[rs], e1 = rp[re];
  const int nr = re - rs, ne = e1 - e0;
  if (ne > kSpmvNnzCap) {
    TA s = 0;
    for (int e = e0 + (int)threadIdx.x; e < e1; e += kBlock) {
      const TA t = (TA)va[e] * (TA)ldv(x, ci[e]);
      s = s + t;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) stv(y, rs, s);
    return;
  }
  // The thread's row bounds, then its <= kSpmvEpt entries' indices and values, then their
  // gathers: every load of a phase in flight at once.
  const bool hasRow = (int)threadIdx.x < nr;
  int ra = 0, rb = 0;
  if (hasRow) {
    ra = rp[rs + threadIdx.x] - e0;
    rb = rp[rs + threadIdx.x + 1] - e0;
  }
  spmv_products<TV, TX, TA>(ci + e0, va + e0, ne, x, prod);
  __syncthreads();
  if (hasRow) {
    TA s = 0;
    for (int e = ra; e < rb; ++e) s = s + prod[e];
    stv_t<true>(y, rs + threadIdx.x, s);
  }
}

// The same product without a precomputed row partition: workgroup g takes the fixed rows
// [256 g, 256 g + 256) of n.  A row of at most kSpmvNnzCap entries is summed by its thread over
// its products in column order (the sum k_spmv forms from LDS, so the bits are the same); a
// longer row is reduced by the whole workgroup with k_spmv's strided order and block_sum.
__global__ __launch_bounds__(kBlock) void k_spmv_rows(int n, const int *__restrict__ rp,
                                                      const int *__restrict__ ci,
                                                      const double *__restrict__ va,
                                                      const double *__restrict__ x,
                                                      double *__restrict__ y) {
  const int r0 = blockIdx.x * kBlock;
  const int r = r0 + (int)threadIdx.x;
  if (r < n) {
    const int a = rp[r], b = rp[r + 1];
    if (b - a <= kSpmvNnzCap) {
      double s = 0.0;
      for (int e = a; e < b; ++e) {
        const double t = va[e] * x[ci[e]];
        s = s + t;
      }
      y[r] = s;
    }
  }
  const int rEnd = r0 + kBlock < n ? r0 + kBlock : n;
  for (int q = r0; q < rEnd; ++q) {  // uniform over the workgroup
    const int e0 = rp[q], e1 = rp[q + 1];
    if (e1 - e0 <= kSpmvNnzCap) continue;
    double s = 0.0;
    for (int e = e0 + (int)threadIdx.x; e < e1; e += kBlock) {
      const double t = va[e] * x[ci[e]];
      s = s + t;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) y[q] = s;
  }
}

// r = b - A x in FP64 with a fixed-order sum of squares (the refinement residual).  Block
// partials are per SpMV row block.
__global__ __launch_bounds__(kBlock) void k_resid64(const int *__restrict__ rp,
                                                    const int *__restrict__ ci,
                                                    const double *__restrict__ va,
                                                    const int *__restrict__ br, int nb,
                                                    const double *__restrict__ x,
                                                    const double *__restrict__ b,
                                                    double *__restrict__ r, double *partials,
                                                    unsigned *counter, double *scal) {
  __shared__ double prod[kSpmvNnzCap];
  const int rs = br[blockIdx.x], re = br[blockIdx.x + 1];
  const int e0 = rp[rs], e1 = rp[re];
  const int nr = re - rs, ne = e1 - e0;
  double ss = 0.0;
  if (ne > kSpmvNnzCap) {
    double s = 0.0;
    for (int e = e0 + (int)threadIdx.x; e < e1; e += kBlock) {
      const double t = va[e] * x[ci[e]];
      s = s + t;
    }
    s = block_sum(s);
    const double v = b[rs] - s;
    if (threadIdx.x == 0) {
      r[rs] = v;
      ss = v * v;
    }
  } else {
    const bool hasRow = (int)threadIdx.x < nr;
    int a = 0, bb = 0;
    double bv = 0.0;
    if (hasRow) {
      a = rp[rs + threadIdx.x] - e0;
      bb = rp[rs + threadIdx.x + 1] - e0;
      bv = b[rs + threadIdx.x];
    }
    spmv_products<double, double, double>(ci + e0, va + e0, ne, x, prod);
    __syncthreads();
    if (hasRow) {
      const int row = rs + threadIdx.x;
      double s = 0.0;
      for (int e = a; e < bb; ++e) s = s + prod[e];
      const double v = bv - s;
      r[row] = v;
      ss = v * v;
    }
  }
  const double bs = block_sum(ss);
  if (threadIdx.x == 0) put_partial(partials + blockIdx.x, bs);
  if (!arrive_last(counter, blockIdx.x, (unsigned)nb)) return;
  if (threadIdx.x >= 64) return;
  const double tot = wave_sum_partials(partials, nb);
  if (threadIdx.x == 0) {
    scal[0] = tot;
    scal[1] = 1.0 / sqrt(tot);
    scal[2] = sqrt(tot);
  }
}

inline int grid_for(long long nv) {
  long long g = (nv + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > 4096) g = 4096;
  return (int)g;
}

// Streaming read of nu 16-byte units: each thread keeps eight loads in flight over a grid-stride
// walk.  The loaded words are folded into one value that is stored only if it equals a constant,
// which keeps every load without writing anything in practice.
__global__ __launch_bounds__(kBlock) void k_probe_read(const uint4 *__restrict__ a, long long nu,
                                                       unsigned *sink) {
  constexpr int U = 8;
  unsigned x = 0u;
  const long long stride = (long long)gridDim.x * kBlock * U;
  for (long long u = (long long)blockIdx.x * kBlock * U + threadIdx.x; u < nu; u += stride) {
    uint4 v[U];
#pragma unroll
    for (int i = 0; i < U; ++i)
      v[i] = u + (long long)i * kBlock < nu ? a[u + (long long)i * kBlock] : make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < U; ++i) x ^= v[i].x ^ v[i].y ^ v[i].z ^ v[i].w;
  }
  if (x == 0x9e3779b9u) sink[0] = x;
}

}  // namespace

// =============================================================================================
// Launchers
// =============================================================================================
void probe_read(const void *buf, long long units16, unsigned *sink, hipStream_t s) {
  hipLaunchKernelGGL(k_probe_read, dim3(2048), dim3(kBlock), 0, s,
                     static_cast<const uint4 *>(buf), units16, sink);
}

int spmv_partition(const int *rp, int n, int *br) {
  int nb = 0;
  int r = 0;
  br[0] = 0;
  while (r < n) {
    const int start = r;
    if (rp[r + 1] - rp[r] > kSpmvNnzCap) {
      r = r + 1;
    } else {
      while (r < n && r - start < kBlock && rp[r + 1] - rp[start] <= kSpmvNnzCap) ++r;
    }
    br[++nb] = r;
  }
  return nb;
}

void spmv(const CsrDev &A, const double *x, double *y, hipStream_t s) {
  if (A.nBlocks == 0) return;
  hipLaunchKernelGGL((k_spmv<double, double, double>), dim3(A.nBlocks), dim3(kBlock), 0, s,
                     A.rowPtr, A.colInd, A.val, A.blockRows, x, y);
}

void spmv32(const CsrDev &A, const float *x, float *y, hipStream_t s) {
  if (A.nBlocks == 0) return;
  hipLaunchKernelGGL((k_spmv<float, float, float>), dim3(A.nBlocks), dim3(kBlock), 0, s,
                     A.rowPtr, A.colInd, A.val32, A.blockRows, x, y);
}

void spmv16(const CsrDev &A, const bf16 *x, bf16 *y, hipStream_t s) {
  if (A.nBlocks == 0) return;
  hipLaunchKernelGGL((k_spmv<float, bf16, float>), dim3(A.nBlocks), dim3(kBlock), 0, s,
                     A.rowPtr, A.colInd, A.val32, A.blockRows, x, y);
}

void residual64(const CsrDev &A, const double *x, const double *b, double *r, RedWork w,
                double *scal, hipStream_t s) {
  if (A.nBlocks == 0) return;
  // Partials beyond the kMaxPartials slots of the reduction workspace are not needed: the
  // caller sizes the workspace for max(nBlocks, kMaxPartials).
  hipLaunchKernelGGL(k_resid64, dim3(A.nBlocks), dim3(kBlock), 0, s, A.rowPtr, A.colInd, A.val,
                     A.blockRows, A.nBlocks, x, b, r, (double *)w.partials, w.counters, scal);
}

template <typename TS>
void fill(TS *p, long long count, double v, hipStream_t s) {
  const long long nv = units<TS>(count);
  hipLaunchKernelGGL((k_fill<TS>), dim3(grid_for(nv)), dim3(kBlock), 0, s, p, nv, v);
}

void fill_int(int *p, long long count, int v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_int, dim3(grid_for(count)), dim3(kBlock), 0, s, p, count, v);
}

template <typename TS>
void multidot(const TS *B, long long ld, int p, const TS *x, long long n, double *out, int ostride,
              RedWork w, const int *flag, hipStream_t s) {
  if (p <= 0) return;
  using TA = typename Acc<TS>::type;
  const int G = red_blocks<TS>(n);
  const long long nv = units<TS>(n);
  // Column groups of at most kDotCols columns (56 for FP64, whose 64-wide accumulators would
  // leave one wavefront per SIMD), balanced; a group's width is rounded up to one of 1, 2, 4,
  // 8, 12, 16, 20, 24 or a multiple of 8 (the kernel template), so x is read once per group and
  // at most three loads per unit alias a column already read.
  const int widest = sizeof(TS) == 8 ? kDotCols - 8 : kDotCols;
  const int groups = (p + widest - 1) / widest;
  const int per = (p + groups - 1) / groups;
  const int cb = per <= 2    ? per
                 : per <= 4  ? 4
                 : per <= 24 ? std::max(8, (per + 3) / 4 * 4)
                             : (per + 7) / 8 * 8;
  TA *part = (TA *)w.partials;
#define DMR_MDOT(CBV)                                                                      \
  hipLaunchKernelGGL((k_multidot<TS, CBV>), dim3(G, groups), dim3(kBlock), 0, s, B, ld, p, x, \
                     nv, G, part, w.counters, out, ostride, flag)
  switch (cb) {
    case 1: DMR_MDOT(1); break;
    case 2: DMR_MDOT(2); break;
    case 4: DMR_MDOT(4); break;
    case 8: DMR_MDOT(8); break;
    case 12: DMR_MDOT(12); break;
    case 16: DMR_MDOT(16); break;
    case 20: DMR_MDOT(20); break;
    case 24: DMR_MDOT(24); break;
    case 32: DMR_MDOT(32); break;
    case 40: DMR_MDOT(40); break;
    case 48: DMR_MDOT(48); break;
    case 56: DMR_MDOT(56); break;
    default: DMR_MDOT(64); break;
  }
  static_assert(kDotCols == 64, "multidot launcher covers group widths up to 64");
#undef DMR_MDOT
}

template <typename TS>
void combo(int mode, const TS *in, TS *out, const TS *B1, int p1, const double *c1, const TS *B2,
           int p2, const double *c2, long long ld, long long n, double beta, int red, const TS *g,
           const Fin &fin, RedWork w, const int *flag, hipStream_t s, const double *betaDev) {
  using TA = typename Acc<TS>::type;
  const int G = red_blocks<TS>(n);
  const long long nv = units<TS>(n);
  TA *part = (TA *)w.partials;
  unsigned *cnt = w.counters;
#define DMR_COMBO(M, R)                                                                         \
  hipLaunchKernelGGL((k_combo<TS, M, R>), dim3(G), dim3(kBlock), 0, s, in, out, B1, p1, c1, B2, \
                     p2, c2, ld, nv, beta, g, G, part, cnt, fin, flag, betaDev)
  if (mode == COMB_SUB) {
    switch (red) {
      case 0: DMR_COMBO(COMB_SUB, 0); break;
      case 1: DMR_COMBO(COMB_SUB, 1); break;
      default: DMR_COMBO(COMB_SUB, 3); break;
    }
  } else {
    switch (red) {
      case 0: DMR_COMBO(COMB_ACC, 0); break;
      case 1: DMR_COMBO(COMB_ACC, 1); break;
      default: DMR_COMBO(COMB_ACC, 3); break;
    }
  }
#undef DMR_COMBO
}

template <typename TS>
void combo_multi(TS *out, int J, const TS *B1, int p1, const double *C1, const TS *B2, int p2,
                 const double *C2, int ldc, long long ld, long long n, double *scal, RedWork w,
                 hipStream_t s) {
  using TA = typename Acc<TS>::type;
  constexpr int JB = 64 / Unit<TS>::N;  // accumulator registers per thread stay at 64 values
  const int G = red_blocks<TS>(n);
  const int groups = (J + JB - 1) / JB;
  const size_t lds = sizeof(TA) * (size_t)(p1 + p2) * JB;
  hipLaunchKernelGGL((k_combo_multi<TS, JB>), dim3(G, groups), dim3(kBlock), lds, s, out, J, B1,
                     p1, C1, B2, p2, C2, ldc, ld, units<TS>(n), G, (TA *)w.partials, w.counters,
                     scal);
}

// Most units per thread of k_orth: its two register copies of the thread's units hold at most
// 2 x 64 values (bf16: 8 units; 16 would not fit in registers).
template <typename TS> constexpr int kOrthMaxUpt = 64 / Unit<TS>::N < 16 ? 64 / Unit<TS>::N : 16;

template <typename TS>
int orth(const TS *g, TS *dst, const TS *A, int pa, double *ha, const TS *B, int pb, double *hb,
         long long ld, long long n, int red, const Fin &fin, RedWork w, unsigned *bar,
         unsigned epoch0, int *err, const int *flag, hipStream_t s) {
  if (pa + pb > kMaxCoef || pa > kMaxCoef / 2 || pb > kMaxCoef / 2) return 0;
  const int G = red_blocks<TS>(n);
  const long long nv = units<TS>(n);
  const long long per = (long long)G * kBlock;
  const long long upt = (nv + per - 1) / per;
  OrthArgs a{g,  dst, A,  pa,     ha,  B,   pb,  hb,   ld, nv,
             G,  w.partials, bar, epoch0, red, fin, flag, err};
  void *args[] = {&a};
  hipError_t e;
  if (upt <= 1)
    e = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_orth<TS, 1>), dim3(G),
                                   dim3(kBlock), args, 0, s);
  else if (upt <= 2)
    e = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_orth<TS, 2>), dim3(G),
                                   dim3(kBlock), args, 0, s);
  else if (upt <= 4)
    e = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_orth<TS, 4>), dim3(G),
                                   dim3(kBlock), args, 0, s);
  else if (upt <= 8)
    e = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_orth<TS, 8>), dim3(G),
                                   dim3(kBlock), args, 0, s);
  else if (upt <= 16 && kOrthMaxUpt<TS> >= 16)
    e = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_orth<TS, kOrthMaxUpt<TS>>),
                                   dim3(G), dim3(kBlock), args, 0, s);
  else
    return 0;
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return (pa > 0) + (pb > 0) + 1;
}

template <typename TS>
void scale(TS *v, long long n, const double *scal, const int *flag, hipStream_t s) {
  const long long nv = units<TS>(n);
  hipLaunchKernelGGL((k_scale<TS>), dim3(grid_for(nv)), dim3(kBlock), 0, s, v, nv, scal, flag);
}

template <typename TS>
void scale_cols(TS *v, int J, long long ld, long long n, const double *scal, hipStream_t s) {
  if (J <= 0) return;
  const long long nv = units<TS>(n);
  hipLaunchKernelGGL((k_scale_cols<TS>), dim3(grid_for(nv), J), dim3(kBlock), 0, s, v, ld, nv,
                     scal);
}

template <typename TD, typename TSrc>
void convert(TD *dst, const TSrc *src, long long n, hipStream_t s) {
  hipLaunchKernelGGL((k_convert<TD, TSrc>), dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, n);
}

template <typename TS>
void add_to(double *x, const TS *d, long long n, hipStream_t s) {
  const long long nv = units<TS>(n);
  hipLaunchKernelGGL((k_add_to<TS>), dim3(grid_for(nv)), dim3(kBlock), 0, s, x, d, nv);
}

void cycle_begin(double *H, int m, int *flag, unsigned *bar, hipStream_t s) {
  const int count = (m + 1) * m;
  hipLaunchKernelGGL(k_cycle_begin, dim3((count + kBlock - 1) / kBlock), dim3(kBlock), 0, s, H,
                     count, flag, bar);
}

template <typename TS>
void breakdown_fixup(TS *P, long long ld, long long n, int o, int d, int m, double *H,
                     const int *flag, hipStream_t s) {
  (void)n;
  const long long nv = units<TS>(ld) * (m + 1);
  hipLaunchKernelGGL((k_breakdown_fixup<TS>), dim3(grid_for(nv)), dim3(kBlock), 0, s, P, ld, o, d,
                     m, H, flag);
}

// ---- explicit instantiations ----
#define DMR_INST(TS)                                                                              \
  template void fill<TS>(TS *, long long, double, hipStream_t);                                   \
  template void multidot<TS>(const TS *, long long, int, const TS *, long long, double *, int,    \
                             RedWork, const int *, hipStream_t);                                  \
  template void combo<TS>(int, const TS *, TS *, const TS *, int, const double *, const TS *, int, \
                          const double *, long long, long long, double, int, const TS *,          \
                          const Fin &, RedWork, const int *, hipStream_t, const double *);        \
  template void combo_multi<TS>(TS *, int, const TS *, int, const double *, const TS *, int,      \
                                const double *, int, long long, long long, double *, RedWork,     \
                                hipStream_t);                                                     \
  template void scale<TS>(TS *, long long, const double *, const int *, hipStream_t);             \
  template int orth<TS>(const TS *, TS *, const TS *, int, double *, const TS *, int, double *,  \
                        long long, long long, int, const Fin &, RedWork, unsigned *, unsigned,    \
                        int *, const int *, hipStream_t);                                         \
  template void scale_cols<TS>(TS *, int, long long, long long, const double *, hipStream_t);     \
  template void add_to<TS>(double *, const TS *, long long, hipStream_t);                         \
  template void breakdown_fixup<TS>(TS *, long long, long long, int, int, int, double *,          \
                                    const int *, hipStream_t);
DMR_INST(double)
DMR_INST(float)
DMR_INST(bf16)
#undef DMR_INST
template void convert<float, double>(float *, const double *, long long, hipStream_t);
template void convert<double, float>(double *, const float *, long long, hipStream_t);
template void convert<bf16, double>(bf16 *, const double *, long long, hipStream_t);
template void convert<double, bf16>(double *, const bf16 *, long long, hipStream_t);

}  // namespace dmr

// =============================================================================================
// In-place CSR product without allocation (dmrCsrApplyStatic in gDMRC.h).  The product is staged
// in a static device array, `scratch`, of kStaticRows doubles, then copied over x.  Host code
// uses the array through the address hipGetSymbolAddress returns for the current device (the
// array's name is a host-side stand-in), looked up once per device on its first call.
// =============================================================================================
namespace {
constexpr long long kStaticRows = 1LL << 22;  // 32 MiB per GPU
constexpr int kStaticDevices = 64;
}  // namespace

static __device__ __align__(256) double scratch[kStaticRows];

namespace {
std::mutex g_scratchMu;
double *g_scratchPtr[kStaticDevices] = {};
}  // namespace

extern "C" int dmrCsrApplyStatic(int n, int nnz, const int *rowPtr, const int *colInd,
                                 const double *val, double *x, hipStream_t stream) {
  dmr::clear_error();
  if (n < 1 || nnz < 0 || !rowPtr || !colInd || !val || !x)
    return dmr::set_error(DMR_ERR_ARG, "dmrCsrApplyStatic: invalid arguments");
  if ((long long)n > kStaticRows)
    return dmr::set_error(DMR_ERR_ARG,
                          "dmrCsrApplyStatic: n exceeds the static staging array (4194304 rows)");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kStaticDevices)
    return dmr::set_error(DMR_ERR_HIP, "dmrCsrApplyStatic: no current device");
  double *y = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchPtr[dev]) {
      void *ptr = nullptr;
      const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !ptr)
        return dmr::set_error(DMR_ERR_HIP, (std::string("dmrCsrApplyStatic: hipGetSymbolAddress: ") +
                                            hipGetErrorString(e)).c_str());
      g_scratchPtr[dev] = static_cast<double *>(ptr);
    }
    y = g_scratchPtr[dev];
  }
  const unsigned grid = (unsigned)((n + dmr::kBlock - 1) / dmr::kBlock);
  hipLaunchKernelGGL(dmr::k_spmv_rows, dim3(grid), dim3(dmr::kBlock), 0, stream, n, rowPtr,
                     colInd, val, (const double *)x, y);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(x, y, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, stream);
  if (e != hipSuccess)
    return dmr::set_error(DMR_ERR_HIP, (std::string("dmrCsrApplyStatic: ") +
                                        hipGetErrorString(e)).c_str());
  return DMR_OK;
}
