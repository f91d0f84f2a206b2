// Host driver of the block (multi-right-hand-side) DMR and its C ABI (dmrBlock* in gDMR.h).
//
// s <= 32 right-hand sides of one operator share one block Krylov space and one deflation
// subspace Z (n x k).  The cycle is the DMR cycle of the single-vector solver (gDMR.cu:1790-2582
// restated for blocks) with every vector operation turned into a block one:
//   deflation image  A Z = W R_Z             SpMM + Cholesky QR (twice)
//   start block      V_0 S_0 = (I - W W^T) R  gemm_tn + gemm_nn + Cholesky QR
//   block step j     U = A V_j; C = [W V_0..V_j]^T U; U -= [W V_0..V_j] C; U = V_{j+1} S_{j+1}
//                    (one-stage block classical Gram-Schmidt: the projection and the update are
//                    tall-skinny GEMMs on MFMA; the normalisation is CholQR2)
//   least squares    T = P^T R (full projection); min ||T - H Y|| column by column on the block
//                    Hessenberg H (lower bandwidth s)
//   updates          R -= P (H Y); X_a += [Z V] Y diag(rho); rho_i *= ||R_i||; R_i /= ||R_i||
//   refresh          Z <- normalize([Z V_0 .. V_{kr-1}] C), C the k smallest right singular
//                    vectors of the leading (k + (kr+1) s) x (k + kr s) block of H: the block
//                    analogue of genDeflatDMR (gDMR.cu:1583-1704) on the widest leading subspace
//                    whose one-workgroup Jacobi SVD fits in LDS (DMR_BLOCK_KR caps kr)
// in FP32 storage and MFMA arithmetic (FP32 accumulation), inside FP64 iterative refinement per
// column: R = B - A X in FP64, inner block DMR on R / ||R||, X += ||R|| D.  The Hessenberg least
// squares, the Cholesky factors and the SVD are FP64.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "dmr_block.h"
#include "gDMR.h"

namespace {

constexpr double kSvdTol = 1e-14;  // Jacobi stopping tolerance (DMR_BLOCK_SVD_TOL overrides)
constexpr int kSvdMaxSweeps = 128;
constexpr double kDepTol = 1e-12;  // Cholesky pivot / max pivot below which a column is dependent
constexpr int kMaxEvents = 8192;
constexpr int kStallCycles = 25;  // inner cycles without 0.1% progress before a refinement restart

#define BCHK(call)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return dmr::set_error(DMR_ERR_HIP, (std::string(#call) + ": " + hipGetErrorString(e_)).c_str()); \
  } while (0)

}  // namespace

struct dmrBlockContext {
  int device = 0, n = 0, mb = 0, k = 0, sMax = 0;
  long long ld = 0;
  int rowsMax = 0, colsMax = 0, G = 0;
  dmr::CsrDev csr;
  float *ownedVal32 = nullptr;
  bool val32Stale = false;
  bool haveOp = false;
  // Per-column FP64 operators of the refinement residual (dmrBlockSetColumnOperators): column c
  // uses colVal + c * colLd; colCount columns at most.  innerMean: the FP32 inner operator is
  // their entrywise mean, recomputed at the start of every solve.
  const double *colVal = nullptr;
  long long colLd = 0;
  int colCount = 0;
  bool innerMean = false;
  void *big = nullptr, *small = nullptr;
  // P: the basis [W V_0 V_1 ...], FP32 or (bfBasis) bf16 storage in the same allocation.  Ub: the
  // newest block in FP32 while it is orthogonalised (the product A V_j, or A Z, or the start
  // block), written into its slot of P by the last normalisation.
  float *P = nullptr, *Z = nullptr, *Zn = nullptr, *R = nullptr, *Xa = nullptr, *Ub = nullptr;
  bool bfBasis = false;
  double *R64 = nullptr;
  double *H = nullptr, *T = nullptr, *Hw = nullptr, *Tw = nullptr, *Y = nullptr;
  float *C32 = nullptr, *Yr32 = nullptr, *ZH32 = nullptr, *Sinv32 = nullptr, *Cz32 = nullptr;
  float *Cpair = nullptr;  // coefficients of the paired residual and solution update
  float *part = nullptr;
  size_t partFloats = 0;
  double *G64 = nullptr, *Sacc = nullptr, *Csvd = nullptr, *sig = nullptr;
  double *rho = nullptr, *q = nullptr, *scal = nullptr, *bn = nullptr, *rn = nullptr,
         *f = nullptr, *zq = nullptr, *cpart = nullptr;
  int *flags = nullptr;  // [0] dependent column seen this cycle, [1] Jacobi sweeps
  double *gcSeg = nullptr;   // gram_chol's segment sums
  unsigned *gcCtr = nullptr;  // gram_chol's arrival counter (zero between launches)
  double *host = nullptr;  // pinned staging
  bool hasZ = false;
  int krMax = 64;  // cap on the leading block steps of the subspace refresh
  double svdTol = kSvdTol;
  int cholPasses = 0;  // Cholesky QR passes per block normalisation; 0: by storage type
  bool fusedGram = true;  // first Gram matrix from the update's epilogue where possible
  bool gramChol = true;   // its reduce and the Cholesky step in one launch (gram_chol)
  int spmmForce = 0;      // SpMM variant pinned by DMR_BLOCK_SPMM (0: the tuned one)
  bool foldT = true;      // R's projection in the last block step's pass where possible
  bool svdSide = true;    // the refresh's SVD on the side stream (DMR_BLOCK_SVD_SIDE=0: inline)
  // Launch choices pinned over the tuned ones (DMR_BLOCK_NN_BASIS: the update over a bf16 basis,
  // 4, 8 or 16 for the FP32 matrix cores, 100 or 101 for the bf16 ones; DMR_BLOCK_TN_VAR: the
  // projection variant, 2 or 3 for a bf16 basis); 0: as tuned.
  int nnPin = 0, tnVarPin = 0;
  // Launch shapes chosen by the autotuner (per block width s), and whether it ran.
  struct Tuning {
    int tnRowsProj = 512, tnVarProj = 0;  // gemm_tn for projections (p > s)
    int tnRowsGram = 512, tnVarGram = 0;  // gemm_tn for Gram matrices (p = s)
    int nnUnroll = 8;                     // gemm_nn loads in flight per wavefront (KiB)
    int nnBasis = 8;                      // the same for a bf16 basis; >= 100: bf16 matrix cores
    int spmmCols = 8;                     // SpMM columns per thread
    int s = 0;
  } tune;
  bool autotune = true;
  // Captured restart cycles, keyed by (s, deflated, which buffer holds Z).
  bool graphs = true;
  hipStream_t gst = nullptr;
  // The subspace refresh's SVD runs on a side stream, overlapped with the least squares and the
  // residual and solution updates.
  hipStream_t side = nullptr;
  hipEvent_t evFork = nullptr, evJoin = nullptr;
  // Hand-off between the caller's stream and gst, on which every captured cycle is replayed.
  hipEvent_t evIn = nullptr, evOut = nullptr;
  struct Cap {
    int s, deflate, parity;
    hipGraphExec_t exec;
    bool refreshed;
    int matvecs;
  };
  std::vector<Cap> caps;
  float *zbuf0 = nullptr;
  double innerTol = 1e-5;
  int maxOuter = 30;
  dmrStats st{};
  std::vector<double> hist;
  struct Prof {
    bool on = false;
    double ms[DMR_NPHASES] = {};
    double by[DMR_NPHASES] = {};
    double fl[DMR_NPHASES] = {};
    long long cnt[DMR_NPHASES] = {};
    std::vector<hipEvent_t> ev;
    std::vector<int> phase;
    std::vector<double> bytes, flops;
    int used = 0;
  } prof;
};

namespace {

using dmr::blk::kMaxRhs;

// Rows per gemm_tn workgroup (DMR_BLOCK_TN_ROWS overrides; multiple of 32).
int tn_rows_target() {
  static const int v = [] {
    const char *e = std::getenv("DMR_BLOCK_TN_ROWS");
    const int x = e ? std::atoi(e) : 0;
    return x >= 32 ? (x + 31) / 32 * 32 : 512;
  }();
  return v;
}

template <typename F>
void timedf(dmrBlockContext *c, hipStream_t s, int phase, double bytes, double flops, F &&f) {
  auto &p = c->prof;
  if (!p.on || p.used >= kMaxEvents) {
    f();
    return;
  }
  if (p.ev.empty()) {
    p.ev.resize(2 * kMaxEvents);
    for (auto &e : p.ev) hipEventCreate(&e);
    p.phase.resize(kMaxEvents);
    p.bytes.resize(kMaxEvents);
    p.flops.resize(kMaxEvents);
  }
  const int i = p.used++;
  hipEventRecord(p.ev[2 * i], s);
  f();
  hipEventRecord(p.ev[2 * i + 1], s);
  p.phase[i] = phase;
  p.bytes[i] = bytes;
  p.flops[i] = flops;
}

template <typename F>
void timed(dmrBlockContext *c, hipStream_t s, int phase, double bytes, F &&f) {
  timedf(c, s, phase, bytes, 0.0, f);
}

void prof_collect(dmrBlockContext *c) {
  auto &p = c->prof;
  for (int i = 0; i < p.used; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.ev[2 * i], p.ev[2 * i + 1]) == hipSuccess) {
      p.ms[p.phase[i]] += ms;
      p.by[p.phase[i]] += p.bytes[i];
      p.fl[p.phase[i]] += p.flops[i];
      ++p.cnt[p.phase[i]];
    }
  }
  p.used = 0;
}

struct Sizes {
  size_t big, small;
};

Sizes sizes(long long n, int mb, int k, int sMax, int G, long long ld) {
  const int rowsMax = k + (mb + 1) * sMax, colsMax = k + mb * sMax;
  const int pPadMax = (rowsMax + 31) & ~31;
  const int cRed = colsMax;
  Sizes z;
  z.big = sizeof(float) * (size_t)ld * (size_t)(rowsMax + 2 * k + 2 * sMax + std::max(sMax, k)) +
          sizeof(double) * (size_t)ld * sMax;
  z.small = sizeof(double) * ((size_t)rowsMax * colsMax * 2 + (size_t)rowsMax * 32 * 2 +
                              (size_t)colsMax * 32 + 32 * 32 * 2 +
                              (size_t)cRed * 32 + cRed + 2 + 8 * 64 +
                              (size_t)dmr::blk::colnorm_blocks(n) * 32) +
            sizeof(float) * ((size_t)rowsMax * 32 * 3 + (size_t)colsMax * 32 + 32 * 32 +
                             (size_t)cRed * 32 + (size_t)G * pPadMax * 32) +
            sizeof(int) * 64 + 64 * 32 +
            sizeof(double) * 256 * dmr::blk::kGramCholSegs + sizeof(unsigned) * 64;
  return z;
}

void drop_graphs(dmrBlockContext *c) {
  for (auto &g : c->caps) hipGraphExecDestroy(g.exec);
  c->caps.clear();
}

void ctx_release(dmrBlockContext *c) {
  if (!c) return;
  hipSetDevice(c->device);
  hipDeviceSynchronize();
  drop_graphs(c);
  if (c->gst) hipStreamDestroy(c->gst);
  if (c->side) hipStreamDestroy(c->side);
  if (c->evFork) hipEventDestroy(c->evFork);
  if (c->evJoin) hipEventDestroy(c->evJoin);
  if (c->evIn) hipEventDestroy(c->evIn);
  if (c->evOut) hipEventDestroy(c->evOut);
  if (c->big) hipFree(c->big);
  if (c->small) hipFree(c->small);
  if (c->ownedVal32) hipFree(c->ownedVal32);
  if (c->host) hipHostFree(c->host);
  for (auto &e : c->prof.ev) hipEventDestroy(e);
  delete c;
}

// Carves aligned sub-arrays out of a pool.
struct Carver {
  char *p;
  template <typename T> T *take(size_t count) {
    T *r = reinterpret_cast<T *>(p);
    p += (sizeof(T) * count + 255) / 256 * 256;
    return r;
  }
};

int ctx_alloc(dmrBlockContext *c) {
  const long long ld = c->ld;
  const int k = c->k, s = c->sMax, mb = c->mb;
  Sizes z = sizes(c->n, mb, k, s, c->G, ld);
  z.small += 64 * 256;  // alignment slack of the carver
  z.big += 16 * 256;
  BCHK(hipMalloc(&c->big, z.big));
  BCHK(hipMemset(c->big, 0, z.big));
  BCHK(hipMalloc(&c->small, z.small));
  BCHK(hipMemset(c->small, 0, z.small));
  BCHK(hipHostMalloc((void **)&c->host, sizeof(double) * 4 * 64, hipHostMallocDefault));
  Carver b{(char *)c->big};
  c->P = b.take<float>((size_t)ld * c->rowsMax);
  c->Z = b.take<float>((size_t)ld * std::max(k, 1));
  c->Zn = b.take<float>((size_t)ld * std::max(k, 1));
  c->R = b.take<float>((size_t)ld * s);
  c->Xa = b.take<float>((size_t)ld * s);
  c->R64 = b.take<double>((size_t)ld * s);
  c->Ub = b.take<float>((size_t)ld * std::max(s, std::max(k, 1)));
  Carver m{(char *)c->small};
  const int rowsMax = c->rowsMax, colsMax = c->colsMax;
  const int pPadMax = (rowsMax + 31) & ~31;
  const int cRed = colsMax;
  c->H = m.take<double>((size_t)rowsMax * colsMax);
  c->Hw = m.take<double>((size_t)rowsMax * colsMax);
  c->T = m.take<double>((size_t)rowsMax * 32);
  c->Tw = m.take<double>((size_t)rowsMax * 32);
  c->Y = m.take<double>((size_t)colsMax * 32);
  c->G64 = m.take<double>(32 * 32);
  c->Sacc = m.take<double>(32 * 32);
  c->Csvd = m.take<double>((size_t)cRed * 32);
  c->sig = m.take<double>(cRed + 2);
  c->rho = m.take<double>(64);
  c->q = m.take<double>(64);
  c->scal = m.take<double>(64);
  c->bn = m.take<double>(64);
  c->rn = m.take<double>(64);
  c->f = m.take<double>(64);
  c->zq = m.take<double>(64);
  c->cpart = m.take<double>((size_t)dmr::blk::colnorm_blocks(c->n) * 32);
  c->C32 = m.take<float>((size_t)rowsMax * 32);
  c->ZH32 = m.take<float>((size_t)rowsMax * 32);
  c->Cpair = m.take<float>((size_t)rowsMax * 32);
  c->Yr32 = m.take<float>((size_t)colsMax * 32);
  c->Sinv32 = m.take<float>(32 * 32);
  c->Cz32 = m.take<float>((size_t)cRed * 32);
  c->partFloats = (size_t)c->G * pPadMax * 32;
  c->part = m.take<float>(c->partFloats);
  c->flags = m.take<int>(64);
  c->gcSeg = m.take<double>((size_t)256 * dmr::blk::kGramCholSegs);
  c->gcCtr = m.take<unsigned>(64);
  return DMR_OK;
}

int tn_blocks(const dmrBlockContext *c, int rowsTarget) {
  return dmr::blk::gemm_tn_blocks(c->n, dmr::blk::kMaxTnBlocks, rowsTarget);
}

using dmr::bf16;

// Bytes the fixed-order sum of G partials of a p-row product reads (16 or 32 columns each).
double reduce_bytes(int G, int p, int s) { return 4.0 * G * p * (s <= 16 ? 16 : 32); }

// C = P^T U with the tuned shape of its class (Gram matrix or projection): the partials timed
// under `phase` (bytes and flops of the product), their fixed-order sum under DMR_PH_REDUCE.
template <typename TP>
void tn(dmrBlockContext *c, int phase, double bytes, double flops, const TP *P, int p,
        const float *U, int s, double *C64, int ldc, float *C32, bool gram, hipStream_t st) {
  const auto &t = c->tune;
  const int G = tn_blocks(c, gram ? t.tnRowsGram : t.tnRowsProj);
  timedf(c, st, phase, bytes, flops, [&] {
    dmr::blk::gemm_tn(P, c->ld, p, U, c->ld, s, c->n, G, gram ? t.tnVarGram : t.tnVarProj,
                      c->part, C64, ldc, C32, st, false);
  });
  timed(c, st, DMR_PH_REDUCE, reduce_bytes(G, p, s),
        [&] { dmr::blk::gemm_tn_reduce(c->part, G, p, s, C64, ldc, C32, st); });
}

template <typename TP, typename TO>
void nn(dmrBlockContext *c, const TP *P, int p, const float *C, const float *In, float a,
        float b, TO *Out, int s, hipStream_t st) {
  const int un = std::is_same<TP, bf16>::value ? c->tune.nnBasis : c->tune.nnUnroll;
  dmr::blk::gemm_nn(P, c->ld, p, C, In, c->ld, a, b, Out, c->ld, s, c->n, un, st);
}

// Partial Gram matrices from the epilogue of a bf16-basis update (see gemm_nn): how many, or 0
// when this block width, launch shape or partial buffer does not allow it.
int gram_parts(const dmrBlockContext *c, int s) {
  if (!c->fusedGram || !c->bfBasis || s > 16 || c->tune.nnBasis < 100) return 0;
  const int G = dmr::blk::gemm_nn_gram_blocks(c->n, c->tune.nnBasis);
  return (size_t)G * 1024 <= c->partFloats ? G : 0;
}

// Out = a In + b P C over a bf16 basis, writing the partial Gram matrices of Out (gram_parts).
void nn_gram(dmrBlockContext *c, const bf16 *P, int p, const float *C, const float *In, float a,
             float b, float *Out, int s, hipStream_t st) {
  dmr::blk::gemm_nn(P, c->ld, p, C, In, c->ld, a, b, Out, c->ld, s, c->n, c->tune.nnBasis, st,
                    c->part);
}

template <typename TX>
void spmm(dmrBlockContext *c, const TX *X, float *Y, int s, hipStream_t st) {
  dmr::blk::spmm32(c->csr, X, c->ld, Y, c->ld, s, c->spmmForce ? c->spmmForce : c->tune.spmmCols,
                   st);
}

// Process-wide table of autotuned shapes, keyed by the problem shape.
struct TuneKey {
  long long n, nnz;
  int s, k, mb, bf;
  bool operator==(const TuneKey &o) const {
    return n == o.n && nnz == o.nnz && s == o.s && k == o.k && mb == o.mb && bf == o.bf;
  }
};
std::mutex g_tuneMu;
std::vector<std::pair<TuneKey, dmrBlockContext::Tuning>> g_tuned;

// Persisted shapes: a text file (dmrBlockTuningFile, else DMR_TUNE_FILE) of one line per tuned
// shape, "arch n nnz s k mb bf  tnRowsProj tnVarProj tnRowsGram tnVarGram nnUnroll nnBasis
// spmmCols", read into the table before the first tuning of the process (lines of other device
// architectures are skipped) and appended to after each new tuning.  Every process that uses
// the file launches the same kernels for a shape, so its solves are bit-reproducible across
// processes as well.
std::string g_tuneFile;
bool g_tuneFileSet = false;
bool g_tuneLoaded = false;

std::string device_arch() {
  int d = 0;
  hipDeviceProp_t p{};
  if (hipGetDevice(&d) != hipSuccess || hipGetDeviceProperties(&p, d) != hipSuccess) return "";
  std::string a = p.gcnArchName;
  return a.substr(0, a.find(':'));
}

const std::string &tune_file() {
  if (!g_tuneFileSet) {
    g_tuneFileSet = true;
    if (const char *e = std::getenv("DMR_TUNE_FILE")) g_tuneFile = e;
  }
  return g_tuneFile;
}

// Caller holds g_tuneMu.
void tune_file_load_locked() {
  if (g_tuneLoaded) return;
  g_tuneLoaded = true;
  const std::string &path = tune_file();
  if (path.empty()) return;
  FILE *f = std::fopen(path.c_str(), "r");
  if (!f) return;
  const std::string arch = device_arch();
  char line[512];
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    char a[64];
    TuneKey k{};
    dmrBlockContext::Tuning t;
    if (std::sscanf(line, "%63s %lld %lld %d %d %d %d %d %d %d %d %d %d %d", a, &k.n, &k.nnz, &k.s,
                    &k.k, &k.mb, &k.bf, &t.tnRowsProj, &t.tnVarProj, &t.tnRowsGram,
                    &t.tnVarGram, &t.nnUnroll, &t.nnBasis, &t.spmmCols) != 14 ||
        arch != a)
      continue;
    t.s = k.s;
    bool have = false;
    for (auto &e : g_tuned) have = have || e.first == k;
    if (!have) g_tuned.emplace_back(k, t);
  }
  std::fclose(f);
}

// Caller holds g_tuneMu.
void tune_file_append_locked(const TuneKey &k, const dmrBlockContext::Tuning &t) {
  const std::string &path = tune_file();
  if (path.empty()) return;
  FILE *f = std::fopen(path.c_str(), "a");
  if (!f) return;
  std::fprintf(f, "%s %lld %lld %d %d %d %d %d %d %d %d %d %d %d\n", device_arch().c_str(), k.n,
               k.nnz, k.s, k.k, k.mb, k.bf, t.tnRowsProj, t.tnVarProj, t.tnRowsGram, t.tnVarGram,
               t.nnUnroll, t.nnBasis, t.spmmCols);
  std::fclose(f);
}

// Times each candidate launch shape on the handle's own buffers (HIP events, best of 3 after a
// warm-up launch) and keeps the fastest: gemm_tn rows per workgroup and variant, separately for
// projections over the full basis and for s x s Gram matrices; the gemm_nn load-ahead depth; the
// SpMM columns per thread.  Results are cached per shape for the process, so every handle of a
// shape launches the same kernels (and computes the same bits).  The buffers it writes (Xa, the
// partials) are re-initialised by the solve that follows.
template <typename TP>
void autotune_t(dmrBlockContext *c, TP *P, int s, hipStream_t st) {
  constexpr bool bf = std::is_same<TP, bf16>::value;
  const TuneKey key{c->n, c->csr.nnz, s, c->k, c->mb, bf ? 1 : 0};
  {
    std::lock_guard<std::mutex> lk(g_tuneMu);
    tune_file_load_locked();
    for (auto &e : g_tuned)
      if (e.first == key) {
        c->tune = e.second;
        return;
      }
  }
  dmrBlockContext::Tuning t;
  t.s = s;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  auto best_of = [&](auto &&launch) {
    launch();
    float best = 1e30f;
    for (int r = 0; r < 3; ++r) {
      hipEventRecord(e0, st);
      launch();
      hipEventRecord(e1, st);
      hipEventSynchronize(e1);
      float ms = 0.f;
      hipEventElapsedTime(&ms, e0, e1);
      best = std::min(best, ms);
    }
    return best;
  };
  const int pFull = c->k + (c->mb + 1) * s;
  // Rows per workgroup; a larger range means fewer partials for the fixed-order reduce timed
  // with each candidate (at n = 2^21 the targets up to 2048 all hit the 1024-workgroup cap).
  const int rowsCand[] = {256, 512, 1024, 2048, 4096, 8192};
  float bp = 1e30f, bg = 1e30f;
  for (int rows : rowsCand)
    for (int var = 0; var < 4; ++var) {
      const int G = tn_blocks(c, rows);
      // A bf16 basis: the FP32 (2) or bf16 (3) matrix-core kernels; FP32 Gram matrices: 0, 1, 2.
      const float tp = (bf ? var < 2 : var == 3) ? 1e30f : best_of([&] {
        dmr::blk::gemm_tn(P, c->ld, pFull, c->R, c->ld, s, c->n, G, var, c->part, nullptr, 0,
                          c->C32, st);
      });
      if (tp < bp) {
        bp = tp;
        t.tnRowsProj = rows;
        t.tnVarProj = var;
      }
      const float tg = var == 3 ? 1e30f : best_of([&] {
        dmr::blk::gemm_tn(c->R, c->ld, s, c->R, c->ld, s, c->n, G, var, c->part, c->G64, s,
                          nullptr, st);
      });
      if (tg < bg) {
        bg = tg;
        t.tnRowsGram = rows;
        t.tnVarGram = var;
      }
    }
  float bn = 1e30f;
  for (int un : {4, 8, 16, 100, 101}) {
    if (!bf && un >= 100) continue;
    const float tt = best_of([&] {
      dmr::blk::gemm_nn(P, c->ld, pFull, c->C32, c->Xa, c->ld, 1.0f, 0.0f, c->Xa, c->ld, s,
                        c->n, un, st);
    });
    if (tt < bn) {
      bn = tt;
      t.nnUnroll = t.nnBasis = un;
    }
  }
  if (bf) {
    // FP32 operands of a bf16-basis cycle (normalisations, subspace products): p = s columns.
    float bq = 1e30f;
    for (int un : {4, 8, 16}) {
      const float tt = best_of([&] {
        dmr::blk::gemm_nn(c->Ub, c->ld, s, c->C32, c->Xa, c->ld, 1.0f, 0.0f, c->Xa, c->ld, s,
                          c->n, un, st);
      });
      if (tt < bq) {
        bq = tt;
        t.nnUnroll = un;
      }
    }
  }
  float bs = 1e30f;
  for (int cg : {4, 8, 16, 104, 108, 116, 208, 216}) {
    if (bf && cg < 100) continue;  // bf16 vectors: the up-front variant only
    const float tt = best_of([&] {
      dmr::blk::spmm32(c->csr, P, c->ld, c->Xa, c->ld, s, cg, st);
    });
    if (tt < bs) {
      bs = tt;
      t.spmmCols = cg;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  c->tune = t;
  std::lock_guard<std::mutex> lk(g_tuneMu);
  g_tuned.emplace_back(key, t);
  tune_file_append_locked(key, t);
}

void autotune(dmrBlockContext *c, int s, hipStream_t st) {
  if (c->bfBasis)
    autotune_t(c, reinterpret_cast<bf16 *>(c->P), s, st);
  else
    autotune_t(c, c->P, s, st);
}

// Two Cholesky QR passes on the sc columns of U (in place).  The combined triangular factor goes
// to Hdst (ld ldh) when given.
// gramParts > 0: the first pass's Gram matrix is already in c->part as that many partials
// (written by the update that produced U).
template <typename TO>
void cholqr2(dmrBlockContext *c, float *U, int sc, double *Hdst, int ldh, TO *dst,
             hipStream_t st, int gramParts = 0) {
  const long long n = c->n;
  // A bf16-stored block is orthogonal only to bf16 rounding (about 4e-3), so one Cholesky QR
  // pass, orthogonal to about kappa^2 * 6e-8, is enough there; FP32 storage keeps two.
  const int passes = c->cholPasses ? c->cholPasses : (c->bfBasis ? 1 : 2);
  for (int pass = 0; pass < passes; ++pass) {
    double *hd = pass == passes - 1 ? Hdst : nullptr;
    if (pass == 0 && gramParts > 0 && c->gramChol) {
      timed(c, st, DMR_PH_REDUCE, reduce_bytes(gramParts, sc, sc), [&] {
        dmr::blk::gram_chol(c->part, gramParts, sc, kDepTol, c->Sinv32, c->Sacc, 1, hd, ldh,
                            c->flags, c->gcSeg, c->gcCtr, st);
      });
    } else {
      if (pass == 0 && gramParts > 0) {
        timed(c, st, DMR_PH_REDUCE, reduce_bytes(gramParts, sc, sc),
              [&] { dmr::blk::gram_reduce(c->part, gramParts, sc, c->G64, st); });
      } else {
        tn(c, DMR_PH_ORTH, 4.0 * n * sc, 2.0 * n * sc * sc, U, sc, U, sc, c->G64, sc, nullptr,
           true, st);
      }
      timed(c, st, DMR_PH_ORTH, 0.0, [&] {
        dmr::blk::chol_qr(c->G64, sc, kDepTol, c->Sinv32, c->Sacc, pass == 0, hd, ldh, c->flags,
                          st);
      });
    }
    timedf(c, st, DMR_PH_ORTH, 4.0 * n * sc + (double)sizeof(TO) * n * sc,
           2.0 * n * sc * sc, [&] {
      if (pass < passes - 1) {
        if (!dmr::blk::scale_block(U, c->ld, c->Sinv32, U, c->ld, sc, n, st))
          nn(c, U, sc, c->Sinv32, nullptr, 0.0f, 1.0f, U, sc, st);
      } else if (!dmr::blk::scale_block(U, c->ld, c->Sinv32, dst, c->ld, sc, n, st)) {
        nn(c, U, sc, c->Sinv32, nullptr, 0.0f, 1.0f, dst, sc, st);
      }
    });
  }
}

double spmm_bytes(const dmrBlockContext *c, int s) {
  return (double)c->csr.nnz * 8.0 + 4.0 * (c->n + 1) + 8.0 * (double)c->n * s;
}

// One restart cycle on the normalised residual block c->R (s columns), with the basis stored as
// TP (float or bf16).  Returns whether the refreshed subspace was written to c->Zn.
template <typename TP>
bool cycle_t(dmrBlockContext *c, TP *P, int s, bool deflate, hipStream_t st) {
  const long long n = c->n, ld = c->ld;
  const int k = c->k, mb = c->mb;
  const int p0 = deflate ? k : 0;
  const int rows = p0 + (mb + 1) * s, cols = p0 + mb * s;
  float *U = c->Ub;
  const int gp = gram_parts(c, s);  // Gram matrices from the update's epilogue, when possible
  // R's projection folded into the last block step (bf16 basis, matrix-core projection,
  // s <= 16: U and R as the two 16-column tiles of one pass).
  const bool foldT = std::is_same<TP, bf16>::value && s <= 16 && mb >= 1 &&
                     c->tune.tnVarProj == 3 && c->foldT;
  const double vb = 4.0 * (double)n;              // one FP32 column
  const double pb = (double)sizeof(TP) * (double)n;  // one basis column
  dmr::fill<double>(c->H, (long long)rows * cols, 0.0, st);
  dmr::blk::zero_ints(c->flags, 2, st);
  // The subspace refresh uses the k smallest right singular vectors of the leading block of H
  // over [Z V_0 .. V_{kr-1}] (the widest whose SVD fits in LDS, at least k columns).  That block
  // is final once block step kr - 1 is done, so its SVD runs on the side stream from there on,
  // overlapped with the remaining block steps, the least squares and the updates.
  bool refreshed = false;
  int kr = 0, cR = 0, rR = 0;
  if (k > 0) {
    kr = deflate ? 1 : std::max(1, (k + s - 1) / s);
    while (kr < std::min(mb, c->krMax) &&
           dmr::svd_lds(p0 + (kr + 2) * s, p0 + (kr + 1) * s) <= 160 * 1024)
      ++kr;
    if (kr <= mb) {
      cR = p0 + kr * s;
      rR = cR + s;
      refreshed = cR >= k && dmr::svd_lds(rR, cR) <= 160 * 1024;
    }
  }
  hipStream_t ss = c->svdSide ? c->side : st;
  auto fork_svd = [&] {
    if (c->svdSide) {
      hipEventRecord(c->evFork, st);
      hipStreamWaitEvent(c->side, c->evFork, 0);
    }
    // The SVD reads the leading rR x cR block of H in place: the steps still running on the
    // main stream write only columns cR and beyond.
    timed(c, ss, DMR_PH_SVD, 0.0, [&] {
      dmr::svd_smallest_rect(c->H, rR, rows, cR, k, c->svdTol, kSvdMaxSweeps, 1, c->Csvd,
                             c->sig, c->flags + 1, ss);
      dmr::blk::pack_coef(c->Csvd, cR, k, c->Cz32, ss);
    });
    if (c->svdSide) hipEventRecord(c->evJoin, c->side);
  };
  if (deflate) {
    timedf(c, st, DMR_PH_SPMV, spmm_bytes(c, k), 2.0 * c->csr.nnz * k,
          [&] { spmm(c, c->Z, U, k, st); });
    c->st.matvecs += k;
    cholqr2(c, U, k, c->H, rows, P, st);
  }
  TP *V0 = P + (long long)p0 * ld;
  if (deflate) {
    tn(c, DMR_PH_GEMVT, pb * k + vb * s, 2.0 * n * k * s, P, k, c->R, s, nullptr, 0, c->C32,
       false, st);
    timedf(c, st, DMR_PH_UPDATE, pb * k + vb * 2 * s, 2.0 * n * k * s, [&] {
      if constexpr (std::is_same<TP, bf16>::value) {
        if (gp) {
          nn_gram(c, P, k, c->C32, c->R, 1.0f, -1.0f, U, s, st);
          return;
        }
      }
      nn(c, P, k, c->C32, c->R, 1.0f, -1.0f, U, s, st);
    });
  } else {
    hipMemcpyAsync(U, c->R, sizeof(float) * ld * s, hipMemcpyDeviceToDevice, st);
  }
  cholqr2(c, U, s, (double *)nullptr, 0, V0, st, deflate ? gp : 0);
  for (int j = 0; j < mb; ++j) {
    const TP *Vj = P + (long long)(p0 + j * s) * ld;
    const int pj = p0 + (j + 1) * s;
    TP *slot = P + (long long)pj * ld;
    double *Hcol = c->H + (long long)(p0 + j * s) * rows;
    timedf(c, st, DMR_PH_SPMV, spmm_bytes(c, s) - (4.0 - sizeof(TP)) * n * s,
           2.0 * c->csr.nnz * s, [&] { spmm(c, Vj, U, s, st); });
    c->st.matvecs += s;
    // The last block step also projects the cycle's residual block R on P[:, 0:pj] in the same
    // pass (to T), leaving only the last block's projection for the end of the cycle.
    const bool withT = foldT && j == mb - 1;
    bool paired = false;
    if constexpr (std::is_same<TP, bf16>::value) {
      if (withT) {
        paired = true;
        const int G = tn_blocks(c, c->tune.tnRowsProj);
        timedf(c, st, DMR_PH_GEMVT, pb * pj + vb * s * 2, 2.0 * n * pj * s * 2, [&] {
          dmr::blk::gemm_tn_pair(P, ld, pj, U, c->R, ld, s, n, G, c->part, Hcol, rows, c->C32,
                                 c->T, rows, st, false);
        });
        timed(c, st, DMR_PH_REDUCE, reduce_bytes(G, pj, 32), [&] {
          dmr::blk::gemm_tn_reduce(c->part, G, pj, s, Hcol, rows, c->C32, st, true, c->T, rows);
        });
      }
    }
    if (!paired)
      tn(c, DMR_PH_GEMVT, pb * pj + vb * s, 2.0 * n * pj * s, P, pj, U, s, Hcol, rows, c->C32,
         false, st);
    timedf(c, st, DMR_PH_UPDATE, pb * pj + vb * 2 * s, 2.0 * n * pj * s, [&] {
      if constexpr (std::is_same<TP, bf16>::value) {
        if (gp) {
          nn_gram(c, P, pj, c->C32, U, 1.0f, -1.0f, U, s, st);
          return;
        }
      }
      nn(c, P, pj, c->C32, U, 1.0f, -1.0f, U, s, st);
    });
    cholqr2(c, U, s, Hcol + pj, rows, slot, st, gp);
    if (refreshed && j == kr - 1) fork_svd();
  }
  // Projection of the residual block on P (the part the last block step left), least squares,
  // residual and solution updates.
  if (foldT) {
    tn(c, DMR_PH_GEMVT, pb * s + vb * s, 2.0 * n * s * s, P + (long long)cols * ld, s, c->R, s,
       c->T + cols, rows, nullptr, false, st);
  } else {
    tn(c, DMR_PH_GEMVT, pb * rows + vb * s, 2.0 * n * rows * s, P, rows, c->R, s, c->T, rows,
       nullptr, false, st);
  }
  timed(c, st, DMR_PH_LSQ, 0.0, [&] {
    dmr::blk::lsq_block(c->H, rows, cols, s, c->T, s, c->rho, c->Hw, c->Tw, c->Y, c->Yr32,
                        c->ZH32, st);
  });
  // R -= P (H Y) and Xa += [V_0 .. V_{mb-1}] Y: over a bf16 basis with the matrix-core update,
  // one pass over P for both (gemm_nn_pair); else two.
  bool paired = false;
  if constexpr (std::is_same<TP, bf16>::value) paired = s <= 16 && c->tune.nnBasis >= 100;
  if (paired) {
    if constexpr (std::is_same<TP, bf16>::value) {
      timedf(c, st, DMR_PH_COMBO, pb * rows + vb * 4 * s, 2.0 * n * (rows + cols - p0) * s, [&] {
        dmr::blk::pack_pair(c->ZH32, c->Yr32, p0, cols, rows, s, c->Cpair, st);
        dmr::blk::gemm_nn_pair(P, ld, rows, c->Cpair, c->R, c->Xa, ld, s, n, c->tune.nnBasis,
                               st);
      });
    }
  } else {
    timedf(c, st, DMR_PH_COMBO, pb * rows + vb * 2 * s, 2.0 * n * rows * s, [&] {
      nn(c, P, rows, c->ZH32, c->R, 1.0f, -1.0f, c->R, s, st);
    });
  }
  timed(c, st, DMR_PH_SCALE, vb * s,
        [&] { dmr::blk::colnorms<float>(c->R, ld, s, n, c->cpart, c->q, 1, st); });
  if (deflate)
    timedf(c, st, DMR_PH_ZGEMM, vb * (k + 2 * s), 2.0 * n * k * s, [&] {
      nn(c, c->Z, k, c->Yr32, c->Xa, 1.0f, 1.0f, c->Xa, s, st);
    });
  if (!paired)
    timedf(c, st, DMR_PH_COMBO, pb * mb * s + vb * 2 * s, 2.0 * n * mb * s * s, [&] {
      nn(c, V0, mb * s, c->Yr32 + (long long)p0 * 32, c->Xa, 1.0f, 1.0f, c->Xa, s, st);
    });
  // Subspace refresh, part 2: Z <- normalize([Z V_0 .. V_{kr-1}] C) once the SVD has joined.
  if (refreshed) {
    if (c->svdSide) hipStreamWaitEvent(st, c->evJoin, 0);
    const int kv = cR - p0;  // kr * s Krylov columns
    timed(c, st, DMR_PH_ZGEMM, vb * (p0 + k + 2 * k) + pb * kv, [&] {
      if (deflate) {
        nn(c, c->Z, k, c->Cz32, nullptr, 0.0f, 1.0f, c->Zn, k, st);
        nn(c, V0, kv, c->Cz32 + (long long)p0 * 32, c->Zn, 1.0f, 1.0f, c->Zn, k, st);
      } else {
        nn(c, V0, kv, c->Cz32, nullptr, 0.0f, 1.0f, c->Zn, k, st);
      }
      dmr::blk::colnorms<float>(c->Zn, ld, k, n, c->cpart, c->zq, 1, st);
      dmr::blk::recip(c->zq, c->zq, k, nullptr, st);
      dmr::scale_cols<float>(c->Zn, k, ld, n, c->zq, st);
    });
  }
  dmr::blk::track(c->rho, c->q, c->scal, s, st);
  timed(c, st, DMR_PH_SCALE, vb * 2 * s,
        [&] { dmr::scale_cols<float>(c->R, s, ld, n, c->scal, st); });
  return refreshed;
}

bool cycle(dmrBlockContext *c, int s, bool deflate, hipStream_t st) {
  return c->bfBasis ? cycle_t(c, reinterpret_cast<bf16 *>(c->P), s, deflate, st)
                    : cycle_t(c, c->P, s, deflate, st);
}

// One restart cycle as a replay of its captured graph.  The cycle's launch sequence depends only
// on (s, deflated, which of the two subspace buffers holds Z) once the shapes are tuned, so each
// such variant is captured once (on the handle's own stream gst, in thread-local capture mode).
// Replays also run on gst, ordered after the caller's stream by an event and the caller's stream
// after them, so a handle that callers drive from several streams and threads (the serving
// loop's slots) always replays its graphs on the stream they were captured on.
int cycle_graph(dmrBlockContext *c, int s, bool deflate, hipStream_t st, bool *refreshed) {
  const int parity = c->Z == c->zbuf0 ? 0 : 1;
  const dmrBlockContext::Cap *cap = nullptr;
  for (const auto &g : c->caps)
    if (g.s == s && g.deflate == (int)deflate && g.parity == parity) cap = &g;
  if (!cap) {
    const int mv0 = c->st.matvecs;
    BCHK(hipStreamBeginCapture(c->gst, hipStreamCaptureModeThreadLocal));
    const bool r = cycle(c, s, deflate, c->gst);
    hipGraph_t g = nullptr;
    BCHK(hipStreamEndCapture(c->gst, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    BCHK(e);
    c->caps.push_back({s, (int)deflate, parity, ex, r, c->st.matvecs - mv0});
    c->st.matvecs = mv0;
    cap = &c->caps.back();
  }
  BCHK(hipEventRecord(c->evIn, st));
  BCHK(hipStreamWaitEvent(c->gst, c->evIn, 0));
  BCHK(hipGraphLaunch(cap->exec, c->gst));
  BCHK(hipEventRecord(c->evOut, c->gst));
  BCHK(hipStreamWaitEvent(st, c->evOut, 0));
  c->st.matvecs += cap->matvecs;
  *refreshed = cap->refreshed;
  return DMR_OK;
}

int block_solve(dmrBlockContext *c, int mode, int s, double *X, long long ldx, const double *B,
                long long ldb, hipStream_t st, double tol, int maxCycles, double *relRes) {
  const long long n = c->n, ld = c->ld;
  double *hb = c->host, *hr = c->host + 64, *hrho = c->host + 128, *hf = c->host + 192;
  if (c->innerMean) {
    dmr::blk::mean_values(c->colVal, c->colLd, c->colCount, c->csr.nnz, c->ownedVal32, st);
    c->val32Stale = false;
  } else if (c->val32Stale) {
    dmr::convert<float, double>(c->ownedVal32, c->csr.val, c->csr.nnz, st);
    c->val32Stale = false;
  }
  if (c->autotune && c->tune.s != s) autotune(c, s, st);
  if (c->bfBasis && c->nnPin) c->tune.nnBasis = c->nnPin;
  if (c->tnVarPin && (c->bfBasis ? c->tnVarPin >= 2 : c->tnVarPin <= 2))
    c->tune.tnVarProj = c->tnVarPin;
  dmr::blk::colnorms<double>(B, ldb, s, n, c->cpart, c->bn, 1, st);
  BCHK(hipMemcpyAsync(hb, c->bn, sizeof(double) * s, hipMemcpyDeviceToHost, st));
  bool useZ = mode != 0 && c->hasZ;
  bool cycled = false;
  double bmax = 0.0;
  // maxCycles bounds the restart cycles of the whole solve, summed over the refinement steps (as
  // the single-vector mixed solve bounds them by matRank), so a slowly converging solve returns
  // DMR_ERR_NOCONV with its best iterate instead of running maxOuter inner solves of maxCycles.
  int cyclesLeft = maxCycles;
  for (int outer = 0;; ++outer) {
    const double valBytes = 8.0 * c->csr.nnz * (c->colVal ? s : 1);
    timed(c, st, DMR_PH_RESID, valBytes + 4.0 * c->csr.nnz + 4.0 * (n + 1) + 24.0 * n * s, [&] {
      dmr::blk::resid_block(c->csr, c->colVal, c->colLd, X, ldx, B, ldb, c->R64, ld, s, st);
    });
    c->st.matvecs += s;
    dmr::blk::colnorms<double>(c->R64, ld, s, n, c->cpart, c->rn, 1, st);
    BCHK(hipMemcpyAsync(hr, c->rn, sizeof(double) * s, hipMemcpyDeviceToHost, st));
    BCHK(hipStreamSynchronize(st));
    if (c->prof.on) prof_collect(c);
    bool done = true;
    double worst = 0.0;
    for (int j = 0; j < s; ++j) {
      const double rel = hb[j] > 0.0 ? hr[j] / hb[j] : hr[j];
      relRes[j] = rel;
      worst = std::max(worst, rel);
      bmax = std::max(bmax, hb[j]);
      if (!(rel <= tol)) done = false;
    }
    c->st.beta = bmax;
    c->st.rho = worst;
    if (done) return DMR_OK;
    if (outer >= c->maxOuter || cyclesLeft <= 0) return DMR_ERR_NOCONV;
    if (!std::isfinite(worst)) return DMR_ERR_RUNAWAY;
    // Inner tolerances per column, and the scale of each residual column to unit norm.
    std::vector<double> tolIn(s);
    for (int j = 0; j < s; ++j) {
      hf[j] = hr[j] > 0.0 ? 1.0 / hr[j] : 0.0;
      const double want = hr[j] > 0.0 ? 0.5 * tol * (hb[j] > 0.0 ? hb[j] : 1.0) / hr[j] : 1.0;
      tolIn[j] = std::min(0.5, std::max(c->innerTol, want));
    }
    BCHK(hipMemcpyAsync(c->f, hf, sizeof(double) * s, hipMemcpyHostToDevice, st));
    dmr::blk::scale_to32(c->R64, ld, c->f, c->R, ld, s, n, st);
    dmr::fill<float>(c->Xa, ld * s, 0.0, st);
    dmr::blk::recip(c->f, c->scal, s, c->rho, st);  // rho = 1
    // Stagnation exit: an inner solve whose worst estimate has not improved by 0.1% for
    // kStallCycles cycles returns to the refinement, which recomputes the FP64 residual and
    // restarts (bounded by maxOuter), instead of running to maxCycles.
    double bestEst = HUGE_VAL;
    int stall = 0;
    while (cyclesLeft > 0) {
      --cyclesLeft;
      const bool deflate = c->k > 0 && c->hasZ && (useZ || cycled);
      bool refreshed = false;
      if (c->graphs && !c->prof.on) {
        const int rc = cycle_graph(c, s, deflate, st, &refreshed);
        if (rc != DMR_OK) return rc;
      } else {
        refreshed = cycle(c, s, deflate, st);
      }
      cycled = true;
      BCHK(hipMemcpyAsync(hrho, c->rho, sizeof(double) * s, hipMemcpyDeviceToHost, st));
      BCHK(hipMemcpyAsync(hf + 32, c->flags, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
      BCHK(hipStreamSynchronize(st));
      if (c->prof.on) prof_collect(c);
      const int *fl = reinterpret_cast<const int *>(hf + 32);
      ++c->st.cycles;
      if (fl[0]) ++c->st.breakdowns;
      // Dependent block columns are zero basis vectors: the refresh never selects them, so it is
      // kept on such cycles too.
      if (refreshed) {
        std::swap(c->Z, c->Zn);
        c->hasZ = true;
        c->st.svdSweeps += fl[1];
        ++c->st.svdCalls;
      }
      bool innerDone = true;
      double est = 0.0;
      for (int j = 0; j < s; ++j) {
        if (!(hrho[j] <= tolIn[j])) innerDone = false;
        est = std::max(est, hrho[j] * (hb[j] > 0.0 ? hr[j] / hb[j] : hr[j]));
      }
      c->hist.push_back(est);
      if (!std::isfinite(est)) return DMR_ERR_RUNAWAY;
      if (innerDone) break;
      if (est < 0.999 * bestEst) {
        bestEst = est;
        stall = 0;
      } else if (++stall >= kStallCycles) {
        break;
      }
    }
    // X += ||R_j|| D_j.
    dmr::blk::add_scaled(X, ldx, c->Xa, ld, c->rn, s, n, st);
    useZ = true;
  }
}

}  // namespace

extern "C" {

size_t dmrBlockFootprint(int elements, int blockSteps, int deflatDim, int maxRhs) {
  if (elements < 1 || blockSteps < 1 || deflatDim < 0 || maxRhs < 1) return 0;
  const long long ld = dmr::basis_ld(elements);
  const int G = dmr::blk::gemm_tn_blocks(elements, dmr::blk::kMaxTnBlocks, 256);
  const Sizes z = sizes(elements, blockSteps, deflatDim, maxRhs, G, ld);
  return z.big + z.small + 80 * 256;
}

int dmrBlockCreate(dmrBlockHandle *h, int device, int elements, int blockSteps, int deflatDim,
                   int maxRhs) {
  dmr::clear_error();
  if (!h) return dmr::set_error(DMR_ERR_ARG, "null handle pointer");
  *h = nullptr;
  if (elements < 1 || blockSteps < 1 || deflatDim < 0 || deflatDim > kMaxRhs || maxRhs < 1 ||
      maxRhs > kMaxRhs)
    return dmr::set_error(DMR_ERR_ARG, "block DMR needs n >= 1, steps >= 1, 0 <= k <= 32, "
                                       "1 <= maxRhs <= 32");
  if (deflatDim + (blockSteps + 1) * maxRhs > dmr::blk::kMaxCols)
    return dmr::set_error(DMR_ERR_ARG, "k + (blockSteps + 1) * maxRhs exceeds 384 basis columns");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return dmr::set_error(DMR_ERR_HIP, "no such device");
  BCHK(hipSetDevice(device));
  dmrBlockContext *c = new dmrBlockContext();
  c->device = device;
  c->n = elements;
  c->mb = blockSteps;
  c->k = deflatDim;
  c->sMax = maxRhs;
  c->ld = dmr::basis_ld(elements);
  c->rowsMax = deflatDim + (blockSteps + 1) * maxRhs;
  c->colsMax = deflatDim + blockSteps * maxRhs;
  // Partials for the most row ranges any tuned shape uses (256 rows per workgroup).
  c->G = dmr::blk::gemm_tn_blocks(elements, dmr::blk::kMaxTnBlocks, 256);
  c->tune.tnRowsProj = c->tune.tnRowsGram = tn_rows_target();
  if (const char *e = std::getenv("DMR_BLOCK_GRAPHS")) c->graphs = e[0] != '0';
  if (const char *e = std::getenv("DMR_BLOCK_TUNE")) c->autotune = e[0] != '0';
  dmr::blk::prepare();
  if (const char *e = std::getenv("DMR_BLOCK_CHOLQR_PASSES")) {
    const int v = std::atoi(e);
    if (v == 1 || v == 2) c->cholPasses = v;
  }
  if (const char *e = std::getenv("DMR_BLOCK_FUSED_GRAM")) c->fusedGram = e[0] != '0';
  if (const char *e = std::getenv("DMR_BLOCK_GRAM_CHOL")) c->gramChol = e[0] != '0';
  if (const char *e = std::getenv("DMR_BLOCK_FOLD_T")) c->foldT = e[0] != '0';
  if (const char *e = std::getenv("DMR_BLOCK_SVD_SIDE")) c->svdSide = e[0] != '0';
  if (const char *e = std::getenv("DMR_BLOCK_NN_BASIS")) {
    const int v = std::atoi(e);
    if (v == 4 || v == 8 || v == 16 || v == 100 || v == 101) c->nnPin = v;
  }
  if (const char *e = std::getenv("DMR_BLOCK_TN_VAR")) {
    const int v = std::atoi(e);
    if (v >= 0 && v <= 3) c->tnVarPin = v;
  }
  if (const char *e = std::getenv("DMR_BLOCK_SPMM")) {
    const int v = std::atoi(e);
    for (int ok : {4, 8, 16, 104, 108, 116, 208, 216})
      if (v == ok) c->spmmForce = v;
  }
  if (const char *e = std::getenv("DMR_BLOCK_SVD_TOL")) {
    const double v = std::atof(e);
    if (v > 0.0 && v < 1.0) c->svdTol = v;
  }
  if (const char *e = std::getenv("DMR_BLOCK_KR")) {
    const int v = std::atoi(e);
    if (v >= 1) c->krMax = v;
  }
  int rc = ctx_alloc(c);
  if (rc == DMR_OK && (hipStreamCreateWithFlags(&c->gst, hipStreamNonBlocking) != hipSuccess ||
                       hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
                       hipEventCreateWithFlags(&c->evFork, hipEventDisableTiming) != hipSuccess ||
                       hipEventCreateWithFlags(&c->evJoin, hipEventDisableTiming) != hipSuccess ||
                       hipEventCreateWithFlags(&c->evIn, hipEventDisableTiming) != hipSuccess ||
                       hipEventCreateWithFlags(&c->evOut, hipEventDisableTiming) != hipSuccess))
    rc = dmr::set_error(DMR_ERR_HIP, "stream creation failed");
  if (rc != DMR_OK) {
    ctx_release(c);
    return rc;
  }
  c->zbuf0 = c->Z;
  *h = c;
  return DMR_OK;
}

int dmrBlockDestroy(dmrBlockHandle h) {
  if (!h) return dmr::set_error(DMR_ERR_ARG, "null handle");
  ctx_release(h);
  return DMR_OK;
}

int dmrBlockSetOperatorCsr(dmrBlockHandle h, int nnz, const int *rowPtr, const int *colInd,
                           const double *val, const float *val32) {
  dmr::clear_error();
  if (!h || !rowPtr || !colInd || !val || nnz < 0)
    return dmr::set_error(DMR_ERR_ARG, "invalid CSR arguments");
  BCHK(hipSetDevice(h->device));
  drop_graphs(h);
  h->csr.n = h->n;
  h->csr.nnz = nnz;
  h->csr.rowPtr = rowPtr;
  h->csr.colInd = colInd;
  h->csr.val = val;
  h->csr.blockRows = nullptr;
  h->csr.nBlocks = 0;
  h->colVal = nullptr;
  h->colLd = 0;
  h->colCount = 0;
  h->innerMean = false;
  if (!val32) {
    if (h->ownedVal32) hipFree(h->ownedVal32);
    h->ownedVal32 = nullptr;
    BCHK(hipMalloc(&h->ownedVal32, sizeof(float) * std::max(nnz, 1)));
    h->csr.val32 = h->ownedVal32;
    h->val32Stale = true;
  } else {
    h->csr.val32 = val32;
    h->val32Stale = false;
  }
  h->haveOp = true;
  return DMR_OK;
}

int dmrBlockSetOperatorValues(dmrBlockHandle h, const double *val, const float *val32) {
  dmr::clear_error();
  if (!h || !h->haveOp || !val) return dmr::set_error(DMR_ERR_STATE, "no CSR operator to update");
  if ((val32 ? val32 : h->ownedVal32) != h->csr.val32) drop_graphs(h);
  h->csr.val = val;
  h->innerMean = false;
  if (val32) {
    h->csr.val32 = val32;
    h->val32Stale = false;
  } else {
    if (!h->ownedVal32) {
      BCHK(hipMalloc(&h->ownedVal32, sizeof(float) * std::max(h->csr.nnz, 1)));
    }
    h->csr.val32 = h->ownedVal32;
    h->val32Stale = true;
  }
  return DMR_OK;
}

int dmrBlockSetColumnOperators(dmrBlockHandle h, const double *val, long long ldv, int count,
                               int innerMean) {
  dmr::clear_error();
  if (!h || !h->haveOp) return dmr::set_error(DMR_ERR_STATE, "no CSR operator to extend");
  if (!val) {
    h->colVal = nullptr;
    h->colLd = 0;
    h->colCount = 0;
    if (h->innerMean) {
      h->innerMean = false;
      h->val32Stale = h->csr.val32 == h->ownedVal32;
    }
    return DMR_OK;
  }
  if (count < 1 || count > h->sMax || ldv < h->csr.nnz)
    return dmr::set_error(DMR_ERR_ARG, "column operators need 1 <= count <= maxRhs, ldv >= nnz");
  BCHK(hipSetDevice(h->device));
  h->colVal = val;
  h->colLd = ldv;
  h->colCount = count;
  h->innerMean = innerMean != 0;
  if (h->innerMean) {
    if (!h->ownedVal32)
      BCHK(hipMalloc(&h->ownedVal32, sizeof(float) * std::max(h->csr.nnz, 1)));
    if (h->csr.val32 != h->ownedVal32) drop_graphs(h);
    h->csr.val32 = h->ownedVal32;
  }
  return DMR_OK;
}

int dmrBlockSetPrecision(dmrBlockHandle h, int precision) {
  dmr::clear_error();
  if (!h) return dmr::set_error(DMR_ERR_ARG, "null handle");
  if (precision != DMR_PREC_FP32 && precision != DMR_PREC_BF16)
    return dmr::set_error(DMR_ERR_ARG, "block basis storage is DMR_PREC_FP32 or DMR_PREC_BF16");
  const bool bf = precision == DMR_PREC_BF16;
  if (bf != h->bfBasis) {
    BCHK(hipSetDevice(h->device));
    drop_graphs(h);
    h->bfBasis = bf;
    h->tune.s = 0;  // launch shapes are tuned per storage type
  }
  return DMR_OK;
}

int dmrBlockSetRefinement(dmrBlockHandle h, double innerTol, int maxOuter) {
  if (!h || !(innerTol > 0.0 && innerTol < 1.0) || maxOuter < 1)
    return dmr::set_error(DMR_ERR_ARG, "invalid refinement settings");
  h->innerTol = innerTol;
  h->maxOuter = maxOuter;
  return DMR_OK;
}

int dmrBlockSolve(dmrBlockHandle h, int mode, int nrhs, double *X, long long ldx, const double *B,
                  long long ldb, hipStream_t stream, double solTol, int maxCycles,
                  double *relRes) {
  dmr::clear_error();
  if (!h) return dmr::set_error(DMR_ERR_STATE, "block solver not initialised");
  if (!h->haveOp) return dmr::set_error(DMR_ERR_STATE, "no operator set");
  if (nrhs < 1 || nrhs > h->sMax || !X || !B || !relRes || ldx < h->n || ldb < h->n ||
      maxCycles < 1 || !(solTol >= 0.0))
    return dmr::set_error(DMR_ERR_ARG, "invalid block solve arguments");
  if (h->colVal && nrhs > h->colCount)
    return dmr::set_error(DMR_ERR_ARG, "more right-hand sides than column operators");
  BCHK(hipSetDevice(h->device));
  h->st = dmrStats{};
  h->hist.clear();
  const auto t0 = std::chrono::steady_clock::now();
  const int rc = block_solve(h, mode, nrhs, X, ldx, B, ldb, stream, solTol, maxCycles, relRes);
  h->st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  h->st.status = rc;
  if (rc != DMR_OK && dmrGetLastError() == DMR_OK)
    dmr::set_error(rc, rc == DMR_ERR_NOCONV ? "refinement limit reached" : "residual diverged");
  return rc;
}

int dmrBlockGetStats(dmrBlockHandle h, dmrStats *st, double *residHist, int cap) {
  if (!h) return -1;
  if (st) *st = h->st;
  const int cnt = (int)h->hist.size();
  if (residHist)
    for (int i = 0; i < std::min(cap, cnt); ++i) residHist[i] = h->hist[i];
  return cnt;
}

int dmrBlockHasZ(dmrBlockHandle h) { return h && h->hasZ ? 1 : 0; }

int dmrBlockSetOptions(dmrBlockHandle h, int graphs, int autotune) {
  if (!h) return dmr::set_error(DMR_ERR_ARG, "null handle");
  if (graphs >= 0) {
    if (!graphs) drop_graphs(h);
    h->graphs = graphs != 0;
  }
  if (autotune >= 0) h->autotune = autotune != 0;
  return DMR_OK;
}

int dmrBlockTuningFile(const char *path) {
  std::lock_guard<std::mutex> lk(g_tuneMu);
  g_tuneFile = path ? path : "";
  g_tuneFileSet = true;
  g_tuneLoaded = false;
  return DMR_OK;
}

int dmrBlockGetTuning(dmrBlockHandle h, int *out) {
  if (!h || !out) return dmr::set_error(DMR_ERR_ARG, "null argument");
  const auto &t = h->tune;
  out[0] = t.tnRowsProj;
  out[1] = t.tnVarProj;
  out[2] = t.tnRowsGram;
  out[3] = t.tnVarGram;
  out[4] = h->bfBasis ? t.nnBasis : t.nnUnroll;
  out[5] = t.spmmCols;
  out[6] = (int)h->caps.size();
  return DMR_OK;
}

int dmrBlockImportZ(dmrBlockHandle h, const double *hostZ) {
  dmr::clear_error();
  if (!h || !hostZ) return dmr::set_error(DMR_ERR_ARG, "null argument");
  if (h->k == 0) return DMR_OK;
  BCHK(hipSetDevice(h->device));
  std::vector<float> tmp((size_t)h->n * h->k);
  for (size_t i = 0; i < tmp.size(); ++i) tmp[i] = (float)hostZ[i];
  BCHK(hipMemcpy2D(h->Z, sizeof(float) * h->ld, tmp.data(), sizeof(float) * h->n,
                   sizeof(float) * h->n, h->k, hipMemcpyHostToDevice));
  h->hasZ = true;
  return DMR_OK;
}

int dmrBlockExportZ(dmrBlockHandle h, double *hostZ) {
  dmr::clear_error();
  if (!h || !hostZ) return dmr::set_error(DMR_ERR_ARG, "null argument");
  if (h->k == 0) return DMR_OK;
  BCHK(hipSetDevice(h->device));
  BCHK(hipDeviceSynchronize());
  std::vector<float> tmp((size_t)h->n * h->k);
  BCHK(hipMemcpy2D(tmp.data(), sizeof(float) * h->n, h->Z, sizeof(float) * h->ld,
                   sizeof(float) * h->n, h->k, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < tmp.size(); ++i) hostZ[i] = tmp[i];
  return DMR_OK;
}

int dmrBlockSetProfiling(dmrBlockHandle h, int on) {
  if (!h) return dmr::set_error(DMR_ERR_ARG, "null handle");
  h->prof.on = on != 0;
  for (int i = 0; i < DMR_NPHASES; ++i) {
    h->prof.ms[i] = 0.0;
    h->prof.by[i] = 0.0;
    h->prof.fl[i] = 0.0;
    h->prof.cnt[i] = 0;
  }
  h->prof.used = 0;
  return DMR_OK;
}

int dmrBlockGetProfile(dmrBlockHandle h, double *ms, double *bytes, double *flops,
                       long long *launches) {
  if (!h) return dmr::set_error(DMR_ERR_ARG, "null handle");
  for (int i = 0; i < DMR_NPHASES; ++i) {
    if (ms) ms[i] = h->prof.ms[i];
    if (bytes) bytes[i] = h->prof.by[i];
    if (flops) flops[i] = h->prof.fl[i];
    if (launches) launches[i] = h->prof.cnt[i];
  }
  return DMR_NPHASES;
}

}  // extern "C"
