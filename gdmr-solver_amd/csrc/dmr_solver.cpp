// Host driver of the DMR solver: contexts, the restart cycle, the FP64 solve, the
// mixed-precision refinement solve, and the C ABI of gDMR.h.
//
// The cycle follows the reference solver (seanMolesky/gDMR-solver gDMR.cu) step for step:
//   solverDMR        gDMR.cu:2638-2901  -> solve_fp64 / run_cycles (restart loop, guards)
//   arnoldiSetupDMR  gDMR.cu:2586-2632  -> start-vector choice incl. the stagnation rule
//   arnoldiDMR       gDMR.cu:1790-2332  -> deflation-image QR of A Z, then CGS Krylov steps
//   deflatorDMR      gDMR.cu:1710-1787  -> projection against W before the Krylov CGS
//   updateTrgCoeffs  gDMR.cu:2334-2408  -> t = P^T r (full projection, not beta e1)
//   updateAppResDMR  gDMR.cu:2411-2582  -> LSQ, residual and solution updates, stagnation flag
//   genDeflatDMR     gDMR.cu:1583-1704  -> Z <- normalize([V Z] * smallest right sing. vectors)
// on one device with the basis stored contiguously: P holds m+1 columns, W = P[0..k-1] when the
// cycle is deflated, and the Krylov vectors follow at slot o = (d > 0 ? k : 0).  The
// multi-device coefficient reordering (reorderHessCoeffs) and partial-vector sums are identities
// on one device and do not exist here.  Every step is enqueued on the caller's stream; the host
// synchronises once per cycle to read the residual ratio that drives the restart decisions.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "dmr_internal.h"
#include "gDMR.h"

namespace {

thread_local int g_err = DMR_OK;
thread_local std::string g_errmsg;
thread_local bool g_op_failed = false;
thread_local std::string g_op_msg;

int set_err(int code, const std::string &msg) {
  g_err = code;
  g_errmsg = msg;
  return code;
}
void clear_err() {
  g_err = DMR_OK;
  g_errmsg.clear();
}

#define HIPCHK(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess)                                                               \
      return set_err(DMR_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));   \
  } while (0)

// Fixed-order reductions of the FP64 path use the tolerance below for the Jacobi SVD: the FP64
// analogue of the reference's FP32 gesvdj tolerance (gDMR.cu:88), at most 128 sweeps
// (gDMR.cu:89).  The stagnation threshold 0.995 is gDMR.cu:2506.
constexpr double kSvdTol = 1e-14;
constexpr int kSvdMaxSweeps = 128;
// The FP32 and bf16 inner solves of the mixed path use the reference's own FP32 tolerance
// (svTolDMR = 2e-7, gDMR.cu:88): their basis carries FP32 rounding, so sweeps past it add nothing.
constexpr double kSvdTolLow = 2e-7;
constexpr int kMaxCycleGraphs = 16;
// Largest relative error of an implicit deflation image (checked on its last column) that the
// next cycle accepts instead of recomputing the image (DMR_IMAGE_TOL overrides).  The check's
// own FP32 product of a smallest-singular direction cancels to about 1e-7 times the operator's
// conditioning, so a bound near the FP32 rounding rejects sound images (1e-4 rejected nearly all
// at 128^3); 1e-2 keeps them and still catches a chain whose error grows (measured on the 2-D
// 512^2 operator at m = 80, which diverges without the check).
double image_tol() {
  static const double v = [] {
    const char *e = std::getenv("DMR_IMAGE_TOL");
    const double x = e ? std::atof(e) : 0.0;
    return x > 0.0 ? x : 1e-2;
  }();
  return v;
}
constexpr double kStagnation = 0.995;

long long pad_ld(long long n) { return dmr::basis_ld(n); }

template <typename TS> struct Work {
  TS *P = nullptr, *Z = nullptr, *Zn = nullptr, *r = nullptr, *xa = nullptr, *g = nullptr,
     *u = nullptr;
  TS *Wn = nullptr;  // image of the refreshed subspace (FP32 inner solves)
  bool hasZ = false;
};

}  // namespace

namespace dmr {

namespace {
long long env_ll(const char *name, long long dflt, long long lo, long long hi) {
  const char *v = std::getenv(name);
  if (!v || !*v) return dflt;
  char *end = nullptr;
  const long long x = std::strtoll(v, &end, 10);
  if (end == v || x < lo || x > hi) return dflt;
  return x;
}
}  // namespace

bool env_flag(const char *name, bool dflt) { return env_ll(name, dflt ? 1 : 0, 0, 1) != 0; }

int set_error(int code, const char *msg) { return set_err(code, msg ? msg : ""); }
void clear_error() { clear_err(); }

int red_blocks_cap() {
  static const int cap = (int)env_ll("DMR_RED_BLOCKS", kDefaultRedBlocks, 1, kMaxRedBlocks);
  return cap;
}

long long basis_ld(long long n) {
  static const long long pad = env_ll("DMR_LD_PAD", 512, 0, 1 << 20) / 256 * 256;
  return (n + 255) / 256 * 256 + pad;
}

}  // namespace dmr

enum OpKind { OP_NONE = 0, OP_CSR = 1, OP_FN = 2, OP_LEGACY = 3 };

struct dmrContext {
  int device = 0;
  int n = 0, m = 0, k = 0, prec = DMR_PREC_FP64;
  long long ld = 0;
  // operator
  int opKind = OP_NONE;
  dmr::CsrDev csr;
  int *ownedBlockRows = nullptr;
  float *ownedVal32 = nullptr;
  bool val32Stale = false;  // ownedVal32 is converted from csr.val on the next solve's stream
  dmrOperatorFn opFn = nullptr;
  void *opCtx = nullptr;
  void (*legacy)(void) = nullptr;
  double *ext = nullptr;  // the caller's source vector during a legacy-operator solve
  // memory
  void *bigPool = nullptr;
  void *smallPool = nullptr;
  double *hostStatus = nullptr;  // pinned
  // Side stream (default priority) on which the Hessenberg SVD of a cycle runs while the main
  // stream does the least squares and the residual / solution updates.
  hipStream_t side = nullptr;
  hipEvent_t evH = nullptr, evSvd = nullptr;
  // Captured restart cycles (CSR operators, profiling off; DMR_GRAPHS=1): each variant of a
  // cycle's launch sequence is captured once on the context's own stream cs and replayed there,
  // after an event on the caller's stream; the caller's stream then waits on an event after the
  // replay.  At most kMaxCycleGraphs are kept (oldest dropped); all are dropped when the CSR
  // pattern changes.
  struct CycleGraph {
    int d, fromResidual, implicitImg, zParity;
    double tol;
    const void *val, *val32;  // the operator values the capture read
    hipGraphExec_t exec;
    long long matvecs;
  };
  bool graphs = false;
  hipStream_t cs = nullptr;
  hipEvent_t evIn = nullptr, evOut = nullptr;
  std::vector<CycleGraph> cycleGraphs;
  double *rhoDev = nullptr;  // rho of the current cycle (read by the solution update)
  Work<double> w64;
  Work<float> w32;
  Work<dmr::bf16> w16;  // bf16 basis storage (DMR_PREC_BF16)
  double *x64 = nullptr, *b64 = nullptr, *r64 = nullptr;  // refinement vectors (mixed)
  double *H = nullptr, *t = nullptr, *y = nullptr, *z = nullptr, *C = nullptr, *sig = nullptr,
         *cv = nullptr, *scal = nullptr, *scal2 = nullptr, *scalZ = nullptr, *one = nullptr;
  int *flag = nullptr, *sweeps = nullptr;
  unsigned *bar = nullptr;  // grid-barrier words of the fused orthogonalisation
  int *devErr = nullptr;    // set by a device-side wait that exceeded its bound
  unsigned epoch = 0;       // barrier phases completed in the current cycle
  bool fused = false;       // fused cooperative orthogonalisation (DMR_FUSED=1 enables)
  // Deflated Krylov steps project against [W V] in one Gram-Schmidt stage instead of the
  // reference's deflate-then-CGS; default for mixed precision (DMR_ONE_STAGE overrides).
  bool oneStage = false;
  // FP32 inner solves take the image A Z of a refreshed subspace from the Arnoldi relation of the
  // cycle that refreshed it (W = P Q, R from the QR of H C) instead of k operator applications
  // and a Gram-Schmidt QR (DMR_IMPLICIT_IMAGE=0 turns it off).  imgValid: Wn, imgQ / imgR and
  // scalZ hold the image of the current Z under the current operator values.
  bool implicitImage = false;
  bool imgValid = false;
  // Krylov steps at the end of a cycle left out of its subspace refresh, so that the refresh SVD
  // overlaps them: 0 in the FP64 path (the reference's refresh), DMR_REFRESH_TAIL (default 2)
  // in the FP32 and bf16 inner solves.
  int refreshTail = 0;
  double *imgQ = nullptr, *imgR = nullptr, *imgScal = nullptr, *imgCol = nullptr,
         *imgChk = nullptr;
  int *imgOk = nullptr;
  dmr::RedWork red{};
  dmr::RedWork redOp{};  // sized to the SpMV block count of the operator
  void *redOpPool = nullptr;
  // persistent state (kryFlagCountDMR persists across solves, gDMR.cu:890)
  int kryLoopFlag = 0;
  int kryFlagCount = 0;
  // refinement settings
  double innerTol = 1e-5;
  int maxOuter = 60;
  // statistics of the last solve
  dmrStats st{};
  std::vector<double> hist;
  // per-phase HIP-event profile of the restart cycles (off unless requested)
  struct Prof {
    bool on = false;
    std::vector<hipEvent_t> ev;
    std::vector<int> phase;
    std::vector<double> bytes;
    int used = 0;
    double ms[DMR_NPHASES] = {};
    double by[DMR_NPHASES] = {};
    long long cnt[DMR_NPHASES] = {};
  } prof;
};

namespace {

size_t footprint(int n, int m, int k, int prec) {
  const long long ld = pad_ld(n);
  size_t big = sizeof(double) * (size_t)ld * (size_t)(m + 1 + 2 * k + 4);
  if (prec != DMR_PREC_FP64) {
    big += (prec == DMR_PREC_BF16 ? 2 : 4) * (size_t)ld * (size_t)(m + 1 + 3 * k + 4);
    big += sizeof(double) * (size_t)ld * 3;
  }
  const size_t small = sizeof(double) * ((size_t)(m + 1) * m * 2 + 8 * (m + 2) + 64 + 2 * k +
                                         (size_t)(m + 1 + k) * k + 2 * k + 16) +
                       sizeof(double) * dmr::kMaxPartials + sizeof(unsigned) * dmr::kCounters +
                       4096;
  return big + small;
}

int ctx_alloc(dmrContext *c) {
  const long long ld = c->ld;
  const int n = c->n, m = c->m, k = c->k;
  (void)n;
  size_t big = sizeof(double) * (size_t)ld * (size_t)(m + 1 + 2 * k + 4);
  size_t big32 = 0;
  const size_t es = c->prec == DMR_PREC_BF16 ? 2 : 4;  // inner basis element size
  if (c->prec != DMR_PREC_FP64) {
    big32 = es * (size_t)ld * (size_t)(m + 1 + 3 * k + 4) + sizeof(double) * ld * 3;
    big = 0;  // the FP64 Krylov basis is not needed by a mixed-precision context
  }
  const size_t total = big + big32;
  HIPCHK(hipMalloc(&c->bigPool, total));
  HIPCHK(hipMemset(c->bigPool, 0, total));
  char *p = (char *)c->bigPool;
  if (c->prec == DMR_PREC_FP64) {
    double *d = (double *)p;
    c->w64.P = d;
    d += ld * (m + 1);
    c->w64.Z = d;
    d += ld * k;
    c->w64.Zn = d;
    d += ld * k;
    c->w64.r = d;
    d += ld;
    c->w64.xa = d;
    d += ld;
    c->w64.g = d;
    d += ld;
    c->w64.u = d;
    d += ld;
  } else {
    double *d = (double *)p;
    c->x64 = d;
    d += ld;
    c->b64 = d;
    d += ld;
    c->r64 = d;
    d += ld;
    auto carve = [&](auto &w, auto *f) {
      w.P = f;
      f += ld * (m + 1);
      w.Z = f;
      f += ld * k;
      w.Zn = f;
      f += ld * k;
      w.r = f;
      f += ld;
      w.xa = f;
      f += ld;
      w.g = f;
      f += ld;
      w.u = f;
      f += ld;
      w.Wn = f;
    };
    if (c->prec == DMR_PREC_BF16)
      carve(c->w16, (dmr::bf16 *)d);
    else
      carve(c->w32, (float *)d);
  }
  // small arrays
  const size_t nsmall = (size_t)(m + 1) * m   // H
                        + (m + 2) * 4         // t, y, z, cv
                        + (size_t)m * (k + 1) // C
                        + (m + 2)             // sig
                        + 16 + 16 + 64 + 8    // scal, scal2, scalZ, one
                        + (size_t)(m + 1 + k) * k + 2 * k + 8  // imgQ, imgR, imgScal,
                                                               // imgCol, imgChk
                        + 8;                                   // rhoDev
  const size_t smallBytes = sizeof(double) * (nsmall + dmr::kMaxPartials) +
                            sizeof(unsigned) * dmr::kCounters + sizeof(int) * 64;
  HIPCHK(hipMalloc(&c->smallPool, smallBytes));
  HIPCHK(hipMemset(c->smallPool, 0, smallBytes));
  double *s = (double *)c->smallPool;
  c->H = s;
  s += (size_t)(m + 1) * m;
  c->t = s;
  s += m + 2;
  c->y = s;
  s += m + 2;
  c->z = s;
  s += m + 2;
  c->cv = s;
  s += m + 2;
  c->C = s;
  s += (size_t)m * (k + 1);
  c->sig = s;
  s += m + 2;
  c->scal = s;
  s += 16;
  c->scal2 = s;
  s += 16;
  c->scalZ = s;
  s += 64;
  c->one = s;
  s += 8;
  c->imgQ = s;
  s += (size_t)(m + 1) * k;
  c->imgR = s;
  s += (size_t)k * k;
  c->imgScal = s;
  s += k;
  c->imgCol = s;
  s += k;
  c->imgChk = s;
  s += 8;
  c->rhoDev = s;
  s += 8;
  c->red.partials = s;
  s += dmr::kMaxPartials;
  c->red.counters = (unsigned *)s;
  int *ip = (int *)(c->red.counters + dmr::kCounters);
  c->flag = ip;
  c->sweeps = ip + 1;
  c->bar = reinterpret_cast<unsigned *>(ip + 2);
  c->devErr = ip + 4;
  c->imgOk = ip + 5;
  // The fused cooperative step kernel is bitwise equivalent but measured slower than the
  // separate kernels at 128^3 (2.2 vs ~4 TB/s), so it is opt-in.
  c->fused = dmr::env_flag("DMR_FUSED", false);
  c->oneStage = dmr::env_flag("DMR_ONE_STAGE", c->prec != DMR_PREC_FP64);
  c->refreshTail = c->prec == DMR_PREC_FP64
                       ? 0
                       : (int)dmr::env_ll("DMR_REFRESH_TAIL", 2, 0, m);
  c->implicitImage = c->prec == DMR_PREC_FP32 && k > 0 &&
                     dmr::image_qr_lds(m, k) <= 160 * 1024 &&
                     dmr::env_flag("DMR_IMPLICIT_IMAGE", true);
  const double one = 1.0;
  HIPCHK(hipMemcpy(c->one, &one, sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipHostMalloc((void **)&c->hostStatus, 64 * sizeof(double), hipHostMallocDefault));
  HIPCHK(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&c->evH, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&c->evSvd, hipEventDisableTiming));
  HIPCHK(hipStreamCreateWithFlags(&c->cs, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&c->evIn, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&c->evOut, hipEventDisableTiming));
  c->graphs = dmr::env_flag("DMR_GRAPHS", false);
  // LDS limits of the one-workgroup kernels, set before any cycle can be captured.
  dmr::prepare_dense();
  return DMR_OK;
}

void ctx_free(dmrContext *c) {
  if (!c) return;
  hipSetDevice(c->device);
  hipDeviceSynchronize();
  if (c->bigPool) hipFree(c->bigPool);
  if (c->smallPool) hipFree(c->smallPool);
  if (c->redOpPool) hipFree(c->redOpPool);
  for (auto e : c->prof.ev) hipEventDestroy(e);
  if (c->ownedBlockRows) hipFree(c->ownedBlockRows);
  if (c->ownedVal32) hipFree(c->ownedVal32);
  if (c->hostStatus) hipHostFree(c->hostStatus);
  if (c->evH) hipEventDestroy(c->evH);
  if (c->evSvd) hipEventDestroy(c->evSvd);
  if (c->side) hipStreamDestroy(c->side);
  for (auto &g : c->cycleGraphs) hipGraphExecDestroy(g.exec);
  if (c->evIn) hipEventDestroy(c->evIn);
  if (c->evOut) hipEventDestroy(c->evOut);
  if (c->cs) hipStreamDestroy(c->cs);
  delete c;
}

// Captured cycles hold the operator's device pointers: dropped when they change.
void drop_cycle_graphs(dmrContext *c) {
  for (auto &g : c->cycleGraphs) hipGraphExecDestroy(g.exec);
  c->cycleGraphs.clear();
}

int validate_shape(int n, int m, int k) {
  if (n < 1) return set_err(DMR_ERR_ARG, "elements must be >= 1");
  if (m < 1 || m > 120) return set_err(DMR_ERR_ARG, "basisDim must be in [1, 120]");
  if (k < 0 || k >= m) return set_err(DMR_ERR_ARG, "deflatDim must be in [0, basisDim)");
  if (k > 64) return set_err(DMR_ERR_ARG, "deflatDim must be <= 64");
  return DMR_OK;
}

// ---- operator application on a typed basis ----
struct Op64 {
  dmrContext *c;
  hipStream_t s;
  int operator()(const double *src, double *dst) const {
    switch (c->opKind) {
      case OP_CSR:
        dmr::spmv(c->csr, src, dst, s);
        return DMR_OK;
      case OP_FN:
        c->opFn(src, dst, s, c->opCtx);
        break;
      case OP_LEGACY:
        // In-place callback on the caller's source vector (testLS.cu:97-128 contract).
        if (hipMemcpyAsync(c->ext, src, sizeof(double) * c->n, hipMemcpyDeviceToDevice, s) !=
            hipSuccess)
          return set_err(DMR_ERR_HIP, "copy to operator vector failed");
        c->legacy();
        if (hipMemcpyAsync(dst, c->ext, sizeof(double) * c->n, hipMemcpyDeviceToDevice, s) !=
            hipSuccess)
          return set_err(DMR_ERR_HIP, "copy from operator vector failed");
        break;
      default:
        return set_err(DMR_ERR_STATE, "no operator set");
    }
    if (g_op_failed) {
      g_op_failed = false;
      return set_err(DMR_ERR_OPERATOR, "operator reported failure: " + g_op_msg);
    }
    return DMR_OK;
  }
};

struct Op32 {
  dmrContext *c;
  hipStream_t s;
  int operator()(const float *src, float *dst) const {
    dmr::spmv32(c->csr, src, dst, s);
    return DMR_OK;
  }
};

template <typename TS> Work<TS> &work_of(dmrContext *c);
template <> Work<double> &work_of<double>(dmrContext *c) { return c->w64; }
template <> Work<float> &work_of<float>(dmrContext *c) { return c->w32; }
template <> Work<dmr::bf16> &work_of<dmr::bf16>(dmrContext *c) { return c->w16; }

struct Op16 {
  dmrContext *c;
  hipStream_t s;
  int operator()(const dmr::bf16 *src, dmr::bf16 *dst) const {
    dmr::spmv16(c->csr, src, dst, s);
    return DMR_OK;
  }
};

template <typename TS> struct LowOp;
template <> struct LowOp<float> { using type = Op32; };
template <> struct LowOp<dmr::bf16> { using type = Op16; };

dmr::Fin fin_none() { return dmr::Fin{dmr::FIN_NONE, nullptr, nullptr, 0.0, 0, nullptr}; }

// Runs f (which enqueues kernels on s) bracketed by HIP events when profiling is on; the elapsed
// times are folded into the per-phase totals by prof_collect after the stream is synchronised.
template <typename F>
int timed(dmrContext *c, hipStream_t s, int phase, double bytes, F &&f) {
  auto &p = c->prof;
  if (!p.on) return f();
  if ((size_t)(2 * p.used + 2) > p.ev.size()) {
    for (int i = 0; i < 64; ++i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return set_err(DMR_ERR_HIP, "hipEventCreate failed");
      p.ev.push_back(e);
    }
  }
  if (hipEventRecord(p.ev[2 * p.used], s) != hipSuccess)
    return set_err(DMR_ERR_HIP, "hipEventRecord failed");
  const int rc = f();
  if (hipEventRecord(p.ev[2 * p.used + 1], s) != hipSuccess)
    return set_err(DMR_ERR_HIP, "hipEventRecord failed");
  if ((int)p.phase.size() <= p.used) {
    p.phase.resize(p.used + 1);
    p.bytes.resize(p.used + 1);
  }
  p.phase[p.used] = phase;
  p.bytes[p.used] = bytes;
  ++p.used;
  return rc;
}

void prof_collect(dmrContext *c) {
  auto &p = c->prof;
  for (int i = 0; i < p.used; ++i) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p.ev[2 * i], p.ev[2 * i + 1]) == hipSuccess) {
      p.ms[p.phase[i]] += ms;
      p.by[p.phase[i]] += p.bytes[i];
      ++p.cnt[p.phase[i]];
    }
  }
  p.used = 0;
}

// The restart loop of solverDMR (gDMR.cu:2814-2874) from a normalised residual in w.r, the
// current approximation in w.xa, tracked residual norm rho and ||b|| = beta.  Returns DMR_OK on
// convergence, DMR_ERR_NOCONV / DMR_ERR_RUNAWAY on the reference's failure exits, or an error.
template <typename TS, typename Op>
int run_cycles(dmrContext *c, const Op &op0, hipStream_t s0, int startDef, double tol,
               int matRank, double beta, double &rho, int &grdItr, double histScale) {
  Work<TS> &w = work_of<TS>(c);
  // A cycle's launches go to s: the caller's stream s0, or, when cycles are captured, the
  // context's own stream cs (the operator launches there too).
  const bool useGraph = c->graphs && !c->prof.on && !c->fused && c->opKind == OP_CSR;
  hipStream_t s = useGraph ? c->cs : s0;
  Op op = op0;
  op.s = s;
  const long long n = c->n, ld = c->ld;
  const int m = c->m, k = c->k;
  const dmr::RedWork red = c->red;
  // Algorithmic bytes: one length-n vector pass, and one SpMV (CSR, int32 indices).
  const double vb = (double)sizeof(TS) * (double)n;
  const double spmvBytes = (double)c->csr.nnz * (double)(sizeof(TS) + 4) + 4.0 * (n + 1) +
                           2.0 * (double)sizeof(TS) * (double)n;
  while (rho > tol * beta) {
    if (grdItr > matRank) return DMR_ERR_NOCONV;
    if (rho / beta > 1.0) return DMR_ERR_RUNAWAY;
    const int d = (grdItr == 0 && startDef == 0) ? 0 : k;
    const int o = d > 0 ? k : 0;
    // ---- arnoldiSetupDMR (gDMR.cu:2586-2632).  The period 2*matRank/(m-k) divides by zero in
    // the reference when it is 0; a zero period here means "always restart from r".
    const int period = (m - k) > 0 ? (2 * matRank) / (m - k) : 0;
    const bool fromResidual =
        c->kryLoopFlag == 0 || period == 0 || (c->kryFlagCount % period) == 0;
    TS *Po = w.P + (long long)o * ld;
    const bool implicitImg = d > 0 && c->implicitImage && c->imgValid;
    // ---- genDeflatDMR, first half (gDMR.cu:1583-1704): the SVD of the Hessenberg matrix runs on
    // the side stream, overlapped with the work that does not need it.  It takes the leading
    // (cols + 1) x cols block, cols = d + lead: the whole matrix (lead = m - d) in the FP64 path;
    // in the FP32 and bf16 inner solves the block that is final refreshTail Krylov steps before
    // the end of the cycle, so that the SVD runs beside those steps (the refreshed subspace then
    // combines Z and the first lead Krylov vectors).
    const bool refresh = k > 0;
    const int lead = (m - d - c->refreshTail >= 1 && d + m - d - c->refreshTail >= k)
                         ? m - d - c->refreshTail
                         : m - d;
    const int cols = d + lead;
    // rho of this cycle for the solution update, read on the device (the host's value).
    c->hostStatus[20] = rho;
    if (hipMemcpyAsync(c->rhoDev, c->hostStatus + 20, sizeof(double), hipMemcpyHostToDevice,
                       s0) != hipSuccess)
      return set_err(DMR_ERR_HIP, "rho upload failed");
    // The cycle's launches, from the start vector to the refreshed subspace.
    auto enqueue = [&]() -> int {
    if (fromResidual) {
      if (hipMemcpyAsync(Po, w.r, sizeof(TS) * ld, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return set_err(DMR_ERR_HIP, "setup copy failed");
    } else {
      const TS *src = w.P + (long long)(o + (m - k) / 2) * ld;
      if (src != Po &&
          hipMemcpyAsync(Po, src, sizeof(TS) * ld, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return set_err(DMR_ERR_HIP, "setup copy failed");
    }
    // ---- arnoldiDMR (gDMR.cu:1790-2332)
    dmr::cycle_begin(c->H, m, c->flag, c->bar, s);
    c->epoch = 0;
    // Two-stage classical Gram-Schmidt of g into dst (deflatorDMR then the Krylov projection,
    // kept as two reductions as in the reference): u = g; [hA = A^T u; u -= A hA];
    // [hB = B^T u; u -= B hB]; norms; dst = u / ||u||.  One fused cooperative kernel when the
    // grid fits the device, else the same arithmetic as separate kernels.
    auto orthogonalize = [&](const TS *g, TS *dst, const TS *A, int pa, double *ha, const TS *B,
                             int pb, double *hb, int redm, const dmr::Fin &fin,
                             const int *skip) {
      if (c->fused) {
        int ph = 0;
        timed(c, s, DMR_PH_ORTH, vb * (2.0 + 2.0 * pa + 2.0 * pb), [&] {
          ph = dmr::orth<TS>(g, dst, A, pa, ha, B, pb, hb, ld, n, redm, fin, red, c->bar, c->epoch,
                             c->devErr, skip, s);
          return 0;
        });
        if (ph > 0) {
          c->epoch += (unsigned)ph;
          return;
        }
      }
      const TS *src = g;
      if (pa > 0) {
        timed(c, s, DMR_PH_GEMVT, vb * (pa + 1), [&] {
          dmr::multidot<TS>(A, ld, pa, g, n, ha, 1, red, skip, s);
          return 0;
        });
        timed(c, s, DMR_PH_UPDATE, vb * (pa + 2), [&] {
          dmr::combo<TS>(dmr::COMB_SUB, g, w.u, A, pa, ha, nullptr, 0, nullptr, ld, n, 0.0, 0,
                         nullptr, fin_none(), red, skip, s);
          return 0;
        });
        src = w.u;
      }
      if (pb > 0)
        timed(c, s, DMR_PH_GEMVT, vb * (pb + 1), [&] {
          dmr::multidot<TS>(B, ld, pb, src, n, hb, 1, red, skip, s);
          return 0;
        });
      timed(c, s, DMR_PH_UPDATE, vb * (pb + 2 + (redm & 2 ? 1 : 0)), [&] {
        dmr::combo<TS>(dmr::COMB_SUB, src, dst, B, pb, hb, nullptr, 0, nullptr, ld, n, 0.0, redm,
                       g, fin, red, skip, s);
        return 0;
      });
      timed(c, s, DMR_PH_SCALE, vb * 2, [&] {
        dmr::scale<TS>(dst, n, c->scal + 1, skip, s);
        return 0;
      });
    };
    // Deflation image: A Z = W R by CGS (gDMR.cu:1823-2031), or, in an FP32 inner solve whose Z
    // was refreshed by the previous cycle under the same operator, from that cycle's Arnoldi
    // relation: W = P Q (in Wn) and R diag(1 / ||z~||) (image_qr at the previous cycle's end).
    if (implicitImg) {
      const int rcImg = timed(c, s, DMR_PH_COMBO, vb * 2.0 * k, [&] {
        if (hipMemcpyAsync(w.P, w.Wn, sizeof(TS) * (size_t)ld * k, hipMemcpyDeviceToDevice, s) !=
            hipSuccess)
          return set_err(DMR_ERR_HIP, "image copy failed");
        dmr::image_fill(c->H, m, k, c->imgR, c->scalZ, s);
        return 0;
      });
      if (rcImg) return rcImg;
    }
    for (int i = 0; i < (implicitImg ? 0 : d); ++i) {
      int rc = timed(c, s, DMR_PH_SPMV, spmvBytes, [&] { return op(w.Z + (long long)i * ld, w.g); });
      if (rc) return rc;
      ++c->st.matvecs;
      double *hcol = c->H + (long long)i * (m + 1);
      dmr::Fin fin{dmr::FIN_ARNOLDI, c->scal, hcol + i, 0.0, 0, nullptr};
      orthogonalize(w.g, w.P + (long long)i * ld, nullptr, 0, nullptr, w.P, i, hcol, 3, fin,
                    nullptr);
    }
    // Start vector orthogonalised against W (deflatorDMR(0, 0), gDMR.cu:2033-2060).
    if (d > 0) {
      dmr::Fin fin{dmr::FIN_NORM, c->scal, nullptr, 0.0, 0, nullptr};
      orthogonalize(Po, Po, nullptr, 0, nullptr, w.P, k, c->cv, 1, fin, nullptr);
    }
    bool forked = false;
    auto fork_refresh = [&]() -> int {
      forked = true;
      if (hipEventRecord(c->evH, s) != hipSuccess ||
          hipStreamWaitEvent(c->side, c->evH, 0) != hipSuccess)
        return set_err(DMR_ERR_HIP, "side stream fork failed");
      const double svdTol = std::is_same<TS, double>::value ? kSvdTol : kSvdTolLow;
      timed(c, c->side, DMR_PH_SVD, 0.0, [&] {
        dmr::svd_smallest_rect(c->H, cols + 1, m + 1, cols, k, svdTol, kSvdMaxSweeps, 0, c->C,
                               c->sig, c->sweeps, c->side);
        if (c->implicitImage)
          dmr::image_qr(c->H, m + 1, cols, c->C, k, c->imgQ, c->imgR, c->imgOk, c->side);
        return 0;
      });
      if (hipEventRecord(c->evSvd, c->side) != hipSuccess)
        return set_err(DMR_ERR_HIP, "side stream join failed");
      return 0;
    };
    // Krylov steps (gDMR.cu:2069-2328): deflatorDMR(1, j) puts W^T g into H[0:k, d+j-1], then
    // CGS against V_0..V_{j-1} puts h into H[d:d+j, d+j-1]; H[d+j, d+j-1] = v_j . g.
    for (int j = 1; j <= m - d; ++j) {
      int rc = timed(c, s, DMR_PH_SPMV, spmvBytes,
                     [&] { return op(w.P + (long long)(o + j - 1) * ld, w.g); });
      if (rc) return rc;
      ++c->st.matvecs;
      double *hcol = c->H + (long long)(d + j - 1) * (m + 1);
      TS *vj = w.P + (long long)(o + j) * ld;
      dmr::Fin fin{dmr::FIN_ARNOLDI, c->scal, hcol + d + j, tol, j, c->flag};
      if (d > 0 && c->oneStage)
        // [W V_0..V_{j-1}] = P[0..k+j-1] is contiguous: one projection and one update instead of
        // deflate-then-CGS (the inner FP32 solve of the mixed path; H receives the same entries).
        orthogonalize(w.g, vj, nullptr, 0, nullptr, w.P, k + j, hcol, 3, fin, c->flag);
      else
        orthogonalize(w.g, vj, w.P, d > 0 ? k : 0, hcol, Po, j, hcol + d, 3, fin, c->flag);
      if (refresh && j == lead && lead < m - d) {
        rc = fork_refresh();
        if (rc) return rc;
      }
    }
    dmr::breakdown_fixup<TS>(w.P, ld, n, o, d, m, c->H, c->flag, s);
    if (refresh && !forked) {
      const int rc = fork_refresh();
      if (rc) return rc;
    }
    // ---- updateAppResDMR (gDMR.cu:2411-2582)
    timed(c, s, DMR_PH_GEMVT, vb * (m + 2), [&] {
      dmr::multidot<TS>(w.P, ld, m + 1, w.r, n, c->t, 1, red, nullptr, s);
      return 0;
    });
    timed(c, s, DMR_PH_LSQ, 0.0, [&] {
      dmr::lsq(c->H, m, c->t, c->y, c->z, s);
      return 0;
    });
    {
      // r <- r - ([V] z[d..m] + [W] z[0..d-1]), q = ||r||.
      dmr::Fin fin{dmr::FIN_NORM, c->scal2, nullptr, 0.0, 0, nullptr};
      timed(c, s, DMR_PH_COMBO, vb * (m + 3), [&] {
        dmr::combo<TS>(dmr::COMB_ACC, w.r, w.r, Po, m + 1 - o, c->z + d, w.P, d, c->z, ld, n,
                       -1.0, 1, nullptr, fin, red, nullptr, s);
        return 0;
      });
      timed(c, s, DMR_PH_SCALE, vb * 2, [&] {
        dmr::scale<TS>(w.r, n, c->scal2 + 1, nullptr, s);
        return 0;
      });
      // x_a <- x_a + rho * ([V] y[d..m-1] + [Z] y[0..d-1]).
      timed(c, s, DMR_PH_COMBO, vb * (m + 2), [&] {
        dmr::combo<TS>(dmr::COMB_ACC, w.xa, w.xa, Po, m - d, c->y + d, w.Z, d, c->y, ld, n, 0.0,
                       0, nullptr, fin_none(), red, nullptr, s, c->rhoDev);
        return 0;
      });
    }
    // ---- genDeflatDMR, second half: Z_new = normalize([V Z] C) into Zn, adopted below unless
    // the cycle broke down.
    if (refresh) {
      if (hipStreamWaitEvent(s, c->evSvd, 0) != hipSuccess)
        return set_err(DMR_ERR_HIP, "side stream join failed");
      timed(c, s, DMR_PH_ZGEMM, vb * (cols + k) + vb * 2 * k, [&] {
        dmr::combo_multi<TS>(w.Zn, k, Po, lead, c->C + d, w.Z, d, c->C, cols, ld, n, c->scalZ,
                             red, s);
        dmr::scale_cols<TS>(w.Zn, k, ld, n, c->scalZ, s);
        return 0;
      });
      if (c->implicitImage) {
        timed(c, s, DMR_PH_ZGEMM, vb * (cols + 1 + k), [&] {
          dmr::combo_multi<TS>(w.Wn, k, w.P, cols + 1, c->imgQ, nullptr, 0, nullptr, cols + 1, ld,
                               n, c->imgScal, red, s);
          return 0;
        });
        // The relation carries the rounding of the cycle's basis, and in a chain of implicit
        // images the error of each one feeds the next (through the Z rows of C, divided by
        // ||z~||): on an ill-conditioned operator it can grow from cycle to cycle.  So every
        // implicit image is checked on its last column, whose singular value is the smallest:
        // e = A z_{k-1} - W R[:, k-1] / ||z~_{k-1}||, and a relative error above image_tol()
        // makes the next cycle recompute the image (one SpMV and one k-column update here,
        // against k SpMVs and a Gram-Schmidt QR saved per cycle).
        int rcChk = timed(c, s, DMR_PH_SPMV, spmvBytes,
                          [&] { return op(w.Zn + (long long)(k - 1) * ld, w.g); });
        if (rcChk) return rcChk;
        ++c->st.matvecs;
        dmr::image_check_prep(c->imgR, k, c->scalZ, c->imgCol, s);
        dmr::Fin finChk{dmr::FIN_STORE, c->imgChk, nullptr, 0.0, 0, nullptr};
        timed(c, s, DMR_PH_COMBO, vb * (k + 2), [&] {
          dmr::combo<TS>(dmr::COMB_SUB, w.g, w.u, w.Wn, k, c->imgCol, nullptr, 0, nullptr, ld, n,
                         0.0, 1, nullptr, finChk, red, nullptr, s);
          return 0;
        });
        dmr::image_check_final(c->imgCol, k, c->imgChk, image_tol(), c->imgOk, s);
      }
    }
    return 0;
    };
    if (useGraph) {
      // Replay of this variant's captured cycle (captured now if it is new).  What the launch
      // sequence depends on: d, the start vector's source, the implicit image, which buffer holds
      // Z, and the breakdown tolerance; the operator's pointers are fixed while graphs exist.
      const int zParity = w.Z < w.Zn ? 0 : 1;
      const void *val = c->csr.val, *val32 = c->csr.val32;
      const dmrContext::CycleGraph *g = nullptr;
      for (const auto &e : c->cycleGraphs)
        if (e.d == d && e.fromResidual == (int)fromResidual && e.implicitImg == (int)implicitImg &&
            e.zParity == zParity && e.tol == tol && e.val == val && e.val32 == val32)
          g = &e;
      if (!g) {
        const long long mv0 = c->st.matvecs;
        if (hipStreamBeginCapture(c->cs, hipStreamCaptureModeThreadLocal) != hipSuccess)
          return set_err(DMR_ERR_HIP, "cycle capture failed");
        const int rcE = enqueue();
        hipGraph_t graph = nullptr;
        const hipError_t ee = hipStreamEndCapture(c->cs, &graph);
        if (rcE) {
          if (graph) hipGraphDestroy(graph);
          return rcE;
        }
        if (ee != hipSuccess) return set_err(DMR_ERR_HIP, "cycle capture failed");
        hipGraphExec_t ex = nullptr;
        const hipError_t ei = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (ei != hipSuccess) return set_err(DMR_ERR_HIP, "cycle graph instantiation failed");
        if ((int)c->cycleGraphs.size() >= kMaxCycleGraphs) {
          hipGraphExecDestroy(c->cycleGraphs.front().exec);
          c->cycleGraphs.erase(c->cycleGraphs.begin());
        }
        c->cycleGraphs.push_back({d, (int)fromResidual, (int)implicitImg, zParity, tol, val,
                                  val32, ex, c->st.matvecs - mv0});
        c->st.matvecs = mv0;
        g = &c->cycleGraphs.back();
      }
      if (hipEventRecord(c->evIn, s0) != hipSuccess ||
          hipStreamWaitEvent(c->cs, c->evIn, 0) != hipSuccess ||
          hipGraphLaunch(g->exec, c->cs) != hipSuccess)
        return set_err(DMR_ERR_HIP, "cycle graph launch failed");
      c->st.matvecs += g->matvecs;
    } else {
      const int rcE = enqueue();
      if (rcE) return rcE;
    }
    if (hipMemcpyAsync(c->hostStatus, c->scal2, 4 * sizeof(double), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipMemcpyAsync(c->hostStatus + 8, c->flag, sizeof(int), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipMemcpyAsync(c->hostStatus + 9, c->devErr, sizeof(int), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipMemcpyAsync(c->hostStatus + 10, c->sweeps, sizeof(int), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        (c->implicitImage && refresh &&
         hipMemcpyAsync(c->hostStatus + 11, c->imgOk, sizeof(int), hipMemcpyDeviceToHost, s) !=
             hipSuccess))
      return set_err(DMR_ERR_HIP, "status readback failed");
    if (useGraph && (hipEventRecord(c->evOut, c->cs) != hipSuccess ||
                     hipStreamWaitEvent(s0, c->evOut, 0) != hipSuccess))
      return set_err(DMR_ERR_HIP, "stream hand-off failed");
    if (hipStreamSynchronize(s) != hipSuccess) return set_err(DMR_ERR_HIP, "stream sync failed");
    if (c->prof.on) prof_collect(c);
    const double q = c->hostStatus[2];
    int bflag, derr, sweeps;
    std::memcpy(&bflag, c->hostStatus + 8, sizeof(int));
    std::memcpy(&derr, c->hostStatus + 9, sizeof(int));
    std::memcpy(&sweeps, c->hostStatus + 10, sizeof(int));
    if (derr) return set_err(DMR_ERR_HIP, "device grid barrier exceeded its wait bound");
    if (bflag) ++c->st.breakdowns;
    // Z <- the refreshed subspace (genDeflatDMR).  After a breakdown the basis beyond the
    // breakdown step is zero and H is padded with identity columns, so singular vectors from the
    // padding combine zero vectors (the reference then divides by sqrt(0)); such a refresh is
    // discarded and the previous subspace is kept.
    int imgOk = 0;
    if (c->implicitImage && refresh) std::memcpy(&imgOk, c->hostStatus + 11, sizeof(int));
    c->imgValid = c->implicitImage && refresh && bflag == 0 && imgOk == 1;
    if (refresh && bflag == 0) {
      std::swap(w.Z, w.Zn);
      w.hasZ = true;
      if (sweeps > 0) {
        c->st.svdSweeps += sweeps;
        ++c->st.svdCalls;
      }
    }
    if (q > kStagnation) {
      c->kryLoopFlag = 1;
      ++c->kryFlagCount;
    } else {
      c->kryLoopFlag = 0;
    }
    rho = q * rho;
    c->hist.push_back(rho / beta * histScale);
    ++c->st.cycles;
    ++grdItr;
  }
  if (c->prof.on) {
    if (hipStreamSynchronize(s) != hipSuccess) return set_err(DMR_ERR_HIP, "stream sync failed");
    prof_collect(c);
  }
  return DMR_OK;
}

double solve_fp64(dmrContext *c, int mode, double *x, const double *b, hipStream_t s, double tol,
                  int matRank, int *numIts) {
  Work<double> &w = c->w64;
  const long long n = c->n, ld = c->ld;
  const int k = c->k;
  *numIts = matRank * matRank;  // gDMR.cu:2646
  // mode != 0 deflates the first cycle with the stored subspace; with none stored (no previous
  // solve, or only cycles that broke down) the first cycle runs undeflated as for mode 0.
  const int startDef = (mode != 0 && w.hasZ) ? k : 0;
  c->kryLoopFlag = 0;
  const dmr::RedWork red = c->red;
  Op64 op{c, s};
  if (c->opKind == OP_LEGACY) c->ext = x;
  // r = b, beta = ||b|| (gDMR.cu:2681-2707); x_a = x0 (gDMR.cu:2709).
  if (hipMemcpyAsync(w.r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s) != hipSuccess ||
      hipMemcpyAsync(w.xa, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s) != hipSuccess) {
    set_err(DMR_ERR_HIP, "input copy failed");
    return 1.0;
  }
  dmr::multidot<double>(w.r, ld, 1, w.r, n, c->scal2 + 4, 1, red, nullptr, s);
  // A x0 through the operator; the legacy callback works in place on x (gDMR.cu:2732).
  int rc;
  if (c->opKind == OP_LEGACY) {
    c->legacy();
    rc = DMR_OK;
    if (g_op_failed) {
      g_op_failed = false;
      rc = set_err(DMR_ERR_OPERATOR, "operator reported failure: " + g_op_msg);
    } else if (hipMemcpyAsync(w.g, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s) !=
               hipSuccess) {
      rc = set_err(DMR_ERR_HIP, "copy from operator vector failed");
    }
  } else {
    rc = op(x, w.g);
  }
  if (rc) return 1.0;
  ++c->st.matvecs;
  // r = r - A x0, rho = ||r|| (gDMR.cu:2748-2783).
  {
    dmr::Fin fin{dmr::FIN_NORM, c->scal, nullptr, 0.0, 0, nullptr};
    dmr::combo<double>(dmr::COMB_SUB, w.r, w.r, w.g, 1, c->one, nullptr, 0, nullptr, ld, n, 0.0,
                       1, nullptr, fin, red, nullptr, s);
  }
  if (hipMemcpyAsync(c->hostStatus, c->scal, 3 * sizeof(double), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipMemcpyAsync(c->hostStatus + 4, c->scal2 + 4, sizeof(double), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_err(DMR_ERR_HIP, "setup readback failed");
    return 1.0;
  }
  const double beta = std::sqrt(c->hostStatus[4]);
  double rho = std::sqrt(c->hostStatus[0]);
  c->st.beta = beta;
  if (rho < beta * tol) {
    // Initial guess already within tolerance (gDMR.cu:2785-2796): the reference returns the
    // absolute residual norm here and leaves numIts at matRank^2.
    hipMemcpyAsync(x, w.xa, sizeof(double) * n, hipMemcpyDeviceToDevice, s);
    hipStreamSynchronize(s);
    c->st.status = -1;
    c->st.rho = rho;
    return rho;
  }
  dmr::scale<double>(w.r, n, c->scal + 1, nullptr, s);
  int grdItr = 0;
  rc = run_cycles<double>(c, op, s, startDef, tol, matRank, beta, rho, grdItr, 1.0);
  c->st.rho = rho;
  if (rc == DMR_ERR_NOCONV || rc == DMR_ERR_RUNAWAY) {
    *numIts = grdItr;
    c->st.status = rc;
    set_err(rc, rc == DMR_ERR_NOCONV ? "cycle limit (matRank) reached" : "residual runaway");
    return rho / beta;
  }
  if (rc) return rho / beta;
  *numIts = grdItr - 1;  // gDMR.cu:2876
  c->st.status = 0;
  // Export the solution (gDMR.cu:2885).
  if (hipMemcpyAsync(x, w.xa, sizeof(double) * n, hipMemcpyDeviceToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_err(DMR_ERR_HIP, "solution export failed");
  }
  return rho / beta;
}

// Mixed precision: FP64 iterative refinement around an inner DMR whose Krylov basis, deflation
// vectors and SpMV are FP32 with FP32-accumulated reductions; the Hessenberg LSQ / QR / SVD stay
// FP64.  Each outer step computes r = b - A x with the FP64 operator, solves A d = r with the
// inner DMR to a relative tolerance max(innerTol, 0.5 * tol * ||b|| / ||r||) starting from the
// recycled FP32 subspace, and adds d to x.  Returns the FP64-checked ||b - A x|| / ||b||.
// Mixed precision: FP64 refinement around an inner DMR whose basis is stored as TS (FP32 or
// bf16) with FP32 SpMV values and FP32-accumulated reductions.
template <typename TS>
double solve_mixed(dmrContext *c, int mode, double *x, const double *b, hipStream_t s, double tol,
                   int matRank, int *numIts) {
  Work<TS> &w = work_of<TS>(c);
  const long long n = c->n, ld = c->ld;
  const int k = c->k;
  *numIts = 0;
  if (c->opKind != OP_CSR) {
    set_err(DMR_ERR_STATE, "mixed precision needs the built-in CSR operator");
    return 1.0;
  }
  const dmr::RedWork red = c->red;
  if (c->val32Stale) {
    // FP32 copy of the operator values, ordered on the solve stream (no device-wide sync, so
    // solves on other streams keep running).
    dmr::convert<float, double>(c->ownedVal32, c->csr.val, c->csr.nnz, s);
    c->val32Stale = false;
  }
  if (hipMemcpyAsync(c->b64, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s) != hipSuccess ||
      hipMemcpyAsync(c->x64, x, sizeof(double) * n, hipMemcpyDeviceToDevice, s) != hipSuccess) {
    set_err(DMR_ERR_HIP, "input copy failed");
    return 1.0;
  }
  dmr::multidot<double>(c->b64, ld, 1, c->b64, n, c->scal2 + 4, 1, red, nullptr, s);
  if (hipMemcpyAsync(c->hostStatus + 4, c->scal2 + 4, sizeof(double), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_err(DMR_ERR_HIP, "readback failed");
    return 1.0;
  }
  const double beta = std::sqrt(c->hostStatus[4]);
  c->st.beta = beta;
  typename LowOp<TS>::type op{c, s};
  int totalCycles = 0;
  double relres = 1.0;
  c->kryLoopFlag = 0;
  // The operator values may have changed since the last solve: the first deflated cycle
  // recomputes the image of Z.
  c->imgValid = false;
  bool useZ = mode != 0 && w.hasZ;
  for (int outer = 0;; ++outer) {
    dmr::residual64(c->csr, c->x64, c->b64, c->r64, c->redOp, c->scal, s);
    ++c->st.matvecs;
    if (hipMemcpyAsync(c->hostStatus, c->scal, 3 * sizeof(double), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_err(DMR_ERR_HIP, "residual readback failed");
      return relres;
    }
    const double rn = c->hostStatus[2];
    relres = rn / beta;
    if (relres <= tol || beta == 0.0) break;
    if (outer >= c->maxOuter || totalCycles > matRank) {
      c->st.status = DMR_ERR_NOCONV;
      set_err(DMR_ERR_NOCONV, "refinement did not reach the tolerance");
      break;
    }
    // r32 = r / ||r||, x_a = 0.
    dmr::scale<double>(c->r64, n, c->scal + 1, nullptr, s);
    dmr::convert<TS, double>(w.r, c->r64, n, s);
    dmr::fill<TS>(w.xa, ld, 0.0, s);
    double tolIn = std::max(c->innerTol, 0.5 * tol / relres);
    if (tolIn > 0.5) tolIn = 0.5;
    double rho = rn;
    int grdItr = 0;
    const int startDef = useZ ? k : 0;
    const int rc = run_cycles<TS>(c, op, s, startDef, tolIn, matRank, rn, rho, grdItr,
                                  rn / beta);
    totalCycles += grdItr;
    if (rc != DMR_OK && rc != DMR_ERR_NOCONV && rc != DMR_ERR_RUNAWAY) return relres;
    useZ = (k > 0) && w.hasZ;
    dmr::add_to<TS>(c->x64, w.xa, n, s);
  }
  *numIts = totalCycles;
  if (c->st.status == 0) c->st.status = relres <= tol ? 0 : DMR_ERR_NOCONV;
  c->st.rho = relres * beta;
  if (hipMemcpyAsync(x, c->x64, sizeof(double) * n, hipMemcpyDeviceToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    set_err(DMR_ERR_HIP, "solution export failed");
  return relres;
}

dmrContext *g_ctx = nullptr;

int create_ctx(dmrContext **out, int device, int n, int m, int k, int prec) {
  *out = nullptr;
  int rc = validate_shape(n, m, k);
  if (rc) return rc;
  if (prec != DMR_PREC_FP64 && prec != DMR_PREC_FP32 && prec != DMR_PREC_BF16)
    return set_err(DMR_ERR_ARG, "unknown precision");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_err(DMR_ERR_HIP, "no HIP device");
  if (device < 0 || device >= ndev) return set_err(DMR_ERR_ARG, "device out of range");
  HIPCHK(hipSetDevice(device));
  dmrContext *c = new dmrContext();
  c->device = device;
  c->n = n;
  c->m = m;
  c->k = k;
  c->prec = prec;
  c->ld = pad_ld(n);
  rc = ctx_alloc(c);
  if (rc) {
    ctx_free(c);
    return rc;
  }
  *out = c;
  return DMR_OK;
}

}  // namespace

// =============================================================================================
// C ABI
// =============================================================================================
extern "C" {

int dmrGetLastError(void) { return g_err; }
const char *dmrGetLastErrorString(void) { return g_errmsg.c_str(); }
void dmrRaiseOperatorError(const char *msg) {
  g_op_failed = true;
  g_op_msg = msg ? msg : "";
}

int dmrReductionBlocks(int elements) { return dmr::red_blocks<double>(elements); }

size_t dmrFootprint(int elements, int basisDim, int deflatDim, int precision) {
  return footprint(elements, basisDim, deflatDim, precision);
}

int dmrCreate(dmrHandle *h, int device, int elements, int basisDim, int deflatDim, int precision) {
  clear_err();
  if (!h) return set_err(DMR_ERR_ARG, "null handle pointer");
  return create_ctx(h, device, elements, basisDim, deflatDim, precision);
}

int dmrDestroy(dmrHandle h) {
  if (g_ctx == h) g_ctx = nullptr;
  ctx_free(h);
  return DMR_OK;
}

int dmrSetOperatorCsr(dmrHandle h, int nnz, const int *rowPtr, const int *colInd,
                      const double *val, const float *val32) {
  clear_err();
  if (!h || !rowPtr || !colInd || !val || nnz < 0)
    return set_err(DMR_ERR_ARG, "invalid CSR arguments");
  HIPCHK(hipSetDevice(h->device));
  drop_cycle_graphs(h);
  const int n = h->n;
  std::vector<int> rp(n + 1);
  HIPCHK(hipMemcpy(rp.data(), rowPtr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  if (rp[0] != 0 || rp[n] != nnz) return set_err(DMR_ERR_ARG, "rowPtr inconsistent with nnz");
  std::vector<int> br(n + 1);
  const int nb = dmr::spmv_partition(rp.data(), n, br.data());
  if (h->ownedBlockRows) hipFree(h->ownedBlockRows);
  h->ownedBlockRows = nullptr;
  HIPCHK(hipMalloc(&h->ownedBlockRows, sizeof(int) * (nb + 1)));
  HIPCHK(hipMemcpy(h->ownedBlockRows, br.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice));
  if (h->redOpPool) hipFree(h->redOpPool);
  h->redOpPool = nullptr;
  const size_t rbytes = sizeof(double) * (size_t)std::max(nb, 1) + sizeof(unsigned) * dmr::kCounters;
  HIPCHK(hipMalloc(&h->redOpPool, rbytes));
  HIPCHK(hipMemset(h->redOpPool, 0, rbytes));
  h->redOp.partials = h->redOpPool;
  h->redOp.counters = (unsigned *)((char *)h->redOpPool + sizeof(double) * std::max(nb, 1));
  h->csr.n = n;
  h->csr.nnz = nnz;
  h->csr.rowPtr = rowPtr;
  h->csr.colInd = colInd;
  h->csr.val = val;
  h->csr.blockRows = h->ownedBlockRows;
  h->csr.nBlocks = nb;
  h->csr.val32 = val32;
  if (h->prec != DMR_PREC_FP64 && !val32) {
    if (h->ownedVal32) hipFree(h->ownedVal32);
    h->ownedVal32 = nullptr;
    HIPCHK(hipMalloc(&h->ownedVal32, sizeof(float) * std::max(nnz, 1)));
    h->csr.val32 = h->ownedVal32;
    h->val32Stale = true;
  } else {
    h->val32Stale = false;
  }
  h->opKind = OP_CSR;
  return DMR_OK;
}

int dmrSetOperatorValues(dmrHandle h, const double *val, const float *val32) {
  clear_err();
  if (!h || h->opKind != OP_CSR || !val) return set_err(DMR_ERR_STATE, "no CSR operator to update");
  bool stale = false;
  if (h->prec != DMR_PREC_FP64 && !val32) {
    if (!h->ownedVal32) return set_err(DMR_ERR_STATE, "no FP32 value buffer");
    val32 = h->ownedVal32;
    stale = true;
  }
  h->csr.val = val;
  h->csr.val32 = val32;
  h->val32Stale = stale;
  return DMR_OK;
}

int dmrSetOperatorFn(dmrHandle h, dmrOperatorFn op, void *ctx) {
  clear_err();
  if (!h || !op) return set_err(DMR_ERR_ARG, "null operator");
  h->opFn = op;
  h->opCtx = ctx;
  h->opKind = OP_FN;
  return DMR_OK;
}

int dmrSetOperatorLegacy(dmrHandle h, void (*linComp)(void)) {
  clear_err();
  if (!h || !linComp) return set_err(DMR_ERR_ARG, "null operator");
  h->legacy = linComp;
  h->opKind = OP_LEGACY;
  return DMR_OK;
}

double dmrSolve(dmrHandle h, int mode, double *x, const double *b, hipStream_t stream,
                double solTol, int matRank, int *numIts) {
  clear_err();
  int dummy = 0;
  if (!numIts) numIts = &dummy;
  if (!h) {
    set_err(DMR_ERR_STATE, "solver not initialised");
    return 1.0;
  }
  if (!x || !b || matRank < 0 || !(solTol >= 0.0)) {
    set_err(DMR_ERR_ARG, "invalid solve arguments");
    return 1.0;
  }
  if (h->opKind == OP_NONE) {
    set_err(DMR_ERR_STATE, "no operator set");
    return 1.0;
  }
  if (hipSetDevice(h->device) != hipSuccess) {
    set_err(DMR_ERR_HIP, "hipSetDevice failed");
    return 1.0;
  }
  g_op_failed = false;
  h->st = dmrStats{};
  h->hist.clear();
  const auto t0 = std::chrono::steady_clock::now();
  double res;
  if (h->prec == DMR_PREC_FP64)
    res = solve_fp64(h, mode, x, b, stream, solTol, matRank, numIts);
  else if (h->prec == DMR_PREC_BF16)
    res = solve_mixed<dmr::bf16>(h, mode, x, b, stream, solTol, matRank, numIts);
  else
    res = solve_mixed<float>(h, mode, x, b, stream, solTol, matRank, numIts);
  h->st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res;
}

namespace {
// The recycled subspace of a handle in its storage precision: device pointer, element size,
// and its hasZ flag.
struct ZView {
  void *Z;
  size_t es;
  bool *has;
};
ZView zview(dmrHandle h) {
  if (h->prec == DMR_PREC_FP64) return {h->w64.Z, sizeof(double), &h->w64.hasZ};
  if (h->prec == DMR_PREC_BF16) return {h->w16.Z, 2, &h->w16.hasZ};
  return {h->w32.Z, sizeof(float), &h->w32.hasZ};
}
// Host conversions to and from the storage precision (bf16: round to nearest even).
void to_storage(const double *src, size_t cnt, int prec, void *dst) {
  if (prec == DMR_PREC_FP64) {
    std::memcpy(dst, src, sizeof(double) * cnt);
  } else if (prec == DMR_PREC_BF16) {
    uint16_t *d = (uint16_t *)dst;
    for (size_t i = 0; i < cnt; ++i) {
      const float f = (float)src[i];
      uint32_t u;
      std::memcpy(&u, &f, 4);
      d[i] = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
    }
  } else {
    float *d = (float *)dst;
    for (size_t i = 0; i < cnt; ++i) d[i] = (float)src[i];
  }
}
void from_storage(const void *src, size_t cnt, int prec, double *dst) {
  if (prec == DMR_PREC_FP64) {
    std::memcpy(dst, src, sizeof(double) * cnt);
  } else if (prec == DMR_PREC_BF16) {
    const uint16_t *s16 = (const uint16_t *)src;
    for (size_t i = 0; i < cnt; ++i) {
      const uint32_t u = (uint32_t)s16[i] << 16;
      float f;
      std::memcpy(&f, &u, 4);
      dst[i] = f;
    }
  } else {
    const float *s32 = (const float *)src;
    for (size_t i = 0; i < cnt; ++i) dst[i] = s32[i];
  }
}
}  // namespace

int dmrHasZ(dmrHandle h) {
  if (!h) return 0;
  return *zview(h).has ? 1 : 0;
}

int dmrImportZ(dmrHandle h, const double *hostZ) {
  clear_err();
  if (!h || !hostZ) return set_err(DMR_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(h->device));
  if (h->k == 0) return DMR_OK;
  const ZView z = zview(h);
  std::vector<char> tmp(z.es * (size_t)h->n * h->k);
  to_storage(hostZ, (size_t)h->n * h->k, h->prec, tmp.data());
  HIPCHK(hipMemcpy2D(z.Z, z.es * h->ld, tmp.data(), z.es * h->n, z.es * h->n, h->k,
                     hipMemcpyHostToDevice));
  *z.has = true;
  return DMR_OK;
}

int dmrExportZ(dmrHandle h, double *hostZ) {
  clear_err();
  if (!h || !hostZ) return set_err(DMR_ERR_ARG, "null argument");
  HIPCHK(hipSetDevice(h->device));
  if (h->k == 0) return DMR_OK;
  HIPCHK(hipDeviceSynchronize());
  const ZView z = zview(h);
  std::vector<char> tmp(z.es * (size_t)h->n * h->k);
  HIPCHK(hipMemcpy2D(tmp.data(), z.es * h->n, z.Z, z.es * h->ld, z.es * h->n, h->k,
                     hipMemcpyDeviceToHost));
  from_storage(tmp.data(), (size_t)h->n * h->k, h->prec, hostZ);
  return DMR_OK;
}

size_t dmrSubspaceBytes(dmrHandle h) {
  if (!h) return 0;
  return zview(h).es * (size_t)h->ld * (size_t)h->k;
}

int dmrSubspaceToDevice(dmrHandle h, void *dst, hipStream_t stream) {
  if (!h || !dst) return set_err(DMR_ERR_ARG, "null argument");
  if (h->k == 0) return DMR_OK;
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipMemcpyAsync(dst, zview(h).Z, dmrSubspaceBytes(h), hipMemcpyDeviceToDevice, stream));
  return DMR_OK;
}

int dmrSubspaceFromDevice(dmrHandle h, const void *src, hipStream_t stream) {
  if (!h || !src) return set_err(DMR_ERR_ARG, "null argument");
  if (h->k == 0) return DMR_OK;
  const ZView z = zview(h);
  HIPCHK(hipSetDevice(h->device));
  HIPCHK(hipMemcpyAsync(z.Z, src, dmrSubspaceBytes(h), hipMemcpyDeviceToDevice, stream));
  *z.has = true;
  return DMR_OK;
}

int dmrGetStagnationCount(dmrHandle h) { return h ? h->kryFlagCount : 0; }

int dmrSetStagnationCount(dmrHandle h, int count) {
  if (!h || count < 0) return set_err(DMR_ERR_ARG, "invalid stagnation count");
  h->kryFlagCount = count;
  return DMR_OK;
}

int dmrGetStatsH(dmrHandle h, dmrStats *st, double *residHist, int cap) {
  if (!h) return -1;
  if (st) *st = h->st;
  const int cnt = (int)h->hist.size();
  if (residHist) {
    for (int i = 0; i < std::min(cap, cnt); ++i) residHist[i] = h->hist[i];
  }
  return cnt;
}

int dmrSetRefinement(dmrHandle h, double innerTol, int maxOuter) {
  if (!h || !(innerTol > 0.0) || maxOuter < 1) return set_err(DMR_ERR_ARG, "invalid settings");
  h->innerTol = innerTol;
  h->maxOuter = maxOuter;
  return DMR_OK;
}

int dmrSetProfiling(dmrHandle h, int on) {
  if (!h) return set_err(DMR_ERR_ARG, "null handle");
  h->prof.on = on != 0;
  for (int i = 0; i < DMR_NPHASES; ++i) {
    h->prof.ms[i] = 0.0;
    h->prof.by[i] = 0.0;
    h->prof.cnt[i] = 0;
  }
  h->prof.used = 0;
  return DMR_OK;
}

int dmrGetProfile(dmrHandle h, double *ms, double *bytes, long long *launches) {
  if (!h) return set_err(DMR_ERR_ARG, "null handle");
  for (int i = 0; i < DMR_NPHASES; ++i) {
    if (ms) ms[i] = h->prof.ms[i];
    if (bytes) bytes[i] = h->prof.by[i];
    if (launches) launches[i] = h->prof.cnt[i];
  }
  return DMR_NPHASES;
}

// ---- reference global-state interface ----
void initDMR(int blocks, int threadsPerBlock, int *devList, int numDevs, int elements,
             int basisDim, int deflatDim, void (*linComp)(void)) {
  (void)blocks;
  (void)threadsPerBlock;
  clear_err();
  if (!devList || numDevs != 1) {
    set_err(DMR_ERR_ARG, "exactly one device is supported (numDevs == 1)");
    return;
  }
  if (!linComp) {
    set_err(DMR_ERR_ARG, "null operator callback");
    return;
  }
  if (g_ctx) {
    ctx_free(g_ctx);
    g_ctx = nullptr;
  }
  dmrContext *c = nullptr;
  if (create_ctx(&c, devList[0], elements, basisDim, deflatDim, DMR_PREC_FP64)) return;
  c->legacy = linComp;
  c->opKind = OP_LEGACY;
  g_ctx = c;
}

void initDMRx(int device, int elements, int basisDim, int deflatDim, dmrOperatorFn op,
              void *ctx) {
  clear_err();
  if (!op) {
    set_err(DMR_ERR_ARG, "null operator");
    return;
  }
  if (g_ctx) {
    ctx_free(g_ctx);
    g_ctx = nullptr;
  }
  dmrContext *c = nullptr;
  if (create_ctx(&c, device, elements, basisDim, deflatDim, DMR_PREC_FP64)) return;
  c->opFn = op;
  c->opCtx = ctx;
  c->opKind = OP_FN;
  g_ctx = c;
}

void freeDMR(void) {
  ctx_free(g_ctx);
  g_ctx = nullptr;
}

double solverDMR(int mode, double *extSrcVec, double *extImgVec, hipStream_t extStream,
                 int extDevNum, double solTol, int matRank, int *numIts) {
  if (!g_ctx) {
    set_err(DMR_ERR_STATE, "solverDMR called before initDMR");
    return 1.0;
  }
  if (extDevNum != g_ctx->device) {
    set_err(DMR_ERR_ARG, "operator device must be the solver device");
    return 1.0;
  }
  return dmrSolve(g_ctx, mode, extSrcVec, extImgVec, extStream, solTol, matRank, numIts);
}

void impSVDMR(double *cMem) {
  if (!g_ctx) {
    set_err(DMR_ERR_STATE, "impSVDMR called before initDMR");
    return;
  }
  dmrImportZ(g_ctx, cMem);
}

void expSVDMR(double *cMem) {
  if (!g_ctx) {
    set_err(DMR_ERR_STATE, "expSVDMR called before initDMR");
    return;
  }
  dmrExportZ(g_ctx, cMem);
}

int dmrGetStats(dmrStats *st, double *residHist, int cap) {
  if (!g_ctx) return -1;
  return dmrGetStatsH(g_ctx, st, residHist, cap);
}

dmrHandle dmrGlobalHandle(void) { return g_ctx; }

}  // extern "C"
