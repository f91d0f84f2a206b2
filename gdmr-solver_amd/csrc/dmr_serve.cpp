// Serving loop: worker slots with their own HIP streams, one solver handle per operator family
// (its recycled subspace stays resident in HBM between requests), admission by free HBM with
// eviction of idle families, per-request timing.
//
// Scheduling rule: a family runs one request at a time, in submission order; requests of
// different families run concurrently in different slots.  A family's recycled state (the
// deflation subspace Z and the persistent stagnation counter) is therefore handed along the
// family's own request sequence only, so each result is a function of that sequence and not of
// what else is being served.  Together with the fixed-order device reductions this makes results
// bitwise the same whether a family is served alone or in a mix.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "gDMR.h"
#include "gDMRServe.h"

namespace {

// Largest restart-cycle budget a request may carry (its matRank).
constexpr int kMaxRequestCycles = 100000;

struct Family {
  int n = 0, nnz = 0, m = 0, k = 0, prec = DMR_PREC_FP64;
  // Block family: block handle, block steps, widest block, staging of a batched launch.
  bool block = false;
  int mb = 0, sMax = 0;
  dmrBlockHandle bh = nullptr;
  double *stage = nullptr;  // 2 * n * sMax doubles: B then X
  const int *rowPtr = nullptr;
  const int *colInd = nullptr;
  size_t footprint = 0;
  double innerTol = 0.0;  // 0: the handle's default
  int maxOuter = 0;
  // Guarded by the server mutex.
  bool busy = false;         // a slot is running one of its requests, or evicting its handle
  dmrHandle h = nullptr;     // resident handle (holds Z in HBM), or null
  double lastUse = 0.0;
  // Recycled state of an evicted family (host copy of Z, stagnation counter).
  std::vector<double> hostZ;
  bool hostHasZ = false;
  int flagCount = 0;
};

struct Request {
  long long id = 0;
  int family = 0;
  const double *val = nullptr;
  std::vector<const double *> b;
  std::vector<double *> x;
  double tol = 0.0;
  int matRank = 0;
  double submitTime = 0.0;
};

}  // namespace

struct dmrServer {
  int device = 0;
  size_t reserve = 0;
  std::chrono::steady_clock::time_point t0;
  std::vector<std::unique_ptr<Family>> families;
  std::mutex mu;                   // everything below, and the Family fields marked above
  std::condition_variable cvWork;  // work queued / a family became free / stop
  std::condition_variable cvDone;  // a request completed / memory released
  std::deque<Request> queue;
  std::vector<dmrRequestResult> results;
  long long nextId = 1;
  int running = 0;
  long long submitted = 0, completed = 0;
  size_t pendingBytes = 0;  // admitted allocations not yet visible in hipMemGetInfo
  int handles = 0;
  bool stop = false;
  std::vector<std::thread> workers;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

size_t free_hbm(int device) {
  size_t fr = 0, tot = 0;
  hipSetDevice(device);
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
  return fr;
}

bool resident(const Family *F) { return F->block ? F->bh != nullptr : F->h != nullptr; }

void release_block(Family *F) {
  if (F->bh) dmrBlockDestroy(F->bh);
  F->bh = nullptr;
  if (F->stage) hipFree(F->stage);
  F->stage = nullptr;
}

// Releases the handle of an idle family, keeping its recycled state on the host.  Called with the
// family marked busy by the caller (so no slot picks it meanwhile) and the server mutex released.
void evict(dmrServer *S, Family *F) {
  if (F->block) {
    if (F->k > 0 && dmrBlockHasZ(F->bh)) {
      F->hostZ.resize((size_t)F->n * F->k);
      F->hostHasZ = dmrBlockExportZ(F->bh, F->hostZ.data()) == DMR_OK;
    } else {
      F->hostHasZ = false;
    }
    release_block(F);
    std::lock_guard<std::mutex> lk(S->mu);
    --S->handles;
    return;
  }
  if (F->k > 0 && dmrHasZ(F->h)) {
    F->hostZ.resize((size_t)F->n * F->k);
    F->hostHasZ = dmrExportZ(F->h, F->hostZ.data()) == DMR_OK;
  } else {
    F->hostHasZ = false;
  }
  F->flagCount = dmrGetStagnationCount(F->h);
  dmrDestroy(F->h);
  F->h = nullptr;
  std::lock_guard<std::mutex> lk(S->mu);
  --S->handles;
}

// Creates the family's handle once the device has room for it (footprint + reserve free, counting
// admissions in flight), evicting the least recently used idle family while it does not.
int admit(dmrServer *S, Family *F, const double *val) {
  {
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
      const size_t fr = free_hbm(S->device);
      const size_t avail = fr > S->pendingBytes ? fr - S->pendingBytes : 0;
      if (avail >= F->footprint + S->reserve) {
        S->pendingBytes += F->footprint;
        break;
      }
      Family *victim = nullptr;
      for (auto &G : S->families)
        if (G.get() != F && !G->busy && resident(G.get()) &&
            (!victim || G->lastUse < victim->lastUse))
          victim = G.get();
      if (victim) {
        victim->busy = true;
        lk.unlock();
        evict(S, victim);
        lk.lock();
        victim->busy = false;
        S->cvWork.notify_all();
        continue;
      }
      // Nothing idle to evict: wait for running requests to release memory.  A request that
      // cannot fit even on an otherwise empty device fails instead of waiting forever.
      if (S->handles == 0 && S->pendingBytes == 0) return DMR_ERR_ALLOC;
      S->cvDone.wait_for(lk, std::chrono::milliseconds(20));
    }
  }
  int rc = DMR_OK;
  if (F->block) {
    rc = dmrBlockCreate(&F->bh, S->device, F->n, F->mb, F->k, F->sMax);
    if (rc == DMR_OK)
      rc = dmrBlockSetOperatorCsr(F->bh, F->nnz, F->rowPtr, F->colInd, val, nullptr);
    if (rc == DMR_OK &&
        hipMalloc(&F->stage, sizeof(double) * 2 * (size_t)F->n * F->sMax) != hipSuccess)
      rc = DMR_ERR_ALLOC;
    if (rc == DMR_OK) rc = dmrBlockSetPrecision(F->bh, F->prec);
    if (rc == DMR_OK && F->hostHasZ) rc = dmrBlockImportZ(F->bh, F->hostZ.data());
    if (rc == DMR_OK && (F->innerTol > 0.0 || F->maxOuter > 0))
      rc = dmrBlockSetRefinement(F->bh, F->innerTol > 0.0 ? F->innerTol : 1e-5,
                                 F->maxOuter > 0 ? F->maxOuter : 30);
    if (rc != DMR_OK) release_block(F);
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->pendingBytes -= F->footprint;
      if (F->bh) ++S->handles;
    }
    S->cvDone.notify_all();
    if (F->bh) {
      F->hostZ.clear();
      F->hostZ.shrink_to_fit();
      F->hostHasZ = false;
    }
    return rc;
  }
  rc = dmrCreate(&F->h, S->device, F->n, F->m, F->k, F->prec);
  if (rc == DMR_OK) rc = dmrSetOperatorCsr(F->h, F->nnz, F->rowPtr, F->colInd, val, nullptr);
  if (rc == DMR_OK && F->hostHasZ) rc = dmrImportZ(F->h, F->hostZ.data());
  if (rc == DMR_OK) rc = dmrSetStagnationCount(F->h, F->flagCount);
  if (rc == DMR_OK && (F->innerTol > 0.0 || F->maxOuter > 0))
    rc = dmrSetRefinement(F->h, F->innerTol > 0.0 ? F->innerTol : 1e-5,
                          F->maxOuter > 0 ? F->maxOuter : 60);
  if (rc != DMR_OK && F->h) {
    dmrDestroy(F->h);
    F->h = nullptr;
  }
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->pendingBytes -= F->footprint;
    if (F->h) ++S->handles;
  }
  S->cvDone.notify_all();
  if (F->h) {
    F->hostZ.clear();
    F->hostZ.shrink_to_fit();
    F->hostHasZ = false;
  }
  return rc;
}

// A worker slot: takes the oldest queued request whose family is free, runs its right-hand sides
// in order on the slot's stream, and frees the family for its next request.
// Solves one launch of a block family: the right-hand sides of every request in `batch` as one
// block (gathered into the family's staging block, solved, scattered back).
void run_block(Family *F, std::vector<Request> &batch, hipStream_t stream,
               std::vector<dmrRequestResult> &res) {
  int total = 0;
  for (auto &r : batch) total += (int)r.b.size();
  const size_t nb = sizeof(double) * (size_t)F->n;
  double *B = F->stage, *X = F->stage + (size_t)F->n * F->sMax;
  int rc = dmrBlockSetOperatorValues(F->bh, batch[0].val, nullptr);
  int col = 0;
  for (auto &r : batch)
    for (size_t i = 0; i < r.b.size() && rc == DMR_OK; ++i, ++col) {
      if (hipMemcpyAsync(B + (size_t)col * F->n, r.b[i], nb, hipMemcpyDeviceToDevice, stream) !=
              hipSuccess ||
          hipMemcpyAsync(X + (size_t)col * F->n, r.x[i], nb, hipMemcpyDeviceToDevice, stream) !=
              hipSuccess)
        rc = DMR_ERR_HIP;
    }
  const bool have = dmrBlockHasZ(F->bh) != 0;
  std::vector<double> rel(total, 0.0);
  if (rc == DMR_OK) {
    // The launch's cycle budget: the first request's (at most kMaxRequestCycles), summed over
    // the refinement steps.
    const int maxCycles = std::min(batch[0].matRank, kMaxRequestCycles);
    rc = dmrBlockSolve(F->bh, have ? 1 : 0, total, X, F->n, B, F->n, stream, batch[0].tol,
                       maxCycles, rel.data());
  }
  dmrStats st{};
  dmrBlockGetStats(F->bh, &st, nullptr, 0);
  col = 0;
  for (size_t q = 0; q < batch.size(); ++q) {
    auto &r = batch[q];
    res[q].recycled = have ? 1 : 0;
    res[q].cycles = st.cycles;
    res[q].batch = total;
    for (size_t i = 0; i < r.b.size(); ++i, ++col) {
      res[q].maxRelRes = std::max(res[q].maxRelRes, rel[col]);
      if (hipMemcpyAsync(r.x[i], X + (size_t)col * F->n, nb, hipMemcpyDeviceToDevice, stream) !=
              hipSuccess &&
          rc == DMR_OK)
        rc = DMR_ERR_HIP;
    }
  }
  if (hipStreamSynchronize(stream) != hipSuccess && rc == DMR_OK) rc = DMR_ERR_HIP;
  for (auto &r : res) r.status = rc;
}

// A worker slot: takes the oldest queued request whose family is free (for a block family also
// the later queued requests it can batch with), runs it on the slot's stream, and frees the
// family for its next launch.
void worker(dmrServer *S, int slot) {
  hipSetDevice(S->device);
  hipStream_t stream = nullptr;
  hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
  for (;;) {
    std::vector<Request> batch;
    Family *F = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      auto ready = [&] {
        for (auto it = S->queue.begin(); it != S->queue.end(); ++it)
          if (!S->families[it->family]->busy) return it;
        return S->queue.end();
      };
      S->cvWork.wait(lk, [&] { return (S->stop && S->queue.empty()) || ready() != S->queue.end(); });
      if (S->stop && S->queue.empty()) break;
      auto it = ready();
      F = S->families[it->family].get();
      batch.push_back(std::move(*it));
      it = S->queue.erase(it);
      if (F->block) {
        // Later requests of the family with the same operator values and tolerance join the
        // launch while the block has room (in submission order).
        int total = (int)batch[0].b.size();
        while (it != S->queue.end()) {
          const Request &r = *it;
          if (r.family == batch[0].family && r.val == batch[0].val && r.tol == batch[0].tol &&
              total + (int)r.b.size() <= F->sMax) {
            total += (int)r.b.size();
            batch.push_back(std::move(*it));
            it = S->queue.erase(it);
          } else {
            ++it;
          }
        }
      }
      F->busy = true;
      S->running += (int)batch.size();
    }
    std::vector<dmrRequestResult> res(batch.size());
    const double t0 = S->now();
    for (size_t q = 0; q < batch.size(); ++q) {
      res[q].id = batch[q].id;
      res[q].family = batch[q].family;
      res[q].nrhs = (int)batch[q].b.size();
      res[q].slot = slot;
      res[q].submitTime = batch[q].submitTime;
      res[q].startTime = t0;
      res[q].batch = res[q].nrhs;
    }
    int rc = resident(F) ? DMR_OK : admit(S, F, batch[0].val);
    if (rc != DMR_OK) {
      for (auto &r : res) r.status = rc;
    } else if (F->block) {
      run_block(F, batch, stream, res);
    } else {
      Request &req = batch[0];
      dmrRequestResult &r0 = res[0];
      rc = dmrSetOperatorValues(F->h, req.val, nullptr);
      if (rc == DMR_OK) {
        const bool have = dmrHasZ(F->h) != 0;
        r0.recycled = have ? 1 : 0;
        for (size_t i = 0; i < req.b.size() && rc == DMR_OK; ++i) {
          int its = 0;
          const int mode = (have || i > 0) ? 1 : 0;
          const double r = dmrSolve(F->h, mode, req.x[i], req.b[i], stream, req.tol, req.matRank,
                                    &its);
          rc = dmrGetLastError();
          dmrStats st{};
          dmrGetStatsH(F->h, &st, nullptr, 0);
          r0.cycles += st.cycles;
          r0.maxRelRes = std::max(r0.maxRelRes, r);
        }
        // The next request of this family may run on another slot's stream.
        if (hipStreamSynchronize(stream) != hipSuccess && rc == DMR_OK) rc = DMR_ERR_HIP;
      }
      r0.status = rc;
    }
    const double t1 = S->now();
    {
      std::lock_guard<std::mutex> lk(S->mu);
      F->lastUse = t1;
      F->busy = false;
      for (auto &r : res) {
        r.endTime = t1;
        S->results.push_back(r);
      }
      S->running -= (int)batch.size();
      S->completed += (long long)batch.size();
    }
    S->cvDone.notify_all();
    S->cvWork.notify_all();
  }
  hipStreamDestroy(stream);
}

}  // namespace

extern "C" {

int dmrServerCreate(dmrServerHandle *s, int device, int slots, size_t hbmReserve) {
  if (!s || slots < 1 || slots > 64) return DMR_ERR_ARG;
  *s = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return DMR_ERR_HIP;
  dmrServer *S = new dmrServer();
  S->device = device;
  S->reserve = hbmReserve;
  S->t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < slots; ++i) S->workers.emplace_back(worker, S, i);
  *s = S;
  return DMR_OK;
}

int dmrServerAddFamily(dmrServerHandle s, int n, int nnz, const int *rowPtr, const int *colInd,
                       int basisDim, int deflatDim, int precision, int *family) {
  if (!s || !rowPtr || !colInd || n < 1 || nnz < 0 || !family) return DMR_ERR_ARG;
  if (basisDim < 1 || deflatDim < 0 || deflatDim >= basisDim) return DMR_ERR_ARG;
  if (precision != DMR_PREC_FP64 && precision != DMR_PREC_FP32 && precision != DMR_PREC_BF16)
    return DMR_ERR_ARG;
  auto F = std::make_unique<Family>();
  F->n = n;
  F->nnz = nnz;
  F->rowPtr = rowPtr;
  F->colInd = colInd;
  F->m = basisDim;
  F->k = deflatDim;
  F->prec = precision;
  // Handle allocation plus its CSR block partition, FP32 values and reduction workspace.
  F->footprint = dmrFootprint(n, basisDim, deflatDim, precision) + (size_t)n * 4 / 64 +
                 (precision != DMR_PREC_FP64 ? (size_t)nnz * 4 : 0) + (8u << 20);
  std::lock_guard<std::mutex> lk(s->mu);
  s->families.push_back(std::move(F));
  *family = (int)s->families.size() - 1;
  return DMR_OK;
}

int dmrServerAddBlockFamily(dmrServerHandle s, int n, int nnz, const int *rowPtr,
                            const int *colInd, int blockSteps, int deflatDim, int maxRhs,
                            int *family) {
  if (!s || !rowPtr || !colInd || n < 1 || nnz < 0 || !family) return DMR_ERR_ARG;
  if (blockSteps < 1 || deflatDim < 0 || deflatDim > 32 || maxRhs < 1 || maxRhs > 32 ||
      deflatDim + (blockSteps + 1) * maxRhs > 384)
    return DMR_ERR_ARG;
  auto F = std::make_unique<Family>();
  F->block = true;
  F->n = n;
  F->nnz = nnz;
  F->rowPtr = rowPtr;
  F->colInd = colInd;
  F->mb = blockSteps;
  F->k = deflatDim;
  F->sMax = maxRhs;
  F->prec = DMR_PREC_FP32;
  F->footprint = dmrBlockFootprint(n, blockSteps, deflatDim, maxRhs) +
                 sizeof(double) * 2 * (size_t)n * maxRhs + (size_t)nnz * 4 + (8u << 20);
  std::lock_guard<std::mutex> lk(s->mu);
  s->families.push_back(std::move(F));
  *family = (int)s->families.size() - 1;
  return DMR_OK;
}

int dmrServerSetRefinement(dmrServerHandle s, int family, double innerTol, int maxOuter) {
  if (!s || !(innerTol > 0.0 && innerTol < 1.0) || maxOuter < 1) return DMR_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  if (family < 0 || family >= (int)s->families.size()) return DMR_ERR_ARG;
  Family *F = s->families[family].get();
  if (F->busy || resident(F)) return DMR_ERR_STATE;  // fixed once the family has run
  F->innerTol = innerTol;
  F->maxOuter = maxOuter;
  return DMR_OK;
}

int dmrServerSetBlockPrecision(dmrServerHandle s, int family, int precision) {
  if (!s || (precision != DMR_PREC_FP32 && precision != DMR_PREC_BF16)) return DMR_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  if (family < 0 || family >= (int)s->families.size()) return DMR_ERR_ARG;
  Family *F = s->families[family].get();
  if (!F->block) return DMR_ERR_ARG;
  if (F->busy || resident(F)) return DMR_ERR_STATE;  // fixed once the family has run
  F->prec = precision;
  return DMR_OK;
}

int dmrServerSubmit(dmrServerHandle s, int family, const double *val, int nrhs,
                    const double *const *b, double *const *x, double solTol, int matRank,
                    long long *id) {
  if (!s || !val || nrhs < 1 || !b || !x || matRank < 1 || !(solTol >= 0.0)) return DMR_ERR_ARG;
  for (int i = 0; i < nrhs; ++i)
    if (!b[i] || !x[i]) return DMR_ERR_ARG;
  Request r;
  r.family = family;
  r.val = val;
  r.b.assign(b, b + nrhs);
  r.x.assign(x, x + nrhs);
  r.tol = solTol;
  r.matRank = matRank;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (family < 0 || family >= (int)s->families.size()) return DMR_ERR_ARG;
    if (s->stop) return DMR_ERR_STATE;
    r.id = s->nextId++;
    r.submitTime = s->now();
    if (id) *id = r.id;
    s->queue.push_back(std::move(r));
    ++s->submitted;
  }
  s->cvWork.notify_all();
  return DMR_OK;
}

int dmrServerPoll(dmrServerHandle s, dmrRequestResult *out, int cap) {
  if (!s || !out || cap < 1) return 0;
  std::lock_guard<std::mutex> lk(s->mu);
  const int cnt = std::min(cap, (int)s->results.size());
  std::copy(s->results.begin(), s->results.begin() + cnt, out);
  s->results.erase(s->results.begin(), s->results.begin() + cnt);
  return cnt;
}

int dmrServerWait(dmrServerHandle s, dmrRequestResult *out, int cap, double timeoutSec) {
  if (!s || !out || cap < 1) return 0;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cvDone.wait_for(lk, std::chrono::duration<double>(std::max(timeoutSec, 0.0)),
                       [&] { return !s->results.empty(); });
  }
  return dmrServerPoll(s, out, cap);
}

int dmrServerDrain(dmrServerHandle s) {
  if (!s) return DMR_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cvDone.wait(lk, [&] { return s->completed == s->submitted; });
  return DMR_OK;
}

double dmrServerNow(dmrServerHandle s) { return s ? s->now() : 0.0; }

int dmrServerStats(dmrServerHandle s, int *queued, int *running, int *handles, size_t *hbmFree) {
  if (!s) return DMR_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (queued) *queued = (int)s->queue.size();
    if (running) *running = s->running;
    if (handles) *handles = s->handles;
  }
  if (hbmFree) *hbmFree = free_hbm(s->device);
  return DMR_OK;
}

int dmrServerDestroy(dmrServerHandle s) {
  if (!s) return DMR_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cvWork.notify_all();
  s->cvDone.notify_all();
  for (auto &t : s->workers) t.join();
  for (auto &F : s->families) {
    if (F->h) dmrDestroy(F->h);
    release_block(F.get());
  }
  delete s;
  return DMR_OK;
}

}  // extern "C"
