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

struct Family {
  int n = 0, nnz = 0, m = 0, k = 0, prec = DMR_PREC_FP64;
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

// Releases the handle of an idle family, keeping its recycled state on the host.  Called with the
// family marked busy by the caller (so no slot picks it meanwhile) and the server mutex released.
void evict(dmrServer *S, Family *F) {
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
        if (G.get() != F && !G->busy && G->h && (!victim || G->lastUse < victim->lastUse))
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
  int rc = dmrCreate(&F->h, S->device, F->n, F->m, F->k, F->prec);
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
void worker(dmrServer *S, int slot) {
  hipSetDevice(S->device);
  hipStream_t stream = nullptr;
  hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
  for (;;) {
    Request req;
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
      req = std::move(*it);
      S->queue.erase(it);
      F = S->families[req.family].get();
      F->busy = true;
      ++S->running;
    }
    dmrRequestResult res{};
    res.id = req.id;
    res.family = req.family;
    res.nrhs = (int)req.b.size();
    res.slot = slot;
    res.submitTime = req.submitTime;
    res.startTime = S->now();
    int rc = F->h ? DMR_OK : admit(S, F, req.val);
    if (rc == DMR_OK) rc = dmrSetOperatorValues(F->h, req.val, nullptr);
    if (rc == DMR_OK) {
      const bool have = dmrHasZ(F->h) != 0;
      res.recycled = have ? 1 : 0;
      for (size_t i = 0; i < req.b.size() && rc == DMR_OK; ++i) {
        int its = 0;
        const int mode = (have || i > 0) ? 1 : 0;
        const double r = dmrSolve(F->h, mode, req.x[i], req.b[i], stream, req.tol, req.matRank,
                                  &its);
        rc = dmrGetLastError();
        dmrStats st{};
        dmrGetStatsH(F->h, &st, nullptr, 0);
        res.cycles += st.cycles;
        res.maxRelRes = std::max(res.maxRelRes, r);
      }
      // The next request of this family may run on another slot's stream.
      if (hipStreamSynchronize(stream) != hipSuccess && rc == DMR_OK) rc = DMR_ERR_HIP;
    }
    res.status = rc;
    res.endTime = S->now();
    {
      std::lock_guard<std::mutex> lk(S->mu);
      F->lastUse = res.endTime;
      F->busy = false;
      S->results.push_back(res);
      --S->running;
      ++S->completed;
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
  if (precision != DMR_PREC_FP64 && precision != DMR_PREC_FP32) return DMR_ERR_NOTBUILT;
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

int dmrServerSetRefinement(dmrServerHandle s, int family, double innerTol, int maxOuter) {
  if (!s || !(innerTol > 0.0 && innerTol < 1.0) || maxOuter < 1) return DMR_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  if (family < 0 || family >= (int)s->families.size()) return DMR_ERR_ARG;
  Family *F = s->families[family].get();
  if (F->busy || F->h) return DMR_ERR_STATE;  // settings are fixed once the family has run
  F->innerTol = innerTol;
  F->maxOuter = maxOuter;
  return DMR_OK;
}

int dmrServerSubmit(dmrServerHandle s, int family, const double *val, int nrhs,
                    const double *const *b, double *const *x, double solTol, int matRank,
                    long long *id) {
  if (!s || !val || nrhs < 1 || !b || !x || matRank < 0 || !(solTol >= 0.0)) return DMR_ERR_ARG;
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
  for (auto &F : s->families)
    if (F->h) dmrDestroy(F->h);
  delete s;
  return DMR_OK;
}

}  // extern "C"
