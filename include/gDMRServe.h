/*
 * gDMRServe.h -- request-serving loop over libgDMR.so (no counterpart in the reference, whose
 * only driver solves one sequence of systems in a loop, progDMR.c:166-174).
 *
 * A server owns `slots` worker threads, each with its own HIP stream.  Callers register operator
 * families (a CSR sparsity pattern plus m, k and precision) and submit requests (the family's
 * coefficient values for this request and 1..n right-hand sides).  Each family owns one solver
 * handle whose recycled deflation subspace stays resident in HBM from request to request, so
 * recycling carries across the family's stream of requests (the serving form of
 * impSVDMR/expSVDMR, gDMR.cu:431-477).
 *
 * Scheduling: continuous, first-come first-served over families.  A free slot takes the oldest
 * queued request whose family is not running; a family runs one request at a time in submission
 * order, and requests of different families overlap on the device.  A family's handle is created
 * when the device has its footprint (dmrFootprint) plus hbmReserve free; while it does not, the
 * least recently used idle family is evicted (its subspace and stagnation counter move to host
 * memory and come back, bit for bit, when it next runs).  A family's results therefore depend on
 * its own request sequence only: they are bitwise the same served alone or in any mix (block
 * families batch queued requests into one launch; see dmrServerAddBlockFamily).
 *
 * All device pointers stay owned by the caller and must stay valid until the request's result
 * has been returned.  The server's streams do not wait on the caller's: a request's values,
 * right-hand sides and initial guesses must be complete on the device when it is submitted, and
 * its solutions are complete when its result is returned.  Every entry point is thread-safe.
 */
#ifndef GDMR_SERVE_H
#define GDMR_SERVE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dmrServer *dmrServerHandle;

typedef struct {
  long long id;        /* request id returned by dmrServerSubmit                 */
  int family;          /* operator family                                        */
  int nrhs;            /* right-hand sides solved                                */
  int status;          /* 0, or the first DMR_ERR_* code of its solves           */
  int cycles;          /* restart cycles summed over its right-hand sides        */
  int slot;            /* worker slot that ran it                                */
  int recycled;        /* 1 if it started from the family's resident subspace    */
  double maxRelRes;    /* largest relative residual returned by its solves       */
  double submitTime;   /* server clock (s) at submission                         */
  double startTime;    /* server clock (s) when a slot admitted it               */
  double endTime;      /* server clock (s) at completion                         */
  int batch;           /* right-hand sides in the launch that solved it (>= nrhs) */
} dmrRequestResult;

/* slots: worker threads (concurrent requests).  hbmReserve: bytes of HBM kept free. */
int dmrServerCreate(dmrServerHandle *s, int device, int slots, size_t hbmReserve);
/* rowPtr / colInd: device CSR pattern (int32, n+1 / nnz), shared by all requests of the family. */
int dmrServerAddFamily(dmrServerHandle s, int n, int nnz, const int *rowPtr, const int *colInd,
                       int basisDim, int deflatDim, int precision, int *family);
/* A block family: its requests run on a block DMR handle (dmrBlockCreate: FP32 block Krylov
 * space inside FP64 refinement, recycled subspace of k <= 32 vectors) and the requests of the
 * family that are queued when a slot takes it are solved together as one block of up to maxRhs
 * right-hand sides (continuous batching: requests that arrive while the family runs join its
 * next launch).  A batched result depends on which requests shared its launch, so it agrees with
 * the same request served alone to the solve tolerance, not bit for bit. */
int dmrServerAddBlockFamily(dmrServerHandle s, int n, int nnz, const int *rowPtr,
                            const int *colInd, int blockSteps, int deflatDim, int maxRhs,
                            int *family);
/* Basis storage of a block family (dmrBlockSetPrecision: DMR_PREC_FP32, the default, or
 * DMR_PREC_BF16), before its first request.  No counterpart in the reference. */
int dmrServerSetBlockPrecision(dmrServerHandle s, int family, int precision);
/* Refinement settings of a mixed-precision family (dmrSetRefinement), before its first request. */
int dmrServerSetRefinement(dmrServerHandle s, int family, double innerTol, int maxOuter);
/* val: device values (nnz doubles) of this request's operator; b[i], x[i] (i < nrhs): device
 * vectors of n doubles (x[i] is x0 on entry, the solution on completion).  matRank (>= 1)
 * bounds the restart cycles: of each right-hand side's solve for a plain family (the matRank of
 * dmrSolve), of the whole launch for a block family (the maxCycles of dmrBlockSolve, summed over
 * its refinement steps, at most 100000).  A request that spends it completes with
 * DMR_ERR_NOCONV. */
int dmrServerSubmit(dmrServerHandle s, int family, const double *val, int nrhs,
                    const double *const *b, double *const *x, double solTol, int matRank,
                    long long *id);
/* Completed requests (up to cap) without waiting; returns the number written. */
int dmrServerPoll(dmrServerHandle s, dmrRequestResult *out, int cap);
/* Waits up to timeoutSec for at least one completion; returns the number written. */
int dmrServerWait(dmrServerHandle s, dmrRequestResult *out, int cap, double timeoutSec);
/* Blocks until every submitted request has completed (results stay pollable). */
int dmrServerDrain(dmrServerHandle s);
/* Seconds since the server was created (the clock of dmrRequestResult). */
double dmrServerNow(dmrServerHandle s);
/* Requests queued and running, solver handles alive, free HBM bytes. */
int dmrServerStats(dmrServerHandle s, int *queued, int *running, int *handles, size_t *hbmFree);
/* Runs what is still queued, stops the workers and frees every handle. */
int dmrServerDestroy(dmrServerHandle s);

#ifdef __cplusplus
}
#endif

#endif /* GDMR_SERVE_H */
