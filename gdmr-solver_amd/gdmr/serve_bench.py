"""Serving load generator over the gDMRServe.h loop: Poisson arrivals, p50/p99 latency.

Workload (serving): three operator families
  * 2-D convection-diffusion, 5-point, 512^2  (n = 262,144)
  * 3-D convection-diffusion, 7-point, 64^3   (n = 262,144)
  * 3-D convection-diffusion, 7-point, 160^3  (n = 4,096,000)
each with a recycled deflation subspace of k = 20 kept resident in HBM, m = 80 (block
families: 20 block steps per cycle), tol 1e-10.  Long cycles pay on the 2-D family, whose
operator is the worst conditioned: measured with FP32 inner solves, 136 restart cycles per
solve at m = 40 against 21.5 at m = 80.
A request picks a family uniformly, carries that family's coefficients with per-cell U(+-1%)
jitter (one of `--draws` pre-generated draws per family, resident on the device) and 1..8
right-hand sides (uniform; N(0,1) vectors from a resident pool of 8 per family).

Phases: one warm-up request per family (handle creation, code objects); a closed-loop burst that
measures capacity (requests/s with the queue never empty); then an open loop of `--requests`
Poisson arrivals at rate lambda (default 0.8 x capacity).  Latency = completion - submission on
the server clock.  Solutions of a sample of requests are checked on the host in FP64.

Run: python -m gdmr.serve_bench [--requests N] [--rate R] [--slots S] [--prec fp64|mixed] [--block]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

import gdmr

POOL_RHS = 8


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def _families(small: bool):
    if small:
        return [("cd2d-128", lambda t: gdmr.conv_diff_2d(128, jitter=0.01, seed=41, stream=t)),
                ("cd3d-24", lambda t: gdmr.conv_diff_3d(24, jitter=0.01, seed=42, stream=t)),
                ("cd3d-40", lambda t: gdmr.conv_diff_3d(40, jitter=0.01, seed=43, stream=t))]
    return [("cd2d-512", lambda t: gdmr.conv_diff_2d(512, jitter=0.01, seed=41, stream=t)),
            ("cd3d-64", lambda t: gdmr.conv_diff_3d(64, jitter=0.01, seed=42, stream=t)),
            ("cd3d-160", lambda t: gdmr.conv_diff_3d(160, jitter=0.01, seed=43, stream=t))]


class Load:
    def __init__(self, torch, args):
        self.torch, self.args = torch, args
        self.fams = []
        t0 = time.time()
        fams = list(enumerate(_families(args.small)))
        if args.families:
            keep = args.families.split(",")
            fams = [f for f in fams if f[1][0] in keep]
        for fi, (name, gen) in fams:
            draws = [gen(j) for j in range(args.draws)]
            A0 = draws[0]
            dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
            B = [gdmr.rand_normal(500 + fi, j, A0.n) for j in range(POOL_RHS)]
            self.fams.append(dict(name=name, n=A0.n, nnz=A0.nnz, host=draws,
                                  rp=dev(A0.rowptr), ci=dev(A0.col),
                                  vals=[dev(A.val) for A in draws], B=B,
                                  b=[dev(b) for b in B]))
        torch.cuda.synchronize()
        _log(f"families {[(f['name'], f['n'], f['nnz']) for f in self.fams]} "
             f"generated in {time.time() - t0:.1f}s")
        prec = gdmr.PREC_FP64 if args.prec == "fp64" else gdmr.PREC_FP32
        self.srv = gdmr.Server(slots=args.slots)
        self.fam_id = []
        for f in self.fams:
            if args.block:
                fid = self.srv.add_block_family(f["n"], f["nnz"], f["rp"].data_ptr(),
                                                f["ci"].data_ptr(), args.block_steps, args.k,
                                                args.block_rhs, keep=(f["rp"], f["ci"]))
                if args.block_prec == "bf16":
                    self.srv.set_block_precision(fid, gdmr.PREC_BF16)
                    self.srv.set_refinement(fid, 3e-2 if args.inner_tol == 1e-4 else
                                            args.inner_tol, 60)
                else:
                    self.srv.set_refinement(fid, args.inner_tol, 30)
            else:
                fid = self.srv.add_family(f["n"], f["nnz"], f["rp"].data_ptr(),
                                          f["ci"].data_ptr(), args.m, args.k, prec,
                                          keep=(f["rp"], f["ci"]))
                if prec != gdmr.PREC_FP64:
                    self.srv.set_refinement(fid, args.inner_tol, 60)
            self.fam_id.append(fid)
        self.rng = np.random.default_rng(args.seed)
        self.live = {}  # id -> (spec, xs)
        self.done = []  # (result, spec)
        self.checked = []

    def spec(self, fam=None):
        r = self.rng
        fi = int(r.integers(len(self.fams))) if fam is None else fam
        return dict(fam=fi, draw=int(r.integers(self.args.draws)),
                    rhs=[int(i) for i in r.permutation(POOL_RHS)[:int(r.integers(1, 9))]])

    def submit(self, sp):
        torch, f = self.torch, self.fams[sp["fam"]]
        xs = [torch.zeros(f["n"], dtype=torch.float64, device="cuda") for _ in sp["rhs"]]
        # The server's streams do not wait on torch's: the zero fill must be complete.
        torch.cuda.current_stream().synchronize()
        rid = self.srv.submit(self.fam_id[sp["fam"]], f["vals"][sp["draw"]].data_ptr(),
                              [f["b"][i].data_ptr() for i in sp["rhs"]],
                              [x.data_ptr() for x in xs], self.args.tol, self.args.max_cycles)
        self.live[rid] = (sp, xs)
        return rid

    def collect(self, results, check_budget):
        for res in results:
            sp, xs = self.live.pop(res.id)
            self.done.append((res, sp))
            _log(f"request {res.id}: family {self.fams[sp['fam']]['name']}, {res.nrhs} RHS in a "
                 f"launch of {res.batch}, {res.cycles} cycles, status {res.status}, worst "
                 f"residual {res.maxRelRes:.2e}, latency {res.endTime - res.submitTime:.2f}s")
            fcount = sum(1 for c in self.checked if c[0] == sp["fam"])
            if check_budget and fcount < check_budget:
                f = self.fams[sp["fam"]]
                A = f["host"][sp["draw"]]
                worst = 0.0
                for i, x in zip(sp["rhs"], xs):
                    xh = x.cpu().numpy()
                    r = float(np.linalg.norm(f["B"][i] - A.matvec(xh)) / np.linalg.norm(f["B"][i]))
                    worst = r if (r != r or r > worst) else worst  # a NaN stays NaN
                _log(f"request {res.id}: FP64 check on the host, worst relative residual {worst:.2e}")
                self.checked.append((sp["fam"], worst))

    def run_until(self, n_done, check_budget=0, timeout=3600.0):
        t0 = time.time()
        while len(self.done) < n_done:
            self.collect(self.srv.wait(timeout=1.0), check_budget)
            if time.time() - t0 > timeout:
                raise RuntimeError("serving run timed out")


def _pct(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else None


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--requests", type=int, default=60, help="open-loop requests (2000: full)")
    ap.add_argument("--burst", type=int, default=24, help="closed-loop capacity requests")
    ap.add_argument("--rate", type=float, default=0.0, help="arrivals/s (0: 0.8 x capacity)")
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--m", type=int, default=80)
    ap.add_argument("--k", type=int, default=20)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--prec", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--inner-tol", type=float, default=1e-4)
    ap.add_argument("--max-cycles", type=int, default=4000,
                    help="restart-cycle budget of one launch (the request's matRank): a launch "
                         "that spends it returns DMR_ERR_NOCONV")
    ap.add_argument("--draws", type=int, default=4)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--check", type=int, default=2, help="requests per family checked in FP64")
    ap.add_argument("--small", action="store_true", help="small families (quick runs)")
    ap.add_argument("--families", default="",
                    help="comma list of the families to serve (default: all three)")
    ap.add_argument("--block", action="store_true",
                    help="block families: multi-RHS requests as one block solve, and queued "
                         "requests with the same operator batched into one launch")
    ap.add_argument("--block-steps", type=int, default=20)
    ap.add_argument("--block-rhs", type=int, default=16)
    ap.add_argument("--block-prec", choices=("fp32", "bf16"), default="fp32",
                    help="basis storage of block families (bf16: inner tolerance 3e-2 unless "
                         "--inner-tol is given)")
    args = ap.parse_args(argv)
    import torch
    assert torch.cuda.is_available(), "the serving benchmark needs a HIP device"
    L = Load(torch, args)
    srv = L.srv
    # Warm-up: one request per family.
    for fi in range(len(L.fams)):
        L.submit(L.spec(fam=fi))
    L.run_until(len(L.fams))
    L.done.clear()
    _log("warm-up done")
    # Closed loop: the whole burst queued at once.
    for _ in range(args.burst):
        L.submit(L.spec())
    L.run_until(args.burst)
    burst = [r for r, _ in L.done]
    cap_span = max(r.endTime for r in burst) - min(r.submitTime for r in burst)
    capacity = args.burst / cap_span
    cap_solves = sum(r.nrhs for r in burst) / cap_span
    _log(f"closed loop: {args.burst} requests in {cap_span:.2f}s -> {capacity:.3f} req/s, "
         f"{cap_solves:.2f} solves/s")
    L.done.clear()
    # Open loop: Poisson arrivals.
    lam = args.rate if args.rate > 0 else 0.8 * capacity
    arrivals = np.cumsum(L.rng.exponential(1.0 / lam, args.requests))
    t_start = time.perf_counter()
    nxt = 0
    last_note = 0
    while nxt < args.requests or len(L.done) < args.requests:
        now = time.perf_counter() - t_start
        while nxt < args.requests and arrivals[nxt] <= now:
            L.submit(L.spec())
            nxt += 1
        wait = (arrivals[nxt] - now) if nxt < args.requests else 1.0
        L.collect(srv.wait(timeout=max(min(wait, 1.0), 0.0)), args.check)
        if len(L.done) - last_note >= 10:
            last_note = len(L.done)
            _log(f"open loop: {len(L.done)}/{args.requests} done, queue {srv.stats()['queued']}")
    wall = time.perf_counter() - t_start
    res = [r for r, _ in L.done]
    mean_batch = float(np.mean([r.batch for r in res])) if res else 0.0
    lat = [1e3 * (r.endTime - r.submitTime) for r in res]
    svc = [1e3 * (r.endTime - r.startTime) for r in res]
    per_fam = {}
    for fi, f in enumerate(L.fams):
        rs = [r for r, sp in L.done if sp["fam"] == fi]
        lf = [1e3 * (r.endTime - r.submitTime) for r in rs]
        per_fam[f["name"]] = dict(requests=len(rs), solves=sum(r.nrhs for r in rs),
                                  p50_ms=_pct(lf, 50), p99_ms=_pct(lf, 99),
                                  cycles_per_solve=(sum(r.cycles for r in rs) /
                                                    max(1, sum(r.nrhs for r in rs))))
    status_ok = all(r.status == 0 for r in res)
    not_converged = sum(1 for r in res if r.status != 0)
    ws = [c[1] for c in L.checked]
    worst = (float("nan") if any(w != w for w in ws) else max(ws)) if ws else None
    line = {
        "metric": "serving requests/s at 1e-10 (FP64-checked sample), p50/p99 latency",
        "value": len(res) / wall,
        "unit": "requests/s",
        "n_gpus": 1,
        "requests": len(res),
        "solves_per_s": sum(r.nrhs for r in res) / wall,
        "arrival_rate": lam,
        "closed_loop_capacity": capacity,
        "closed_loop_solves_per_s": cap_solves,
        "p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99),
        "service_p50_ms": _pct(svc, 50), "service_p99_ms": _pct(svc, 99),
        "families": per_fam,
        "all_status_ok": status_ok,
        "requests_not_converged": not_converged,
        "max_cycles_per_launch": args.max_cycles,
        "mean_rhs_per_launch": mean_batch,
        "checked_requests": len(L.checked),
        "max_true_rel_residual": worst,
        "dtype": "f64" if (args.prec == "fp64" and not args.block) else "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": "serving: 3 families (2-D 512^2, 3-D 64^3, 3-D 160^3 conv-diff), "
                               f"Poisson arrivals, 1-8 RHS, "
                               + (f"block steps {args.block_steps} (<= {args.block_rhs} RHS), "
                                  if args.block else f"m={args.m} ")
                               + f"k={args.k} tol {args.tol:g}"
                               + (" [small families]" if args.small else ""),
                   "slots": args.slots, "prec": "block" if args.block else args.prec,
                   "basis": args.block_prec if args.block else
                   ("f64" if args.prec == "fp64" else "fp32")},
    }
    srv.close()
    print(json.dumps(line), flush=True)
    return line


if __name__ == "__main__":
    main()
