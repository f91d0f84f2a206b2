"""Batched multi-right-hand-side benchmark: block DMR against sequential FP64 DMR.

Workload (batched): 3-D anisotropic diffusion, 7-point,
96^3 (n = 884,736), coefficients (1, 0.1, 0.01); 32 simultaneous N(0,1) right-hand sides sharing
one deflation subspace of k = 32; tol 1e-10 checked in FP64 per column.

Legs:
  * block  -- dmrBlockSolve: one block Krylov space for the 32 columns (block Gram-Schmidt as
              tall-skinny GEMMs on the matrix cores, FP32 or bf16 basis storage (--prec) and FP32
              accumulation) inside FP64 refinement; blockSteps = 10, so k + 11 * 32 = 384 basis
              columns.  Two blocks of 32 run back to back: the first from scratch, the second
              recycling the subspace.
  * fp64   -- the plain FP64 DMR (m = 40, k = 32), one right-hand side at a time with the
              subspace recycled (mode = i), on the first `--fp64-sample` columns.
Reported: solves/s of each leg, the speed-up, per-column true residuals, the block solution's
error against the FP64 one, per-phase milliseconds, GB/s and matrix-core TFLOP/s.

Run: python -m gdmr.batch_bench [--grid N] [--rhs S] [--fp64-sample K]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

import gdmr


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--grid", type=int, default=96)
    ap.add_argument("--rhs", type=int, default=32)
    ap.add_argument("--block-steps", type=int, default=10)
    ap.add_argument("--k", type=int, default=32)
    ap.add_argument("--m", type=int, default=40, help="basis of the FP64 single-vector leg")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--inner-tol", type=float, default=None,
                    help="inner block-solve tolerance (default 1e-5 for fp32, 3e-2 for bf16)")
    ap.add_argument("--prec", default="fp32", choices=("fp32", "bf16"),
                    help="block basis storage")
    ap.add_argument("--fp64-sample", type=int, default=12)
    ap.add_argument("--seed", type=int, default=96)
    ap.add_argument("--no-profile", action="store_true")
    return ap.parse_args(argv)


def run(args) -> dict:
    """Runs both legs; returns the result record (printed as one JSON line by main)."""
    import torch
    assert torch.cuda.is_available(), "the batched benchmark needs a HIP device"
    N, s, tol = args.grid, args.rhs, args.tol
    A = gdmr.aniso_3d(N, coef=(1.0, 0.1, 0.01))
    n = A.n
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    rp, ci, va = dev(A.rowptr), dev(A.col), dev(A.val)
    blocks = [np.stack([gdmr.rand_normal(args.seed, 1000 * b + j, n) for j in range(s)])
              for b in range(2)]
    Bd = [dev(B) for B in blocks]
    _log(f"aniso {N}^3 n={n} nnz={A.nnz}, {s} right-hand sides per block")

    # Warm-up on a throwaway handle (code objects, first touch).
    w = gdmr.BlockSolver(n, args.block_steps, args.k, s)
    w.set_precision(gdmr.PREC_BF16 if args.prec == "bf16" else gdmr.PREC_FP32)
    w.set_csr(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), A.nnz)
    xw = torch.zeros(s, n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    w.solve(xw.data_ptr(), n, Bd[0].data_ptr(), n, s, 1e-4, max_cycles=4)
    w.close()

    inner = args.inner_tol if args.inner_tol is not None else (3e-2 if args.prec == "bf16"
                                                               else 1e-5)
    prec = gdmr.PREC_BF16 if args.prec == "bf16" else gdmr.PREC_FP32
    bs = gdmr.BlockSolver(n, args.block_steps, args.k, s)
    bs.set_precision(prec)
    bs.set_csr(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), A.nnz)
    bs.set_refinement(inner, 60)
    legs = []
    prof = None
    for b in range(2):
        X = torch.zeros(s, n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rc, rel = bs.solve(X.data_ptr(), n, Bd[b].data_ptr(), n, s, tol, max_cycles=1000,
                           mode=b)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st, _ = bs.stats()
        Xh = X.cpu().numpy()
        tr = [float(np.linalg.norm(blocks[b][j] - A.matvec(Xh[j])) / np.linalg.norm(blocks[b][j]))
              for j in range(s)]
        legs.append(dict(rc=rc, seconds=dt, solves_per_s=s / dt, cycles=st.cycles,
                         breakdowns=st.breakdowns, max_true_rel_residual=float(np.max(tr)), X=Xh))
        _log(f"block {b}: rc {rc}, {st.cycles} cycles, {dt:.2f}s -> {s / dt:.2f} solves/s, "
             f"max true residual {max(tr):.2e}")
    tuning = bs.tuning()
    if not args.no_profile:
        # Phase profile on a third block (recycled; HIP events around every launch, so it runs
        # the launches eagerly rather than as graph replays): not part of the timed numbers.
        B3 = dev(np.stack([gdmr.rand_normal(args.seed, 3000 + j, n) for j in range(s)]))
        X = torch.zeros(s, n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        bs.set_profiling(True)
        bs.solve(X.data_ptr(), n, B3.data_ptr(), n, s, tol, max_cycles=1000, mode=1)
        prof = bs.profile()
        bs.set_profiling(False)
    bs.close()

    # FP64 single-vector DMR on a sample of the first block's columns, subspace recycled.
    sv = gdmr.Solver(n, args.m, args.k)
    sv.set_csr(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), A.nnz)
    x64, t64, cyc64 = [], 0.0, []
    for j in range(args.fp64_sample):
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        b = Bd[0][j].contiguous()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sv.solve(x.data_ptr(), b.data_ptr(), tol, 1000000, mode=j)
        torch.cuda.synchronize()
        t64 += time.perf_counter() - t0
        st, _ = sv.stats()
        cyc64.append(st.cycles)
        x64.append(x.cpu().numpy())
    sv.close()
    fp64_rate = args.fp64_sample / t64
    errs = [float(np.linalg.norm(legs[0]["X"][j] - x64[j]) / np.linalg.norm(x64[j]))
            for j in range(args.fp64_sample)]
    _log(f"fp64 single: cycles {cyc64}, {fp64_rate:.2f} solves/s; block vs fp64 solution error "
         f"max {float(np.max(errs)):.2e}")
    out = {
        "metric": "batched solves/s @1e-10 (FP64-checked), block DMR vs sequential FP64 DMR",
        "value": legs[1]["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "dtype": "bf16+fp32+fp64" if args.prec == "bf16" else "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": f"3-D anisotropic diffusion 7-pt {N}^3 (1, 0.1, 0.01), {s} RHS, "
                               f"block steps {args.block_steps}, k={args.k}, tol {tol:g}, "
                               f"{args.prec} basis, inner tol {inner:g}",
                   "n": n, "nnz": A.nnz},
        "block_first": {k: v for k, v in legs[0].items() if k != "X"},
        "block_recycled": {k: v for k, v in legs[1].items() if k != "X"},
        "fp64_sequential": {"solves_per_s": fp64_rate, "cycles": cyc64,
                            "sample": args.fp64_sample, "m": args.m, "k": args.k},
        "speedup_vs_fp64": legs[1]["solves_per_s"] / fp64_rate,
        "speedup_first_block_vs_fp64": legs[0]["solves_per_s"] / fp64_rate,
        "max_rel_err_vs_fp64_solution": float(np.max(errs)),
        "tuning": tuning,
    }
    if prof:
        out["phase_ms"] = {p: round(v[0], 2) for p, v in prof.items() if v[2]}
        out["phase_GBps"] = {p: round(v[1] / (v[0] * 1e-3) / 1e9, 1)
                             for p, v in prof.items() if v[0] > 0 and v[1] > 0}
        out["phase_TFLOPs"] = {p: round(v[3] / (v[0] * 1e-3) / 1e12, 2)
                               for p, v in prof.items() if v[0] > 0 and v[3] > 0}
    return out


def main(argv=None):
    out = run(parse(argv))
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
