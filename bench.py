"""Benchmark: solves/sec at 1e-10 relative residual (FP64-checked) on the 128^3 sequence.

Workload: 3-D convection-diffusion, 7-point, 128^3
(n = 2,097,152, nnz = 14,581,760), CSR FP64, a sequence of 16 systems whose per-cell diffusion
is jittered U(-1%, +1%) per step (seeded), N(0,1) right-hand sides, m = 50, k = 20, tol 1e-10,
deflation subspace recycled across the sequence (mode = 0 for the first system, 1 after).

Three legs run on the same sequence, each in a fresh solver context:
  * fp64     -- the plain FP64 DMR, one system after the other (the parity-checked path);
  * mixed    -- FP32 Krylov basis and SpMV with FP32-accumulated reductions inside FP64 iterative
                refinement; Hessenberg LSQ / SVD in FP64; one system after the other;
  * seqblock -- mixed + batched: the sequence as block solves of at most 16 consecutive systems
                (balanced sizes; 16 systems: one block, 20: two of 10), one after the other on
                one handle that recycles the deflation subspace from block to block.  Column t is
                refined in FP64 against its own operator A_t; the inner block DMR (block
                Gram-Schmidt as tall-skinny GEMMs on the matrix cores, round(16 sqrt(16 / s))
                block steps for blocks of s systems, recycled subspace k = 20, basis stored as
                bf16 and multiplied on the bf16 matrix cores with FP32 operands split into
                hi + lo parts, FP32 accumulation) shares the entrywise mean of its block's A_t.
                Every system's latency is its block's.
A step is one solve of one system.  `value` is the solves/sec of the fastest leg whose every
system reaches 1e-10 checked in FP64.  Every solve's true residual ||b - A x|| / ||b|| (with its
own operator) is recomputed on the host in FP64 after the timed region.

The line also carries the batched form of the metric ("batched"): 3-D anisotropic diffusion 96^3
with 32 right-hand sides of one operator solved as one block (block DMR on the matrix cores, bf16
basis storage inside FP64 refinement) against the plain FP64 DMR solving them one at a time
(gdmr.batch_bench), and the FP64 host oracle's rate on the same sequence ("cpu_baseline").
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gdmr-solver_amd"))

PROFILE_SYSTEMS = 4  # systems re-solved with per-launch profiling after the timed sequence
METRIC = ("solves/sec @1e-10 rel residual (FP64-checked), mixed+batched vs FP64; p50/p99")
HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_sequence(N, steps, seed):
    import gdmr
    mats, rhs = [], []
    for t in range(steps):
        mats.append(gdmr.conv_diff_3d(N, nu0=0.1, beta=(1.0, 0.5, 0.25), jitter=0.01, seed=seed,
                                      stream=t))
        rhs.append(gdmr.rand_normal(seed, 1000 + t, mats[-1].n))
    return mats, rhs


def run_leg(name, prec, mats, rhs, dev_vals, dev_vals32, rp, ci, m, k, tol, warmup, profile,
            inner_tol=1e-5):
    import torch

    import gdmr
    n = mats[0].n
    # Warm-up on a throwaway context (first-touch, code objects), then a fresh context for the
    # timed sequence so the recycled subspace starts from nothing.
    for w in range(warmup):
        s = gdmr.Solver(n, m, k, precision=prec)
        s.set_csr(rp.data_ptr(), ci.data_ptr(), dev_vals[0].data_ptr(), mats[0].nnz,
                  dev_vals32[0].data_ptr() if prec != gdmr.PREC_FP64 else 0)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        b = torch.from_numpy(rhs[0]).to("cuda")
        s.solve(x.data_ptr(), b.data_ptr(), tol, 100000, mode=0)
        s.close()
        log(f"[{name}] warm-up {w} done")
    s = gdmr.Solver(n, m, k, precision=prec)
    s.set_csr(rp.data_ptr(), ci.data_ptr(), dev_vals[0].data_ptr(), mats[0].nnz,
              dev_vals32[0].data_ptr() if prec != gdmr.PREC_FP64 else 0)
    if prec != gdmr.PREC_FP64:
        s.set_refinement(inner_tol)
    bs = [torch.from_numpy(b).to("cuda") for b in rhs]
    xs = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in rhs]
    lat, cycles, reported = [], [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(len(mats)):
        ts = time.perf_counter()
        s.set_values(dev_vals[t].data_ptr(),
                     dev_vals32[t].data_ptr() if prec != gdmr.PREC_FP64 else 0)
        res, its = s.solve(xs[t].data_ptr(), bs[t].data_ptr(), tol, 100000, mode=1 if t else 0)
        lat.append(time.perf_counter() - ts)
        st, _ = s.stats()
        cycles.append(st.cycles)
        reported.append(res)
        log(f"[{name}] system {t}: {lat[-1]*1e3:.1f} ms, cycles {st.cycles}, residual {res:.3e}, "
            f"status {st.status}, svd sweeps/call {st.svdSweeps / max(st.svdCalls, 1):.1f}")
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    prof = None
    if profile:
        # Phase profile (HIP events around every launch, which slows the launches down) on a
        # re-solve of the first systems after the timed sequence; not part of the timed numbers.
        s.set_profiling(True)
        xp = torch.zeros(n, dtype=torch.float64, device="cuda")
        for t in range(min(PROFILE_SYSTEMS, len(mats))):
            s.set_values(dev_vals[t].data_ptr(),
                         dev_vals32[t].data_ptr() if prec != gdmr.PREC_FP64 else 0)
            xp.zero_()
            torch.cuda.synchronize()
            s.solve(xp.data_ptr(), bs[t].data_ptr(), tol, 100000, mode=1)
        prof = s.profile()
        s.set_profiling(False)
    s.close()
    true = []
    for t, A in enumerate(mats):
        x = xs[t].cpu().numpy()
        true.append(float(np.linalg.norm(rhs[t] - A.matvec(x)) / np.linalg.norm(rhs[t])))
    xs_host = [x.cpu().numpy() for x in xs]
    return dict(total=total, lat=lat, cycles=cycles, reported=reported, true=true, prof=prof,
                x=xs_host)


def run_block1_leg(mats, rhs, dev_vals, rp, ci, m, k, tol, warmup, inner_tol=1e-5):
    """The sequence through the block solver with one right-hand side per system: block steps
    m - k (the same basis size per cycle), recycled subspace of k vectors, FP32 inside FP64
    refinement."""
    import torch

    import gdmr
    n, mb = mats[0].n, m - k
    for w in range(warmup):
        s = gdmr.BlockSolver(n, mb, k, 1)
        s.set_csr(rp.data_ptr(), ci.data_ptr(), dev_vals[0].data_ptr(), mats[0].nnz)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        b = torch.from_numpy(rhs[0]).to("cuda")
        torch.cuda.synchronize()
        s.solve(x.data_ptr(), n, b.data_ptr(), n, 1, tol, max_cycles=100000)
        s.close()
        log(f"[block1] warm-up {w} done")
    s = gdmr.BlockSolver(n, mb, k, 1)
    s.set_csr(rp.data_ptr(), ci.data_ptr(), dev_vals[0].data_ptr(), mats[0].nnz)
    s.set_refinement(inner_tol, 60)
    bs = [torch.from_numpy(b).to("cuda") for b in rhs]
    xs = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in rhs]
    lat, cycles, reported = [], [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(len(mats)):
        ts = time.perf_counter()
        s.set_values(dev_vals[t].data_ptr())
        rc, rel = s.solve(xs[t].data_ptr(), n, bs[t].data_ptr(), n, 1, tol, max_cycles=100000,
                          mode=1 if t else 0)
        lat.append(time.perf_counter() - ts)
        st, _ = s.stats()
        cycles.append(st.cycles)
        reported.append(float(rel[0]))
        log(f"[block1] system {t}: {lat[-1]*1e3:.1f} ms, cycles {st.cycles}, residual {rel[0]:.3e}, "
            f"status {rc}")
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    s.close()
    true = []
    for t, A in enumerate(mats):
        x = xs[t].cpu().numpy()
        true.append(float(np.linalg.norm(rhs[t] - A.matvec(x)) / np.linalg.norm(rhs[t])))
    return dict(total=total, lat=lat, cycles=cycles, reported=reported, true=true, prof=None,
                x=[x.cpu().numpy() for x in xs])


def seq_groups(count, block_max):
    """Consecutive systems split into ceil(count / block_max) blocks of balanced size."""
    nb = -(-count // block_max)
    base, extra = divmod(count, nb)
    groups, t = [], 0
    for g in range(nb):
        size = base + (1 if g < extra else 0)
        groups.append((t, t + size))
        t += size
    return groups


def seq_block_steps(s, k, max_cols=384):
    """Block steps per cycle for blocks of s systems: round(16 sqrt(16 / s)), the measured best
    on the 128^3 sequence (16 steps for blocks of 16, 20 for blocks of 10), capped so that the
    basis k + (steps + 1) s fits max_cols columns."""
    steps = int(round(16.0 * (16.0 / s) ** 0.5))
    return max(1, min(steps, (max_cols - k) // s - 1))


def run_seqblock_leg(mats, rhs, dev_vals, rp, ci, mb, k, tol, warmup, inner_tol, profile,
                     prec="bf16", block_max=16, concurrent=1):
    """The sequence as block solves of at most block_max systems each (consecutive systems,
    balanced sizes): column t is system t with its own FP64 operator A_t in the refinement
    residual, while the inner FP32 block DMR (block Gram-Schmidt on the matrix cores) shares the
    entrywise mean of its block's A_t (dmrBlockSetColumnOperators).  concurrent = 1: one handle
    solves the blocks one after the other and recycles the deflation subspace from block to
    block.  concurrent = c > 1: the blocks are dealt to c handles, each with its own stream and
    host thread, which run at the same time (each recycling its subspace over its own blocks), so
    one block's latency-bound steps (least squares, Cholesky factors, partial sums, the host's
    per-cycle decision) overlap another's streaming kernels.  Every system's latency is its
    block's solve time."""
    import threading

    import torch

    import gdmr
    n, nnz = mats[0].n, mats[0].nnz
    groups = seq_groups(len(mats), block_max)
    smax = max(g1 - g0 for g0, g1 in groups)
    concurrent = max(1, min(concurrent, len(groups)))
    vals = torch.stack(dev_vals).contiguous()  # count x nnz, system t's operator at row t
    B = torch.stack([torch.from_numpy(b) for b in rhs]).to("cuda").contiguous()

    def make():
        h = gdmr.BlockSolver(n, mb, k, smax)
        h.set_csr(rp.data_ptr(), ci.data_ptr(), vals[0].data_ptr(), nnz)
        h.set_refinement(inner_tol, 60)
        h.set_precision(gdmr.PREC_BF16 if prec == "bf16" else gdmr.PREC_FP32)
        return h

    def solve_groups(h, X, gids, recycled, stream=0):
        out = {}
        for i, gi in enumerate(gids):
            g0, g1 = groups[gi]
            h.set_column_operators(vals[g0].data_ptr(), nnz, g1 - g0, inner_mean=True)
            ts = time.perf_counter()
            rc, rel = h.solve(X[g0].data_ptr(), n, B[g0].data_ptr(), n, g1 - g0, tol,
                              max_cycles=100000, mode=1 if (i or recycled) else 0, stream=stream)
            dt = time.perf_counter() - ts
            st, _ = h.stats()
            out[gi] = dict(seconds=dt, rc=rc, rel=rel, cycles=st.cycles)
        return out

    def solve_all(hs, X, recycled, streams=None):
        """Blocks dealt round-robin to the handles; with streams, one thread per handle."""
        deal = [list(range(j, len(groups), len(hs))) for j in range(len(hs))]
        if streams is None:
            out = {}
            for h, gids in zip(hs, deal):
                out.update(solve_groups(h, X, gids, recycled))
            return [out[gi] for gi in range(len(groups))]
        res = [None] * len(hs)

        def work(j):
            res[j] = solve_groups(hs[j], X, deal[j], recycled, streams[j].cuda_stream)
            streams[j].synchronize()

        th = [threading.Thread(target=work, args=(j,)) for j in range(len(hs))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        out = {}
        for r in res:
            out.update(r)
        return [out[gi] for gi in range(len(groups))]

    # Warm-up on throwaway handles, one after the other (it also tunes the launch shapes, which
    # are cached per shape for the process), then fresh handles for the timed sequence.
    for w in range(warmup):
        h = make()
        X = torch.zeros(len(mats), n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        solve_all([h], X, False)
        h.close()
        log(f"[seqblock] warm-up {w} done")
    hs = [make() for _ in range(concurrent)]
    streams = [torch.cuda.Stream() for _ in hs] if concurrent > 1 else None
    X = torch.zeros(len(mats), n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    blocks = solve_all(hs, X, False, streams)
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    for extra in hs[1:]:
        extra.close()
    h = hs[0]
    for (g0, g1), b in zip(groups, blocks):
        log(f"[seqblock] systems {g0}..{g1 - 1} as one block: {b['seconds'] * 1e3:.1f} ms, rc "
            f"{b['rc']}, {b['cycles']} cycles, max residual {b['rel'].max():.3e}")
    tuning = h.tuning()
    prof = None
    if profile:
        # Phase profile (HIP events around every launch, eager launches instead of graph
        # replays) on a re-solve of the same blocks from the recycled subspace; not part of the
        # timed numbers.
        h.set_profiling(True)
        Xp = torch.zeros(len(mats), n, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        solve_all([h], Xp, True)
        prof = h.profile()
        h.set_profiling(False)
        tot = sum(v[0] for v in prof.values())
        log(f"[seqblock] profiled re-solve: {tot:.1f} ms of event time; per phase (ms, launches, "
            "algorithmic TB/s): " + ", ".join(
                f"{p} {v[0]:.1f}/{v[2]}/{v[1] / (v[0] * 1e9):.2f}" if v[1] > 0 else
                f"{p} {v[0]:.1f}/{v[2]}" for p, v in prof.items() if v[2] > 0))
    h.close()
    Xh = X.cpu().numpy()
    true = [float(np.linalg.norm(rhs[t] - A.matvec(Xh[t])) / np.linalg.norm(rhs[t]))
            for t, A in enumerate(mats)]
    lat, cycles, reported = [], [], []
    for (g0, g1), b in zip(groups, blocks):
        lat += [b["seconds"]] * (g1 - g0)
        cycles += [b["cycles"]] * (g1 - g0)
        reported += list(b["rel"])
    return dict(total=total, lat=lat, cycles=cycles, reported=reported, true=true, prof=prof,
                x=[Xh[t] for t in range(len(mats))], block=True, tuning=tuning,
                blocks=[g1 - g0 for g0, g1 in groups], concurrent=concurrent)


def summarize(leg, tol):
    lat = np.array(leg["lat"])
    return {
        "solves_per_s": len(lat) / leg["total"],
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3),
        "cycles": leg["cycles"],
        "max_true_rel_residual": float(np.max(leg["true"])),  # a NaN propagates
        "all_converged": bool(np.max(leg["true"]) <= tol),
    }


MFMA_FP32_PEAK = 157.3e12
MFMA_BF16_PEAK = 2.5e15


def kernel_of(phase, leg, tuning):
    """The rocprof name of the one kernel a single-kernel phase launches in a leg, or None."""
    if leg == "seqblock":
        if phase == "update":
            return "k_gemm_nn_b<1, float, %d>" % (2 if tuning["nn_unroll"] >= 101 else 1)
        if phase == "gemvt":
            return "k_gemm_tn16b"
        if phase == "spmv":
            return "k_spmm_lds" if tuning["spmm_cols"] >= 200 else "k_spmm_w"
        return None
    return {"spmv": "k_spmv", "gemvt": "k_multidot", "update": "k_combo"}.get(phase)


def roofline(prof, leg, tuning=None, bf16=False, n=0, s=0):
    """The dominant streaming kernel of the headline leg: of the phases that launch one kernel
    each (SpMV or SpMM, projection, update), the one with the most measured time; in the seqblock
    leg always the block update (one kernel instance per process; the projection, a family of
    instances by column count, takes within a few percent of its time and is reported in
    phase_GBps).  achieved = its algorithmic bytes / its time (HIP events around each launch on
    the solve stream, in the profiled re-solve)."""
    cands = [p for p in ("spmv", "gemvt", "update")
             if prof.get(p) and prof[p][2] > 0 and kernel_of(p, leg, tuning)]
    best = max(cands, key=lambda p: prof[p][0])
    if leg == "seqblock" and "update" in cands:
        best = "update"
    ms, by, cnt = prof[best][:3]
    fl = prof[best][3] if len(prof[best]) > 3 else None
    kernel = kernel_of(best, leg, tuning)
    ach = by / (ms * 1e-3) if ms > 0 else 0.0
    out = {"bound": "hbm", "kernel": kernel, "phase": best,
           "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
           "frac": ach / HBM_PEAK, "traffic": None, "launches": cnt,
           "avg_launch_us": ms * 1e3 / max(cnt, 1),
           "bytes_per_launch": by / max(cnt, 1),
           "phase_ms": {p: round(prof[p][0], 3) for p in cands}}
    # HBM traffic per launch of the same kernel from the calibrated FETCH_SIZE / WRITE_SIZE
    # passes committed under profiles/ (rocprofv3 --pmc cannot run inside this process).  Used
    # only when it was measured on this kernel at this shape (n, block width, algorithmic bytes
    # per launch within 2%) and is not below the algorithmic bytes.
    tf_path = os.path.join(ROOT, "profiles", "roofline_traffic.json")
    if os.path.exists(tf_path):
        rec = json.load(open(tf_path)).get(kernel)
        bpl = out["bytes_per_launch"]
        if (rec and rec.get("n") == n and rec.get("s") == s and
                abs(rec["bytes_per_launch"] - bpl) <= 0.02 * bpl and
                rec["traffic_bytes_per_launch"] >= 0.99 * rec["bytes_per_launch"]):
            out["traffic"] = rec["traffic_bytes_per_launch"]
            out["traffic_source"] = rec["source"]
            out["traffic_over_algorithmic"] = rec["traffic_bytes_per_launch"] / bpl
    if fl:
        # The useful flops of the contraction.  The bf16 kernels issue each product twice (hi
        # and lo parts of the FP32 operand) on the bf16 matrix cores.
        tf = fl / (ms * 1e-3)
        peak = MFMA_BF16_PEAK if bf16 else MFMA_FP32_PEAK
        issued = 2.0 * tf if bf16 else tf
        out["mfma"] = {"unit": "bf16" if bf16 else "fp32", "useful_tflops": tf / 1e12,
                       "issued_tflops": issued / 1e12, "peak_tflops": peak / 1e12,
                       "frac": issued / peak}
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mats, rhs, m, k, tol, gpu_cycles_per_solve, sample_cycles, sample_cycles_1t):
    """The FP64 host oracle on system 0 of the same sequence, capped at a few restart cycles,
    at the host's core count and at one thread; solves/sec is extrapolated with the GPU FP64
    leg's mean cycles per solve.  The host's DRAM bandwidth (triad) is measured beside it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import gdmr
    import oracle
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = int(os.environ.get("OMP_NUM_THREADS", cores))
    A = mats[0]
    runs = {}
    for threads, cyc_cap in ((cores, sample_cycles), (1, sample_cycles_1t)):
        oracle.set_threads(threads)
        bw = oracle.stream_triad(1 << 26, 3)
        ora = oracle.Oracle(A, m, k, gdmr.reduction_blocks(A.n))
        t0 = time.perf_counter()
        ora.solve(rhs[0], tol=tol, mat_rank=cyc_cap - 1, mode=0)
        dt = time.perf_counter() - t0
        cyc = ora.counts()[0]
        ora.close()
        per_cycle = dt / max(cyc, 1)
        runs[threads] = dict(seconds=dt, cycles=cyc, per_cycle=per_cycle, dram_GBps=bw / 1e9,
                             solves_per_s=1.0 / (per_cycle * gpu_cycles_per_solve))
        log(f"[cpu] {threads} threads: {cyc} cycles in {dt:.1f} s ({per_cycle:.2f} s/cycle), "
            f"DRAM triad {bw / 1e9:.1f} GB/s")
    oracle.set_threads(cores)
    r, r1 = runs[cores], runs[1]
    return {"value": r["solves_per_s"], "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle, system 0 of the 128^3 sequence, first {r['cycles']} restart "
                      f"cycles in {r['seconds']:.1f} s ({r['per_cycle']:.2f} s/cycle, OpenMP "
                      f"{cores} threads); solves/s = 1/(s per cycle x {gpu_cycles_per_solve:.1f} "
                      f"mean cycles per solve of the GPU FP64 leg)",
            "one_thread": {"value": r1["solves_per_s"], "cycles": r1["cycles"],
                           "seconds": r1["seconds"], "s_per_cycle": r1["per_cycle"]},
            "cpu_model": cpu_model(),
            "dram_triad_GBps": {"cores": r["dram_GBps"], "one_thread": r1["dram_GBps"]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--m", type=int, default=50)
    ap.add_argument("--k", type=int, default=20)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--legs", default="fp64,mixed,seqblock",
                    help="comma list of fp64, mixed (FP32 basis), seqblock (the sequence as one "
                         "block), bf16 (bf16 basis storage) and block1 (the block solver with one "
                         "right-hand side)")
    ap.add_argument("--seq-block-steps", type=int, default=0,
                    help="block steps per cycle of the seqblock leg (0: by block width, "
                         "round(16 sqrt(16 / s)): 16 for blocks of 16, 20 for blocks of 10)")
    ap.add_argument("--seq-k", type=int, default=20, help="recycled subspace of the seqblock leg")
    ap.add_argument("--seq-inner-tol", type=float, default=3e-2,
                    help="relative tolerance of each inner block solve of the seqblock leg")
    ap.add_argument("--seq-prec", default="bf16", choices=("bf16", "fp32"),
                    help="basis storage of the seqblock leg")
    ap.add_argument("--seq-block-max", type=int, default=16,
                    help="most systems per block of the seqblock leg (the sequence is split into "
                         "balanced blocks of at most this many)")
    ap.add_argument("--seq-concurrent", type=int, default=1,
                    help="blocks of the seqblock leg solved at the same time, each on its own "
                         "handle, stream and host thread")
    ap.add_argument("--cpu-cycles", type=int, default=5,
                    help="restart cycles of the CPU baseline sample at the host's core count")
    ap.add_argument("--cpu-cycles-1t", type=int, default=2,
                    help="restart cycles of the one-thread CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tune-file", default=os.path.join(ROOT, "gdmr-solver_amd", "tuning",
                                                        "gfx950.txt"),
                    help="persisted launch shapes of the block autotuner (read, and appended to "
                         "for shapes it lacks); 'none': tune in this process only")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--inner-tol", type=float, default=1e-5,
                    help="relative tolerance of each inner FP32 solve of the mixed leg")
    ap.add_argument("--bf16-inner-tol", type=float, default=1e-2,
                    help="relative tolerance of each inner solve of the bf16 leg")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched multi-RHS leg (96^3 anisotropic, 32 RHS)")
    ap.add_argument("--serve", action="store_true",
                    help="run the serving load generator (gdmr.serve_bench) instead; the "
                         "remaining arguments go to it")
    ap.add_argument("--batched", action="store_true",
                    help="run the batched multi-RHS benchmark (gdmr.batch_bench) instead; the "
                         "remaining arguments go to it")
    args, rest = ap.parse_known_args()
    if args.serve or args.batched:
        if args.gpus != 1:
            print("bench.py: this build measures one GPU (--gpus must be 1)", file=sys.stderr)
            sys.exit(2)
        from gdmr import batch_bench, serve_bench
        (serve_bench if args.serve else batch_bench).main(rest)
        return
    if rest:
        ap.error(f"unrecognized arguments: {' '.join(rest)}")
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    import torch

    import gdmr
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    tune_file = None if args.tune_file == "none" else args.tune_file
    if tune_file:
        os.makedirs(os.path.dirname(tune_file), exist_ok=True)
    gdmr.set_tuning_file(tune_file)
    log(f"block launch shapes: {tune_file or 'tuned in this process'}")
    N, m, k, tol = args.grid, args.m, args.k, args.tol
    t0 = time.time()
    mats, rhs = build_sequence(N, args.steps, args.seed)
    log(f"generated {len(mats)} systems n={mats[0].n} nnz={mats[0].nnz} in {time.time()-t0:.1f}s")
    rp = torch.from_numpy(mats[0].rowptr).to("cuda")
    ci = torch.from_numpy(mats[0].col).to("cuda")
    dev_vals = [torch.from_numpy(A.val).to("cuda") for A in mats]
    dev_vals32 = [v.float().contiguous() for v in dev_vals]
    torch.cuda.synchronize()
    legs = {}
    for name in args.legs.split(","):
        if name == "seqblock":
            if args.seq_block_steps <= 0:
                smax = max(g1 - g0 for g0, g1 in seq_groups(len(mats), args.seq_block_max))
                args.seq_block_steps = seq_block_steps(smax, args.seq_k)
            legs[name] = run_seqblock_leg(mats, rhs, dev_vals, rp, ci, args.seq_block_steps,
                                          args.seq_k, tol, args.warmup, args.seq_inner_tol,
                                          not args.no_profile, args.seq_prec,
                                          args.seq_block_max, args.seq_concurrent)
            continue
        if name == "block1":
            legs[name] = run_block1_leg(mats, rhs, dev_vals, rp, ci, m, k, tol, args.warmup,
                                        args.inner_tol)
            continue
        prec = {"fp64": gdmr.PREC_FP64, "mixed": gdmr.PREC_FP32, "bf16": gdmr.PREC_BF16}[name]
        inner = args.bf16_inner_tol if name == "bf16" else args.inner_tol
        legs[name] = run_leg(name, prec, mats, rhs, dev_vals, dev_vals32, rp, ci, m, k, tol,
                             args.warmup, not args.no_profile, inner)
    out_legs = {nm: summarize(lg, tol) for nm, lg in legs.items()}
    if "fp64" in legs:
        for nm in [nm for nm in legs if nm != "fp64"]:
            errs = [float(np.linalg.norm(a - b) / np.linalg.norm(b))
                    for a, b in zip(legs[nm]["x"], legs["fp64"]["x"])]
            out_legs[nm]["max_rel_err_vs_fp64_solution"] = float(np.max(errs))
    # Headline: the fastest leg whose every solve reached the tolerance (FP64-checked).
    ok = [nm for nm in out_legs if out_legs[nm]["all_converged"]] or list(out_legs)
    head = max(ok, key=lambda nm: out_legs[nm]["solves_per_s"])
    hs = out_legs[head]
    line = {
        "metric": METRIC,
        "value": hs["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hs["mean_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "fp32+fp64", "bf16": "bf16+fp32+fp64",
                  "block1": "fp32+fp64",
                  "seqblock": ("bf16+fp32+fp64" if args.seq_prec == "bf16" else
                               "fp32+fp64")}[head],
        "data": "synthetic",
        "config": {"workload": f"3-D conv-diff 7-pt {N}^3 sequence of {args.steps} jittered "
                               f"systems, m={m} k={k} tol {tol:g}, Z recycled",
                   "n": mats[0].n, "nnz": mats[0].nnz, "headline_leg": head,
                   "tuning_file": (os.path.relpath(tune_file, ROOT) if tune_file else None)},
        "p50_ms": hs["p50_ms"],
        "p99_ms": hs["p99_ms"],
        "legs": out_legs,
    }
    if "fp64" in out_legs:
        for nm in [nm for nm in out_legs if nm != "fp64"]:
            line[f"speedup_{nm}_vs_fp64"] = (out_legs[nm]["solves_per_s"] /
                                             out_legs["fp64"]["solves_per_s"])
    if "seqblock" in legs:
        line["config"]["seqblock"] = {"blocks": legs["seqblock"]["blocks"],
                                      "concurrent": legs["seqblock"]["concurrent"],
                                      "block_steps": args.seq_block_steps, "k": args.seq_k,
                                      "inner_tol": args.seq_inner_tol, "basis": args.seq_prec,
                                      "tuning": legs["seqblock"]["tuning"]}
    line["latency_ms"] = {nm: {"p50": out_legs[nm]["p50_ms"], "p99": out_legs[nm]["p99_ms"]}
                          for nm in out_legs}
    if legs[head]["prof"] is not None:
        line["roofline"] = roofline(legs[head]["prof"], head, legs[head].get("tuning"),
                                    head == "seqblock" and args.seq_prec == "bf16",
                                    mats[0].n, max(legs[head].get("blocks", [1])))
        # The box's practical HBM read ceiling, measured here: a flat 16-byte streaming read of
        # 4 GiB (sixteen times the memory-side cache; one launch of about 0.7 ms, so the launch
        # overhead is small), best of 10.
        import torch

        import gdmr
        probe = torch.empty(4 << 30, dtype=torch.uint8, device="cuda")
        ceil = gdmr.probe_read_bandwidth(probe.data_ptr(), probe.numel(), 10)
        del probe
        line["roofline"]["ceiling"] = {
            "read_GBps": ceil, "frac_of_peak": ceil * 1e9 / HBM_PEAK,
            "probe": "flat 16-byte streaming read of 4 GiB, 8 loads in flight per thread, "
                     "best of 10 (dmrProbeReadBandwidth)"}
        line["roofline"]["frac_of_ceiling"] = line["roofline"]["achieved"] / ceil
        log(f"[roofline] {line['roofline']['kernel']}: {line['roofline']['achieved']:.0f} GB/s, "
            f"{100 * line['roofline']['frac']:.1f}% of 8 TB/s, "
            f"{100 * line['roofline']['frac_of_ceiling']:.1f}% of the measured read ceiling "
            f"{ceil:.0f} GB/s")
        line["phase_ms"] = {nm: {p: round(v[0], 3) for p, v in lg["prof"].items()}
                            for nm, lg in legs.items() if lg["prof"]}
        line["phase_GBps"] = {nm: {p: round(v[1] / (v[0] * 1e-3) / 1e9, 1)
                                   for p, v in lg["prof"].items() if v[0] > 0 and v[1] > 0}
                              for nm, lg in legs.items() if lg["prof"]}
        line["phase_bytes_per_launch"] = {
            nm: {p: round(v[1] / v[2]) for p, v in lg["prof"].items() if v[1] > 0 and v[2] > 0}
            for nm, lg in legs.items() if lg["prof"]}
    if not args.no_batched:
        # The batched form of the metric: 32 right-hand sides of one operator as one block solve
        # (block Gram-Schmidt on the matrix cores, FP32 inside FP64 refinement) against the plain
        # FP64 DMR solving them one at a time.
        from gdmr import batch_bench
        bargs = batch_bench.parse(["--no-profile", "--prec", "bf16"])
        br = batch_bench.run(bargs)
        line["batched"] = {
            "workload": br["config"]["workload"],
            "solves_per_s": br["value"],
            "first_block_solves_per_s": br["block_first"]["solves_per_s"],
            "fp64_solves_per_s": br["fp64_sequential"]["solves_per_s"],
            "fp64_sample": br["fp64_sequential"]["sample"],
            "speedup_vs_fp64": br["speedup_vs_fp64"],
            "speedup_first_block_vs_fp64": br["speedup_first_block_vs_fp64"],
            "max_true_rel_residual": float(np.max([br["block_first"]["max_true_rel_residual"],
                                                   br["block_recycled"]["max_true_rel_residual"]])),
            "max_rel_err_vs_fp64_solution": br["max_rel_err_vs_fp64_solution"],
            "tuning": br["tuning"],
        }
    if not args.no_cpu:
        ref = out_legs.get("fp64", hs)
        cyc = float(np.mean(ref["cycles"]))
        line["cpu_baseline"] = cpu_baseline(mats, rhs, m, k, tol, cyc, args.cpu_cycles,
                                            args.cpu_cycles_1t)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
