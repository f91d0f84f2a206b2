"""Serving loop (gDMRServe.h): per-family resident subspaces, continuous batching over slots,
admission by free HBM with eviction.

The invariant under test is the serving rule: a family's results depend on its own request
sequence only.  Each family's requests are served once alone (one slot, no other family) and once
in a mix (several slots, three families interleaved, one of them mixed precision); the solutions
must agree bit for bit, and each must meet its tolerance in a true FP64 residual.  Eviction moves a
family's recycled state to the host and back; that must not change a bit either.  Block families
batch queued requests into one launch; there the bar is agreement within a stated tolerance.
"""
import numpy as np
import pytest

import gdmr

pytestmark = pytest.mark.gpu

REQS = 4  # requests per family


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _families():
    """(name, generator(t) -> CsrMatrix, m, k, precision, tol, nrhs per request)."""
    return [
        ("cd2d-96", lambda t: gdmr.conv_diff_2d(96, jitter=0.01, seed=11, stream=t), 30, 10,
         gdmr.PREC_FP64, 1e-10, (1, 2, 1, 3)),
        ("cd3d-24", lambda t: gdmr.conv_diff_3d(24, jitter=0.01, seed=12, stream=t), 24, 8,
         gdmr.PREC_FP64, 1e-10, (2, 1, 1, 2)),
        ("an3d-20-fp32", lambda t: gdmr.aniso_3d(20, jitter=0.01, seed=13, stream=t), 24, 8,
         gdmr.PREC_FP32, 1e-10, (1, 1, 2, 1)),
    ]


class _Served:
    """Device inputs of every request of the given families, and a server run over them."""

    def __init__(self, torch, fams):
        self.torch = torch
        self.fams = fams
        self.pattern, self.reqs = [], []
        for fi, (name, gen, m, k, prec, tol, nrhs) in enumerate(fams):
            A0 = gen(0)
            self.pattern.append((A0, _dev(torch, A0.rowptr), _dev(torch, A0.col)))
            for t in range(REQS):
                A = gen(t)
                assert np.array_equal(A.col, A0.col)
                B = [gdmr.rand_normal(100 + fi, 10 * t + j, A.n) for j in range(nrhs[t])]
                self.reqs.append(dict(fam=fi, t=t, A=A, val=_dev(torch, A.val), B=B,
                                      b=[_dev(torch, b) for b in B]))

    def run(self, slots, order, hbm_reserve=1 << 28):
        """Submits requests in `order` (indices into self.reqs) to a fresh server; returns
        {request index: (result, [x arrays])}."""
        torch = self.torch
        srv = gdmr.Server(slots=slots, hbm_reserve=hbm_reserve)
        try:
            fam_id = {}
            for fi in sorted({self.reqs[i]["fam"] for i in order}):
                name, gen, m, k, prec, tol, nrhs = self.fams[fi]
                A0, rp, ci = self.pattern[fi]
                fam_id[fi] = srv.add_family(A0.n, A0.nnz, rp.data_ptr(), ci.data_ptr(), m, k,
                                            prec, keep=(rp, ci))
                if prec != gdmr.PREC_FP64:
                    srv.set_refinement(fam_id[fi], 1e-5, 60)
            xs, ids = {}, {}
            for i in order:
                r = self.reqs[i]
                tol = self.fams[r["fam"]][5]
                xs[i] = [torch.zeros(r["A"].n, dtype=torch.float64, device="cuda")
                         for _ in r["b"]]
                # The server's streams do not wait on torch's: inputs must be complete.
                torch.cuda.synchronize()
                rid = srv.submit(fam_id[r["fam"]], r["val"].data_ptr(),
                                 [b.data_ptr() for b in r["b"]], [x.data_ptr() for x in xs[i]],
                                 tol, 10 * r["A"].n, keep=(r["val"], r["b"], xs[i]))
                ids[rid] = i
            srv.drain()
            results = srv.wait(timeout=0.0, cap=4 * len(order))
            assert len(results) == len(order)
            stats = srv.stats()
        finally:
            srv.close()
        out = {}
        for res in results:
            i = ids[res.id]
            out[i] = (res, [x.cpu().numpy() for x in xs[i]])
        return out, stats


def _check_residuals(served, out):
    for i, (res, xs) in out.items():
        r = served.reqs[i]
        tol = served.fams[r["fam"]][5]
        assert res.status == 0, (i, res.status)
        for B, x in zip(r["B"], xs):
            tr = np.linalg.norm(B - r["A"].matvec(x)) / np.linalg.norm(B)
            assert tr <= tol * 1.05, (i, tr)


def test_alone_vs_mix_bitwise(torch_gpu):
    fams = _families()
    served = _Served(torch_gpu, fams)
    per_fam = {fi: [i for i, r in enumerate(served.reqs) if r["fam"] == fi]
               for fi in range(len(fams))}
    alone = {}
    for fi, idx in per_fam.items():
        out, _ = served.run(1, idx)
        alone.update(out)
    _check_residuals(served, alone)
    # Interleave the families' requests (submission order within a family kept).
    mix_order = [per_fam[fi][t] for t in range(REQS) for fi in per_fam]
    mix, stats = served.run(4, mix_order)
    _check_residuals(served, mix)
    assert stats["queued"] == 0 and stats["running"] == 0
    for i in range(len(served.reqs)):
        ra, xa = alone[i]
        rm, xm = mix[i]
        r = served.reqs[i]
        print(f"{fams[r['fam']][0]} req {r['t']}: cycles alone {ra.cycles} mix {rm.cycles}, "
              f"recycled {ra.recycled}/{rm.recycled}, slot {rm.slot}", flush=True)
        assert ra.cycles == rm.cycles
        assert ra.recycled == rm.recycled == (1 if r["t"] > 0 else 0)
        for a, b in zip(xa, xm):
            assert np.array_equal(a, b), f"request {i} differs alone vs in the mix"
    # Requests of different families overlapped in time.
    spans = [(mix[i][0].startTime, mix[i][0].endTime, served.reqs[i]["fam"]) for i in mix]
    overlap = any(s1 < e2 and s2 < e1 and f1 != f2
                  for (s1, e1, f1) in spans for (s2, e2, f2) in spans)
    assert overlap
    # Recycling pays on the 2-D family: a later request needs fewer cycles per right-hand side
    # than the first one (one right-hand side, fresh start).
    first = alone[per_fam[0][0]][0]
    assert first.nrhs == 1
    later = [alone[i][0] for i in per_fam[0][1:]]
    assert min(r.cycles / r.nrhs for r in later) < first.cycles


def test_eviction_keeps_the_chain(torch_gpu):
    """Two families whose handles do not fit together: every admission evicts the other family
    (subspace and stagnation counter to host memory and back), and the solutions are bitwise the
    ones of the family served alone with room to spare."""
    torch = torch_gpu
    fams = [
        ("cd3d-40", lambda t: gdmr.conv_diff_3d(40, jitter=0.01, seed=21, stream=t), 30, 10,
         gdmr.PREC_FP64, 1e-10, (1, 1, 1, 1)),
        ("an3d-40", lambda t: gdmr.aniso_3d(40, jitter=0.01, seed=22, stream=t), 30, 10,
         gdmr.PREC_FP64, 1e-10, (1, 1, 1, 1)),
    ]
    served = _Served(torch, fams)
    per_fam = {fi: [i for i, r in enumerate(served.reqs) if r["fam"] == fi] for fi in (0, 1)}
    alone = {}
    for fi, idx in per_fam.items():
        out, _ = served.run(1, idx)
        alone.update(out)
    n = 40 ** 3
    fp = gdmr.footprint(n, 30, 10) + n * 4 // 64 + (8 << 20)
    torch.cuda.synchronize()
    free, _ = torch.cuda.mem_get_info()
    reserve = free - int(1.6 * fp)
    assert reserve > 0
    order = [per_fam[fi][t] for t in range(REQS) for fi in (0, 1)]
    mix, stats = served.run(2, order, hbm_reserve=reserve)
    _check_residuals(served, mix)
    assert stats["handles"] <= 1
    for i in order:
        for a, b in zip(alone[i][1], mix[i][1]):
            assert np.array_equal(a, b), f"request {i} changed by eviction"
        assert alone[i][0].cycles == mix[i][0].cycles


XTOL_BATCH = 1e-6  # relative solution agreement, alone vs batched (both at 1e-10 residual)


@pytest.mark.parametrize("basis", ["fp32", "bf16"])
def test_block_family_batches_within_tolerance(torch_gpu, basis):
    """Block families solve queued requests that share an operator as one block launch
    (continuous batching).  Served alone (one request at a time: no batching) and in a mix (every
    request of three families queued at once: requests batch), every solution meets its tolerance
    in FP64 and the alone and mixed solutions agree to XTOL_BATCH; with an FP32 or a bf16 block
    basis."""
    torch = torch_gpu
    tol = 1e-10
    A1 = gdmr.conv_diff_3d(24, jitter=0.01, seed=51, stream=0)
    A2 = [gdmr.aniso_3d(20, jitter=0.01, seed=52, stream=t) for t in range(2)]
    A3 = gdmr.conv_diff_2d(64, jitter=0.01, seed=53, stream=0)
    pats = [(A1, _dev(torch, A1.rowptr), _dev(torch, A1.col)),
            (A2[0], _dev(torch, A2[0].rowptr), _dev(torch, A2[0].col)),
            (A3, _dev(torch, A3.rowptr), _dev(torch, A3.col))]
    vals = [[_dev(torch, A1.val)], [_dev(torch, A.val) for A in A2], [_dev(torch, A3.val)]]
    mats = [[A1], A2, [A3]]
    reqs = []  # (family, value draw, host RHS list)
    for q, nr in enumerate((1, 3, 2, 1, 2, 3)):
        reqs.append((0, 0, [gdmr.rand_normal(60, 10 * q + j, A1.n) for j in range(nr)]))
    for q in range(4):
        reqs.append((1, q % 2, [gdmr.rand_normal(61, 10 * q + j, A2[0].n) for j in range(2)]))
    for q in range(2):
        reqs.append((2, 0, [gdmr.rand_normal(62, q, A3.n)]))
    bdev = [[_dev(torch, b) for b in r[2]] for r in reqs]

    def serve(slots, one_by_one):
        srv = gdmr.Server(slots=slots, hbm_reserve=1 << 28)
        try:
            fams = [srv.add_block_family(A1.n, A1.nnz, pats[0][1].data_ptr(), pats[0][2].data_ptr(),
                                         6, 16, 16),
                    srv.add_block_family(A2[0].n, A2[0].nnz, pats[1][1].data_ptr(),
                                         pats[1][2].data_ptr(), 6, 8, 8),
                    srv.add_family(A3.n, A3.nnz, pats[2][1].data_ptr(), pats[2][2].data_ptr(), 30,
                                   10)]
            if basis == "bf16":
                for f in fams[:2]:
                    srv.set_block_precision(f, gdmr.PREC_BF16)
                    srv.set_refinement(f, 3e-2, 60)
                with pytest.raises(gdmr.SolverError):
                    srv.set_block_precision(fams[2], gdmr.PREC_BF16)  # not a block family
            xs, ids, results = {}, {}, []
            for i, (f, d, B) in enumerate(reqs):
                xs[i] = [torch.zeros(len(b), dtype=torch.float64, device="cuda") for b in B]
                torch.cuda.synchronize()
                rid = srv.submit(fams[f], vals[f][d].data_ptr(), [b.data_ptr() for b in bdev[i]],
                                 [x.data_ptr() for x in xs[i]], tol, 100000)
                ids[rid] = i
                if one_by_one:
                    srv.drain()
            srv.drain()
            results = srv.wait(timeout=0.0, cap=64)
        finally:
            srv.close()
        assert len(results) == len(reqs)
        return {ids[r.id]: (r, [x.cpu().numpy() for x in xs[ids[r.id]]]) for r in results}

    alone = serve(1, True)
    mix = serve(3, False)
    batched = 0
    for i, (f, d, B) in enumerate(reqs):
        ra, xa = alone[i]
        rm, xm = mix[i]
        assert ra.status == 0 and rm.status == 0, (i, ra.status, rm.status)
        assert ra.batch == ra.nrhs
        batched += rm.batch > rm.nrhs
        A = mats[f][d]
        for b, a, m in zip(B, xa, xm):
            for x in (a, m):
                tr = np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)
                assert tr <= tol * 1.05, (i, tr)
            err = np.linalg.norm(a - m) / np.linalg.norm(a)
            print(f"request {i} (family {f}, batch {rm.batch}): alone vs mix {err:.2e}",
                  flush=True)
            assert err <= XTOL_BATCH, (i, err)
    assert batched > 0


def test_block_family_request_budget(torch_gpu):
    """A request's matRank is the cycle budget of its block launch: a slowly converging family
    with a small budget completes with DMR_ERR_NOCONV after at most that many cycles (the server
    stays live), and the family's next request with a large budget converges."""
    torch = torch_gpu
    A = gdmr.conv_diff_2d(32, jitter=0.01, seed=81)
    rp, ci, va = _dev(torch, A.rowptr), _dev(torch, A.col), _dev(torch, A.val)
    B = [gdmr.rand_normal(82, j, A.n) for j in range(3)]
    bd = [_dev(torch, b) for b in B]
    srv = gdmr.Server(slots=2, hbm_reserve=1 << 28)
    try:
        f = srv.add_block_family(A.n, A.nnz, rp.data_ptr(), ci.data_ptr(), 4, 8, 8)
        srv.set_block_precision(f, gdmr.PREC_BF16)
        srv.set_refinement(f, 3e-2, 60)
        out = []
        for budget in (3, 5000):
            xs = [torch.zeros(A.n, dtype=torch.float64, device="cuda") for _ in B]
            torch.cuda.synchronize()
            srv.submit(f, va.data_ptr(), [b.data_ptr() for b in bd], [x.data_ptr() for x in xs],
                       1e-10, budget)
            srv.drain()
            res = srv.wait(timeout=0.0, cap=4)
            assert len(res) == 1
            out.append((res[0], [x.cpu().numpy() for x in xs]))
    finally:
        srv.close()
    (r0, x0), (r1, x1) = out
    print(f"budget 3: status {r0.status} cycles {r0.cycles}; budget 5000: status {r1.status} "
          f"cycles {r1.cycles} max rel {r1.maxRelRes:.2e}", flush=True)
    assert r0.status == gdmr.DMR_ERR_NOCONV and r0.cycles <= 3
    assert r1.status == 0
    for b, x in zip(B, x1):
        assert np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b) <= 1e-10 * 1.05


def test_block_family_slots_match_one_thread_replay(torch_gpu):
    """A bf16 block family served from four slots (each launch on whichever slot's thread and
    stream picks it up) gives, launch by launch, the cycle counts and the solutions bit for bit of
    the same launches replayed in order on one thread and one stream through a fresh block handle
    (the server batches requests by timing, so the replay takes each launch's composition from the
    served results)."""
    torch = torch_gpu
    draws = [gdmr.conv_diff_3d(40, jitter=0.01, seed=71, stream=t) for t in range(2)]
    A0 = draws[0]
    rp, ci = _dev(torch, A0.rowptr), _dev(torch, A0.col)
    vals = [_dev(torch, A.val) for A in draws]
    pool = [_dev(torch, gdmr.rand_normal(72, j, A0.n)) for j in range(6)]
    rng = np.random.default_rng(73)
    specs = [(int(w // 3) % 2, [int(i) for i in rng.permutation(6)[:int(rng.integers(1, 6))]])
             for w in range(15)]
    mb, k, smax, tol, budget = 6, 16, 16, 1e-10, 3000
    srv = gdmr.Server(slots=4, hbm_reserve=1 << 28)
    served = {}
    try:
        f = srv.add_block_family(A0.n, A0.nnz, rp.data_ptr(), ci.data_ptr(), mb, k, smax)
        srv.set_block_precision(f, gdmr.PREC_BF16)
        srv.set_refinement(f, 3e-2, 60)
        ids = {}
        # Three waves of five requests: queued together, so later requests join launches.
        for wave in range(3):
            for q in range(5 * wave, 5 * wave + 5):
                d, cols = specs[q]
                xs = [torch.zeros(A0.n, dtype=torch.float64, device="cuda") for _ in cols]
                torch.cuda.synchronize()
                rid = srv.submit(f, vals[d].data_ptr(), [pool[j].data_ptr() for j in cols],
                                 [x.data_ptr() for x in xs], tol, budget)
                ids[rid] = (q, xs)
            srv.drain()
        for r in srv.wait(timeout=0.0, cap=64):
            q, xs = ids[r.id]
            served[q] = (r, [x.cpu().numpy() for x in xs])
    finally:
        srv.close()
    assert len(served) == len(specs)
    launches = {}
    for q, (r, _) in served.items():
        assert r.status == 0, (q, r.status)
        launches.setdefault((r.slot, r.startTime), []).append(q)
    order = sorted(launches.items(), key=lambda kv: kv[0][1])
    print(f"{len(order)} launches on slots {[key[0] for key, _ in order]}, widths "
          f"{[sum(len(specs[q][1]) for q in qs) for _, qs in order]}", flush=True)
    bs = gdmr.BlockSolver(A0.n, mb, k, smax)
    try:
        first = min(order[0][1])
        bs.set_csr(rp.data_ptr(), ci.data_ptr(), vals[specs[first][0]].data_ptr(), A0.nnz)
        bs.set_precision(gdmr.PREC_BF16)
        bs.set_refinement(3e-2, 60)
        stream = torch.cuda.Stream()
        for (slot, _), qs in order:
            qs = sorted(qs)  # the server batches in submission order
            cols = [j for q in qs for j in specs[q][1]]
            B = torch.stack([pool[j] for j in cols]).contiguous()  # column c at c * n
            X = torch.zeros_like(B)
            torch.cuda.synchronize()
            bs.set_values(vals[specs[qs[0]][0]].data_ptr())
            rc, _ = bs.solve(X.data_ptr(), A0.n, B.data_ptr(), A0.n, len(cols), tol, budget,
                             1 if bs.has_z() else 0, stream=stream.cuda_stream)
            stream.synchronize()
            assert rc == 0
            st, _ = bs.stats()
            Xh = X.cpu().numpy()
            print(f"launch of {qs} on slot {slot}: served {served[qs[0]][0].cycles} cycles, "
                  f"replay {st.cycles}", flush=True)
            c = 0
            for q in qs:
                r, xs = served[q]
                assert r.cycles == st.cycles and r.batch == len(cols)
                for x in xs:
                    assert np.array_equal(x, Xh[c]), f"request {q} differs from the replay"
                    c += 1
    finally:
        bs.close()
    for q, (r, xs) in served.items():
        A = draws[specs[q][0]]
        for j, x in zip(specs[q][1], xs):
            b = pool[j].cpu().numpy()
            assert np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b) <= tol * 1.05


def test_block_handle_driven_from_several_threads(torch_gpu):
    """One bf16 block handle driven in turn by four long-lived threads, each with its own stream
    (the serving loop's slot pattern: captured cycle graphs recorded in one thread and replayed
    from the others), gives the cycle counts and the solutions bit for bit of the same launches
    run on one thread and one stream through a second handle."""
    import queue
    import threading

    torch = torch_gpu
    draws = [gdmr.conv_diff_3d(40, jitter=0.01, seed=74, stream=t) for t in range(2)]
    A0 = draws[0]
    rp, ci = _dev(torch, A0.rowptr), _dev(torch, A0.col)
    vals = [_dev(torch, A.val) for A in draws]
    pool = [_dev(torch, gdmr.rand_normal(75, j, A0.n)) for j in range(8)]
    rng = np.random.default_rng(76)
    launches = [(int(L // 2) % 2, [int(i) for i in rng.permutation(8)[:int(rng.integers(1, 9))]])
                for L in range(12)]
    mb, k, smax, tol, budget = 6, 16, 16, 1e-10, 3000

    def handle():
        bs = gdmr.BlockSolver(A0.n, mb, k, smax)
        bs.set_csr(rp.data_ptr(), ci.data_ptr(), vals[launches[0][0]].data_ptr(), A0.nnz)
        bs.set_precision(gdmr.PREC_BF16)
        bs.set_refinement(3e-2, 60)
        return bs

    def launch(bs, L, stream):
        d, cols = launches[L]
        B = torch.stack([pool[j] for j in cols]).contiguous()  # column c at c * n
        X = torch.zeros_like(B)
        torch.cuda.synchronize()
        bs.set_values(vals[d].data_ptr())
        rc, _ = bs.solve(X.data_ptr(), A0.n, B.data_ptr(), A0.n, len(cols), tol, budget,
                         1 if bs.has_z() else 0, stream=stream.cuda_stream)
        stream.synchronize()
        st, _ = bs.stats()
        return rc, st.cycles, X.cpu().numpy()

    bs = handle()
    threaded = {}
    try:
        jobs = [queue.Queue() for _ in range(4)]
        done = queue.Queue()

        def worker(i):
            torch.cuda.set_device(0)
            stream = torch.cuda.Stream()
            while True:
                L = jobs[i].get()
                if L is None:
                    return
                try:
                    done.put((L, launch(bs, L, stream), threading.get_ident()))
                except Exception as e:  # reported by the main thread
                    done.put((L, e, threading.get_ident()))

        threads = [threading.Thread(target=worker, args=(i,)) for i in range(4)]
        for t in threads:
            t.start()
        try:
            for L in range(len(launches)):
                jobs[L % 4].put(L)  # one launch at a time, each from the next thread
                L2, out, tid = done.get(timeout=120)
                assert L2 == L
                if isinstance(out, Exception):
                    raise out
                threaded[L] = (out, tid)
        finally:
            for q in jobs:
                q.put(None)
            for t in threads:
                t.join(timeout=120)
    finally:
        bs.close()
    assert len({tid for _, tid in threaded.values()}) == 4
    ref = handle()
    try:
        stream = torch.cuda.Stream()
        for L in range(len(launches)):
            rc, cyc, X = launch(ref, L, stream)
            (rc_t, cyc_t, X_t), _ = threaded[L]
            print(f"launch {L} (width {len(launches[L][1])}): threads {cyc_t} cycles, one "
                  f"thread {cyc}", flush=True)
            assert rc == rc_t == 0
            assert cyc == cyc_t
            assert np.array_equal(X, X_t), f"launch {L} differs across threads"
    finally:
        ref.close()
