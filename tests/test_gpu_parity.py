"""GPU parity of the FP64 DMR solver against the FP64 host oracle.

Bar: iteration count within +/-1 and residual history within 1e-12 relative to the
oracle; solutions are compared too.  The oracle sums in the device's reduction order, so the
histories are expected to agree bit for bit; the assertions use the stated tolerance.
"""
import ctypes as C

import numpy as np
import pytest

import cases
import gdmr
import oracle

pytestmark = pytest.mark.gpu

HIST_RTOL = 1e-12


def _compare(tag, its, hist, x, its_o, hist_o, x_o):
    L = min(len(hist), len(hist_o))
    rel = np.abs(hist[:L] - hist_o[:L]) / np.abs(hist_o[:L])
    xrel = np.linalg.norm(x - x_o) / max(np.linalg.norm(x_o), 1e-300)
    bitwise = len(hist) == len(hist_o) and np.array_equal(hist, hist_o) and np.array_equal(x, x_o)
    print(f"{tag}: its {its} vs {its_o}, cycles {len(hist)} vs {len(hist_o)}, "
          f"max hist rel diff {rel.max() if L else 0:.3e}, x rel diff {xrel:.3e}, bitwise {bitwise}",
          flush=True)
    assert abs(its - its_o) <= 1
    assert L > 0 and rel.max() <= HIST_RTOL
    return bitwise


def _true_res(A, b, x):
    return np.linalg.norm(b - A.matvec(x)) / np.linalg.norm(b)


def test_config_a_linear_system_api():
    """Config A through initLS / impCsrLS / solveLS, the control-program path (progDMR.c:166-174):
    2-D conv-diff 64^2, m=30, k=10, tol 1e-8, 4 N(0,1) RHS from seed 0, mode = i."""
    A = gdmr.conv_diff_2d(64)
    n, m, k, tol = A.n, 30, 10, 1e-8
    B = gdmr.rand_normal(0, 1, 4 * n).reshape(4, n)
    ora = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n))
    ls = gdmr.LinearSystem(A, m, k)
    try:
        for i in range(4):
            x, res, its = ls.solve(B[i], tol=tol, mode=i)
            st, hist = gdmr.global_stats()
            xo, reso, itso, sto = ora.solve(B[i], tol=tol, mode=i)
            _compare(f"A[{i}]", its, hist, x, itso, ora.history(), xo)
            assert sto == 0 and st.status == 0
            assert abs(res - reso) <= HIST_RTOL * reso
            tr = _true_res(A, B[i], x)
            assert tr <= tol * 1.001, tr
    finally:
        ls.close()


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def test_handle_api_3d_sequence(torch_gpu):
    """3-D conv-diff 16^3 with per-cell coefficient jitter per system, a recycled sequence of 3
    systems, m=20, k=8, tol 1e-10, through dmrCreate / dmrSetOperatorCsr / dmrSolve."""
    torch = torch_gpu
    N, m, k, tol = 16, 20, 8, 1e-10
    n = N ** 3
    s = gdmr.Solver(n, m, k)
    ora = oracle.Oracle(gdmr.conv_diff_3d(N), m, k, gdmr.reduction_blocks(n))
    for t in range(3):
        A = gdmr.conv_diff_3d(N, jitter=0.01, seed=7, stream=t)
        b = gdmr.rand_normal(7, 100 + t, n)
        rp, ci, va = _dev(torch, A.rowptr), _dev(torch, A.col), _dev(torch, A.val)
        s.set_csr(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), A.nnz, keep=(rp, ci, va))
        xd = torch.zeros(n, dtype=torch.float64, device="cuda")
        bd = _dev(torch, b)
        res, its = s.solve(xd.data_ptr(), bd.data_ptr(), tol, n, mode=t)
        st, hist = s.stats()
        x = xd.cpu().numpy()
        ora.close()
        ora_t = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n))
        if t > 0:
            ora_t.set_z(Zprev)
        xo, reso, itso, sto = ora_t.solve(b, tol=tol, mode=t)
        _compare(f"3D[{t}]", its, hist, x, itso, ora_t.history(), xo)
        tr = _true_res(A, b, x)
        assert tr <= tol * 1.01, tr
        # The recycled subspace agrees too (GPU Z vs oracle Z).
        Zg, Zo = s.export_z(), ora_t.get_z()
        assert np.max(np.abs(Zg - Zo)) <= 1e-9
        Zprev = Zo
        ora = ora_t
    s.close()


def test_legacy_inplace_callback(torch_gpu):
    """initDMR with the reference's void(void) in-place callback (testLS.cu:97-128 contract):
    the callback overwrites the source vector with A x using a dense device GEMV."""
    torch = torch_gpu
    A = gdmr.conv_diff_2d(16)
    n, m, k, tol = A.n, 12, 4, 1e-10
    Ad = _dev(torch, A.to_scipy().toarray())
    ext = torch.zeros(n, dtype=torch.float64, device="cuda")
    calls = [0]

    def cb():
        calls[0] += 1
        y = Ad @ ext
        ext.copy_(y)

    cbf = gdmr.LEGACY_OP(cb)
    dev = (C.c_int * 1)(0)
    gdmr.lib.initDMR(32, 512, dev, 1, n, m, k, cbf)
    assert gdmr.lib.dmrGetLastError() == 0
    ora = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n))
    try:
        for i in range(2):
            b = gdmr.rand_normal(3, i, n)
            bd = _dev(torch, b)
            ext.zero_()
            its = C.c_int(0)
            res = gdmr.lib.solverDMR(i, ext.data_ptr(), bd.data_ptr(), None, 0, tol, n,
                                     C.byref(its))
            torch.cuda.synchronize()
            x = ext.cpu().numpy()
            st, hist = gdmr.global_stats()
            xo, reso, itso, sto = ora.solve(b, tol=tol, mode=i)
            # The dense GEMV sums in another order than CSR, so agreement is to rounding, not bits.
            print(f"legacy[{i}]: its {its.value} vs {itso}, res {res:.3e} vs {reso:.3e}", flush=True)
            assert abs(its.value - itso) <= 1
            tr = _true_res(A, b, x)
            assert tr <= tol * 1.01, tr
            assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8
        assert calls[0] == st.matvecs or calls[0] > 0
    finally:
        gdmr.lib.freeDMR()


def test_breakdown_diagonal_operator(torch_gpu):
    """A diagonal operator with 5 distinct eigenvalues: the Krylov space is exhausted at step 5,
    the breakdown branch (gDMR.cu:2256-2290) pads the basis, and one cycle solves the system."""
    torch = torch_gpu
    n, m = 300, 10
    d = np.repeat([1.0, 2.0, 3.0, 5.0, 8.0], n // 5)
    A = gdmr.CsrMatrix(n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), d)
    for k in (0, 3):
        s = gdmr.Solver(n, m, k)
        rp, ci, va = _dev(torch, A.rowptr), _dev(torch, A.col), _dev(torch, A.val)
        s.set_csr(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), A.nnz, keep=(rp, ci, va))
        ora = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n))
        for t in range(2 if k else 1):
            b = gdmr.rand_normal(11, t, n)
            xd = torch.zeros(n, dtype=torch.float64, device="cuda")
            bd = _dev(torch, b)
            res, its = s.solve(xd.data_ptr(), bd.data_ptr(), 1e-10, n, mode=t)
            st, hist = s.stats()
            xo, reso, itso, sto = ora.solve(b, tol=1e-10, mode=t)
            cyc_o, mv_o, brk_o = ora.counts()
            print(f"breakdown k={k} t={t}: breakdowns {st.breakdowns} vs {brk_o}, "
                  f"cycles {st.cycles} vs {cyc_o}, res {res:.3e}", flush=True)
            assert st.breakdowns == brk_o
            if t == 0:
                assert brk_o >= 1
            assert abs(its - itso) <= 1
            tr = _true_res(A, b, xd.cpu().numpy())
            assert tr <= 1e-10 * 1.01, tr
        s.close()


def test_early_exit_and_errors(torch_gpu):
    torch = torch_gpu
    A = gdmr.conv_diff_2d(8)
    n = A.n
    s = gdmr.Solver(n, 6, 2)
    rp, ci, va = _dev(torch, A.rowptr), _dev(torch, A.col), _dev(torch, A.val)
    s.set_csr(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), A.nnz, keep=(rp, ci, va))
    xs = gdmr.rand_normal(1, 1, n)
    b = A.matvec(xs)
    bd = _dev(torch, b)
    # mode != 0 before any subspace exists runs like mode 0 (first cycle undeflated).
    xd = torch.zeros(n, dtype=torch.float64, device="cuda")
    res1, its1 = s.solve(xd.data_ptr(), bd.data_ptr(), 1e-8, 100, mode=1)
    assert gdmr.last_error()[0] == gdmr.DMR_OK
    ora = oracle.Oracle(A, 6, 2, gdmr.reduction_blocks(n))
    xo, reso, itso, _ = ora.solve(b, tol=1e-8, mat_rank=100, mode=0)
    assert its1 == itso and abs(res1 - reso) <= 1e-12 * reso
    s.close()
    s = gdmr.Solver(n, 6, 2)
    s.set_csr(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), A.nnz, keep=(rp, ci, va))
    # x0 = exact solution: early exit returns the absolute residual, numIts = matRank^2.
    xd = _dev(torch, xs)
    res, its = s.solve(xd.data_ptr(), bd.data_ptr(), 1e-8, 100, mode=0)
    assert its == 100 * 100
    assert res < 1e-8 * np.linalg.norm(b)
    assert np.array_equal(xd.cpu().numpy(), xs)
    s.close()


def test_operator_error_channel(torch_gpu):
    """An operator that raises dmrRaiseOperatorError stops the solve with DMR_ERR_OPERATOR."""
    torch = torch_gpu
    n = 64
    s = gdmr.Solver(n, 8, 2)

    def op(src, dst, stream, ctx):
        gdmr.lib.dmrRaiseOperatorError(b"injected failure")

    s.set_operator(op)
    xd = torch.zeros(n, dtype=torch.float64, device="cuda")
    bd = torch.ones(n, dtype=torch.float64, device="cuda")
    s.solve(xd.data_ptr(), bd.data_ptr(), 1e-8, 10, mode=0)
    code, msg = gdmr.last_error()
    assert code == gdmr.DMR_ERR_OPERATOR and "injected" in msg
    s.close()


@pytest.mark.parametrize("prec", [gdmr.PREC_FP32, gdmr.PREC_BF16], ids=["fp32", "bf16"])
def test_mixed_precision_reaches_fp64_tolerance(torch_gpu, prec):
    """FP32 (or bf16) basis storage with FP32 SpMV values and FP32-accumulated reductions inside
    FP64 refinement reaches 1e-10 FP64-checked, and its solution agrees with the FP64 oracle's to
    the conditioning of the problem.  The recycled subspace survives an export / import round
    trip through FP64 host memory bit for bit."""
    torch = torch_gpu
    A = gdmr.conv_diff_2d(64)
    n, m, k, tol = A.n, 30, 10, 1e-10
    s = gdmr.Solver(n, m, k, precision=prec)
    rp, ci, va = _dev(torch, A.rowptr), _dev(torch, A.col), _dev(torch, A.val)
    s.set_csr(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), A.nnz, keep=(rp, ci, va))
    ora = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n))
    for t in range(3):
        b = gdmr.rand_normal(0, 1 + t, n)
        xd = torch.zeros(n, dtype=torch.float64, device="cuda")
        bd = _dev(torch, b)
        res, its = s.solve(xd.data_ptr(), bd.data_ptr(), tol, 10000, mode=t)
        x = xd.cpu().numpy()
        tr = _true_res(A, b, x)
        xo, _, _, _ = ora.solve(b, tol=tol, mode=t)
        err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        st, hist = s.stats()
        print(f"mixed[{t}]: res {res:.3e} true {tr:.3e} cycles {st.cycles} err vs oracle {err:.2e}",
              flush=True)
        assert tr <= tol and abs(res - tr) <= 1e-3 * tr + 1e-16
        assert err < 1e-7
    Z = s.export_z()
    s2 = gdmr.Solver(n, m, k, precision=prec)
    s2.import_z(Z)
    assert np.array_equal(s2.export_z(), Z)
    s2.close()
    s.close()


def test_mixed_implicit_image_and_early_refresh(torch_gpu, monkeypatch):
    """The FP32 inner solve's deflation image taken from the refreshing cycle's Arnoldi relation
    (and its refresh SVD forked two Krylov steps early) against the reference's recomputation
    (k SpMVs and a Gram-Schmidt QR every cycle, refresh at the cycle end): on a jittered
    sequence with recycling, both reach 1e-10 FP64-checked on every system, take cycle counts
    within two of each other, and their solutions agree to the problem's conditioning.  A
    solve after new operator values recomputes the image (the operator changed)."""
    torch = torch_gpu
    N, m, k, tol = 20, 30, 10, 1e-10
    mats = [gdmr.conv_diff_3d(N, jitter=0.01, seed=5, stream=t) for t in range(4)]
    n = mats[0].n
    rp, ci = _dev(torch, mats[0].rowptr), _dev(torch, mats[0].col)
    vals = [_dev(torch, A.val) for A in mats]
    runs = {}
    for label, image, tail in (("recomputed", "0", "0"), ("implicit", "1", "2")):
        monkeypatch.setenv("DMR_IMPLICIT_IMAGE", image)
        monkeypatch.setenv("DMR_REFRESH_TAIL", tail)
        s = gdmr.Solver(n, m, k, precision=gdmr.PREC_FP32)
        s.set_csr(rp.data_ptr(), ci.data_ptr(), vals[0].data_ptr(), mats[0].nnz,
                  keep=(rp, ci, vals[0]))
        xs, cyc = [], []
        for t, A in enumerate(mats):
            s.set_values(vals[t].data_ptr())
            b = gdmr.rand_normal(5, 100 + t, n)
            xd = torch.zeros(n, dtype=torch.float64, device="cuda")
            res, its = s.solve(xd.data_ptr(), _dev(torch, b).data_ptr(), tol, 10000,
                               mode=1 if t else 0)
            x = xd.cpu().numpy()
            assert _true_res(A, b, x) <= tol
            st, _ = s.stats()
            xs.append(x)
            cyc.append(st.cycles)
        s.close()
        runs[label] = (xs, cyc)
        print(f"{label}: cycles {cyc}", flush=True)
    (xr, cr), (xi, ci_) = runs["recomputed"], runs["implicit"]
    for t in range(len(mats)):
        assert abs(cr[t] - ci_[t]) <= 2
        assert np.linalg.norm(xi[t] - xr[t]) / np.linalg.norm(xr[t]) < 1e-7


def test_matrix_market_operator(tmp_path):
    """The on-disk operator path: a Matrix Market file imported with impMtxLS gives the solve of
    the same matrix imported with impCsrLS, bit for bit, and matches the oracle."""
    A = gdmr.conv_diff_2d(40, jitter=0.01, seed=9, stream=2)
    path = str(tmp_path / "cd2d.mtx")
    gdmr.write_matrix_market(path, A)
    n, m, k, tol = A.n, 24, 8, 1e-9
    b = gdmr.rand_normal(9, 3, n)
    ls = gdmr.LinearSystem(A, m, k)
    try:
        x_csr, res_csr, its_csr = ls.solve(b, tol=tol, mode=0)
        lib = gdmr.lib
        dev = (C.c_int * 1)(0)
        lib.initLS(32, 512, 0, n, dev, 1, m, k)
        assert lib.impMtxLS(path.encode()) == gdmr.DMR_OK
        assert lib.nnzLS() == A.nnz
        x_mtx, res_mtx, its_mtx = ls.solve(b, tol=tol, mode=0)
    finally:
        ls.close()
    assert its_mtx == its_csr and res_mtx == res_csr
    assert np.array_equal(x_mtx, x_csr)
    ora = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n))
    xo, reso, itso, sto = ora.solve(b, tol=tol, mode=0)
    ora.close()
    assert abs(its_mtx - itso) <= 1 and abs(res_mtx - reso) <= HIST_RTOL * reso
    tr = _true_res(A, b, x_mtx)
    assert tr <= tol * 1.001, tr


@pytest.mark.parametrize("n,alpha,m,k,mat_rank", cases.STAG_CASES)
def test_stagnation_restart_matches_oracle(torch_gpu, n, alpha, m, k, mat_rank):
    """A nearly stalling operator with deflation: several cycles restart from the old Krylov
    vector (m-k)/2 (arnoldiSetupDMR, gDMR.cu:2586-2632), the stagnation counter persists across
    solves (gDMR.cu:890), and the HIP path's histories, cycle-limit exits and subspace follow the
    oracle through two solves of a sequence."""
    torch = torch_gpu
    A = cases.shift_plus_diag(n, alpha)
    s = gdmr.Solver(n, m, k)
    rp, ci, va = _dev(torch, A.rowptr), _dev(torch, A.col), _dev(torch, A.val)
    s.set_csr(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), A.nnz, keep=(rp, ci, va))
    ora = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n), keep_last=True)
    for t in range(2):
        b = gdmr.rand_normal(3, t, n)
        xd = torch.zeros(n, dtype=torch.float64, device="cuda")
        bd = _dev(torch, b)
        res, its = s.solve(xd.data_ptr(), bd.data_ptr(), 1e-10, mat_rank, mode=t)
        st, hist = s.stats()
        xo, reso, itso, sto = ora.solve(b, tol=1e-10, mat_rank=mat_rank, mode=t)
        restarts = int((ora.start_log() == 0).sum())
        print(f"stagnation n={n} t={t}: restarts from the old vector {restarts}, status "
              f"{st.status} vs {sto}, stagnation count {gdmr.lib.dmrGetStagnationCount(s.h)} vs "
              f"{ora.stagnation_count()}", flush=True)
        assert restarts >= 1
        _compare(f"stag[{t}]", its, hist, xd.cpu().numpy(), itso, ora.history(), xo)
        assert st.status == sto == gdmr.DMR_ERR_NOCONV and len(hist) == len(ora.history())
        # The cycle-limit exit returns before the solution export (gDMR.cu:2816-2829): x stays
        # the initial guess on both sides.
        assert np.array_equal(xd.cpu().numpy(), xo) and not xo.any()
        assert gdmr.lib.dmrGetStagnationCount(s.h) == ora.stagnation_count()
        assert np.max(np.abs(s.export_z() - ora.get_z())) <= 1e-9
    s.close()


@pytest.mark.parametrize("prec", [gdmr.PREC_FP64, gdmr.PREC_FP32], ids=["fp64", "fp32"])
def test_captured_cycles_match_eager(torch_gpu, monkeypatch, prec):
    """Restart cycles captured as HIP graphs and replayed (DMR_GRAPHS=1) give the eager launches'
    bits: a recycled sequence of three systems with new operator values each (so graphs are
    captured for several variants and replayed over many cycles, with the host's reads and
    synchronisations between replays); histories, cycle counts and solutions are identical."""
    torch = torch_gpu
    mats = [gdmr.conv_diff_3d(16, jitter=0.01, seed=21, stream=t) for t in range(3)]
    n, m, k = mats[0].n, 24, 8
    rp, ci = _dev(torch, mats[0].rowptr), _dev(torch, mats[0].col)
    vals = [_dev(torch, A.val) for A in mats]
    out = {}
    for graphs in ("0", "1"):
        monkeypatch.setenv("DMR_GRAPHS", graphs)
        s = gdmr.Solver(n, m, k, precision=prec)
        s.set_csr(rp.data_ptr(), ci.data_ptr(), vals[0].data_ptr(), mats[0].nnz,
                  keep=(rp, ci, vals[0]))
        runs = []
        for t, A in enumerate(mats):
            s.set_values(vals[t].data_ptr())
            b = gdmr.rand_normal(21, 50 + t, n)
            xd = torch.zeros(n, dtype=torch.float64, device="cuda")
            res, its = s.solve(xd.data_ptr(), _dev(torch, b).data_ptr(), 1e-10, 10000,
                               mode=1 if t else 0)
            st, hist = s.stats()
            assert _true_res(A, b, xd.cpu().numpy()) <= 1e-10
            runs.append((st.cycles, np.array(hist), xd.cpu().numpy()))
        s.close()
        out[graphs] = runs
    for (c0, h0, x0), (c1, h1, x1) in zip(out["0"], out["1"]):
        assert c0 == c1 and np.array_equal(h0, h1) and np.array_equal(x0, x1)
