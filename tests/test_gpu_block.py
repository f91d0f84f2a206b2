"""Block (multi-right-hand-side) DMR on the GPU (dmrBlock* in gDMR.h).

The block solver runs FP32 storage and matrix-core arithmetic inside FP64 iterative refinement,
so it is not bitwise comparable with the FP64 oracle.  The bar, per column:
  * the true FP64 residual ||b - A x|| / ||b|| (recomputed on the host) meets the tolerance;
  * the solution agrees with the FP64 oracle's solution of the same system to the accuracy the
    residual allows: ||x - x_oracle|| / ||x_oracle|| <= XTOL (stated below, measured with margin
    on these systems);
  * results are bitwise reproducible (fixed-order reductions, deterministic MFMA) from run to run.
"""
import numpy as np
import pytest

import gdmr
import oracle

pytestmark = pytest.mark.gpu

XTOL = 1e-6


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _block(torch, A, B, mb, k, smax=None, ld=None, tol=1e-10, mode=0, solver=None, x0=None,
           prec=None, inner_tol=None):
    """Solves A X = B (B: s x n host array, one right-hand side per row) with a block handle."""
    s, n = B.shape
    ld = ld or n
    Bp = np.zeros((s, ld))
    Bp[:, :n] = B
    Bd = _dev(torch, Bp)
    X = np.zeros((s, ld))
    if x0 is not None:
        X[:, :n] = x0
    Xd = _dev(torch, X)
    own = solver is None
    if own:
        solver = gdmr.BlockSolver(n, mb, k, smax or s)
        rp, ci, va = _dev(torch, A.rowptr), _dev(torch, A.col), _dev(torch, A.val)
        solver.set_csr(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), A.nnz, keep=(rp, ci, va))
        if prec is not None:
            solver.set_precision(prec)
        if inner_tol is not None:
            solver.set_refinement(inner_tol, 60)
    torch.cuda.synchronize()
    rc, rel = solver.solve(Xd.data_ptr(), ld, Bd.data_ptr(), ld, s, tol, max_cycles=500, mode=mode)
    st, hist = solver.stats()
    Xh = Xd.cpu().numpy()[:, :n].copy()
    if own:
        solver.close()
    return rc, rel, st, Xh


def _true_res(A, b, x):
    nb = np.linalg.norm(b)
    return np.linalg.norm(b - A.matvec(x)) / (nb if nb > 0 else 1.0)


def test_block_matches_oracle(torch_gpu):
    """Anisotropic 3-D diffusion 16^3 (coefficients 1, 0.1, 0.01), 8 right-hand sides, 4 block
    steps, k = 8: every column meets 1e-10 in FP64 and matches the FP64 oracle's solution."""
    A = gdmr.aniso_3d(16, coef=(1.0, 0.1, 0.01))
    n, s, tol = A.n, 8, 1e-10
    B = np.stack([gdmr.rand_normal(31, j, n) for j in range(s)])
    rc, rel, st, X = _block(torch_gpu, A, B, mb=4, k=8)
    print(f"block: rc {rc} cycles {st.cycles} breakdowns {st.breakdowns} refreshes {st.svdCalls} "
          f"max rel {rel.max():.2e}", flush=True)
    assert rc == 0
    ora = oracle.Oracle(A, 30, 8, gdmr.reduction_blocks(n))
    try:
        for j in range(s):
            tr = _true_res(A, B[j], X[j])
            assert tr <= tol * 1.001, (j, tr)
            assert abs(tr - rel[j]) <= 1e-3 * tol
            xo, reso, itso, sto = ora.solve(B[j], tol=1e-12, mode=j)
            err = np.linalg.norm(X[j] - xo) / np.linalg.norm(xo)
            print(f"  column {j}: true residual {tr:.2e}, error vs oracle {err:.2e}", flush=True)
            assert err <= XTOL, (j, err)
    finally:
        ora.close()


def test_block_bitwise_reproducible(torch_gpu):
    A = gdmr.conv_diff_3d(20, jitter=0.01, seed=5, stream=1)
    B = np.stack([gdmr.rand_normal(32, j, A.n) for j in range(12)])
    r1 = _block(torch_gpu, A, B, mb=5, k=12, smax=16)
    r2 = _block(torch_gpu, A, B, mb=5, k=12, smax=16)
    assert r1[0] == 0 and r2[0] == 0
    assert r1[2].cycles == r2[2].cycles
    assert np.array_equal(r1[3], r2[3])
    assert np.array_equal(r1[1], r2[1])


def test_block_shapes_recycling_and_padding(torch_gpu):
    """Partial block (5 of 32 columns), padded leading dimension, k = 0, and a recycled second
    block started from the subspace of the first (mode 1) on jittered coefficients."""
    torch = torch_gpu
    A0 = gdmr.conv_diff_3d(24, jitter=0.01, seed=8, stream=0)
    A1 = gdmr.conv_diff_3d(24, jitter=0.01, seed=8, stream=1)
    n, tol = A0.n, 1e-10
    B0 = np.stack([gdmr.rand_normal(33, j, n) for j in range(5)])
    B1 = np.stack([gdmr.rand_normal(34, j, n) for j in range(5)])
    # k = 0, padded ld.
    rc, rel, st, X = _block(torch, A0, B0, mb=6, k=0, smax=32, ld=n + 100)
    assert rc == 0 and st.svdCalls == 0
    for j in range(5):
        assert _true_res(A0, B0[j], X[j]) <= tol * 1.001
    # Recycling across two blocks of a sequence (same pattern, new values).
    bs = gdmr.BlockSolver(n, 6, 16, 32)
    rp, ci = _dev(torch, A0.rowptr), _dev(torch, A0.col)
    v0, v1 = _dev(torch, A0.val), _dev(torch, A1.val)
    bs.set_csr(rp.data_ptr(), ci.data_ptr(), v0.data_ptr(), A0.nnz, keep=(rp, ci, v0))
    rc0, rel0, st0, X0 = _block(torch, A0, B0, 6, 16, solver=bs)
    assert rc0 == 0 and bs.has_z()
    bs.set_values(v1.data_ptr(), keep=(v1,))
    rc1, rel1, st1, X1 = _block(torch, A1, B1, 6, 16, solver=bs, mode=1)
    bs.close()
    assert rc1 == 0
    print(f"first block {st0.cycles} cycles, recycled block {st1.cycles} cycles", flush=True)
    for j in range(5):
        assert _true_res(A1, B1[j], X1[j]) <= tol * 1.001


def test_block_dependent_and_zero_columns(torch_gpu):
    """Duplicate right-hand sides make the block rank deficient (dependent columns are dropped
    from the basis); a zero right-hand side has the zero solution; a nonzero initial guess is
    refined from."""
    A = gdmr.conv_diff_2d(48, jitter=0.01, seed=3, stream=2)
    n, tol = A.n, 1e-10
    b = gdmr.rand_normal(35, 0, n)
    c = gdmr.rand_normal(35, 1, n)
    B = np.stack([b, c, b, np.zeros(n), 2.0 * c])
    x0 = np.zeros((5, n))
    x0[1] = 0.5 * gdmr.rand_normal(36, 0, n)
    rc, rel, st, X = _block(torch_gpu, A, B, mb=6, k=8, x0=x0)
    print(f"rc {rc} cycles {st.cycles} breakdowns {st.breakdowns} rel {rel}", flush=True)
    assert rc == 0 and st.breakdowns > 0
    for j in (0, 1, 2, 4):
        assert _true_res(A, B[j], X[j]) <= tol * 1.001, j
    assert np.all(X[3] == 0.0) and rel[3] == 0.0
    assert np.linalg.norm(X[0] - X[2]) <= 1e-8 * np.linalg.norm(X[0])


def test_block_rejects_bad_shapes(torch_gpu):
    with pytest.raises(gdmr.SolverError):
        gdmr.BlockSolver(1000, 12, 32, 32)  # 32 + 13 * 32 > 384 basis columns
    with pytest.raises(gdmr.SolverError):
        gdmr.BlockSolver(1000, 4, 8, 33)
    bs = gdmr.BlockSolver(1000, 4, 8, 8)
    rc, rel = bs.solve(0, 1000, 0, 1000, 4, 1e-8)
    assert rc == gdmr.DMR_ERR_STATE  # no operator
    bs.close()


def test_block_graph_replays_match_eager(torch_gpu):
    """Restart cycles replayed from captured HIP graphs give the eager launches' bits: three
    solves in a row (each replaying its cycle graphs many times, with the eager refinement work
    between the replays) on a graph handle and on an eager handle of the same (tuned) shape."""
    torch = torch_gpu
    A = gdmr.conv_diff_3d(24, jitter=0.01, seed=9, stream=3)
    n, s = A.n, 8
    rp, ci, va = _dev(torch, A.rowptr), _dev(torch, A.col), _dev(torch, A.val)
    hs = []
    for graphs in (1, 0):
        bs = gdmr.BlockSolver(n, 6, 16, s)
        bs.set_csr(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), A.nnz)
        bs.set_options(graphs=graphs)
        hs.append(bs)
    for t in range(3):
        B = np.stack([gdmr.rand_normal(37, 10 * t + j, n) for j in range(s)])
        outs = [_block(torch, A, B, 6, 16, solver=bs, mode=t) for bs in hs]
        (rc_g, rel_g, st_g, X_g), (rc_e, rel_e, st_e, X_e) = outs
        assert rc_g == 0 and rc_e == 0
        assert st_g.cycles == st_e.cycles and st_g.matvecs == st_e.matvecs
        assert np.array_equal(X_g, X_e), f"solve {t}: graph replay differs from eager launches"
        for j in range(s):
            assert _true_res(A, B[j], X_g[j]) <= 1e-10 * 1.001
    tg, te = hs[0].tuning(), hs[1].tuning()
    print(f"tuning {tg}; eager handle {te}", flush=True)
    assert tg["graphs"] > 0 and te["graphs"] == 0
    assert {k: v for k, v in tg.items() if k != "graphs"} == \
        {k: v for k, v in te.items() if k != "graphs"}
    for bs in hs:
        bs.close()


def _column_operator_solve(torch, mats, B, mb, k, inner_mean, tol=1e-10, solver=None,
                           prec=None, inner_tol=None):
    """Solves A_t x_t = b_t for a sequence of systems with one sparsity pattern as one block:
    column t uses its own FP64 operator (dmrBlockSetColumnOperators)."""
    s, n = B.shape
    nnz = mats[0].nnz
    vals = _dev(torch, np.stack([A.val for A in mats]))
    rp, ci = _dev(torch, mats[0].rowptr), _dev(torch, mats[0].col)
    bs = solver or gdmr.BlockSolver(n, mb, k, s)
    bs.set_csr(rp.data_ptr(), ci.data_ptr(), vals[0].data_ptr(), nnz, keep=(rp, ci))
    bs.set_column_operators(vals.data_ptr(), nnz, s, inner_mean=inner_mean, keep=(vals,))
    if prec is not None:
        bs.set_precision(prec)
    if inner_tol is not None:
        bs.set_refinement(inner_tol, 60)
    Bd = _dev(torch, B)
    Xd = torch.zeros(s, n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc, rel = bs.solve(Xd.data_ptr(), n, Bd.data_ptr(), n, s, tol, max_cycles=500)
    st, _ = bs.stats()
    X = Xd.cpu().numpy()
    if solver is None:
        bs.close()
    return rc, rel, st, X


@pytest.mark.parametrize("inner_mean,prec", [(True, "fp32"), (False, "fp32"), (True, "bf16")])
def test_column_operators_solve_a_jittered_sequence(torch_gpu, inner_mean, prec):
    """Six systems of a jittered 3-D convection-diffusion sequence (16^3, per-cell diffusion
    U(+-1%) per system) solved as one block: column t is refined against its own FP64 operator,
    the inner FP32 solve shares the mean operator (or system 0's).  Every column meets 1e-10 with
    its own operator and matches the FP64 oracle's solution of its own system to XTOL."""
    mats = [gdmr.conv_diff_3d(16, jitter=0.01, seed=41, stream=t) for t in range(6)]
    n, s, tol = mats[0].n, len(mats), 1e-10
    B = np.stack([gdmr.rand_normal(42, t, n) for t in range(s)])
    p = gdmr.PREC_BF16 if prec == "bf16" else gdmr.PREC_FP32
    rc, rel, st, X = _column_operator_solve(torch_gpu, mats, B, mb=6, k=8, inner_mean=inner_mean,
                                            prec=p, inner_tol=3e-2 if prec == "bf16" else None)
    print(f"inner_mean {inner_mean} {prec}: rc {rc} cycles {st.cycles} max rel {rel.max():.2e}",
          flush=True)
    assert rc == 0
    for t, A in enumerate(mats):
        tr = _true_res(A, B[t], X[t])
        assert tr <= tol * 1.001, (t, tr)
        assert abs(tr - rel[t]) <= 1e-3 * tol
        # Against another system's operator the residual is far above the tolerance.
        other = mats[(t + 1) % s]
        assert _true_res(other, B[t], X[t]) > 1e3 * tol
        ora = oracle.Oracle(A, 30, 8, gdmr.reduction_blocks(n))
        try:
            xo, _, _, _ = ora.solve(B[t], tol=1e-12, mode=0)
        finally:
            ora.close()
        err = np.linalg.norm(X[t] - xo) / np.linalg.norm(xo)
        print(f"  system {t}: true residual {tr:.2e}, error vs oracle {err:.2e}", flush=True)
        assert err <= XTOL, (t, err)


def test_column_operators_equal_to_the_operator_match_plain_block(torch_gpu):
    """Column operators that all equal the block operator (inner operator unchanged) give the
    plain block solve's bits; clearing them restores the plain behaviour on the same handle."""
    torch = torch_gpu
    A = gdmr.conv_diff_3d(18, jitter=0.01, seed=43, stream=0)
    n, s = A.n, 5
    B = np.stack([gdmr.rand_normal(44, j, n) for j in range(s)])
    rc0, rel0, st0, X0 = _block(torch, A, B, mb=5, k=8)
    rc1, rel1, st1, X1 = _column_operator_solve(torch, [A] * s, B, mb=5, k=8, inner_mean=False)
    assert rc0 == 0 and rc1 == 0
    assert st0.cycles == st1.cycles
    assert np.array_equal(X0, X1) and np.array_equal(rel0, rel1)


def test_column_operators_reject_bad_arguments(torch_gpu):
    torch = torch_gpu
    mats = [gdmr.conv_diff_3d(10, jitter=0.01, seed=45, stream=t) for t in range(3)]
    n, nnz = mats[0].n, mats[0].nnz
    vals = _dev(torch, np.stack([A.val for A in mats]))
    rp, ci = _dev(torch, mats[0].rowptr), _dev(torch, mats[0].col)
    bs = gdmr.BlockSolver(n, 4, 4, 4)
    with pytest.raises(gdmr.SolverError):  # no operator yet
        bs.set_column_operators(vals.data_ptr(), nnz, 3)
    bs.set_csr(rp.data_ptr(), ci.data_ptr(), vals[0].data_ptr(), nnz)
    with pytest.raises(gdmr.SolverError):  # more than maxRhs
        bs.set_column_operators(vals.data_ptr(), nnz, 5)
    with pytest.raises(gdmr.SolverError):  # ldv < nnz
        bs.set_column_operators(vals.data_ptr(), nnz - 1, 3)
    bs.set_column_operators(vals.data_ptr(), nnz, 3)
    B = _dev(torch, np.stack([gdmr.rand_normal(46, j, n) for j in range(4)]))
    X = torch.zeros(4, n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc, _ = bs.solve(X.data_ptr(), n, B.data_ptr(), n, 4, 1e-10)  # 4 columns, 3 operators
    assert rc == gdmr.DMR_ERR_ARG
    bs.set_column_operators(0, 0, 0)  # cleared: a 4-column solve with the block operator
    rc, rel = bs.solve(X.data_ptr(), n, B.data_ptr(), n, 4, 1e-10, max_cycles=500)
    assert rc == 0
    bs.close()


def test_bf16_basis_matches_oracle_and_fp32_block(torch_gpu):
    """bf16 basis storage (GEMMs and SpMM read bf16, compute in FP32 on the matrix cores): the
    anisotropic 16^3 case of test_block_matches_oracle meets 1e-10 in FP64 per column and agrees
    with the FP64 oracle's solutions (and so with the FP32-basis block solve) to XTOL."""
    A = gdmr.aniso_3d(16, coef=(1.0, 0.1, 0.01))
    n, s, tol = A.n, 8, 1e-10
    B = np.stack([gdmr.rand_normal(31, j, n) for j in range(s)])
    rc, rel, st, X = _block(torch_gpu, A, B, mb=4, k=8, prec=gdmr.PREC_BF16, inner_tol=3e-2)
    rc32, rel32, st32, X32 = _block(torch_gpu, A, B, mb=4, k=8)
    print(f"bf16: rc {rc} cycles {st.cycles}; fp32: rc {rc32} cycles {st32.cycles}", flush=True)
    assert rc == 0 and rc32 == 0
    ora = oracle.Oracle(A, 30, 8, gdmr.reduction_blocks(n))
    try:
        for j in range(s):
            tr = _true_res(A, B[j], X[j])
            assert tr <= tol * 1.001, (j, tr)
            xo, _, _, _ = ora.solve(B[j], tol=1e-12, mode=j)
            err = np.linalg.norm(X[j] - xo) / np.linalg.norm(xo)
            err32 = np.linalg.norm(X[j] - X32[j]) / np.linalg.norm(X32[j])
            print(f"  column {j}: true residual {tr:.2e}, vs oracle {err:.2e}, vs fp32 block "
                  f"{err32:.2e}", flush=True)
            assert err <= XTOL and err32 <= XTOL, (j, err, err32)
    finally:
        ora.close()


def test_bf16_basis_dependent_columns_graphs_and_reproducibility(torch_gpu):
    """bf16 basis: dependent and zero right-hand sides (rank-deficient blocks), graph replays
    against eager launches, and bitwise reproducibility."""
    torch = torch_gpu
    A = gdmr.conv_diff_2d(48, jitter=0.01, seed=3, stream=2)
    n, tol = A.n, 1e-10
    b = gdmr.rand_normal(35, 0, n)
    c = gdmr.rand_normal(35, 1, n)
    B = np.stack([b, c, b, np.zeros(n), 2.0 * c])
    rc, rel, st, X = _block(torch, A, B, mb=6, k=8, prec=gdmr.PREC_BF16, inner_tol=3e-2)
    print(f"rc {rc} cycles {st.cycles} breakdowns {st.breakdowns}", flush=True)
    assert rc == 0 and st.breakdowns > 0
    for j in (0, 1, 2, 4):
        assert _true_res(A, B[j], X[j]) <= tol * 1.001, j
    assert np.all(X[3] == 0.0) and rel[3] == 0.0
    A3 = gdmr.conv_diff_3d(24, jitter=0.01, seed=9, stream=3)
    n3, s = A3.n, 8
    rp, ci, va = _dev(torch, A3.rowptr), _dev(torch, A3.col), _dev(torch, A3.val)
    hs = []
    for graphs in (1, 0, 1):
        bs = gdmr.BlockSolver(n3, 6, 16, s)
        bs.set_csr(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), A3.nnz)
        bs.set_precision(gdmr.PREC_BF16)
        bs.set_refinement(3e-2, 60)
        bs.set_options(graphs=graphs)
        hs.append(bs)
    for t in range(2):
        B3 = np.stack([gdmr.rand_normal(37, 10 * t + j, n3) for j in range(s)])
        outs = [_block(torch, A3, B3, 6, 16, solver=bs, mode=t) for bs in hs]
        for rc_, rel_, st_, X_ in outs:
            assert rc_ == 0
            for j in range(s):
                assert _true_res(A3, B3[j], X_[j]) <= 1e-10 * 1.001
        assert np.array_equal(outs[0][3], outs[1][3]), "graph replay differs from eager launches"
        assert np.array_equal(outs[0][3], outs[2][3]), "not reproducible"
    for bs in hs:
        bs.close()


def test_precision_rejects_fp64(torch_gpu):
    bs = gdmr.BlockSolver(1000, 4, 8, 8)
    with pytest.raises(gdmr.SolverError):
        bs.set_precision(gdmr.PREC_FP64)
    bs.set_precision(gdmr.PREC_BF16)
    bs.set_precision(gdmr.PREC_FP32)
    bs.close()


def test_inner_stagnation_returns_to_refinement(torch_gpu):
    """A cyclic shift with unit-vector right-hand sides e_0, e_300: every Krylov block is
    orthogonal to the right-hand sides until the space wraps around, so restarted block GMRES
    makes exactly no progress.  The solve ends with DMR_ERR_NOCONV at the refinement limit: each
    inner solve gives up after 25 cycles without progress instead of running to the cycle cap."""
    n = 600
    A = gdmr.CsrMatrix(n, np.arange(n + 1, dtype=np.int32),
                       ((np.arange(n) + 1) % n).astype(np.int32), np.ones(n))
    B = np.zeros((2, n))
    B[0, 0] = 1.0
    B[1, 300] = 1.0
    solver = gdmr.BlockSolver(n, 4, 4, 2)
    rp, ci, va = _dev(torch_gpu, A.rowptr), _dev(torch_gpu, A.col), _dev(torch_gpu, A.val)
    solver.set_csr(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), A.nnz, keep=(rp, ci, va))
    solver.set_refinement(1e-2, 3)
    Bd = _dev(torch_gpu, B)
    Xd = torch_gpu.zeros(2, n, dtype=torch_gpu.float64, device="cuda")
    torch_gpu.cuda.synchronize()
    rc, rel = solver.solve(Xd.data_ptr(), n, Bd.data_ptr(), n, 2, 1e-10, max_cycles=1000)
    st, hist = solver.stats()
    solver.close()
    print(f"rc {rc} cycles {st.cycles} rel {rel}", flush=True)
    assert rc == gdmr.DMR_ERR_NOCONV
    assert st.cycles <= 4 * 26
    assert np.all(np.isfinite(rel))


def _ragged_rows(N=24, seed=5):
    """conv_diff_3d(N) with extra off-diagonal entries: rows [0, 4096) keep the 7-point stencil,
    every 16th row of [4096, 8192) gets 3 more entries (long rows inside workgroups whose
    nonzeros still fit the LDS stage), every row past 8192 gets 2 more (workgroups that do not
    fit).  Extra values are +-0.005; columns are distinct within a row."""
    A = gdmr.conv_diff_3d(N)
    rng = np.random.default_rng(seed)
    rows_c, rows_v = [], []
    for i in range(A.n):
        c = list(A.col[A.rowptr[i]:A.rowptr[i + 1]])
        v = list(A.val[A.rowptr[i]:A.rowptr[i + 1]])
        extra = 3 if (4096 <= i < 8192 and i % 16 == 0) else (2 if i >= 8192 else 0)
        while extra:
            j = int(rng.integers(A.n))
            if j not in c:
                c.append(j)
                v.append(float(rng.choice([-0.005, 0.005])))
                extra -= 1
        rows_c.append(c)
        rows_v.append(v)
    rp = np.zeros(A.n + 1, dtype=np.int32)
    rp[1:] = np.cumsum([len(c) for c in rows_c])
    return gdmr.CsrMatrix(A.n, rp, np.concatenate(rows_c).astype(np.int32),
                          np.concatenate(rows_v).astype(np.float64))


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_spmm_variants_identical_on_ragged_rows(torch_gpu, monkeypatch, prec):
    """Every SpMM variant (one thread per row with 4, 8 or 16 columns; indices loaded up front;
    indices staged in LDS) sums each row's entries in order, so block solves that differ only in
    the pinned variant give identical bits, on rows of 7 to 10 entries that take every path of
    the staged kernel.  Each solve meets 1e-10 in FP64."""
    A = _ragged_rows()
    n, s = A.n, 4
    B = np.stack([gdmr.rand_normal(77, j, n) for j in range(s)])
    monkeypatch.setenv("DMR_BLOCK_TUNE", "0")
    codes = [104, 108, 116, 208, 216] + ([4, 8, 16] if prec == "fp32" else [])
    ref = None
    for code in codes:
        monkeypatch.setenv("DMR_BLOCK_SPMM", str(code))
        rc, rel, st, X = _block(torch_gpu, A, B, mb=4, k=4,
                                prec=gdmr.PREC_BF16 if prec == "bf16" else None,
                                inner_tol=3e-2 if prec == "bf16" else None)
        worst = max(_true_res(A, B[j], X[j]) for j in range(s))
        print(f"{prec} spmm {code}: rc {rc} cycles {st.cycles} worst true residual {worst:.2e}",
              flush=True)
        assert rc == 0 and worst <= 1e-10 * 1.001
        if ref is None:
            ref = X
        else:
            assert np.array_equal(X, ref), code


@pytest.mark.parametrize("s", [10, 16, 20, 32])
def test_lsq_register_path_matches_window_path(torch_gpu, monkeypatch, s):
    """The block least squares on its register path (one thread per column of [H | T], on a
    16-row window for bandwidths up to 16 and a 32-row one up to 32) and on its LDS-window path
    (DMR_LSQ_WINDOW=1) solve the same systems:
    both meet 1e-10 in FP64 and their solutions agree to the accuracy the residual allows."""
    A = gdmr.conv_diff_3d(16, jitter=0.01, seed=12, stream=0)
    n = A.n
    B = np.stack([gdmr.rand_normal(91, j, n) for j in range(s)])
    X = {}
    for path in ("register", "window"):
        monkeypatch.setenv("DMR_LSQ_WINDOW", "1" if path == "window" else "0")
        rc, rel, st, X[path] = _block(torch_gpu, A, B, mb=6, k=8)
        worst = max(_true_res(A, B[j], X[path][j]) for j in range(s))
        print(f"s {s} {path}: rc {rc} cycles {st.cycles} worst true residual {worst:.2e}",
              flush=True)
        assert rc == 0 and worst <= 1e-10 * 1.001
    err = max(np.linalg.norm(X["register"][j] - X["window"][j]) / np.linalg.norm(X["window"][j])
              for j in range(s))
    print(f"s {s}: max relative difference {err:.2e}", flush=True)
    assert err <= XTOL


@pytest.mark.parametrize("prec", ["fp32", "bf16"])
def test_cycle_budget_bounds_the_whole_solve(torch_gpu, prec):
    """maxCycles bounds the restart cycles of the whole block solve, summed over its refinement
    steps: a solve that needs more cycles than its budget (2-D convection-diffusion 32^2, 4 block
    steps, 1e-10) returns DMR_ERR_NOCONV after at most the budget, with X holding the last refined
    iterate (finite residuals below the right-hand sides'), and the same handle then solves to
    the tolerance with a large budget."""
    torch = torch_gpu
    A = gdmr.conv_diff_2d(32, jitter=0.01, seed=71)
    n, s, budget = A.n, 4, 3
    B = np.stack([gdmr.rand_normal(72, j, n) for j in range(s)])
    solver = gdmr.BlockSolver(n, 4, 8, s)
    rp, ci, va = _dev(torch, A.rowptr), _dev(torch, A.col), _dev(torch, A.val)
    solver.set_csr(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), A.nnz, keep=(rp, ci, va))
    if prec == "bf16":
        solver.set_precision(gdmr.PREC_BF16)
        solver.set_refinement(3e-2, 60)
    else:
        solver.set_refinement(1e-2, 60)
    Bd = _dev(torch, B)
    Xd = torch.zeros(s, n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc, rel = solver.solve(Xd.data_ptr(), n, Bd.data_ptr(), n, s, 1e-10, max_cycles=budget)
    st, hist = solver.stats()
    print(f"{prec}: budget {budget}: rc {rc} cycles {st.cycles} rel {rel.max():.2e}", flush=True)
    assert rc == gdmr.DMR_ERR_NOCONV
    assert st.cycles <= budget
    assert np.all(np.isfinite(rel)) and rel.max() < 1.0
    X = Xd.cpu().numpy()
    for j in range(s):
        assert abs(_true_res(A, B[j], X[j]) - rel[j]) <= 1e-6 * rel[j]
    rc, rel = solver.solve(Xd.data_ptr(), n, Bd.data_ptr(), n, s, 1e-10, max_cycles=5000,
                           mode=1)
    st, hist = solver.stats()
    print(f"{prec}: large budget: rc {rc} cycles {st.cycles} rel {rel.max():.2e}", flush=True)
    assert rc == 0 and rel.max() <= 1e-10
    solver.close()


@pytest.mark.parametrize("s", [16, 20])
def test_bf16_block_fast_paths_on_and_off(torch_gpu, monkeypatch, s):
    """A bf16 block of 16 or 20 columns with the bf16 matrix-core update and projection pinned
    (DMR_BLOCK_NN_BASIS, DMR_BLOCK_TN_VAR), once with the folded residual projection and the Gram
    matrix from the update's epilogue (DMR_BLOCK_FOLD_T, DMR_BLOCK_FUSED_GRAM) and once without,
    and with the bf16 update at 128 and 256 rows per wavefront: every column meets 1e-10 in FP64
    and the solutions agree to XTOL with the plain configuration's.  At s = 20 the folded paths do
    not apply (they need s <= 16): the two-tile kernels and the window least squares run instead."""
    torch = torch_gpu
    A = gdmr.conv_diff_3d(24, jitter=0.01, seed=21)
    n = A.n
    B = np.stack([gdmr.rand_normal(22, j, n) for j in range(s)])
    sols = {}
    for nn_basis in ("100", "101"):
        for fused in ("0", "1"):
            monkeypatch.setenv("DMR_BLOCK_NN_BASIS", nn_basis)
            monkeypatch.setenv("DMR_BLOCK_TN_VAR", "3")
            monkeypatch.setenv("DMR_BLOCK_FOLD_T", fused)
            monkeypatch.setenv("DMR_BLOCK_FUSED_GRAM", fused)
            rc, rel, st, X = _block(torch, A, B, mb=5, k=16, prec=gdmr.PREC_BF16, inner_tol=3e-2)
            print(f"s {s} nn {nn_basis} fused {fused}: rc {rc} cycles {st.cycles} "
                  f"max rel {rel.max():.2e}", flush=True)
            assert rc == 0
            for j in range(s):
                assert _true_res(A, B[j], X[j]) <= 1e-10 * 1.001, j
            sols[(nn_basis, fused)] = X
    base = sols[("100", "0")]
    for key, X in sols.items():
        err = np.max(np.linalg.norm(X - base, axis=1) / np.linalg.norm(base, axis=1))
        print(f"  {key} vs plain: {err:.2e}", flush=True)
        assert err <= XTOL, key


def test_tuning_file_persists_launch_shapes(torch_gpu, tmp_path):
    """Launch shapes from a tuning file are used as they stand (no tuning run), and a shape the
    file lacks is tuned and appended to it; lines of another device architecture are skipped."""
    torch = torch_gpu
    arch = torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    A = gdmr.conv_diff_3d(13, jitter=0.01, seed=31, stream=0)  # a shape no other test tunes
    n, k, mb = A.n, 6, 5
    path = tmp_path / "tune.txt"
    # FP32 basis, s = 4: projection rows 256 / variant 2, Gram rows 512 / variant 0, 4 KiB of
    # update loads in flight, SpMM 8 columns per thread (valid but not what tuning would pick
    # on every run); the same shape under another architecture with different numbers.
    path.write_text(f"# test shapes\nnot-{arch} {n} {A.nnz} 4 {k} {mb} 0 1024 1 1024 1 16 16 16\n"
                    f"{arch} {n} {A.nnz} 4 {k} {mb} 0 256 2 512 0 4 4 8\n")
    gdmr.set_tuning_file(str(path))
    try:
        B = np.stack([gdmr.rand_normal(61, j, n) for j in range(4)])
        solver = gdmr.BlockSolver(n, mb, k, 4)
        rp, ci, va = _dev(torch, A.rowptr), _dev(torch, A.col), _dev(torch, A.val)
        solver.set_csr(rp.data_ptr(), ci.data_ptr(), va.data_ptr(), A.nnz, keep=(rp, ci, va))
        rc, rel, st, X = _block(torch, A, B, mb, k, solver=solver)
        t = solver.tuning()
        print(f"tuning from file: {t}", flush=True)
        assert rc == 0 and np.max([_true_res(A, B[j], X[j]) for j in range(4)]) <= 1e-10 * 1.001
        assert (t["tn_rows_proj"], t["tn_lds_proj"], t["tn_rows_gram"], t["tn_lds_gram"],
                t["nn_unroll"], t["spmm_cols"]) == (256, 2, 512, 0, 4, 8)
        # s = 3 is not in the file: tuned, then appended.
        rc, rel, st, X = _block(torch, A, B[:3], mb, k, solver=solver)
        assert rc == 0
        solver.close()
        lines = [ln.split() for ln in path.read_text().splitlines() if not ln.startswith("#")]
        new = [ln for ln in lines if ln[0] == arch and ln[3] == "3"]
        assert len(lines) == 3 and len(new) == 1
        assert new[0][1:7] == [str(n), str(A.nnz), "3", str(k), str(mb), "0"]
    finally:
        gdmr.set_tuning_file(None)
