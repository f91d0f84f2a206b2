"""The reference's C ABI entry points through ctypes, each against the oracle or the reference
semantics it replaces (gDMR.h:13-34, gDMRC.h:6-29):

* impSVDMR / expSVDMR (gDMR.cu:431-477, the checkpoint of the recycled subspace): a sequence
  interrupted after its first solve (export, freeLS / freeDMR, initLS / initDMR, import) resumes
  bit for bit like the uninterrupted sequence;
* impOptLS / expOptLS (testLS.cu:288-307, the dense operator): round trip, and a solve with the
  dense-imported operator equals the CSR-imported one bit for bit;
* linOptLS (testLS.cu:97-128): x <- A x in place equals the oracle's SpMV bit for bit;
* impImgLS / expImgLS, impSrcLS / expSrcLS: round trips;
* the built control program progDMR (progDMR.c:135-182 of the reference): its printed residuals
  and iteration counts are the oracle's for the same sequence.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import gdmr
import oracle

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


def _ls(A, m, k):
    dev = (C.c_int * 1)(0)
    gdmr.lib.initLS(32, 512, 0, A.n, dev, 1, m, k)
    assert gdmr.lib.impCsrLS(A.rowptr.ctypes.data_as(C.POINTER(C.c_int)),
                             A.col.ctypes.data_as(C.POINTER(C.c_int)), _p(A.val), A.nnz) == 0


def _solve(b, tol, mode, n):
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n)
    gdmr.lib.impImgLS(_p(b))
    gdmr.lib.impSrcLS(_p(x))
    its = C.c_int(0)
    res = gdmr.lib.solveLS(mode, tol, n, C.byref(its))
    gdmr.lib.expSrcLS(_p(x))
    st, hist = gdmr.global_stats()
    return x, res, its.value, hist


def test_subspace_checkpoint_resumes_bitwise():
    """Config A's operator (2-D conv-diff 64^2, m = 30, k = 10, tol 1e-8): solve b0 (mode 0) and
    b1 (mode 1) in one session, and again with the subspace exported after b0, the solver freed
    and re-initialised, and the subspace imported before b1.  The second solve of both runs is
    identical bit for bit, and both equal the oracle's."""
    A = gdmr.conv_diff_2d(64)
    n, m, k, tol = A.n, 30, 10, 1e-8
    B = gdmr.rand_normal(0, 1, 2 * n).reshape(2, n)
    _ls(A, m, k)
    try:
        _solve(B[0], tol, 0, n)
        x1, res1, its1, h1 = _solve(B[1], tol, 1, n)
    finally:
        gdmr.lib.freeLS()
    _ls(A, m, k)
    Zh = np.zeros(k * n)
    try:
        _solve(B[0], tol, 0, n)
        gdmr.lib.expSVDMR(_p(Zh))
        assert gdmr.lib.dmrGetLastError() == 0
    finally:
        gdmr.lib.freeLS()
    _ls(A, m, k)
    try:
        Zback = np.zeros(k * n)
        gdmr.lib.impSVDMR(_p(Zh))
        gdmr.lib.expSVDMR(_p(Zback))
        assert np.array_equal(Zback, Zh)
        x2, res2, its2, h2 = _solve(B[1], tol, 1, n)
    finally:
        gdmr.lib.freeLS()
    print(f"checkpoint: its {its1} vs {its2}, res {res1:.3e} vs {res2:.3e}", flush=True)
    assert its1 == its2 and res1 == res2
    assert np.array_equal(h1, h2) and np.array_equal(x1, x2)
    ora = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n))
    ora.solve(B[0], tol=tol, mode=0)
    xo, reso, itso, _ = ora.solve(B[1], tol=tol, mode=1)
    assert itso == its2 and np.array_equal(ora.history(), h2) and np.array_equal(xo, x2)
    # The exported subspace has unit columns (genDeflatDMR normalises them).
    assert np.allclose(np.linalg.norm(Zh.reshape(k, n), axis=1), 1.0, atol=1e-12)


def test_subspace_import_before_init_is_an_error():
    gdmr.lib.freeDMR()
    Z = np.zeros(16)
    gdmr.lib.impSVDMR(_p(Z))
    assert gdmr.lib.dmrGetLastError() == gdmr.DMR_ERR_STATE
    gdmr.lib.expSVDMR(_p(Z))
    assert gdmr.lib.dmrGetLastError() == gdmr.DMR_ERR_STATE


def test_dense_operator_import_export_and_solve():
    """impOptLS takes the reference's dense column-major n x n operator; expOptLS gives it back;
    a solve with it equals the solve with the same matrix imported as CSR, bit for bit."""
    A = gdmr.conv_diff_2d(24, jitter=0.01, seed=5)
    n, m, k, tol = A.n, 20, 6, 1e-10
    dense = np.asfortranarray(A.to_scipy().toarray())
    b = gdmr.rand_normal(5, 1, n)
    _ls(A, m, k)
    try:
        x_csr, res_csr, its_csr, h_csr = _solve(b, tol, 0, n)
        colmajor = np.ascontiguousarray(dense.T).reshape(-1)  # column c at c * n
        gdmr.lib.impOptLS(_p(colmajor))
        assert gdmr.lib.nnzLS() == A.nnz
        back = np.zeros(n * n)
        gdmr.lib.expOptLS(_p(back))
        assert np.array_equal(back, colmajor)
        x_dn, res_dn, its_dn, h_dn = _solve(b, tol, 0, n)
    finally:
        gdmr.lib.freeLS()
    assert its_dn == its_csr and res_dn == res_csr
    assert np.array_equal(h_dn, h_csr) and np.array_equal(x_dn, x_csr)
    assert np.linalg.norm(b - A.matvec(x_dn)) / np.linalg.norm(b) <= tol * 1.001


def test_lin_opt_ls_and_vector_round_trips():
    """linOptLS overwrites the imported source vector with A x (products in column order, as the
    oracle's SpMV), and impImgLS / expImgLS and impSrcLS / expSrcLS move vectors unchanged."""
    A = gdmr.conv_diff_3d(12, jitter=0.01, seed=6)
    n = A.n
    x = gdmr.rand_normal(6, 2, n)
    b = gdmr.rand_normal(6, 3, n)
    _ls(A, 8, 2)
    try:
        gdmr.lib.impImgLS(_p(b))
        bo = np.zeros(n)
        gdmr.lib.expImgLS(_p(bo))
        assert np.array_equal(bo, b)
        gdmr.lib.impSrcLS(_p(x))
        xo = np.zeros(n)
        gdmr.lib.expSrcLS(_p(xo))
        assert np.array_equal(xo, x)
        gdmr.lib.linOptLS()
        y = np.zeros(n)
        gdmr.lib.expSrcLS(_p(y))
        gdmr.lib.linOptLS()  # applied again to its own output: A (A x)
        y2 = np.zeros(n)
        gdmr.lib.expSrcLS(_p(y2))
    finally:
        gdmr.lib.freeLS()
    assert np.array_equal(y, oracle.spmv(A, x))
    assert np.array_equal(y2, oracle.spmv(A, oracle.spmv(A, x)))


def test_prog_dmr_matches_oracle():
    """The C99 control program, linked with -lgDMR and rpath $ORIGIN, runs config A (2-D conv-diff
    64^2, m = 30, k = 10, tol 1e-8, 4 N(0,1) right-hand sides from seed 0, mode = i): every
    printed residual and iteration count is the oracle's, and every true residual meets tol."""
    prog = os.path.join(gdmr.LIB_DIR, "progDMR")
    out = subprocess.run([prog], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0
    lines = re.findall(r"gDMR solve (\d+): residual (\S+) \(true (\S+)\) after (\d+) iterations",
                       out.stdout)
    assert len(lines) == 4
    A = gdmr.conv_diff_2d(64)
    n = A.n
    B = gdmr.rand_normal(0, 1, 4 * n).reshape(4, n)
    ora = oracle.Oracle(A, 30, 10, gdmr.reduction_blocks(n))
    for i, (idx, res, true, its) in enumerate(lines):
        xo, reso, itso, sto = ora.solve(B[i], tol=1e-8, mode=i)
        assert int(idx) == i and int(its) == itso
        assert res == f"{reso:.3e}"
        assert float(true) <= 1e-8 * 1.001
