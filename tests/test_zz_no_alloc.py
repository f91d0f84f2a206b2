"""The allocation-free in-place operator, dmrCsrApplyStatic (x <- A x with the product staged in
a static device array), against linOptLS, the entry point whose contract it follows
(testLS.cu:97-128): the same bits on two calls in a row, the second on a smaller system, and a
clear error for a system larger than the static array."""
import ctypes as C

import numpy as np
import pytest

import gdmr

pytestmark = pytest.mark.gpu

STATIC_ROWS = 1 << 22


def _with_long_row(A, row):
    """A with row `row` replaced by a dense row (every column, so more entries than one SpMV
    block stages): the long-row path of both kernels."""
    n = A.n
    rows = [A.col[A.rowptr[r]:A.rowptr[r + 1]] for r in range(n)]
    vals = [A.val[A.rowptr[r]:A.rowptr[r + 1]] for r in range(n)]
    rows[row] = np.arange(n, dtype=np.int32)
    vals[row] = gdmr.rand_uniform(5, row, n, -1.0, 1.0)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(c) for c in rows])
    return gdmr.CsrMatrix(n, rp, np.concatenate(rows).astype(np.int32), np.concatenate(vals))


def _lin_opt_ls(A, x):
    ls = gdmr.LinearSystem(A, 8, 2)
    try:
        xs = np.ascontiguousarray(x, np.float64).copy()
        gdmr.lib.impSrcLS(xs.ctypes.data_as(C.POINTER(C.c_double)))
        gdmr.lib.linOptLS()
        out = np.zeros(A.n)
        gdmr.lib.expSrcLS(out.ctypes.data_as(C.POINTER(C.c_double)))
    finally:
        ls.close()
    return out


def _static(torch, A, x):
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    rp, ci, va, xd = dev(A.rowptr), dev(A.col), dev(A.val), dev(x)
    torch.cuda.synchronize()
    rc = gdmr.lib.dmrCsrApplyStatic(A.n, A.nnz, rp.data_ptr(), ci.data_ptr(), va.data_ptr(),
                                    xd.data_ptr(), None)
    torch.cuda.synchronize()
    return rc, xd.cpu().numpy()


def test_static_apply_matches_lin_opt_ls(torch_gpu):
    big = _with_long_row(gdmr.conv_diff_3d(20, jitter=0.01, seed=3), 1234)  # n = 8000
    small = gdmr.conv_diff_2d(33, jitter=0.01, seed=4)                        # n = 1089
    for A in (big, small):
        x = gdmr.rand_normal(6, A.n, A.n)
        ref = _lin_opt_ls(A, x)
        rc, got = _static(torch_gpu, A, x)
        print(f"n {A.n} nnz {A.nnz}: rc {rc}, max |diff| {np.max(np.abs(got - ref)):.3e}",
              flush=True)
        assert rc == gdmr.DMR_OK
        assert np.array_equal(got, ref)
        assert np.allclose(got, A.matvec(x), rtol=1e-13, atol=1e-13)


def test_static_apply_rejects_oversized_systems(torch_gpu):
    torch = torch_gpu
    x = torch.zeros(8, dtype=torch.float64, device="cuda")
    rp = torch.zeros(9, dtype=torch.int32, device="cuda")
    # n above the static array is refused on the host, before any launch.
    rc = gdmr.lib.dmrCsrApplyStatic(STATIC_ROWS + 1, 0, rp.data_ptr(), rp.data_ptr(),
                                    x.data_ptr(), x.data_ptr(), None)
    code, msg = gdmr.last_error()
    assert rc == gdmr.DMR_ERR_ARG and code == gdmr.DMR_ERR_ARG and "static" in msg
    rc = gdmr.lib.dmrCsrApplyStatic(0, 0, rp.data_ptr(), rp.data_ptr(), x.data_ptr(),
                                    x.data_ptr(), None)
    assert rc == gdmr.DMR_ERR_ARG
