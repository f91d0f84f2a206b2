"""Python host mirror of libgDMR.so (ctypes over the C ABI of include/gDMR.h and gDMRC.h).

The solver itself is the native library: a C++ host driver and hand-written HIP kernels for
gfx950.  This module only binds its C entry points for tests, benchmarks and scripted use, in the
shape of the reference interface (seanMolesky/gDMR-solver gDMR.h:13-34, gDMRC.h:6-29):

* the global-state solver API (initDMR / solverDMR / freeDMR / impSVDMR / expSVDMR),
* the linear-system API (initLS / impImgLS / impSrcLS / impCsrLS / solveLS / expSrcLS / freeLS),
* handle-based contexts (``Solver``) for several operator families at once,
* the synthetic problem generators and the counter-based random numbers.

Importing fails loudly when the library has not been built; there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(os.path.dirname(_HERE), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libgDMR.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"libgDMR.so not found at {LIB_PATH}; build it with `make -C gdmr-solver_amd`")

lib = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)

# ---- error codes (gDMR.h) ----
DMR_OK, DMR_ERR_ARG, DMR_ERR_STATE, DMR_ERR_HIP, DMR_ERR_ALLOC = 0, 1, 2, 3, 4
DMR_KGRAMCHOL_WORK_BYTES = 256 + 64 * 256 * 8  # gDMRKernels.h
DMR_ERR_OPERATOR, DMR_ERR_NOCONV, DMR_ERR_RUNAWAY, DMR_ERR_NOTBUILT = 5, 6, 7, 8
PREC_FP64, PREC_FP32, PREC_BF16 = 0, 1, 2

_vp, _ip, _dp, _fp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_float)
LEGACY_OP = C.CFUNCTYPE(None)
OOP_OP = C.CFUNCTYPE(None, _vp, _vp, _vp, _vp)


class Stats(C.Structure):
    _fields_ = [("cycles", C.c_int), ("matvecs", C.c_int), ("breakdowns", C.c_int),
                ("status", C.c_int), ("beta", C.c_double), ("rho", C.c_double),
                ("seconds", C.c_double), ("svdSweeps", C.c_int), ("svdCalls", C.c_int)]


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


# Reference interface.
_sig("initDMR", None, [C.c_int, C.c_int, _ip, C.c_int, C.c_int, C.c_int, C.c_int, LEGACY_OP])
_sig("freeDMR", None, [])
_sig("solverDMR", C.c_double, [C.c_int, _vp, _vp, _vp, C.c_int, C.c_double, C.c_int, _ip])
_sig("impSVDMR", None, [_dp])
_sig("expSVDMR", None, [_dp])
_sig("initLS", None, [C.c_int, C.c_int, C.c_int, C.c_int, _ip, C.c_int, C.c_int, C.c_int])
_sig("freeLS", None, [])
for _n in ("impImgLS", "expImgLS", "impSrcLS", "expSrcLS", "impOptLS", "expOptLS"):
    _sig(_n, None, [_dp])
_sig("solveLS", C.c_double, [C.c_int, C.c_double, C.c_int, _ip])
_sig("linOptLS", None, [])
# Extensions.
_sig("impCsrLS", C.c_int, [_ip, _ip, _dp, C.c_int])
_sig("nnzLS", C.c_int, [])
_sig("dmrGetLastError", C.c_int, [])
_sig("dmrGetLastErrorString", C.c_char_p, [])
_sig("dmrRaiseOperatorError", None, [C.c_char_p])
_sig("initDMRx", None, [C.c_int, C.c_int, C.c_int, C.c_int, OOP_OP, _vp])
_sig("dmrGetStats", C.c_int, [C.POINTER(Stats), _dp, C.c_int])
_sig("dmrReductionBlocks", C.c_int, [C.c_int])
_sig("dmrCreate", C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int])
_sig("dmrDestroy", C.c_int, [_vp])
_sig("dmrSetOperatorCsr", C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp])
_sig("dmrSetOperatorValues", C.c_int, [_vp, _vp, _vp])
_sig("dmrSetOperatorFn", C.c_int, [_vp, OOP_OP, _vp])
_sig("dmrSetOperatorLegacy", C.c_int, [_vp, LEGACY_OP])
_sig("dmrSolve", C.c_double, [_vp, C.c_int, _vp, _vp, _vp, C.c_double, C.c_int, _ip])
_sig("dmrImportZ", C.c_int, [_vp, _dp])
_sig("dmrExportZ", C.c_int, [_vp, _dp])
_sig("dmrHasZ", C.c_int, [_vp])
_sig("dmrGetStatsH", C.c_int, [_vp, C.POINTER(Stats), _dp, C.c_int])
_sig("dmrSetRefinement", C.c_int, [_vp, C.c_double, C.c_int])
_sig("dmrGlobalHandle", _vp, [])
_sig("dmrFootprint", C.c_size_t, [C.c_int, C.c_int, C.c_int, C.c_int])
_sig("dmrSubspaceBytes", C.c_size_t, [_vp])
_sig("dmrSubspaceToDevice", C.c_int, [_vp, _vp, _vp])
_sig("dmrSubspaceFromDevice", C.c_int, [_vp, _vp, _vp])
_sig("dmrServerCreate", C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_size_t])
_sig("dmrServerAddFamily", C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_int, C.c_int,
                                     _ip])
_sig("dmrServerSubmit", C.c_int, [_vp, C.c_int, _vp, C.c_int, C.POINTER(_vp), C.POINTER(_vp),
                                  C.c_double, C.c_int, C.POINTER(C.c_longlong)])
_sig("dmrServerNow", C.c_double, [_vp])
_sig("dmrServerDrain", C.c_int, [_vp])
_sig("dmrServerStats", C.c_int, [_vp, _ip, _ip, _ip, C.POINTER(C.c_size_t)])
_sig("dmrServerDestroy", C.c_int, [_vp])
_sig("dmrServerSetRefinement", C.c_int, [_vp, C.c_int, C.c_double, C.c_int])
_sig("dmrServerSetBlockPrecision", C.c_int, [_vp, C.c_int, C.c_int])
_sig("dmrServerAddBlockFamily", C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, C.c_int, C.c_int,
                                          C.c_int, _ip])
_sig("dmrGetStagnationCount", C.c_int, [_vp])
_sig("dmrSetStagnationCount", C.c_int, [_vp, C.c_int])
_sig("dmrBlockCreate", C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int])
_sig("dmrBlockDestroy", C.c_int, [_vp])
_sig("dmrBlockSetOperatorCsr", C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp])
_sig("dmrBlockSetOperatorValues", C.c_int, [_vp, _vp, _vp])
_sig("dmrBlockSetRefinement", C.c_int, [_vp, C.c_double, C.c_int])
_sig("dmrBlockSetColumnOperators", C.c_int, [_vp, _vp, C.c_longlong, C.c_int, C.c_int])
_sig("dmrBlockSetPrecision", C.c_int, [_vp, C.c_int])
_sig("dmrBlockSolve", C.c_int, [_vp, C.c_int, C.c_int, _vp, C.c_longlong, _vp, C.c_longlong, _vp,
                                C.c_double, C.c_int, _dp])
_sig("dmrBlockGetStats", C.c_int, [_vp, C.POINTER(Stats), _dp, C.c_int])
_sig("dmrBlockHasZ", C.c_int, [_vp])
_sig("dmrBlockImportZ", C.c_int, [_vp, _dp])
_sig("dmrBlockExportZ", C.c_int, [_vp, _dp])
_sig("dmrBlockSetProfiling", C.c_int, [_vp, C.c_int])
_sig("dmrBlockGetProfile", C.c_int, [_vp, _dp, _dp, _dp, C.POINTER(C.c_longlong)])
_sig("dmrBlockFootprint", C.c_size_t, [C.c_int, C.c_int, C.c_int, C.c_int])
_sig("dmrBlockSetOptions", C.c_int, [_vp, C.c_int, C.c_int])
_sig("dmrBlockGetTuning", C.c_int, [_vp, _ip])
_sig("dmrBlockTuningFile", C.c_int, [C.c_char_p])
_sig("dmrSetProfiling", C.c_int, [_vp, C.c_int])
_sig("dmrGetProfile", C.c_int, [_vp, _dp, _dp, C.POINTER(C.c_longlong)])
_sig("dmrCsrApplyStatic", C.c_int, [C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp])
# Kernel-level entry points of the block solver (gDMRKernels.h).
_ll, _i32, _f32, _f64 = C.c_longlong, C.c_int, C.c_float, C.c_double
KBASIS_FP32, KBASIS_BF16 = 0, 1
_sig("dmrKernelTnWorkspace", C.c_size_t, [_ll, _i32, _i32])
_sig("dmrKernelGemmTn", C.c_int, [_i32, _vp, _ll, _i32, _vp, _ll, _i32, _ll, _i32, _i32, _vp, _vp,
                                  _i32, _vp, _vp])
_sig("dmrKernelGemmTnPair", C.c_int, [_vp, _ll, _i32, _vp, _vp, _ll, _i32, _ll, _i32, _vp, _vp,
                                      _i32, _vp, _vp, _i32, _vp])
_sig("dmrKernelGemmNn", C.c_int, [_i32, _vp, _ll, _i32, _vp, _vp, _ll, _f32, _f32, _vp, _ll, _i32,
                                  _ll, _i32, _vp, _vp])
_sig("dmrKernelGramBlocks", C.c_int, [_ll, _i32])
_sig("dmrKernelGramReduce", C.c_int, [_vp, _i32, _i32, _vp, _vp])
_sig("dmrKernelGemmNnPair", C.c_int, [_vp, _ll, _i32, _vp, _vp, _vp, _ll, _i32, _ll, _i32, _vp])
_sig("dmrKernelSpmm", C.c_int, [_i32, _i32, _i32, _vp, _vp, _vp, _vp, _ll, _vp, _ll, _i32, _i32,
                                _vp])
_sig("dmrKernelCholQr", C.c_int, [_vp, _i32, _f64, _vp, _vp, _i32, _vp, _i32, _vp, _vp])
_sig("dmrKernelGramChol", C.c_int, [_vp, _i32, _i32, _f64, _vp, _vp, _vp, _i32, _vp, _vp, _vp])
_sig("dmrKernelScaleBlock", C.c_int, [_vp, _ll, _vp, _vp, _ll, _i32, _i32, _ll, _vp])
_sig("dmrKernelLsqBlock", C.c_int, [_vp, _i32, _i32, _i32, _vp, _i32, _vp, _vp, _vp, _vp, _vp,
                                    _vp, _vp])
_sig("dmrKernelLsqFits", C.c_int, [_i32, _i32, _i32])
_sig("dmrProbeReadBandwidth", C.c_int, [_vp, C.c_size_t, _i32, _vp, _dp])
PHASES = ("spmv", "gemvt", "update", "scale", "combo", "zgemm", "lsq", "svd", "orth", "reduce",
          "resid")
# Problems and random numbers.
_sig("dmrGenConvDiff2D", C.c_longlong, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_ulonglong, C.c_ulonglong, _ip, _ip, _dp])
_sig("dmrGenConvDiff3D", C.c_longlong, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_ulonglong, C.c_ulonglong, _ip, _ip, _dp])
_sig("dmrGenAniso3D", C.c_longlong, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_ulonglong, C.c_ulonglong, _ip, _ip, _dp])
_sig("dmrRandNormal", None, [C.c_ulonglong, C.c_ulonglong, C.c_longlong, _dp])
_sig("dmrRandUniform", None, [C.c_ulonglong, C.c_ulonglong, C.c_longlong, C.c_double, C.c_double,
                              _dp])
_sig("dmrSplitMix64", None, [C.c_ulonglong, C.c_longlong, C.POINTER(C.c_ulonglong)])
_sig("dmrReadMatrixMarket", C.c_int, [C.c_char_p, _ip, _ip, C.POINTER(C.c_longlong), _ip, _ip,
                                      _dp])
_sig("dmrWriteMatrixMarket", C.c_int, [C.c_char_p, C.c_int, _ip, _ip, _dp])
_sig("impMtxLS", C.c_int, [C.c_char_p])

# Every symbol the public headers declare (checked by the CPU test suite).
HEADER_SYMBOLS = (
    "initDMR freeDMR solverDMR impSVDMR expSVDMR dmrGetLastError dmrGetLastErrorString "
    "dmrRaiseOperatorError initDMRx dmrGetStats dmrReductionBlocks dmrCreate dmrDestroy "
    "dmrSetOperatorCsr dmrSetOperatorValues dmrSetOperatorFn dmrSetOperatorLegacy dmrSolve "
    "dmrImportZ dmrExportZ "
    "dmrHasZ dmrGetStatsH dmrSetRefinement dmrGlobalHandle dmrFootprint dmrSetProfiling "
    "dmrGetProfile dmrSubspaceBytes dmrSubspaceToDevice dmrSubspaceFromDevice "
    "dmrGetStagnationCount dmrSetStagnationCount dmrCsrApplyStatic "
    "dmrBlockCreate dmrBlockDestroy dmrBlockSetOperatorCsr dmrBlockSetOperatorValues "
    "dmrBlockSetRefinement dmrBlockSetColumnOperators dmrBlockSetPrecision dmrBlockSolve dmrBlockGetStats dmrBlockHasZ dmrBlockImportZ "
    "dmrBlockExportZ dmrBlockSetProfiling dmrBlockGetProfile dmrBlockFootprint "
    "dmrBlockSetOptions dmrBlockGetTuning dmrBlockTuningFile "
    "dmrServerCreate dmrServerAddFamily dmrServerAddBlockFamily dmrServerSetRefinement "
    "dmrServerSetBlockPrecision "
    "dmrServerSubmit dmrServerPoll "
    "dmrServerWait dmrServerDrain dmrServerNow dmrServerStats dmrServerDestroy "
    "initLS freeLS impImgLS expImgLS impSrcLS expSrcLS impOptLS expOptLS solveLS linOptLS "
    "impCsrLS nnzLS dmrReadMatrixMarket dmrWriteMatrixMarket impMtxLS dmrGenConvDiff2D dmrGenConvDiff3D dmrGenAniso3D dmrRandNormal "
    "dmrRandUniform dmrSplitMix64 "
    "dmrKernelTnWorkspace dmrKernelGemmTn dmrKernelGemmTnPair dmrKernelGemmNn dmrKernelGramBlocks "
    "dmrKernelGramReduce dmrKernelGemmNnPair dmrKernelSpmm dmrKernelCholQr dmrKernelGramChol dmrKernelScaleBlock "
    "dmrKernelLsqBlock dmrKernelLsqFits dmrProbeReadBandwidth").split()


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def set_tuning_file(path: str | None) -> None:
    """Persist the block autotuner's launch shapes in `path` (read before the first tuning of
    the process, appended to after each new one); None turns persistence off."""
    _check(lib.dmrBlockTuningFile(path.encode() if path else None), "dmrBlockTuningFile")


def probe_read_bandwidth(buf_ptr: int, nbytes: int, reps: int = 10) -> float:
    """The device's measured HBM read bandwidth in GB/s (best of `reps` streaming reads of
    `nbytes` of device memory at buf_ptr; see dmrProbeReadBandwidth)."""
    out = C.c_double(0.0)
    rc = lib.dmrProbeReadBandwidth(C.c_void_p(buf_ptr), nbytes, reps, None, C.byref(out))
    if rc != 0:
        raise RuntimeError("dmrProbeReadBandwidth: %s" % last_error()[1])
    return out.value


def last_error() -> tuple[int, str]:
    return lib.dmrGetLastError(), lib.dmrGetLastErrorString().decode()


# ------------------------------------------------------------------------------------------------
# Problems
# ------------------------------------------------------------------------------------------------
@dataclass
class CsrMatrix:
    n: int
    rowptr: np.ndarray  # int32 [n+1]
    col: np.ndarray     # int32 [nnz]
    val: np.ndarray     # float64 [nnz]

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def matvec(self, x: np.ndarray) -> np.ndarray:
        prod = self.val * x[self.col]
        return np.add.reduceat(prod, self.rowptr[:-1]) if self.nnz else np.zeros(self.n)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.n, self.n))


def _gen(fn, n, *args) -> CsrMatrix:
    nnz = fn(*args, None, None, None)
    rp = np.zeros(n + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    va = np.zeros(nnz, np.float64)
    got = fn(*args, _ptr(rp, C.c_int), _ptr(ci, C.c_int), _ptr(va, C.c_double))
    assert got == nnz, (got, nnz)
    return CsrMatrix(n, rp, ci, va)


def conv_diff_2d(N: int, nu0: float = 0.1, beta=(1.0, 0.5), jitter: float = 0.0,
                 seed: int = 0, stream: int = 0) -> CsrMatrix:
    """h^2 (-nu lap u + beta.grad u) on an N x N interior grid, 5-point central differences."""
    return _gen(lib.dmrGenConvDiff2D, N * N, N, nu0, beta[0], beta[1], jitter, seed, stream)


def conv_diff_3d(N: int, nu0: float = 0.1, beta=(1.0, 0.5, 0.25), jitter: float = 0.0,
                 seed: int = 0, stream: int = 0) -> CsrMatrix:
    """h^2 (-div(nu grad u) + beta.grad u) on an N^3 interior grid, 7-point central differences."""
    return _gen(lib.dmrGenConvDiff3D, N ** 3, N, nu0, beta[0], beta[1], beta[2], jitter, seed,
                stream)


def aniso_3d(N: int, coef=(1.0, 0.1, 0.01), jitter: float = 0.0, seed: int = 0,
             stream: int = 0) -> CsrMatrix:
    """h^2 (-(a u_xx + b u_yy + c u_zz)) on an N^3 interior grid, 7-point."""
    return _gen(lib.dmrGenAniso3D, N ** 3, N, coef[0], coef[1], coef[2], jitter, seed, stream)


def rand_normal(seed: int, stream: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    lib.dmrRandNormal(seed, stream, n, _ptr(out, C.c_double))
    return out


def rand_uniform(seed: int, stream: int, n: int, lo: float, hi: float) -> np.ndarray:
    out = np.empty(n, np.float64)
    lib.dmrRandUniform(seed, stream, n, lo, hi, _ptr(out, C.c_double))
    return out


def splitmix64(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.uint64)
    lib.dmrSplitMix64(seed, n, out.ctypes.data_as(C.POINTER(C.c_ulonglong)))
    return out


def reduction_blocks(n: int) -> int:
    """Block count G of the device's fixed-order reductions over n doubles."""
    return lib.dmrReductionBlocks(n)


def footprint(n: int, m: int, k: int, precision: int = PREC_FP64) -> int:
    return int(lib.dmrFootprint(n, m, k, precision))


def read_matrix_market(path: str) -> CsrMatrix:
    """A square Matrix Market file as CSR (dmrReadMatrixMarket)."""
    rows, cols, nnz = C.c_int(), C.c_int(), C.c_longlong()
    p = os.fsencode(path)
    rc = lib.dmrReadMatrixMarket(p, C.byref(rows), C.byref(cols), C.byref(nnz), None, None, None)
    if rc != DMR_OK:
        raise ValueError(f"dmrReadMatrixMarket({path}) failed with code {rc}")
    if rows.value != cols.value:
        raise ValueError(f"{path}: {rows.value} x {cols.value} is not square")
    rp = np.zeros(rows.value + 1, np.int32)
    ci = np.zeros(max(nnz.value, 1), np.int32)
    va = np.zeros(max(nnz.value, 1), np.float64)
    rc = lib.dmrReadMatrixMarket(p, C.byref(rows), C.byref(cols), C.byref(nnz),
                                 _ptr(rp, C.c_int), _ptr(ci, C.c_int), _ptr(va, C.c_double))
    if rc != DMR_OK:
        raise ValueError(f"dmrReadMatrixMarket({path}) failed with code {rc}")
    return CsrMatrix(rows.value, rp, ci[:nnz.value], va[:nnz.value])


def write_matrix_market(path: str, A: CsrMatrix):
    rp = np.ascontiguousarray(A.rowptr, np.int32)
    ci = np.ascontiguousarray(A.col, np.int32)
    va = np.ascontiguousarray(A.val, np.float64)
    rc = lib.dmrWriteMatrixMarket(os.fsencode(path), A.n, _ptr(rp, C.c_int), _ptr(ci, C.c_int),
                                  _ptr(va, C.c_double))
    if rc != DMR_OK:
        raise ValueError(f"dmrWriteMatrixMarket({path}) failed with code {rc}")


# ------------------------------------------------------------------------------------------------
# Handle-based solver (device memory through torch tensors: plumbing only)
# ------------------------------------------------------------------------------------------------
class SolverError(RuntimeError):
    pass


def _check(rc: int, what: str):
    if rc != DMR_OK:
        code, msg = last_error()
        raise SolverError(f"{what} failed: code {rc} ({msg})")


class Solver:
    """One DMR context (n, m, k, precision) on one device, with its recycled subspace Z."""

    def __init__(self, n: int, m: int, k: int, precision: int = PREC_FP64, device: int = 0):
        self.n, self.m, self.k, self.precision, self.device = n, m, k, precision, device
        h = C.c_void_p()
        _check(lib.dmrCreate(C.byref(h), device, n, m, k, precision), "dmrCreate")
        self.h = h
        self._keep = []

    def close(self):
        if self.h:
            lib.dmrDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_csr(self, rowptr_ptr: int, col_ptr: int, val_ptr: int, nnz: int, val32_ptr: int = 0,
                keep=()):
        """Device pointers of a CSR operator (int32 rowptr / col, float64 val, optional float32)."""
        self._keep = list(keep)
        _check(lib.dmrSetOperatorCsr(self.h, nnz, rowptr_ptr, col_ptr, val_ptr, val32_ptr or None),
               "dmrSetOperatorCsr")

    def set_values(self, val_ptr: int, val32_ptr: int = 0, keep=()):
        """New values for the current CSR pattern (device pointers)."""
        self._keep_vals = list(keep)
        _check(lib.dmrSetOperatorValues(self.h, val_ptr, val32_ptr or None), "dmrSetOperatorValues")

    def set_operator(self, fn):
        """fn(src_ptr, dst_ptr, stream_ptr, ctx): out-of-place operator."""
        cb = OOP_OP(fn)
        self._cb = cb
        _check(lib.dmrSetOperatorFn(self.h, cb, None), "dmrSetOperatorFn")

    def set_legacy(self, fn):
        cb = LEGACY_OP(fn)
        self._cb = cb
        _check(lib.dmrSetOperatorLegacy(self.h, cb), "dmrSetOperatorLegacy")

    def set_refinement(self, inner_tol: float, max_outer: int = 60):
        _check(lib.dmrSetRefinement(self.h, inner_tol, max_outer), "dmrSetRefinement")

    def solve(self, x_ptr: int, b_ptr: int, tol: float, mat_rank: int, mode: int = 0,
              stream: int = 0):
        its = C.c_int(0)
        res = lib.dmrSolve(self.h, mode, x_ptr, b_ptr, stream or None, tol, mat_rank,
                           C.byref(its))
        return res, its.value

    def stats(self, cap: int = 4096):
        st = Stats()
        hist = np.zeros(cap)
        cnt = lib.dmrGetStatsH(self.h, C.byref(st), _ptr(hist, C.c_double), cap)
        return st, hist[:min(cnt, cap)].copy()

    def has_z(self) -> bool:
        return bool(lib.dmrHasZ(self.h))

    def set_profiling(self, on: bool = True):
        _check(lib.dmrSetProfiling(self.h, int(on)), "dmrSetProfiling")

    def profile(self) -> dict:
        """{phase: (ms, algorithmic bytes, launches)} accumulated since set_profiling(True)."""
        ms, by = np.zeros(len(PHASES)), np.zeros(len(PHASES))
        cnt = np.zeros(len(PHASES), np.int64)
        lib.dmrGetProfile(self.h, _ptr(ms, C.c_double), _ptr(by, C.c_double),
                          cnt.ctypes.data_as(C.POINTER(C.c_longlong)))
        return {p: (float(ms[i]), float(by[i]), int(cnt[i])) for i, p in enumerate(PHASES)}

    def import_z(self, Z: np.ndarray):
        Z = np.ascontiguousarray(Z, np.float64)
        assert Z.size == self.n * self.k
        _check(lib.dmrImportZ(self.h, _ptr(Z, C.c_double)), "dmrImportZ")

    def export_z(self) -> np.ndarray:
        Z = np.zeros(self.n * self.k)
        _check(lib.dmrExportZ(self.h, _ptr(Z, C.c_double)), "dmrExportZ")
        return Z.reshape(self.k, self.n)


class BlockSolver:
    """Block DMR context: up to max_rhs right-hand sides share one block Krylov space and one
    recycled subspace Z (dmrBlock* in gDMR.h)."""

    def __init__(self, n: int, block_steps: int, k: int, max_rhs: int, device: int = 0):
        self.n, self.mb, self.k, self.max_rhs = n, block_steps, k, max_rhs
        h = C.c_void_p()
        _check(lib.dmrBlockCreate(C.byref(h), device, n, block_steps, k, max_rhs),
               "dmrBlockCreate")
        self.h = h
        self._keep = []

    def close(self):
        if self.h:
            lib.dmrBlockDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_csr(self, rowptr_ptr, col_ptr, val_ptr, nnz, val32_ptr=0, keep=()):
        self._keep = list(keep)
        _check(lib.dmrBlockSetOperatorCsr(self.h, nnz, rowptr_ptr, col_ptr, val_ptr,
                                          val32_ptr or None), "dmrBlockSetOperatorCsr")

    def set_values(self, val_ptr, val32_ptr=0, keep=()):
        self._keep_vals = list(keep)
        _check(lib.dmrBlockSetOperatorValues(self.h, val_ptr, val32_ptr or None),
               "dmrBlockSetOperatorValues")

    def set_column_operators(self, val_ptr, ldv: int, count: int, inner_mean: bool = True,
                             keep=()):
        """Column c solves with the FP64 values at val_ptr + 8 * c * ldv (same CSR pattern);
        the inner FP32 operator is their mean when inner_mean.  val_ptr = 0 clears."""
        self._keep_cols = list(keep)
        _check(lib.dmrBlockSetColumnOperators(self.h, val_ptr or None, ldv, count,
                                              int(inner_mean)), "dmrBlockSetColumnOperators")

    def set_refinement(self, inner_tol: float, max_outer: int = 30):
        _check(lib.dmrBlockSetRefinement(self.h, inner_tol, max_outer), "dmrBlockSetRefinement")

    def set_precision(self, precision: int):
        """Basis storage: PREC_FP32 (default) or PREC_BF16."""
        _check(lib.dmrBlockSetPrecision(self.h, precision), "dmrBlockSetPrecision")

    def solve(self, x_ptr, ldx, b_ptr, ldb, nrhs, tol, max_cycles=1000, mode=0, stream=0):
        """Returns (status code, per-column FP64-checked relative residuals)."""
        rel = np.zeros(nrhs)
        rc = lib.dmrBlockSolve(self.h, mode, nrhs, x_ptr, ldx, b_ptr, ldb, stream or None, tol,
                               max_cycles, _ptr(rel, C.c_double))
        return rc, rel

    def stats(self, cap: int = 4096):
        st = Stats()
        hist = np.zeros(cap)
        cnt = lib.dmrBlockGetStats(self.h, C.byref(st), _ptr(hist, C.c_double), cap)
        return st, hist[:min(cnt, cap)].copy()

    def has_z(self) -> bool:
        return bool(lib.dmrBlockHasZ(self.h))

    def import_z(self, Z: np.ndarray):
        Z = np.ascontiguousarray(Z, np.float64)
        assert Z.size == self.n * self.k
        _check(lib.dmrBlockImportZ(self.h, _ptr(Z, C.c_double)), "dmrBlockImportZ")

    def export_z(self) -> np.ndarray:
        Z = np.zeros(self.n * self.k)
        _check(lib.dmrBlockExportZ(self.h, _ptr(Z, C.c_double)), "dmrBlockExportZ")
        return Z.reshape(self.k, self.n)

    def set_options(self, graphs: int = -1, autotune: int = -1):
        _check(lib.dmrBlockSetOptions(self.h, graphs, autotune), "dmrBlockSetOptions")

    def tuning(self) -> dict:
        out = (C.c_int * 7)()
        _check(lib.dmrBlockGetTuning(self.h, out), "dmrBlockGetTuning")
        return dict(tn_rows_proj=out[0], tn_lds_proj=out[1], tn_rows_gram=out[2],
                    tn_lds_gram=out[3], nn_unroll=out[4], spmm_cols=out[5], graphs=out[6])

    def set_profiling(self, on: bool = True):
        _check(lib.dmrBlockSetProfiling(self.h, int(on)), "dmrBlockSetProfiling")

    def profile(self) -> dict:
        """{phase: (ms, algorithmic bytes, launches, flops)} since set_profiling(True)."""
        ms, by, fl = np.zeros(len(PHASES)), np.zeros(len(PHASES)), np.zeros(len(PHASES))
        cnt = np.zeros(len(PHASES), np.int64)
        lib.dmrBlockGetProfile(self.h, _ptr(ms, C.c_double), _ptr(by, C.c_double),
                               _ptr(fl, C.c_double), cnt.ctypes.data_as(C.POINTER(C.c_longlong)))
        return {p: (float(ms[i]), float(by[i]), int(cnt[i]), float(fl[i]))
                for i, p in enumerate(PHASES)}


def global_stats(cap: int = 4096):
    st = Stats()
    hist = np.zeros(cap)
    cnt = lib.dmrGetStats(C.byref(st), _ptr(hist, C.c_double), cap)
    return st, hist[:max(0, min(cnt, cap))].copy()


# ------------------------------------------------------------------------------------------------
# Linear-system (control-program) interface, host arrays
# ------------------------------------------------------------------------------------------------
class LinearSystem:
    """initLS ... freeLS, as a control program drives them (progDMR.c:135-182)."""

    def __init__(self, A: CsrMatrix, m: int, k: int, device: int = 0):
        self.A, self.n = A, A.n
        dev = (C.c_int * 1)(device)
        lib.initLS(32, 512, device, A.n, dev, 1, m, k)
        rc = lib.impCsrLS(_ptr(A.rowptr, C.c_int), _ptr(A.col, C.c_int), _ptr(A.val, C.c_double),
                          A.nnz)
        if rc != DMR_OK:
            raise SolverError(f"impCsrLS failed: {rc}")

    def solve(self, b: np.ndarray, x0: np.ndarray | None = None, tol: float = 1e-8,
              mat_rank: int | None = None, mode: int = 0):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
        lib.impImgLS(_ptr(b, C.c_double))
        lib.impSrcLS(_ptr(x, C.c_double))
        its = C.c_int(0)
        res = lib.solveLS(mode, tol, self.n if mat_rank is None else mat_rank, C.byref(its))
        lib.expSrcLS(_ptr(x, C.c_double))
        return x, res, its.value

    def close(self):
        lib.freeLS()


# ------------------------------------------------------------------------------------------------
# Serving loop (gDMRServe.h)
# ------------------------------------------------------------------------------------------------
class RequestResult(C.Structure):
    _fields_ = [("id", C.c_longlong), ("family", C.c_int), ("nrhs", C.c_int), ("status", C.c_int),
                ("cycles", C.c_int), ("slot", C.c_int), ("recycled", C.c_int),
                ("maxRelRes", C.c_double), ("submitTime", C.c_double),
                ("startTime", C.c_double), ("endTime", C.c_double), ("batch", C.c_int)]


_sig("dmrServerPoll", C.c_int, [_vp, C.POINTER(RequestResult), C.c_int])
_sig("dmrServerWait", C.c_int, [_vp, C.POINTER(RequestResult), C.c_int, C.c_double])


class Server:
    """Worker slots over solver handles; one resident recycled subspace per operator family."""

    def __init__(self, slots: int = 4, device: int = 0, hbm_reserve: int = 1 << 30):
        h = C.c_void_p()
        _check(lib.dmrServerCreate(C.byref(h), device, slots, hbm_reserve), "dmrServerCreate")
        self.h = h
        self._keep = {}

    def add_family(self, n, nnz, rowptr_ptr, col_ptr, m, k, precision=PREC_FP64,
                   keep=()) -> int:
        fam = C.c_int()
        _check(lib.dmrServerAddFamily(self.h, n, nnz, rowptr_ptr, col_ptr, m, k, precision,
                                      C.byref(fam)), "dmrServerAddFamily")
        self._keep[("family", fam.value)] = keep
        return fam.value

    def add_block_family(self, n, nnz, rowptr_ptr, col_ptr, block_steps, k, max_rhs,
                         keep=()) -> int:
        fam = C.c_int()
        _check(lib.dmrServerAddBlockFamily(self.h, n, nnz, rowptr_ptr, col_ptr, block_steps, k,
                                           max_rhs, C.byref(fam)), "dmrServerAddBlockFamily")
        self._keep[("family", fam.value)] = keep
        return fam.value

    def set_refinement(self, family: int, inner_tol: float, max_outer: int = 60):
        _check(lib.dmrServerSetRefinement(self.h, family, inner_tol, max_outer),
               "dmrServerSetRefinement")

    def set_block_precision(self, family: int, precision: int):
        _check(lib.dmrServerSetBlockPrecision(self.h, family, precision),
               "dmrServerSetBlockPrecision")

    def submit(self, family, val_ptr, b_ptrs, x_ptrs, tol, mat_rank, keep=()) -> int:
        nr = len(b_ptrs)
        bs = (C.c_void_p * nr)(*b_ptrs)
        xs = (C.c_void_p * nr)(*x_ptrs)
        rid = C.c_longlong()
        _check(lib.dmrServerSubmit(self.h, family, val_ptr, nr, bs, xs, tol, mat_rank,
                                   C.byref(rid)), "dmrServerSubmit")
        self._keep[rid.value] = keep
        return rid.value

    def wait(self, timeout: float = 1.0, cap: int = 256):
        buf = (RequestResult * cap)()
        cnt = lib.dmrServerWait(self.h, buf, cap, timeout)
        out = [buf[i] for i in range(cnt)]
        for r in out:
            self._keep.pop(r.id, None)
        return out

    def drain(self):
        lib.dmrServerDrain(self.h)

    def now(self) -> float:
        return lib.dmrServerNow(self.h)

    def stats(self):
        q, r, hd, fr = C.c_int(), C.c_int(), C.c_int(), C.c_size_t()
        lib.dmrServerStats(self.h, C.byref(q), C.byref(r), C.byref(hd), C.byref(fr))
        return dict(queued=q.value, running=r.value, handles=hd.value, hbm_free=fr.value)

    def close(self):
        if self.h:
            lib.dmrServerDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
