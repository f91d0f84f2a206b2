"""ctypes loader of the FP64 host oracle -- TEST INFRASTRUCTURE.

Only tests, the smoke check and the CPU-baseline leg of the benchmark import this module; the
product never does.  See dmr_oracle.cpp for what the oracle restates (seanMolesky/gDMR-solver
gDMR.cu, function by function) and how its operation order is fixed.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"oracle not built: {LIB_PATH} (make -C oracle)")
lib = C.CDLL(LIB_PATH)

_vp, _ip, _dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype, f.argtypes = res, args


_sig("oracle_create", _vp, [C.c_int, C.c_int, C.c_int, C.c_int])
_sig("oracle_destroy", None, [_vp])
_sig("oracle_keep_last", None, [_vp, C.c_int])
_sig("oracle_set_csr", C.c_int, [_vp, _ip, _ip, _dp])
_sig("oracle_set_values", None, [_vp, _dp])
_sig("oracle_solve", C.c_double, [_vp, C.c_int, _dp, _dp, C.c_double, C.c_int, _ip, _ip])
_sig("oracle_history", C.c_int, [_vp, _dp, C.c_int])
_sig("oracle_counts", None, [_vp, _ip, _ip, _ip])
_sig("oracle_has_z", C.c_int, [_vp])
_sig("oracle_get_z", None, [_vp, _dp])
_sig("oracle_set_z", None, [_vp, _dp])
_sig("oracle_last_cycle", C.c_int, [_vp, _dp, _dp, _dp])
_sig("oracle_start_log", C.c_int, [_vp, _ip, C.c_int])
_sig("oracle_stagnation_cycle", C.c_int, [_vp, _dp, _dp])
_sig("oracle_set_stagnation_rule", None, [_vp, C.c_int])
_sig("oracle_get_stagnation_count", C.c_int, [_vp])
_sig("oracle_lsq", None, [_dp, C.c_int, _dp, _dp, _dp])
_sig("oracle_svd", C.c_int, [_dp, C.c_int, C.c_int, C.c_double, C.c_int, _dp, _dp])
_sig("oracle_spmv", None, [C.c_int, _ip, _ip, _dp, _dp, _dp])
_sig("oracle_dot", C.c_double, [_dp, _dp, C.c_longlong, C.c_int])
_sig("oracle_set_threads", None, [C.c_int])
_sig("oracle_get_threads", C.c_int, [])
_sig("oracle_stream_triad", C.c_double, [C.c_longlong, C.c_int])


def set_threads(threads: int):
    """OpenMP threads of the oracle's loops."""
    lib.oracle_set_threads(threads)


def get_threads() -> int:
    return lib.oracle_get_threads()


def stream_triad(n: int = 1 << 26, reps: int = 5) -> float:
    """Host DRAM bandwidth in bytes/s (a = b + 3c over three arrays of n doubles, best of reps)
    with the current thread count."""
    return lib.oracle_stream_triad(n, reps)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def lsq(H: np.ndarray, t: np.ndarray):
    """H: (m+1, m) array; returns y (m+1, y[m] = 0) and z = H y (m+1)."""
    m = H.shape[1]
    Hf = np.asfortranarray(H, np.float64)
    t = np.ascontiguousarray(t, np.float64)
    y = np.zeros(m + 1)
    z = np.zeros(m + 1)
    lib.oracle_lsq(Hf.ctypes.data_as(_dp), m, _d(t), _d(y), _d(z))
    return y, z


def svd_smallest(H: np.ndarray, k: int, tol: float = 1e-14, max_sweeps: int = 128):
    """Returns (C (m, k), sigma (m,), sweeps) as the device kernel computes them."""
    m = H.shape[1]
    Hf = np.asfortranarray(H, np.float64)
    Cm = np.zeros((k, m))  # column i of C is row i here
    sig = np.zeros(m)
    sw = lib.oracle_svd(Hf.ctypes.data_as(_dp), m, k, tol, max_sweeps, _d(Cm), _d(sig))
    return Cm.T.copy(), sig, sw


def spmv(A, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(A.n)
    lib.oracle_spmv(A.n, _i(A.rowptr), _i(A.col), _d(A.val), _d(x), _d(y))
    return y


def dot(a: np.ndarray, b: np.ndarray, G: int) -> float:
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return lib.oracle_dot(_d(a), _d(b), a.size, G)


class Oracle:
    """FP64 host DMR with the same contract as dmrSolve; G is the device reduction geometry."""

    def __init__(self, A, m: int, k: int, G: int, keep_last: bool = False):
        self.A, self.n, self.m, self.k = A, A.n, m, k
        self.h = lib.oracle_create(A.n, m, k, G)
        if not self.h:
            raise ValueError("invalid oracle shape")
        lib.oracle_set_csr(self.h, _i(A.rowptr), _i(A.col), _d(A.val))
        if keep_last:
            lib.oracle_keep_last(self.h, 1)

    def set_values(self, A):
        """New operator values on the same sparsity pattern; the recycled subspace and the
        stagnation counter are kept (a sequence of systems, as dmrSetOperatorValues)."""
        assert A.nnz == self.A.nnz and np.array_equal(A.col, self.A.col)
        self.A = A
        lib.oracle_set_values(self.h, _d(A.val))

    def close(self):
        if self.h:
            lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, b: np.ndarray, x0: np.ndarray | None = None, tol: float = 1e-8,
              mat_rank: int | None = None, mode: int = 0):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
        its, status = C.c_int(0), C.c_int(0)
        res = lib.oracle_solve(self.h, mode, _d(b), _d(x), tol,
                               self.n if mat_rank is None else mat_rank, C.byref(its),
                               C.byref(status))
        return x, res, its.value, status.value

    def history(self) -> np.ndarray:
        cnt = lib.oracle_history(self.h, None, 0)
        out = np.zeros(max(cnt, 1))
        lib.oracle_history(self.h, _d(out), cnt)
        return out[:cnt]

    def counts(self):
        c, mv, b = C.c_int(), C.c_int(), C.c_int()
        lib.oracle_counts(self.h, C.byref(c), C.byref(mv), C.byref(b))
        return c.value, mv.value, b.value

    def has_z(self) -> bool:
        return bool(lib.oracle_has_z(self.h))

    def get_z(self) -> np.ndarray:
        Z = np.zeros(self.n * self.k)
        lib.oracle_get_z(self.h, _d(Z))
        return Z.reshape(self.k, self.n)

    def set_z(self, Z: np.ndarray):
        Z = np.ascontiguousarray(Z, np.float64).reshape(-1)
        lib.oracle_set_z(self.h, _d(Z))

    def last_cycle(self):
        """(P (m+1, n), H (m+1, m), Z (k, n), d) of the last cycle (keep_last=True)."""
        P = np.zeros((self.m + 1) * self.n)
        H = np.zeros((self.m + 1) * self.m)
        Z = np.zeros(max(self.k, 1) * self.n)
        d = lib.oracle_last_cycle(self.h, _d(P), _d(H), _d(Z))
        return (P.reshape(self.m + 1, self.n), H.reshape(self.m, self.m + 1).T.copy(),
                Z[: self.k * self.n].reshape(self.k, self.n), d)

    def start_log(self) -> np.ndarray:
        """Per cycle of the last solve (keep_last=True): 1 if the start vector was the residual,
        0 if it was the old Krylov vector (m-k)/2 (the stagnation restart)."""
        cnt = lib.oracle_start_log(self.h, None, 0)
        out = np.zeros(max(cnt, 1), dtype=np.int32)
        lib.oracle_start_log(self.h, _i(out), cnt)
        return out[:cnt]

    def stagnation_cycle(self):
        """(prevP, P, d) of the most recent stagnation restart of the last solve (keep_last=True):
        the basis before that cycle's setup and after its Arnoldi process, (m+1, n) each; None
        when no cycle restarted from the old vector."""
        prev = np.zeros((self.m + 1) * self.n)
        P = np.zeros((self.m + 1) * self.n)
        d = lib.oracle_stagnation_cycle(self.h, _d(prev), _d(P))
        if d < 0:
            return None
        return prev.reshape(self.m + 1, self.n), P.reshape(self.m + 1, self.n), d

    def set_stagnation_rule(self, on: bool):
        lib.oracle_set_stagnation_rule(self.h, 1 if on else 0)

    def stagnation_count(self) -> int:
        return lib.oracle_get_stagnation_count(self.h)
