"""CPU pins of the FP64 oracle and of the problem generators.

Each check is chosen so a plausible mistake (a dropped term, a wrong sign or index, a transposed
operand, a wrong column picked from the SVD) fails it: independent dense assembly of the PDE
stencils, LAPACK (numpy) least squares and SVD, an independent textbook GMRES(m) for k = 0, the
Arnoldi relation and basis orthogonality of a deflated cycle, the subspace refresh recomputed
from LAPACK's SVD, the stagnation restart's start vector and rule with deflation, breakdown on an
operator with few distinct eigenvalues, exact deflation, and FP64-checked true residuals.
"""
import math

import numpy as np
import pytest

import cases
import gdmr
import oracle


# ------------------------------------------------------------------------------------------------
# Random numbers: published SplitMix64 outputs (Vigna's splitmix64.c / Java SplittableRandom).
# ------------------------------------------------------------------------------------------------
def test_splitmix64_known_answers():
    assert [hex(v) for v in gdmr.splitmix64(0, 3)] == [
        "0xe220a8397b1dcdaf", "0x6e789e6aa1b965f4", "0x6c45d188009454f"]
    assert list(gdmr.splitmix64(1234567, 5)) == [
        6457827717110365317, 3203168211198807973, 9817491932198370423,
        4593380528125082431, 16408922859458223821]


def test_normals_are_standard_and_reproducible():
    a = gdmr.rand_normal(0, 1, 200_001)
    assert np.array_equal(a, gdmr.rand_normal(0, 1, 200_001))
    assert not np.array_equal(a[:100], gdmr.rand_normal(0, 2, 100))
    assert abs(a.mean()) < 0.01 and abs(a.std() - 1.0) < 0.01
    u = gdmr.rand_uniform(3, 4, 100_000, -0.01, 0.01)
    assert u.min() >= -0.01 and u.max() < 0.01 and abs(u.mean()) < 2e-4


# ------------------------------------------------------------------------------------------------
# Generators against an independent dense assembly.
# ------------------------------------------------------------------------------------------------
def dense_convdiff(N, dim, a, b, scale):
    """Dense h^2 (-div(a s grad u) + b.grad u), written from the PDE, not from the generator."""
    h = 1.0 / (N + 1)
    n = N ** dim
    A = np.zeros((n, n))
    idx = np.array(np.unravel_index(np.arange(n), (N,) * dim, order="F")).T  # x fastest
    strides = [N ** d for d in range(dim)]
    for r in range(n):
        for d in range(dim):
            for sgn in (-1, 1):
                inside = 0 <= idx[r, d] + sgn < N
                nb = r + sgn * strides[d]
                face = a[d] * (0.5 * (scale[r] + scale[nb]) if inside else scale[r])
                A[r, r] += face
                if inside:
                    A[r, nb] += -face + sgn * b[d] * h / 2
    return A


@pytest.mark.parametrize("N,jit", [(5, 0.0), (4, 0.05)])
def test_convdiff_2d_matches_dense_assembly(N, jit):
    A = gdmr.conv_diff_2d(N, 0.1, (1.0, 0.5), jitter=jit, seed=3, stream=9)
    s = 1.0 + (gdmr.rand_uniform(3, 9, N * N, -jit, jit) if jit else np.zeros(N * N))
    D = dense_convdiff(N, 2, (0.1, 0.1), (1.0, 0.5), s)
    assert np.allclose(A.to_scipy().toarray(), D, rtol=0, atol=1e-15)
    assert A.nnz == 5 * N * N - 4 * N
    assert all(np.all(np.diff(A.col[A.rowptr[i]:A.rowptr[i + 1]]) > 0) for i in range(A.n))


@pytest.mark.parametrize("N,jit", [(4, 0.0), (3, 0.02)])
def test_convdiff_3d_matches_dense_assembly(N, jit):
    A = gdmr.conv_diff_3d(N, 0.1, (1.0, 0.5, 0.25), jitter=jit, seed=5, stream=2)
    s = 1.0 + (gdmr.rand_uniform(5, 2, N ** 3, -jit, jit) if jit else np.zeros(N ** 3))
    D = dense_convdiff(N, 3, (0.1,) * 3, (1.0, 0.5, 0.25), s)
    assert np.allclose(A.to_scipy().toarray(), D, rtol=0, atol=1e-15)
    assert A.nnz == 7 * N ** 3 - 6 * N ** 2


def test_aniso_and_kron_assembly():
    """Constant coefficients: the stencil equals a Kronecker-sum assembly (scipy.sparse)."""
    import scipy.sparse as sp
    N = 6
    A = gdmr.aniso_3d(N, (1.0, 0.1, 0.01)).to_scipy()
    T = sp.diags([-np.ones(N - 1), 2 * np.ones(N), -np.ones(N - 1)], [-1, 0, 1])
    I = sp.identity(N)
    K = (sp.kron(I, sp.kron(I, T)) * 1.0 + sp.kron(I, sp.kron(T, I)) * 0.1 +
         sp.kron(T, sp.kron(I, I)) * 0.01)
    assert abs(A - K).max() < 1e-15
    # Interior rows of constant-coefficient convection-diffusion sum to zero.
    C2 = gdmr.conv_diff_2d(8).to_scipy()
    rs = np.asarray(C2.sum(axis=1)).ravel().reshape(8, 8, order="F")
    assert np.max(np.abs(rs[1:-1, 1:-1])) < 1e-15


# ------------------------------------------------------------------------------------------------
# Oracle kernels.
# ------------------------------------------------------------------------------------------------
def test_spmv_and_dot():
    A = gdmr.conv_diff_3d(7, jitter=0.01, seed=1)
    x = gdmr.rand_normal(2, 3, A.n)
    y = oracle.spmv(A, x)
    ys = A.to_scipy() @ x
    assert np.max(np.abs(y - ys)) <= 1e-15 * np.max(np.abs(ys)) * 8
    for n in (1, 7, 4096, 100_003):
        a, b = gdmr.rand_normal(4, 5, n), gdmr.rand_normal(6, 7, n)
        exact = math.fsum(a * b)
        for G in (1, 3, gdmr.reduction_blocks(n)):
            d = oracle.dot(a, b, G)
            assert abs(d - exact) <= 2 * n * 2.2e-16 * np.sum(np.abs(a * b)) + 1e-300


def _hessenberg(m, seed, shift=4.0):
    """Random upper Hessenberg (m+1) x m with a diagonal shift that keeps it well conditioned."""
    rng = np.random.default_rng(seed)
    H = np.triu(rng.standard_normal((m + 1, m)), -1)
    H[np.arange(m), np.arange(m)] += shift
    return H


@pytest.mark.parametrize("m", [1, 5, 30, 50])
def test_lsq_matches_lapack(m):
    H = _hessenberg(m, m)
    t = np.random.default_rng(100 + m).standard_normal(m + 1)
    y, z = oracle.lsq(H, t)
    ref = np.linalg.lstsq(H, t, rcond=None)[0]
    kappa = np.linalg.cond(H)
    assert y[m] == 0.0
    assert np.max(np.abs(y[:m] - ref)) <= 1e-14 * kappa * max(1.0, np.max(np.abs(ref)))
    assert np.allclose(z, H @ y[:m], rtol=0, atol=1e-12 * np.abs(H).max() * np.abs(y).max() * m)
    # Normal equations: H^T (t - H y) = 0.
    assert np.max(np.abs(H.T @ (t - H @ y[:m]))) <= 1e-10 * np.linalg.norm(t) * np.linalg.norm(H)


@pytest.mark.parametrize("m,k", [(6, 2), (30, 10), (50, 20), (31, 5)])
def test_svd_matches_lapack(m, k):
    H = _hessenberg(m, 7 * m)
    Cm, sig, sweeps = oracle.svd_smallest(H, k)
    U, S, Vt = np.linalg.svd(H)
    assert 1 <= sweeps < 128
    assert np.max(np.abs(sig - S)) <= 1e-12 * S[0]
    for i in range(k):
        v = Vt[m - k + i]
        c = Cm[:, i]
        assert abs(np.linalg.norm(c) - 1.0) < 1e-12
        gap = min(abs(S[m - k + i] - S[j]) for j in range(m) if j != m - k + i) / S[0]
        if gap > 1e-6:
            assert min(np.linalg.norm(c - v), np.linalg.norm(c + v)) < 1e-9 / gap
        j = np.argmax(np.abs(c))
        assert c[j] > 0
    assert np.max(np.abs(Cm.T @ Cm - np.eye(k))) < 1e-12


# ------------------------------------------------------------------------------------------------
# Whole solver.
# ------------------------------------------------------------------------------------------------
def textbook_gmres(A, b, m, tol, max_cycles):
    """Restarted GMRES(m), classical Gram-Schmidt, LSQ on beta e1 (Saad, Iterative Methods for
    Sparse Linear Systems, Alg. 6.9), written independently of the oracle.  Returns the relative
    residual after each cycle."""
    n = b.size
    x = np.zeros(n)
    beta0 = np.linalg.norm(b)
    hist = []
    for _ in range(max_cycles):
        r = b - A @ x
        beta = np.linalg.norm(r)
        if beta <= tol * beta0:
            break
        V = np.zeros((n, m + 1))
        Hm = np.zeros((m + 1, m))
        V[:, 0] = r / beta
        for j in range(m):
            w = A @ V[:, j]
            h = V[:, :j + 1].T @ w
            w = w - V[:, :j + 1] @ h
            Hm[:j + 1, j] = h
            Hm[j + 1, j] = np.linalg.norm(w)
            V[:, j + 1] = w / Hm[j + 1, j]
        e1 = np.zeros(m + 1)
        e1[0] = beta
        y = np.linalg.lstsq(Hm, e1, rcond=None)[0]
        x = x + V[:, :m] @ y
        hist.append(np.linalg.norm(b - A @ x) / beta0)
        if hist[-1] <= tol:
            break
    return np.array(hist), x


def test_k0_reduces_to_textbook_gmres():
    A = gdmr.conv_diff_2d(20)
    b = gdmr.rand_normal(0, 1, A.n)
    m, tol = 15, 1e-9
    ora = oracle.Oracle(A, m, 0, gdmr.reduction_blocks(A.n))
    x, res, its, st = ora.solve(b, tol=tol)
    h = ora.history()
    ref, xr = textbook_gmres(A.to_scipy(), b, m, tol, 200)
    assert st == 0 and len(h) == len(ref)
    assert np.max(np.abs(h - ref) / ref) < 1e-6
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-7


def test_config_a_sequence_properties():
    """Config A: residual histories never increase (q <= 1), the true FP64 residual meets the
    tolerance and equals the tracked one, and recycling Z cuts the cycle count."""
    A = gdmr.conv_diff_2d(64)
    n, m, k, tol = A.n, 30, 10, 1e-8
    B = gdmr.rand_normal(0, 1, 4 * n).reshape(4, n)
    ora = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n))
    cycles = []
    for i in range(4):
        x, res, its, st = ora.solve(B[i], tol=tol, mode=i)
        h = ora.history()
        true = np.linalg.norm(B[i] - A.matvec(x)) / np.linalg.norm(B[i])
        assert st == 0 and its == len(h) - 1
        assert np.all(np.diff(h) <= 0)
        assert true <= tol and abs(true - res) <= 1e-6 * res
        cycles.append(len(h))
    assert cycles[0] > min(cycles[1:])
    Z = ora.get_z()
    assert np.allclose(np.linalg.norm(Z, axis=1), 1.0, atol=1e-12)


def test_deflated_cycle_arnoldi_relation():
    """A [Z V_0..V_{m-k-1}] = [W V_0..V_{m-k}] H for a deflated cycle, and P = [W V] has
    orthonormal columns up to classical Gram-Schmidt's loss."""
    A = gdmr.conv_diff_2d(24, jitter=0.01, seed=2)
    n, m, k = A.n, 20, 6
    ora = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n), keep_last=True)
    As = A.to_scipy()
    for mode in (0, 1):
        ora.solve(gdmr.rand_normal(1, mode, n), tol=1e-6, mode=mode)
    P, H, Z, d = ora.last_cycle()
    assert d == k
    src = np.vstack([Z, P[k:m]])          # source basis [Z V_0..V_{m-k-1}], m vectors
    lhs = As @ src.T                       # n x m
    rhs = P.T @ H                          # [W V] H, n x m
    assert np.linalg.norm(lhs - rhs) <= 1e-12 * np.linalg.norm(lhs)
    G = P @ P.T
    assert np.max(np.abs(G - np.eye(m + 1))) < 1e-8
    # A wrong block choice breaks the relation: e.g. using P[0..m-1] as the source.
    assert np.linalg.norm(As @ P[:m].T - rhs) > 1e-3 * np.linalg.norm(lhs)


def test_breakdown_on_diagonal_operator():
    n, m = 300, 10
    d = np.repeat([1.0, 2.0, 3.0, 5.0, 8.0], n // 5)
    A = gdmr.CsrMatrix(n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), d)
    for k in (0, 3):
        ora = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n))
        b = gdmr.rand_normal(11, 0, n)
        x, res, its, st = ora.solve(b, tol=1e-10)
        cyc, mv, brk = ora.counts()
        assert st == 0 and brk >= 1 and cyc <= 2
        assert np.linalg.norm(b - d * x) / np.linalg.norm(b) <= 1e-10


def test_exact_deflation_space_helps():
    """Diagonal A with eigenvalues 1..n: with Z = the k smallest eigenvectors the deflated solve
    needs fewer cycles than the undeflated one."""
    n, m, k = 400, 12, 6
    d = np.arange(1, n + 1, dtype=np.float64)
    A = gdmr.CsrMatrix(n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), d)
    b = gdmr.rand_normal(5, 5, n)
    plain = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n))
    plain.solve(b, tol=1e-10, mode=0)
    exact = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n))
    exact.set_z(np.eye(n)[:k])
    x, res, its, st = exact.solve(b, tol=1e-10, mode=1)
    assert st == 0
    assert exact.counts()[0] < plain.counts()[0]
    assert np.linalg.norm(b - d * x) / np.linalg.norm(b) <= 1e-10


def test_stagnation_and_cycle_limit():
    """A cyclic shift stalls GMRES(m) (q = 1 > 0.995 every cycle): the stagnation rule switches
    the start vector and the cycle limit matRank ends the solve (status 6, numIts = cycles)."""
    n, m, k = 64, 8, 0
    A = gdmr.CsrMatrix(n, np.arange(n + 1, dtype=np.int32),
                       ((np.arange(n) + 1) % n).astype(np.int32), np.ones(n))
    b = np.zeros(n)
    b[0] = 1.0
    ora = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n))
    x, res, its, st = ora.solve(b, tol=1e-10, mat_rank=20)
    h = ora.history()
    # Cycle limit (status 6) or, when rounding lifts q a hair above 1, the runaway guard (7).
    assert st in (6, 7) and its == len(h)
    assert np.all(np.abs(h - 1.0) < 1e-12)
    # Both exits return before the solution export (gDMR.cu:2816-2829): x is the initial guess.
    assert not x.any()


def test_early_exit_semantics():
    A = gdmr.conv_diff_2d(8)
    xs = gdmr.rand_normal(9, 9, A.n)
    b = A.matvec(xs)
    ora = oracle.Oracle(A, 6, 2, gdmr.reduction_blocks(A.n))
    x, res, its, st = ora.solve(b, x0=xs, tol=1e-8, mat_rank=50)
    assert st == -1 and its == 2500 and res < 1e-8 * np.linalg.norm(b)
    # mode != 0 with no stored subspace runs exactly like mode 0.
    fresh = oracle.Oracle(A, 6, 2, gdmr.reduction_blocks(A.n))
    x1, res1, its1, st1 = fresh.solve(b, tol=1e-8, mode=1)
    h1 = fresh.history()
    fresh0 = oracle.Oracle(A, 6, 2, gdmr.reduction_blocks(A.n))
    x0, res0, its0, st0 = fresh0.solve(b, tol=1e-8, mode=0)
    assert st1 == 0 and its1 == its0 and np.array_equal(h1, fresh0.history())
    assert np.array_equal(x1, x0)


# ------------------------------------------------------------------------------------------------
# Subspace refresh (genDeflatDMR, gDMR.cu:1583-1704) recomputed independently.
# ------------------------------------------------------------------------------------------------
def _refresh_from_numpy(P, H, Z, d, k, m):
    """Z_new recomputed from a cycle's basis with numpy.linalg.svd (LAPACK): the source basis is
    [Z | V_0 .. V_{m-d-1}] for a deflated cycle (the Krylov slots start at o = k) and
    [V_0 .. V_{m-1}] for an undeflated one; column i uses the right singular vector of the i-th of
    the k smallest singular values of H, in descending order.  Returns the unnormalised
    combinations (k, n), the singular values and the source basis."""
    _, S, Vt = np.linalg.svd(H)
    V = Vt.T
    src = np.vstack([Z, P[k:m]]) if d > 0 else P[:m]
    zt = np.array([src.T @ V[:, m - k + i] for i in range(k)])
    return zt, S, src, V


def _sign_free_err(a, b):
    return min(np.linalg.norm(a - b), np.linalg.norm(a + b))


@pytest.mark.parametrize("N,m,k,seed", [(24, 20, 6, 2), (32, 30, 10, 0)])
def test_refresh_matches_lapack_svd(N, m, k, seed):
    """The oracle's refreshed subspace equals normalize([Z V] Vs[:, m-k:]) computed with LAPACK's
    SVD of the cycle's H, column by column up to sign, for an undeflated first cycle and for a
    deflated cycle.  ||A z~_i|| equals sigma_{m-k+i} for the unnormalised combination (the
    Arnoldi relation carries it), and the check fails for W in place of Z or for the Krylov
    coefficients taken without the offset d."""
    A = gdmr.conv_diff_2d(N, jitter=0.01, seed=seed)
    n = A.n
    As = A.to_scipy()
    # Undeflated: one cycle only (mat_rank = 0 stops before the second).
    ora0 = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n), keep_last=True)
    ora0.solve(gdmr.rand_normal(1, 0, n), tol=1e-14, mat_rank=0, mode=0)
    P, H, Z, d = ora0.last_cycle()
    assert d == 0 and ora0.counts()[0] == 1
    zt, S, _, _ = _refresh_from_numpy(P, H, Z, d, k, m)
    Znew = ora0.get_z()
    for i in range(k):
        assert _sign_free_err(zt[i] / np.linalg.norm(zt[i]), Znew[i]) < 1e-11
    # Deflated: the last cycle of the second solve of a sequence.
    ora = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n), keep_last=True)
    for mode in (0, 1):
        ora.solve(gdmr.rand_normal(1, mode, n), tol=1e-6, mode=mode)
    P, H, Z, d = ora.last_cycle()
    assert d == k
    zt, S, src, V = _refresh_from_numpy(P, H, Z, d, k, m)
    Znew = ora.get_z()
    for i in range(k):
        assert _sign_free_err(zt[i] / np.linalg.norm(zt[i]), Znew[i]) < 1e-11
        sig = S[m - k + i]
        assert abs(np.linalg.norm(As @ zt[i]) - sig) <= 1e-9 * sig
    # Planted mistakes in the recomputation: each misses the oracle's subspace by far.
    zW = P[:m].T @ V[:, m - k]  # W (slots 0..k-1) in place of Z
    assert _sign_free_err(zW / np.linalg.norm(zW), Znew[0]) > 1e-2
    v = V[:, m - k]
    zD = Z.T @ v[:k] + P[k:m].T @ v[:m - k]  # Krylov coefficients without the offset d
    assert _sign_free_err(zD / np.linalg.norm(zD), Znew[0]) > 1e-2
    zS = src.T @ V[:, m - k - 1]  # one singular vector too far (the k+1-th smallest)
    assert _sign_free_err(zS / np.linalg.norm(zS), Znew[0]) > 1e-2


def _pos_qr(M):
    """QR with a positive diagonal of R (the factorisation a Gram-Schmidt QR produces)."""
    Q, R = np.linalg.qr(M)
    sg = np.sign(np.diag(R))
    sg[sg == 0] = 1.0
    return Q * sg, (R.T * sg).T


@pytest.mark.parametrize("N,m,k,seed", [(24, 20, 6, 2), (32, 30, 10, 0)])
def test_deflation_image_from_arnoldi_relation(N, m, k, seed):
    """The FP32 inner solves take the image A Z = W R of a refreshed subspace from the refreshing
    cycle's Arnoldi relation instead of k operator applications and a Gram-Schmidt QR
    (gDMR.cu:1823-2031): with Z~ = [Z V] C and A [Z V] = P H, the QR H C = Q R gives W = P Q and
    R diag(1 / ||z~_i||).  On the FP64 oracle's own cycle (a deflated one) this pair equals the
    positive-diagonal QR of the recomputed A Z to FP64 accuracy, and a wrong coefficient block
    or a basis shifted by one column misses it by far."""
    A = gdmr.conv_diff_2d(N, jitter=0.01, seed=seed)
    n = A.n
    As = A.to_scipy()
    ora = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n), keep_last=True)
    for mode in (0, 1):
        ora.solve(gdmr.rand_normal(1, mode, n), tol=1e-6, mode=mode)
    P, H, Z, d = ora.last_cycle()
    assert d == k
    Znew = ora.get_z()
    zt, S, src, V = _refresh_from_numpy(P, H, Z, d, k, m)
    # Refresh coefficients with the oracle's signs.
    C = np.empty((m, k))
    for i in range(k):
        c = V[:, m - k + i]
        zi = src.T @ c
        if np.linalg.norm(zi / np.linalg.norm(zi) - Znew[i]) > 1e-6:
            c = -c
        C[:, i] = c
    ztn = np.linalg.norm(src.T @ C, axis=0)
    # Explicit image: the positive-diagonal QR of A Z_new.
    We, Re = _pos_qr(np.column_stack([As @ Znew[i] for i in range(k)]))
    # Implicit image from the relation.
    Q, R = _pos_qr(H @ C)
    Wi = P.T @ Q
    Ri = R / ztn[None, :]
    scale = np.abs(Re).max()
    assert np.abs(Wi - We).max() < 1e-9
    assert np.abs(Ri - Re).max() < 1e-9 * scale
    assert np.abs(np.column_stack([As @ Znew[i] for i in range(k)]) - Wi @ Ri).max() < 1e-10 * scale
    # Planted mistakes: the coefficient block one singular vector off, the basis shifted.
    C2 = C.copy()
    C2[:, 0] = V[:, m - k - 1]
    Q2, _ = _pos_qr(H @ C2)
    assert np.abs(P.T @ Q2 - We).max() > 1e-2
    Qs, _ = _pos_qr(H[:m] @ C)
    assert np.abs(P[1:].T @ Qs - We).max() > 1e-2


# ------------------------------------------------------------------------------------------------
# Stagnation restart with deflation (arnoldiSetupDMR, gDMR.cu:2586-2632; flag gDMR.cu:2506-2514).
# ------------------------------------------------------------------------------------------------


@pytest.mark.parametrize("n,alpha,m,k,mat_rank", cases.STAG_CASES)
def test_stagnation_restart_with_deflation(n, alpha, m, k, mat_rank):
    A = cases.shift_plus_diag(n, alpha)
    b = gdmr.rand_normal(3, 0, n)
    ora = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n), keep_last=True)
    x, res, its, st = ora.solve(b, tol=1e-10, mat_rank=mat_rank, mode=0)
    h = ora.history()
    log = ora.start_log()
    assert st == 6 and len(log) == len(h) == mat_rank + 1
    # The start choice of every cycle follows the reference's rule, replayed from the residual
    # history: the flag is q = rho_c / rho_{c-1} > 0.995 (gDMR.cu:2506-2514), the counter
    # persists, and a flagged cycle restarts from the old vector unless count % period == 0.
    period = 2 * mat_rank // (m - k)
    q = h / np.concatenate([[1.0], h[:-1]])
    count, flag = 0, 0
    for c in range(len(h)):
        from_res = flag == 0 or period == 0 or count % period == 0
        assert log[c] == (1 if from_res else 0), c
        if q[c] > 0.995:
            flag, count = 1, count + 1
        else:
            flag = 0
    assert count == ora.stagnation_count() and (log == 0).sum() >= 2
    # The restarted cycle's start vector is the old Krylov vector (m-k)/2 of the slots o = k,
    # orthogonalised against the new W and normalised (deflatorDMR(0, 0)).
    prevP, P, d = ora.stagnation_cycle()
    assert d == k
    W = P[:k]
    slot = k + (m - k) // 2

    def start_from(v):
        v = v - W.T @ (W @ v)
        return v / np.linalg.norm(v)

    assert np.linalg.norm(start_from(prevP[slot]) - P[k]) < 1e-12
    for wrong in (slot - 1, slot + 1, k):  # neighbouring slots and the old residual slot
        assert np.linalg.norm(start_from(prevP[wrong]) - P[k]) > 1e-3
    # The rule changes the residual history (so a wrong slot would show in any comparison).
    off = oracle.Oracle(A, m, k, gdmr.reduction_blocks(n))
    off.set_stagnation_rule(False)
    off.solve(b, tol=1e-10, mat_rank=mat_rank, mode=0)
    assert np.max(np.abs(off.history() - h) / h) > 1e-2


def test_set_values_keeps_the_sequence_state():
    """Changing the operator values (a jittered sequence) keeps the subspace and the stagnation
    counter: solving system 1 after set_values equals a fresh oracle of system 1 given system 0's
    subspace, and differs from solving system 0's operator again."""
    A0 = gdmr.conv_diff_2d(20, jitter=0.01, seed=8, stream=0)
    A1 = gdmr.conv_diff_2d(20, jitter=0.01, seed=8, stream=1)
    n, m, k = A0.n, 16, 5
    b0, b1 = gdmr.rand_normal(8, 0, n), gdmr.rand_normal(8, 1, n)
    seq = oracle.Oracle(A0, m, k, gdmr.reduction_blocks(n))
    seq.solve(b0, tol=1e-10, mode=0)
    Z0 = seq.get_z()
    seq.set_values(A1)
    x1, _, _, _ = seq.solve(b1, tol=1e-10, mode=1)
    fresh = oracle.Oracle(A1, m, k, gdmr.reduction_blocks(n))
    fresh.set_z(Z0)
    xf, _, _, _ = fresh.solve(b1, tol=1e-10, mode=1)
    assert np.array_equal(x1, xf) and np.array_equal(seq.history(), fresh.history())
    assert np.linalg.norm(b1 - A1.matvec(x1)) / np.linalg.norm(b1) <= 1e-10
    assert np.linalg.norm(b1 - A0.matvec(x1)) / np.linalg.norm(b1) > 1e-6
