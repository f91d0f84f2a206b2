"""Each block-solver kernel on random inputs against an FP64 numpy reference of the same operands,
through the kernel-level C ABI (gDMRKernels.h).

Error bounds, stated per element against the absolute-value product |A||B| of the operands:
* bf16 matrix-core products split the FP32 operand into hi + lo bf16 parts with
  |x - hi - lo| <= 2^-17 |x|; bf16 x bf16 products are exact in FP32; the FP32 accumulation over
  at most R terms adds R 2^-24.  Bound: (2^-16 + R 2^-24) |A||B|.
* FP32 matrix-core products: R 2^-23 |A||B|.
A dropped term, a wrong row range or a transposed operand moves an element by a sizeable fraction
of |A||B| / sqrt(n) (about 1e-3 here), far above these bounds.
"""
import ctypes as C

import numpy as np
import pytest

import gdmr

pytestmark = pytest.mark.gpu

N = 50_003          # rows: not a multiple of any tile, so every tail path runs
LD = 50_048         # even leading dimension, padded like the solver's basis
ROWS_PER_BLOCK = 512
EPS32 = 2.0 ** -24


def _rng(seed):
    return np.random.default_rng(seed)


def _bf16_bits(x):
    """float32 array -> bf16 bits (round to nearest even)."""
    b = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)


def _bf16_value(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


class Dev:
    """Device copies of host arrays; keeps the tensors alive."""

    def __init__(self, torch):
        self.torch, self.keep = torch, []

    def put(self, a):
        t = self.torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
        self.keep.append(t)
        return t

    def zeros(self, shape, dtype):
        t = self.torch.zeros(shape, dtype=dtype, device="cuda")
        self.keep.append(t)
        return t


def _cols(rng, n, ncols, ld, bf16=False):
    """A column-major block (ncols x ld storage, rows past n zero) as float32 or bf16 bits, and
    its exact values in FP64 (ncols x n)."""
    x = np.zeros((ncols, ld), np.float32)
    x[:, :n] = rng.standard_normal((ncols, n)).astype(np.float32)
    if bf16:
        bits = _bf16_bits(x)
        return bits, _bf16_value(bits)[:, :n]
    return x, x[:, :n].astype(np.float64)


def _check(got, ref, absref, bound, what):
    err = np.abs(got - ref)
    worst = float(np.max(err / np.maximum(absref, 1e-300)))
    print(f"{what}: max |err| / |A||B| = {worst:.2e} (bound {bound:.2e})", flush=True)
    assert worst <= bound, what


TN_CASES = [(0, 0, 100, 16), (0, 1, 100, 16), (0, 2, 100, 16), (0, 2, 276, 32),
            (1, 3, 20, 10), (1, 3, 276, 16), (1, 3, 100, 20), (1, 3, 100, 32), (1, 2, 100, 32)]


@pytest.mark.parametrize("basis,variant,p,s", TN_CASES)
def test_gemm_tn(torch_gpu, basis, variant, p, s):
    """C = P^T U: every variant of the projection / Gram kernel with its fixed-order reduce."""
    torch = torch_gpu
    rng = _rng(10 * p + s + variant)
    d = Dev(torch)
    Ph, P64 = _cols(rng, N, p, LD, bf16=basis == 1)
    Uh, U64 = _cols(rng, N, s, LD)
    P, U = d.put(Ph), d.put(Uh)
    part = d.zeros(gdmr.lib.dmrKernelTnWorkspace(N, p, ROWS_PER_BLOCK), torch.float32)
    C64 = d.zeros((s, p), torch.float64)     # column-major p x s, ld p
    C32 = d.zeros((p, 32), torch.float32)
    rc = gdmr.lib.dmrKernelGemmTn(basis, P.data_ptr(), LD, p, U.data_ptr(), LD, s, N,
                                  ROWS_PER_BLOCK, variant, part.data_ptr(), C64.data_ptr(), p,
                                  C32.data_ptr(), None)
    torch.cuda.synchronize()
    assert rc == gdmr.DMR_OK
    got = C64.cpu().numpy().T
    ref = P64 @ U64.T
    absref = np.abs(P64) @ np.abs(U64).T
    bound = (2.0 ** -16 if variant == 3 else 0.0) + ROWS_PER_BLOCK * 2 * EPS32
    _check(got, ref, absref, bound, f"gemm_tn basis {basis} variant {variant} p {p} s {s}")
    c32 = C32.cpu().numpy()
    assert np.array_equal(c32[:, :s], got.astype(np.float32))
    assert not np.any(c32[:, s:])


@pytest.mark.parametrize("s", [10, 16])
def test_gemm_tn_pair(torch_gpu, s):
    """P^T U and P^T U2 in one pass over a bf16 P (the last block step's projection with the
    cycle's residual block)."""
    torch = torch_gpu
    rng = _rng(200 + s)
    d = Dev(torch)
    p = 100
    Ph, P64 = _cols(rng, N, p, LD, bf16=True)
    Uh, U64 = _cols(rng, N, s, LD)
    U2h, U264 = _cols(rng, N, s, LD)
    P, U, U2 = d.put(Ph), d.put(Uh), d.put(U2h)
    part = d.zeros(gdmr.lib.dmrKernelTnWorkspace(N, p, ROWS_PER_BLOCK), torch.float32)
    Ca, Cb = d.zeros((s, p), torch.float64), d.zeros((s, p), torch.float64)
    C32 = d.zeros((p, 32), torch.float32)
    rc = gdmr.lib.dmrKernelGemmTnPair(P.data_ptr(), LD, p, U.data_ptr(), U2.data_ptr(), LD, s, N,
                                      ROWS_PER_BLOCK, part.data_ptr(), Ca.data_ptr(), p,
                                      C32.data_ptr(), Cb.data_ptr(), p, None)
    torch.cuda.synchronize()
    assert rc == gdmr.DMR_OK
    bound = 2.0 ** -16 + ROWS_PER_BLOCK * EPS32
    for C, V, tag in ((Ca, U64, "U"), (Cb, U264, "U2")):
        _check(C.cpu().numpy().T, P64 @ V.T, np.abs(P64) @ np.abs(V).T, bound,
               f"gemm_tn_pair {tag} s {s}")
    c32 = C32.cpu().numpy()
    assert np.array_equal(c32[:, :s], Ca.cpu().numpy().T.astype(np.float32))
    assert not np.any(c32[:, s:])


NN_CASES = [(1, 100, 10, True), (1, 101, 16, True), (1, 100, 20, False), (1, 101, 32, False),
            (1, 8, 16, False), (0, 8, 16, False), (0, 16, 32, False)]


@pytest.mark.parametrize("basis,unroll,s,gram", NN_CASES)
def test_gemm_nn(torch_gpu, basis, unroll, s, gram):
    """Out = In - P C (the block update) and, with the Gram epilogue, the partial Gram matrices of
    Out summed in a fixed order: Out^T Out of the stored output."""
    torch = torch_gpu
    rng = _rng(300 + unroll + s)
    d = Dev(torch)
    p = 100
    Ph, P64 = _cols(rng, N, p, LD, bf16=basis == 1)
    Ch = np.zeros((p, 32), np.float32)
    Ch[:, :s] = rng.standard_normal((p, s)).astype(np.float32) * 0.1
    Inh, In64 = _cols(rng, N, s, LD)
    P, Cd, In = d.put(Ph), d.put(Ch), d.put(Inh)
    Out = d.zeros((s, LD), torch.float32)
    G = gdmr.lib.dmrKernelGramBlocks(N, unroll) if gram else 0
    part = d.zeros(max(G, 1) * 32 * 32, torch.float32)
    rc = gdmr.lib.dmrKernelGemmNn(basis, P.data_ptr(), LD, p, Cd.data_ptr(), In.data_ptr(), LD,
                                  1.0, -1.0, Out.data_ptr(), LD, s, N, unroll,
                                  part.data_ptr() if gram else None, None)
    torch.cuda.synchronize()
    assert rc == gdmr.DMR_OK
    C64 = Ch[:, :s].astype(np.float64)
    ref = In64 - (P64.T @ C64).T
    absref = np.abs(In64) + (np.abs(P64).T @ np.abs(C64)).T
    bound = (2.0 ** -16 if unroll >= 100 else 0.0) + p * 2 * EPS32
    o = Out.cpu().numpy()
    _check(o[:, :N].astype(np.float64), ref, absref, bound,
           f"gemm_nn basis {basis} unroll {unroll} s {s}")
    assert not np.any(o[:, N:])  # rows past n are not written
    if gram:
        G64 = d.zeros((s, s), torch.float64)
        assert gdmr.lib.dmrKernelGramReduce(part.data_ptr(), G, s, G64.data_ptr(), None) == 0
        torch.cuda.synchronize()
        O = o[:, :N].astype(np.float64)
        rows_per_wg = 4 * 128 * (2 if unroll == 101 else 1)
        _check(G64.cpu().numpy().T, O @ O.T, np.abs(O) @ np.abs(O).T, rows_per_wg * 2 * EPS32,
               f"gram epilogue unroll {unroll} s {s}")


@pytest.mark.parametrize("unroll,s", [(100, 10), (101, 16)])
def test_gemm_nn_pair(torch_gpu, unroll, s):
    """X1 -= P C[:, 0:16] and X2 -= P C[:, 16:32] in one pass (the cycle end's residual and
    solution updates); columns past s of both blocks are left alone."""
    torch = torch_gpu
    rng = _rng(400 + unroll + s)
    d = Dev(torch)
    p = 120
    Ph, P64 = _cols(rng, N, p, LD, bf16=True)
    Ch = np.zeros((p, 32), np.float32)
    Ch[:, :s] = rng.standard_normal((p, s)).astype(np.float32) * 0.1
    Ch[:, 16:16 + s] = rng.standard_normal((p, s)).astype(np.float32) * 0.1
    X1h, X164 = _cols(rng, N, 16, LD)
    X2h, X264 = _cols(rng, N, 16, LD)
    P, Cd, X1, X2 = d.put(Ph), d.put(Ch), d.put(X1h), d.put(X2h)
    rc = gdmr.lib.dmrKernelGemmNnPair(P.data_ptr(), LD, p, Cd.data_ptr(), X1.data_ptr(),
                                      X2.data_ptr(), LD, s, N, unroll, None)
    torch.cuda.synchronize()
    assert rc == gdmr.DMR_OK
    bound = 2.0 ** -16 + p * 2 * EPS32
    for X, X64, Xh, c0, tag in ((X1, X164, X1h, 0, "X1"), (X2, X264, X2h, 16, "X2")):
        C64 = Ch[:, c0:c0 + s].astype(np.float64)
        got = X.cpu().numpy()
        ref = X64[:s] - (P64.T @ C64).T
        absref = np.abs(X64[:s]) + (np.abs(P64).T @ np.abs(C64)).T
        _check(got[:s, :N].astype(np.float64), ref, absref, bound, f"gemm_nn_pair {tag} s {s}")
        assert np.array_equal(got[s:], Xh[s:])


def _ragged(seed=5):
    """conv_diff_3d(24) with extra entries on some rows (rows of 4 to 10 entries, so the up-front
    and LDS-staged SpMM variants take their long-row loops too)."""
    A = gdmr.conv_diff_3d(24, jitter=0.01, seed=seed)
    n = A.n
    rng = _rng(seed)
    rows_c, rows_v = [], []
    for r in range(n):
        c = list(A.col[A.rowptr[r]:A.rowptr[r + 1]])
        v = list(A.val[A.rowptr[r]:A.rowptr[r + 1]])
        if r % 7 == 0:
            extra = [j for j in rng.choice(n, 3, replace=False) if j not in c]
            c += extra
            v += list(rng.standard_normal(len(extra)))
        elif r % 11 == 0 and len(c) > 4:
            c, v = c[:4], v[:4]
        order = np.argsort(c)
        rows_c.append(np.array(c, np.int32)[order])
        rows_v.append(np.array(v)[order])
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(c) for c in rows_c])
    return gdmr.CsrMatrix(n, rp, np.concatenate(rows_c), np.concatenate(rows_v))


@pytest.mark.parametrize("xbasis", [0, 1], ids=["fp32", "bf16"])
def test_spmm_variants(torch_gpu, xbasis):
    """Y = A X for 13 columns with FP32 values: every variant gives the same bits, within the
    FP32 accumulation bound of the FP64 product."""
    torch = torch_gpu
    A = _ragged()
    n, s = A.n, 13
    ld = n + 64
    rng = _rng(500 + xbasis)
    d = Dev(torch)
    Xh, X64 = _cols(rng, n, s, ld, bf16=xbasis == 1)
    val32 = A.val.astype(np.float32)
    rp, ci, va, X = d.put(A.rowptr), d.put(A.col), d.put(val32), d.put(Xh)
    variants = ([4, 8, 16] if xbasis == 0 else []) + [104, 108, 116, 208, 216]
    outs = {}
    for v in variants:
        Y = d.zeros((s, ld), torch.float32)
        rc = gdmr.lib.dmrKernelSpmm(xbasis, n, A.nnz, rp.data_ptr(), ci.data_ptr(),
                                    va.data_ptr(), X.data_ptr(), ld, Y.data_ptr(), ld, s, v, None)
        torch.cuda.synchronize()
        assert rc == gdmr.DMR_OK, v
        outs[v] = Y.cpu().numpy()[:, :n]
    A64 = gdmr.CsrMatrix(n, A.rowptr, A.col, val32.astype(np.float64))
    Aabs = gdmr.CsrMatrix(n, A.rowptr, A.col, np.abs(val32.astype(np.float64)))
    ref = np.stack([A64.matvec(X64[j]) for j in range(s)])
    absref = np.stack([Aabs.matvec(np.abs(X64[j])) for j in range(s)])
    first = outs[variants[0]]
    for v, y in outs.items():
        assert np.array_equal(y, first), v
    _check(first.astype(np.float64), ref, absref, 10 * 2 * EPS32, f"spmm basis {xbasis}")


@pytest.mark.parametrize("s", [10, 16, 20, 32])
def test_chol_qr(torch_gpu, s):
    """G = S^T S with S upper triangular (numpy's Cholesky factor), S^-1 to FP32 rounding, the
    accumulated factor S2 S1 over two passes, and a dependent column zeroed and flagged."""
    torch = torch_gpu
    rng = _rng(600 + s)
    d = Dev(torch)
    M = rng.standard_normal((300, s))
    G = M.T @ M
    Gd = d.put(np.asfortranarray(G).T.reshape(-1))  # column-major s x s
    Sinv = d.zeros((32, 32), torch.float32)
    Sacc = d.zeros((s, s), torch.float64)
    H = d.zeros((s, s + 3), torch.float64)  # column-major, ld s + 3
    flags = d.zeros(2, torch.int32)
    rc = gdmr.lib.dmrKernelCholQr(Gd.data_ptr(), s, 1e-12, Sinv.data_ptr(), Sacc.data_ptr(), 1,
                                  H.data_ptr(), s + 3, flags.data_ptr(), None)
    torch.cuda.synchronize()
    assert rc == gdmr.DMR_OK and flags.cpu().numpy()[0] == 0
    S = Sacc.cpu().numpy().T  # column-major s x s
    Sref = np.linalg.cholesky(G).T
    assert np.max(np.abs(S - Sref)) <= 1e-12 * np.max(np.abs(Sref))
    assert np.array_equal(np.tril(S, -1), np.zeros((s, s)))
    assert np.array_equal(H.cpu().numpy()[:, :s].T, S)
    Si = Sinv.cpu().numpy()
    assert not np.any(Si[:, s:]) and not np.any(Si[s:, :])
    Sinv_ref = np.linalg.inv(Sref)
    assert np.max(np.abs(Si[:s, :s] - Sinv_ref)) <= 1e-6 * np.max(np.abs(Sinv_ref))
    # Second pass on the Gram matrix of the normalised block: Sacc <- S2 S1.
    Q = M @ Sinv_ref
    G2 = Q.T @ Q
    G2d = d.put(np.asfortranarray(G2).T.reshape(-1))
    rc = gdmr.lib.dmrKernelCholQr(G2d.data_ptr(), s, 1e-12, Sinv.data_ptr(), Sacc.data_ptr(), 0,
                                  None, 0, flags.data_ptr(), None)
    torch.cuda.synchronize()
    S21 = Sacc.cpu().numpy().T
    ref21 = np.linalg.cholesky(G2).T @ Sref
    assert np.max(np.abs(S21 - ref21)) <= 1e-12 * np.max(np.abs(ref21))
    # A dependent column (column 3 = column 1): flagged, its row of S and column of S^-1 zero.
    Md = M.copy()
    Md[:, 3] = Md[:, 1]
    Gdd = d.put(np.asfortranarray(Md.T @ Md).T.reshape(-1))
    rc = gdmr.lib.dmrKernelCholQr(Gdd.data_ptr(), s, 1e-12, Sinv.data_ptr(), Sacc.data_ptr(), 1,
                                  None, 0, flags.data_ptr(), None)
    torch.cuda.synchronize()
    assert flags.cpu().numpy()[0] == 1
    assert not np.any(Sacc.cpu().numpy().T[3]) and not np.any(Sinv.cpu().numpy()[:, 3])


def _gram_partials(rng, G, s, dep=False):
    """G partial Gram matrices B_g^T B_g (B_g: 16 x s) in the update epilogue's layout
    part[(i G + g) 32 + j] (16 rows), and their FP64 sum (s x s, symmetric positive definite;
    with dep, column 3 repeats column 1 in every B_g, so the sum is singular)."""
    B = rng.standard_normal((G, 16, s)).astype(np.float32)
    if dep:
        B[:, :, 3] = B[:, :, 1]
    pg = np.einsum("gki,gkj->gij", B.astype(np.float64), B.astype(np.float64)).astype(np.float32)
    part = np.zeros((16, G, 32), np.float32)
    part[:s, :, :s] = pg.transpose(1, 0, 2)
    return part.reshape(-1), pg.astype(np.float64).sum(axis=0)


@pytest.mark.parametrize("s,G", [(10, 4096), (16, 98), (10, 1), (16, 20011)])
def test_gram_chol(torch_gpu, s, G):
    """The fused Gram reduce + Cholesky step: S from the FP64 sum of the partials against numpy's
    Cholesky factor, S^-1 to FP32 rounding, and the same factor as the two-launch path
    (dmrKernelGramReduce + dmrKernelCholQr) to the rounding of the sum order; the work counter
    is zero after each call, so a second call on the same work buffer (other G) is right too."""
    torch = torch_gpu
    rng = _rng(700 + s + G)
    d = Dev(torch)
    parth, Gref = _gram_partials(rng, G, s)
    part = d.put(parth)
    work = d.zeros(gdmr.DMR_KGRAMCHOL_WORK_BYTES // 4, torch.int32)
    Sinv = d.zeros((32, 32), torch.float32)
    Sacc = d.zeros((s, s), torch.float64)
    H = d.zeros((s, s + 3), torch.float64)
    flags = d.zeros(2, torch.int32)
    rc = gdmr.lib.dmrKernelGramChol(part.data_ptr(), G, s, 1e-12, Sinv.data_ptr(),
                                    Sacc.data_ptr(), H.data_ptr(), s + 3, flags.data_ptr(),
                                    work.data_ptr(), None)
    torch.cuda.synchronize()
    assert rc == gdmr.DMR_OK and flags.cpu().numpy()[0] == 0
    assert work.cpu().numpy()[0] == 0  # the arrival counter is back at zero
    S = Sacc.cpu().numpy().T
    Sref = np.linalg.cholesky(Gref).T
    err = np.max(np.abs(S - Sref)) / np.max(np.abs(Sref))
    print(f"gram_chol s {s} G {G}: max |S - S_ref| / max |S_ref| = {err:.2e}", flush=True)
    assert err <= 1e-12
    assert np.array_equal(np.tril(S, -1), np.zeros((s, s)))
    assert np.array_equal(H.cpu().numpy()[:, :s].T, S)
    Si = Sinv.cpu().numpy()
    assert not np.any(Si[:, s:]) and not np.any(Si[s:, :])
    Sinv_ref = np.linalg.inv(Sref)
    assert np.max(np.abs(Si[:s, :s] - Sinv_ref)) <= 1e-6 * np.max(np.abs(Sinv_ref))
    # The two-launch path on the same partials (for G = 1 both sum the one partial, so the
    # factorisations must give the same bits).
    G64 = d.zeros((s, s), torch.float64)
    Sacc2 = d.zeros((s, s), torch.float64)
    assert gdmr.lib.dmrKernelGramReduce(part.data_ptr(), G, s, G64.data_ptr(), None) == 0
    assert gdmr.lib.dmrKernelCholQr(G64.data_ptr(), s, 1e-12, Sinv.data_ptr(), Sacc2.data_ptr(), 1,
                                    None, 0, flags.data_ptr(), None) == 0
    torch.cuda.synchronize()
    S2 = Sacc2.cpu().numpy().T
    assert np.max(np.abs(S - S2)) <= 1e-13 * np.max(np.abs(S2))
    if G == 1:
        assert np.array_equal(S, S2) and np.array_equal(Si, Sinv.cpu().numpy())
    # A second call on the same work buffer, with a singular sum: flagged, column 3 zeroed.
    G2 = max(G // 2, 1) + 3
    parth2, _ = _gram_partials(rng, G2, s, dep=True)
    part2 = d.put(parth2)
    flags.zero_()
    rc = gdmr.lib.dmrKernelGramChol(part2.data_ptr(), G2, s, 1e-12, Sinv.data_ptr(),
                                    Sacc.data_ptr(), None, 0, flags.data_ptr(), work.data_ptr(),
                                    None)
    torch.cuda.synchronize()
    assert rc == gdmr.DMR_OK and flags.cpu().numpy()[0] == 1
    assert work.cpu().numpy()[0] == 0
    assert not np.any(Sacc.cpu().numpy().T[3]) and not np.any(Sinv.cpu().numpy()[:, 3])


@pytest.mark.parametrize("outbasis", [0, 1], ids=["fp32", "bf16"])
def test_scale_block(torch_gpu, outbasis):
    """Out = U C for the Cholesky QR normalisation (C upper triangular), into FP32 or bf16."""
    torch = torch_gpu
    s = 12
    rng = _rng(700 + outbasis)
    d = Dev(torch)
    Uh, U64 = _cols(rng, N, s, LD)
    Ch = np.zeros((s, 32), np.float32)
    Ch[:, :s] = np.triu(rng.standard_normal((s, s))).astype(np.float32)
    U, Cd = d.put(Uh), d.put(Ch)
    Out = d.zeros((s, LD), torch.float32 if outbasis == 0 else torch.int16)
    rc = gdmr.lib.dmrKernelScaleBlock(U.data_ptr(), LD, Cd.data_ptr(), Out.data_ptr(), LD,
                                      outbasis, s, N, None)
    torch.cuda.synchronize()
    assert rc == gdmr.DMR_OK
    o = Out.cpu().numpy()[:, :N]
    got = o.astype(np.float64) if outbasis == 0 else _bf16_value(o.view(np.uint16))
    C64 = Ch[:, :s].astype(np.float64)
    ref = (U64.T @ C64).T
    absref = (np.abs(U64).T @ np.abs(C64)).T
    bound = s * 2 * EPS32 + (0.0 if outbasis == 0 else 2.0 ** -8)
    _check(got, ref, absref, bound, f"scale_block out {outbasis}")


@pytest.mark.parametrize("k,s,steps", [(20, 10, 16), (20, 16, 16), (32, 32, 10), (20, 6, 20),
                                        (20, 20, 10)])
def test_lsq_block(torch_gpu, k, s, steps):
    """Banded block-Hessenberg least squares (lower bandwidth s): Y per right-hand side against
    numpy.linalg.lstsq (LAPACK), Yr32 = Y diag(rho), ZH32 = H Y.  Bandwidths 6, 10 and 16 run on
    the register path's 16-row window, 20 and 32 on its 32-row window."""
    torch = torch_gpu
    cols, bw = k + steps * s, s
    rows = cols + bw
    assert gdmr.lib.dmrKernelLsqFits(cols, bw, s) == 1
    rng = _rng(800 + s)
    d = Dev(torch)
    Hm = rng.standard_normal((rows, cols))
    i, j = np.indices((rows, cols))
    Hm[i > j + bw] = 0.0
    Hm[np.arange(cols), np.arange(cols)] += 4.0
    T = rng.standard_normal((rows, s))
    rho = rng.uniform(0.5, 2.0, s)
    Hd = d.put(np.asfortranarray(Hm).T.reshape(-1))
    Td = d.put(np.asfortranarray(T).T.reshape(-1))
    rhod = d.put(rho)
    R = d.zeros(cols * cols, torch.float64)
    Qt = d.zeros(cols * 32, torch.float64)
    Y = d.zeros((cols, 32), torch.float64)
    Yr = d.zeros((cols, 32), torch.float32)
    ZH = d.zeros((rows, 32), torch.float32)
    rc = gdmr.lib.dmrKernelLsqBlock(Hd.data_ptr(), rows, cols, bw, Td.data_ptr(), s,
                                    rhod.data_ptr(), R.data_ptr(), Qt.data_ptr(), Y.data_ptr(),
                                    Yr.data_ptr(), ZH.data_ptr(), None)
    torch.cuda.synchronize()
    assert rc == gdmr.DMR_OK
    y = Y.cpu().numpy()[:, :s]
    yref = np.linalg.lstsq(Hm, T, rcond=None)[0]
    err = np.max(np.abs(y - yref)) / np.max(np.abs(yref))
    print(f"lsq_block k {k} s {s} steps {steps}: max rel error {err:.2e}", flush=True)
    assert err <= 1e-11
    assert np.max(np.abs(Yr.cpu().numpy()[:, :s] - y * rho)) <= 2 * EPS32 * np.max(np.abs(y * rho))
    zh = Hm @ y
    assert np.max(np.abs(ZH.cpu().numpy()[:, :s] - zh)) <= 4 * EPS32 * np.max(np.abs(zh))
