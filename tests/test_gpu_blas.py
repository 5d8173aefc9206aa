"""GPU parity of the BLAS-style GEMM (wgemmdd / mfhip_gemm: C = alpha op(A) op(B)
+ beta C, the interface of the reference's xGEMM shim
test/integration/blas/wgemm.f90) against the oracle's or_gemm_blas, bit for bit:
every transpose pair, padded leading dimensions, alpha/beta forms, renorm
intervals, both tile shapes, the quick returns and the argument errors."""
import itertools

import numpy as np
import pytest

import multifloats_amd as mf
from tests import _native as nt
from tests.test_blas_cpu import store

pytestmark = pytest.mark.gpu

ALPHA, BETA = (1.5, 2.0 ** -60), (-0.75, 2.0 ** -58)


def both(ta, tb, m, n, k, alpha, beta, ri, rng, pad_a=3, pad_b=2, pad_c=5):
    A = store(rng, m, k, m + pad_a) if ta in "Nn" else store(rng, k, m, k + pad_a)
    B = store(rng, k, n, k + pad_b) if tb in "Nn" else store(rng, n, k, n + pad_b)
    C0 = store(rng, m, n, m + pad_c)
    want = C0.copy(order="K")
    assert nt.or_gemm_blas(ta, tb, m, n, k, alpha, A, B, beta, want, ri) == 0
    got = mf.gemm(ta, tb, alpha, A, B, beta, C0.copy(order="K"), m=m, n=n, k=k, renorm_interval=ri)
    return got, want, (A, B, C0)


@pytest.mark.parametrize("ta,tb", list(itertools.product("NTC", "NTC")))
def test_gemm_bits_every_transpose_pair(gpu, ta, tb):
    rng = np.random.default_rng(ord(ta) * 131 + ord(tb))
    got, want, _ = both(ta, tb, 37, 45, 29, ALPHA, BETA, 8, rng)
    nt.assert_bits(got, want, ta + tb)  # padding rows of C included (NaN stays NaN)


@pytest.mark.parametrize("shape", [(300, 260, 270), (257, 255, 9), (64, 1024, 33)])
@pytest.mark.parametrize("ri", [8, 0, 3])
def test_gemm_bits_large_tiles_and_intervals(gpu, shape, ri):
    m, n, k = shape
    rng = np.random.default_rng(m + n + k + ri)
    for ta, tb in ("NN", "TT", "TN"):
        got, want, _ = both(ta, tb, m, n, k, ALPHA, BETA, ri, rng)
        nt.assert_bits(got, want, "%s %s ri=%d" % (ta + tb, shape, ri))


@pytest.mark.parametrize("alpha,beta", [((1.0, 0.0), (0.0, 0.0)), ((1.0, 0.0), (1.0, 0.0)),
                                        ((-2.0, 0.0), (0.0, 0.0)), ((0.5, 2.0 ** -55), (3.0, 0.0)),
                                        ((0.0, 0.0), (0.0, 0.0)), ((0.0, 0.0), (1.0, 0.0)),
                                        ((0.0, 0.0), (-1.5, 2.0 ** -56))])
def test_gemm_alpha_beta_forms(gpu, alpha, beta):
    rng = np.random.default_rng(17)
    for ta, tb in ("NN", "TN"):
        got, want, _ = both(ta, tb, 70, 40, 50, alpha, beta, 8, rng)
        nt.assert_bits(got, want, "%s alpha=%s beta=%s" % (ta + tb, alpha, beta))


def test_gemm_identity_epilogue_is_matmul(gpu):
    rng = np.random.default_rng(4)
    A, B = store(rng, 90, 70, 90), store(rng, 70, 110, 70)
    got = mf.gemm("N", "N", 1.0, A, B)
    nt.assert_bits(got, mf.matmul(A, B), "gemm NN alpha=1 beta=0 vs matmul")


def test_gemm_beta_zero_never_reads_c(gpu):
    rng = np.random.default_rng(8)
    A, B = store(rng, 20, 30, 20), store(rng, 30, 10, 30)
    C = nt.cm(np.full((20, 10, 2), np.nan))
    got = mf.gemm("N", "N", ALPHA, A, B, 0.0, C)
    want = nt.cm(np.full((20, 10, 2), np.nan))
    nt.or_gemm_blas("N", "N", 20, 10, 30, ALPHA, A, B, 0.0, want, 8)
    assert not np.isnan(got).any()
    nt.assert_bits(got, want, "beta=0 over NaN C")


def test_gemm_quick_returns_and_k_zero(gpu):
    rng = np.random.default_rng(9)
    A, B = store(rng, 6, 4, 6), store(rng, 4, 5, 4)
    C0 = store(rng, 6, 5, 6)
    assert np.array_equal(mf.gemm("N", "N", 0.0, A, B, 1.0, C0.copy(order="K")), C0)
    assert np.array_equal(mf.gemm("N", "N", 2.0, A, B, 1.0, C0.copy(order="K"), k=0), C0)
    assert np.array_equal(mf.gemm("N", "N", 2.0, A, B, 2.0, C0.copy(order="K"), k=0), 2 * C0)
    assert np.array_equal(mf.gemm("N", "N", 2.0, A, B, 5.0, C0.copy(order="K"), m=0), C0)
    z = mf.gemm("T", "N", 0.0, store(rng, 4, 6, 4), B, 0.0, nt.cm(np.full((6, 5, 2), np.nan)))
    assert (z == 0).all()


def test_gemm_nonfinite_propagates_like_oracle(gpu):
    rng = np.random.default_rng(10)
    m, n, k = 33, 17, 21
    A, B = store(rng, m, k, m), store(rng, k, n, k)
    A[3, 4, 0] = np.inf
    A[7, 2] = np.nan
    B[5, 9, 0] = -np.inf
    C0 = store(rng, m, n, m)
    want = C0.copy(order="K")
    nt.or_gemm_blas("N", "N", m, n, k, ALPHA, A, B, BETA, want, 8)
    nt.assert_bits(mf.gemm("N", "N", ALPHA, A, B, BETA, C0.copy(order="K")), want, "nonfinite")


def test_gemm_argument_errors(gpu):
    rng = np.random.default_rng(11)
    A, B = store(rng, 6, 4, 6), store(rng, 4, 5, 4)
    for kw in (dict(transa="X", m=6, n=5, k=4), dict(transb="q", m=6, n=5, k=4), dict(k=7)):
        C = store(rng, 6, 5, 6)
        args = dict(transa="N", transb="N")
        args.update(kw)
        with pytest.raises((RuntimeError, ValueError)):
            mf.gemm(args.pop("transa"), args.pop("transb"), 1.0, A, B, 0.0, C, **args)
    # lda < m reaches the library: NaN written to C and the error reported
    C = store(rng, 6, 5, 6)
    with pytest.raises(RuntimeError):
        mf.gemm("N", "N", 1.0, store(rng, 6, 4, 6)[:5], B, 0.0, C, m=6)
    assert np.isnan(C).all()
    # a good call afterwards clears the error
    got = mf.gemm("N", "N", 1.0, A, B)
    nt.assert_bits(got, mf.matmul(A, B), "after error")


def test_device_gemm_entry_point(gpu):
    import torch
    torch.cuda.init()
    rng = np.random.default_rng(12)
    m, n, k = 150, 90, 77
    for ta, tb in ("TN", "NT"):
        A = store(rng, k, m, k + 3) if ta == "T" else store(rng, m, k, m + 3)
        B = store(rng, k, n, k + 1) if tb == "N" else store(rng, n, k, n + 1)
        C0 = store(rng, m, n, m + 2)
        want = C0.copy(order="K")
        nt.or_gemm_blas(ta, tb, m, n, k, ALPHA, A, B, BETA, want, 8)

        def dev(x):
            # column-major storage (ld, cols, 2) -> contiguous device buffer in the same byte order
            return torch.from_numpy(np.ascontiguousarray(x.transpose(1, 0, 2))).cuda()

        dA, dB, dC = dev(A), dev(B), dev(C0)
        mf.dev_gemm(ta, tb, m, n, k, ALPHA, dA.data_ptr(), A.shape[0], dB.data_ptr(), B.shape[0], BETA,
                    dC.data_ptr(), C0.shape[0], 8, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        got = dC.cpu().numpy().transpose(1, 0, 2)
        nt.assert_bits(got, want, "device " + ta + tb)



def _tight(S, rows, cols, canary):
    """The ld (cols - 1) + rows elements of column-major storage S that a BLAS
    call on its leading rows x cols block owns, followed by canary elements."""
    ld = S.shape[0]
    flat = np.ascontiguousarray(S.transpose(1, 0, 2)).reshape(-1, 2)[: ld * (cols - 1) + rows]
    return np.concatenate([flat, np.full((canary, 2), 7.0)])


def test_gemm_touches_only_the_callers_span(gpu):
    """A LAPACK-style call on a sub-block C(i0:i0+m, 0:n) of a larger array
    whose buffer ends right after the block's last element: the library may
    read and write only the ldc (n - 1) + m elements that start at C(i0, 0).
    Canary elements follow the buffer, rows outside the block (someone
    else's) must keep their bits, and A and B are tight buffers too."""
    import ctypes
    rng = np.random.default_rng(21)
    m, n, k, ld, canary = 13, 7, 11, 40, 64
    i0 = ld - m
    span = ld * (n - 1) + m
    big = rng.standard_normal((i0 + span + canary, 2))
    before = big.copy()
    A, B = store(rng, m, k, m + 2), store(rng, k, n, k + 3)
    a_buf, b_buf = _tight(A, m, k, canary), _tight(B, k, n, canary)
    P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    L = mf.lib()
    L.wgemmdd(b"N", b"N", m, n, k, mf._f2(ALPHA), P(a_buf), m + 2, P(b_buf), k + 3, mf._f2(BETA),
              P(big[i0:]), ld, 8)
    assert L.mfhip_last_error() == 0
    assert np.array_equal(big[i0 + span:], before[i0 + span:]), "canary behind C"
    assert (a_buf[-canary:] == 7.0).all() and (b_buf[-canary:] == 7.0).all()
    inside = np.zeros(big.shape[0], bool)
    for j in range(n):
        inside[i0 + j * ld: i0 + j * ld + m] = True
    assert np.array_equal(big[~inside], before[~inside]), "rows outside the block"
    C0 = np.stack([before[i0 + j * ld: i0 + j * ld + m] for j in range(n)], axis=1)
    want = nt.cm(C0)
    assert nt.or_gemm_blas("N", "N", m, n, k, ALPHA, A, B, BETA, want, 8) == 0
    got = np.stack([big[i0 + j * ld: i0 + j * ld + m] for j in range(n)], axis=1)
    nt.assert_bits(got, want, "sub-block")


def test_trsm_touches_only_the_callers_span(gpu):
    """wtrsmdd on a tight B buffer (ldb (n - 1) + m elements) and a tight
    triangle: canaries stay, rows below the block keep their bits."""
    import ctypes
    rng = np.random.default_rng(22)
    m, n, ld, canary = 12, 6, 30, 64
    A = store(rng, m, m, m + 4)
    for d in range(m):
        A[d, d] = [4.0 + d, 0.0]  # a well-conditioned triangle
    span = ld * (n - 1) + m
    bbuf = rng.standard_normal((span + canary, 2))
    before = bbuf.copy()
    a_buf = _tight(A, m, m, canary)
    P = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    L = mf.lib()
    L.wtrsmdd(b"L", b"U", b"N", b"N", m, n, mf._f2(1.5), P(a_buf), m + 4, P(bbuf), ld)
    assert L.mfhip_last_error() == 0
    assert np.array_equal(bbuf[span:], before[span:]), "canary behind B"
    assert (a_buf[-canary:] == 7.0).all()
    inside = np.zeros(bbuf.shape[0], bool)
    for j in range(n):
        inside[j * ld: j * ld + m] = True
    assert np.array_equal(bbuf[~inside], before[~inside]), "rows outside the block"
    want = nt.cm(np.stack([before[j * ld: j * ld + m] for j in range(n)], axis=1))
    assert nt.or_trsm("L", "U", "N", "N", m, n, 1.5, A, want) == 0
    got = np.stack([bbuf[j * ld: j * ld + m] for j in range(n)], axis=1)
    nt.assert_bits(got, want, "trsm sub-block")
