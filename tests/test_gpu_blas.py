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
