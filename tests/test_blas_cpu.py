"""The oracle's BLAS-style GEMM (C = alpha op(A) op(B) + beta C, the interface of
the reference's xGEMM shim test/integration/blas/wgemm.f90) pinned on the CPU:
against the oracle's matmul chains bit for bit where the epilogue is the
identity, against the shim's own loops (restated below with the oracle's DD
operators) within the DD rounding bound, and against the shim's argument
checks and quick returns."""
import itertools

import numpy as np
import pytest

from multifloats_amd import datagen
from tests import _native as nt

U = 2.0 ** -104


def store(rng, rows, cols, ld, pad=np.nan):
    """Column-major storage (ld, cols, 2) holding a random rows x cols DD block
    at the top; rows ld > rows are padding."""
    S = np.full((ld, cols, 2), pad)
    S[:rows] = datagen.config0(rng, rows * cols).reshape(cols, rows, 2).transpose(1, 0, 2)
    return nt.cm(S)


def op(S, t, rows, cols):
    """op(X) (rows x cols) of the storage array S."""
    return S[:rows, :cols] if t in "Nn" else S[:cols, :rows].transpose(1, 0, 2)


def col(v, n):
    return np.tile(np.asarray(v, dtype=np.float64).reshape(1, 2), (n, 1))


def wgemm_loops(ta, tb, m, n, k, alpha, A, B, beta, C):
    """wgemm.f90:52-139 restated with the oracle's DD operators (vectorised over
    i). Returns op-result as a new (ld, n, 2) array."""
    C = C.copy(order="K")
    al, be = np.asarray(alpha, dtype=np.float64), np.asarray(beta, dtype=np.float64)
    zero, one = np.zeros(2), np.array([1.0, 0.0])
    if m == 0 or n == 0 or ((np.array_equal(al, zero) or k == 0) and np.array_equal(be, one)):
        return C
    nota, notb = ta in "Nn", tb in "Nn"
    if np.array_equal(al, zero):
        for j in range(n):
            C[:m, j] = 0.0 if np.array_equal(be, zero) else nt.or_binary("mul", col(be, m), C[:m, j])
        return C
    for j in range(n):
        bj = (lambda l: B[l, j]) if notb else (lambda l: B[j, l])
        if nota:
            if np.array_equal(be, zero):
                C[:m, j] = 0.0
            elif not np.array_equal(be, one):
                C[:m, j] = nt.or_binary("mul", col(be, m), C[:m, j])
            for l in range(k):
                temp = nt.or_binary("mul", al.reshape(1, 2), bj(l).reshape(1, 2))
                C[:m, j] = nt.or_binary("add", C[:m, j], nt.or_binary("mul", col(temp, m), A[:m, l]))
        else:
            temp = np.zeros((m, 2))
            for l in range(k):
                temp = nt.or_binary("add", temp, nt.or_binary("mul", A[l, :m], col(bj(l), m)))
            t = nt.or_binary("mul", col(al, m), temp)
            C[:m, j] = t if np.array_equal(be, zero) else nt.or_binary(
                "add", t, nt.or_binary("mul", col(be, m), C[:m, j]))
    return C


def test_identity_epilogue_is_matmul_bits():
    rng = np.random.default_rng(41)
    for (m, n, k), ri in itertools.product([(13, 9, 17), (1, 5, 40), (33, 2, 1)], [8, 0, 3]):
        A, B = store(rng, m, k, m), store(rng, k, n, k)
        C = nt.cm(np.full((m, n, 2), np.nan))  # beta == 0: C is never read
        assert nt.or_gemm_blas("N", "N", m, n, k, 1.0, A, B, 0.0, C, ri) == 0
        nt.assert_bits(C, nt.or_matmul(A, B, ri), "NN alpha=1 beta=0 (%d,%d,%d) ri=%d" % (m, n, k, ri))


@pytest.mark.parametrize("ta,tb", list(itertools.product("NTc", "NtC")))
def test_transposes_and_leading_dims(ta, tb):
    """op(A), op(B) read through their transposes and padded leading dims give
    the same bits as matmul on the logical operands; C's padding rows stay."""
    rng = np.random.default_rng(ord(ta) * 7 + ord(tb))
    m, n, k = 11, 7, 19
    A = store(rng, m, k, m + 3) if ta in "Nn" else store(rng, k, m, k + 5)
    B = store(rng, k, n, k + 2) if tb in "Nn" else store(rng, n, k, n + 4)
    want = nt.or_matmul(op(A, ta, m, k), op(B, tb, k, n), 8)
    C = nt.cm(np.full((m + 6, n, 2), np.nan))
    assert nt.or_gemm_blas(ta, tb, m, n, k, 1.0, A, B, 0.0, C, 8) == 0
    nt.assert_bits(C[:m], want, ta + tb)
    assert np.isnan(C[m:]).all()
    # a leading-dimension slip (lda one short, still legal) breaks the bits
    C2 = nt.cm(np.zeros((m + 6, n, 2)))
    nt.or_gemm_blas(ta, tb, m, n, k, 1.0, A, B, 0.0, C2, 8, lda=A.shape[0] - 1)
    assert not nt.bits_equal(C2[:m], want).all()


@pytest.mark.parametrize("ta,tb", ["NN", "TN", "NT", "TT"])
@pytest.mark.parametrize("alpha,beta", [((1.5, 2.0 ** -60), (-0.75, 2.0 ** -58)), ((1.0, 0.0), (1.0, 0.0)),
                                        ((-2.25, 0.0), (0.0, 0.0)), ((0.5, 0.0), (3.0, 2.0 ** -55))])
def test_epilogue_within_bound_of_wgemm_loops(ta, tb, alpha, beta):
    rng = np.random.default_rng(5)
    m, n, k = 9, 6, 14
    A = store(rng, m, k, m + 1) if ta == "N" else store(rng, k, m, k + 1)
    B = store(rng, k, n, k) if tb == "N" else store(rng, n, k, n + 2)
    C0 = store(rng, m, n, m + 2)
    got = C0.copy(order="K")
    assert nt.or_gemm_blas(ta, tb, m, n, k, alpha, A, B, beta, got, 8) == 0
    loop = wgemm_loops(ta, tb, m, n, k, alpha, A, B, beta, C0)
    opa, opb = op(A, ta, m, k)[..., 0], op(B, tb, k, n)[..., 0]
    mag = abs(alpha[0]) * (np.abs(opa) @ np.abs(opb)) + abs(beta[0]) * np.abs(C0[:m, :, 0])
    diff = np.abs((got[:m, :, 0] - loop[:m, :, 0]) + (got[:m, :, 1] - loop[:m, :, 1]))
    assert (diff <= 8 * (k + 2) * U * mag).all(), (diff / mag).max()
    assert np.array_equal(got[m:], C0[m:], equal_nan=True)
    if beta != (0.0, 0.0):
        # a dropped beta C term is far outside the bound
        slip = C0.copy(order="K")
        nt.or_gemm_blas(ta, tb, m, n, k, alpha, A, B, 0.0, slip, 8)
        d2 = np.abs(slip[:m, :, 0] - loop[:m, :, 0])
        assert (d2 > 8 * (k + 2) * U * mag).any()


def test_info_codes_in_reference_order():
    rng = np.random.default_rng(2)
    A, B = store(rng, 4, 5, 4), store(rng, 5, 3, 5)
    C0 = store(rng, 4, 3, 4)
    cases = [  # (ta, tb, m, n, k, lda, ldb, ldc) -> wgemm.f90 info
        (("X", "N", 4, 3, 5, 4, 5, 4), 1), (("N", "Q", 4, 3, 5, 4, 5, 4), 2),
        (("N", "N", -1, 3, 5, 4, 5, 4), 3), (("N", "N", 4, -1, 5, 4, 5, 4), 4),
        (("N", "N", 4, 3, -1, 4, 5, 4), 5), (("N", "N", 4, 3, 5, 3, 5, 4), 8),
        (("T", "N", 4, 3, 5, 4, 5, 4), 8),   # op(A) = A^T needs lda >= k = 5
        (("N", "N", 4, 3, 5, 4, 4, 4), 10), (("N", "T", 4, 3, 5, 4, 2, 4), 10),
        (("N", "N", 4, 3, 5, 4, 5, 3), 13), (("X", "N", -1, 3, 5, 0, 5, 4), 1),
        (("N", "N", 0, 3, 5, 0, 5, 0), 8),   # max(1, nrowa) even when m == 0
    ]
    for (ta, tb, m, n, k, lda, ldb, ldc), info in cases:
        C = C0.copy(order="K")
        got = nt.or_gemm_blas(ta, tb, m, n, k, 1.0, A, B, 0.0, C, 8, lda=lda, ldb=ldb, ldc=ldc)
        assert got == info, (ta, tb, m, n, k, lda, ldb, ldc)
        assert np.array_equal(C, C0, equal_nan=True)


def test_quick_returns():
    rng = np.random.default_rng(3)
    A, B = store(rng, 4, 5, 4), store(rng, 5, 3, 5)
    C0 = store(rng, 4, 3, 4)
    C = C0.copy(order="K")
    nt.or_gemm_blas("N", "N", 4, 3, 5, 0.0, A, B, 1.0, C, 8)          # alpha = 0, beta = 1
    assert np.array_equal(C, C0)
    nt.or_gemm_blas("N", "N", 4, 3, 0, 2.0, A, B, 1.0, C, 8)          # k = 0, beta = 1
    assert np.array_equal(C, C0)
    nt.or_gemm_blas("N", "N", 0, 3, 5, 2.0, A, B, 7.0, C, 8)          # m = 0
    assert np.array_equal(C, C0)
    N = nt.cm(np.full((4, 3, 2), np.nan))
    nt.or_gemm_blas("N", "N", 4, 3, 5, 0.0, A, B, 0.0, N, 8)          # alpha = beta = 0: zeros
    assert (N == 0).all()
    nt.or_gemm_blas("N", "N", 4, 3, 0, 2.0, A, B, 2.0, C, 8)          # k = 0: C = beta C exactly
    assert np.array_equal(C, 2 * C0)
    C = C0.copy(order="K")
    At = store(rng, 5, 4, 5)
    assert nt.or_gemm_blas("T", "N", 4, 3, 5, 0.0, At, B, (-0.5, 0.0), C, 8) == 0  # alpha = 0: beta C
    assert np.array_equal(C, -0.5 * C0)


def test_library_rejects_bad_arguments_before_any_compute():
    """wgemmdd checks its arguments (wgemm.f90:16-51) before touching the GPU:
    a bad call writes NaN to C and raises, here without a GPU."""
    import multifloats_amd as mf
    rng = np.random.default_rng(12)
    A, B = store(rng, 6, 4, 6), store(rng, 4, 5, 4)
    for ta, tb, lda_rows in (("X", "N", 6), ("N", "Z", 6), ("N", "N", 5)):
        C = store(rng, 6, 5, 6)
        with pytest.raises(RuntimeError):
            mf.gemm(ta, tb, 1.0, A[:lda_rows].copy(order="K") if lda_rows < 6 else A, B, 0.0, C, m=6, n=5, k=4)
        assert np.isnan(C).all()
