"""GPU parity of the triangular solve (wtrsmdd / mfhip_trsm) against the
oracle's restatement of the reference's xTRSM shim
(test/integration/blas/wtrsm.f90), bit for bit: all eight loop nests with unit
and non-unit diagonals, several alphas, block edges (64-pivot blocks), padded
leading dimensions, the zero skips, the quick returns and argument errors."""
import itertools

import numpy as np
import pytest

import multifloats_amd as mf
from multifloats_amd import datagen
from tests import _native as nt
from tests.test_trsm_cpu import tri

pytestmark = pytest.mark.gpu

VARIANTS = list(itertools.product("LR", "UL", "NT", "NU"))


def system(rng, side, uplo, diag, m, n, pad_a=3, pad_b=2):
    s = m if side == "L" else n
    A = tri(rng, s, uplo == "U", diag == "U", s + pad_a, kind="real")
    B = np.full((m + pad_b, n, 2), 5.0)
    B[:m] = datagen.matrix(rng, m, n)
    return A, nt.cm(B)


def check(side, uplo, ta, diag, m, n, alpha, A, B0, what):
    want = B0.copy(order="K")
    assert nt.or_trsm(side, uplo, ta, diag, m, n, alpha, A, want) == 0
    got = mf.trsm(side, uplo, ta, diag, alpha, A, B0.copy(order="K"), m=m, n=n)
    nt.assert_bits(got, want, what)


@pytest.mark.parametrize("side,uplo,ta,diag", VARIANTS)
def test_trsm_bits_every_variant(gpu, side, uplo, ta, diag):
    rng = np.random.default_rng(hash((side, uplo, ta, diag, 1)) % 2 ** 32)
    # 150 pivots = two full 64-blocks and a ragged one; 70 vectors = ragged tiles
    for (m, n), alpha in (((150, 70), (1.5, 2.0 ** -60)), ((70, 150), (1.0, 0.0)), ((33, 9), (-1.0, 0.0))):
        A, B0 = system(rng, side, uplo, diag, m, n)
        check(side, uplo, ta, diag, m, n, alpha, A, B0, "%s%s%s%s m=%d n=%d" % (side, uplo, ta, diag, m, n))


@pytest.mark.parametrize("side,uplo,ta", list(itertools.product("LR", "UL", "NT")))
def test_trsm_bits_larger(gpu, side, uplo, ta):
    rng = np.random.default_rng(hash((side, uplo, ta, 2)) % 2 ** 32)
    m, n = (300, 200) if side == "L" else (200, 300)
    A, B0 = system(rng, side, uplo, "N", m, n, pad_a=5, pad_b=1)
    check(side, uplo, ta, "N", m, n, (0.75, -2.0 ** -57), A, B0, "%s%s%s large" % (side, uplo, ta))


def test_trsm_single_vector_and_single_pivot(gpu):
    rng = np.random.default_rng(3)
    for side, m, n in (("L", 200, 1), ("R", 1, 200), ("L", 1, 50), ("R", 50, 1), ("L", 64, 64), ("R", 65, 65)):
        for uplo, ta in itertools.product("UL", "NT"):
            A, B0 = system(rng, side, uplo, "N", m, n)
            check(side, uplo, ta, "N", m, n, 2.0, A, B0, "%s%s%s m=%d n=%d" % (side, uplo, ta, m, n))


def test_trsm_zero_skips_match_oracle(gpu):
    """Zero pivots and zero coefficients next to infinities: the skips decide
    between finite results and NaN, and the GPU must skip exactly where the
    shim does (including pivots that only become zero after earlier updates)."""
    rng = np.random.default_rng(4)
    m, n = 130, 40
    for side, uplo, ta in itertools.product("LR", "UL", "NT"):
        s = m if side == "L" else n
        A, B0 = system(rng, side, uplo, "N", m, n)
        # zero coefficients and an infinity in the referenced triangle
        for _ in range(s):
            i, j = sorted(rng.integers(0, s, size=2))
            if i != j:
                if uplo == "L":
                    i, j = j, i
                A[i, j] = 0.0
        i, j = (1, s - 2) if uplo == "U" else (s - 2, 1)
        A[i, j] = (np.inf, 0.0)
        B0[:m, 3] = 0.0      # a zero column (pivots zero for side L)
        B0[5, :] = 0.0       # a zero row (pivots zero for side R)
        B0[7, 11] = (np.inf, 0.0)
        check(side, uplo, ta, "N", m, n, 1.0, A, B0, "skips %s%s%s" % (side, uplo, ta))



def test_trsm_signed_zeros_and_zero_leading_terms(gpu):
    """Updates whose two leading terms are both zero take dd_sub's special
    case (the sign of the zero, the low parts subtracted as they are): -0.0
    entries, entries (0, tiny) and zero coefficients, mixed into ordinary
    ones, must give the shim's bits in every nest."""
    rng = np.random.default_rng(5)
    m, n = 96, 70
    for side, uplo, ta in itertools.product("LR", "UL", "NT"):
        s = m if side == "L" else n
        A, B0 = system(rng, side, uplo, "N", m, n)
        for _ in range(s // 2):
            i, j = sorted(rng.integers(0, s, size=2))
            if i != j:
                if uplo == "L":
                    i, j = j, i
                A[i, j] = (0.0, 0.0) if rng.random() < 0.5 else (-0.0, 0.0)
        idx = rng.integers(0, m, size=40), rng.integers(0, n, size=40)
        for t, (i, j) in enumerate(zip(*idx)):
            B0[i, j] = (-0.0, 0.0) if t % 3 == 0 else (0.0, 2.0 ** -1060) if t % 3 == 1 else (-0.0, -2.0 ** -1070)
        check(side, uplo, ta, "N", m, n, 1.0, A, B0, "zeros %s%s%s" % (side, uplo, ta))

def test_trsm_alpha_zero_and_quick_returns(gpu):
    rng = np.random.default_rng(5)
    A, B0 = system(rng, "L", "U", "N", 20, 10)
    B = B0.copy(order="K")
    B[0, 0] = np.nan
    got = mf.trsm("L", "U", "N", "N", 0.0, A, B, m=20, n=10)
    assert (got[:20] == 0).all() and (got[20:] == 5.0).all()
    got = mf.trsm("L", "U", "N", "N", 2.0, A, B0.copy(order="K"), m=0, n=10)
    assert np.array_equal(got, B0)


def test_trsm_argument_errors(gpu):
    rng = np.random.default_rng(6)
    A, B0 = system(rng, "L", "U", "N", 20, 10)
    for args in (("X", "U", "N", "N"), ("L", "Q", "N", "N"), ("L", "U", "Z", "N"), ("L", "U", "N", "Y")):
        B = B0.copy(order="K")
        with pytest.raises(RuntimeError):
            mf.trsm(*args, 1.0, A, B, m=20, n=10)
        assert np.isnan(B[:20]).all()
    got = mf.trsm("L", "U", "N", "N", 1.0, A, B0.copy(order="K"), m=20, n=10)
    assert not np.isnan(got).any()


def test_device_trsm_entry_point(gpu):
    import torch
    torch.cuda.init()
    rng = np.random.default_rng(7)
    m, n = 140, 90
    for side, uplo, ta in (("L", "L", "N"), ("R", "U", "T")):
        A, B0 = system(rng, side, uplo, "N", m, n)
        want = B0.copy(order="K")
        nt.or_trsm(side, uplo, ta, "N", m, n, (1.25, 0.0), A, want)

        def dev(x):
            return torch.from_numpy(np.ascontiguousarray(x.transpose(1, 0, 2))).cuda()

        dA, dB = dev(A), dev(B0)
        wb = mf.trsm_workspace_bytes(side, m, n)
        work = torch.empty(wb, dtype=torch.uint8, device="cuda")
        mf.dev_trsm(side, uplo, ta, "N", m, n, (1.25, 0.0), dA.data_ptr(), A.shape[0], dB.data_ptr(), B0.shape[0],
                    work.data_ptr(), wb, torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        nt.assert_bits(dB.cpu().numpy().transpose(1, 0, 2), want, "device " + side + uplo + ta)
