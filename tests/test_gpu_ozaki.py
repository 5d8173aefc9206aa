"""GPU tests of the Ozaki INT8 split-precision DD GEMM: elementwise error
against the exact accumulator, |C - AB|_ij <= 2^-100 (|A||B|)_ij."""
import numpy as np
import pytest

import multifloats_amd as mf
from multifloats_amd import datagen
from tests import _native as nt

pytestmark = pytest.mark.gpu

BOUND = 2.0 ** -100


def mats(rng, m, k, n, spread):
    e = spread if spread else None
    return datagen.matrix(rng, m, k, e), datagen.matrix(rng, k, n, e)


def all_err(A, B, C):
    m, n = A.shape[0], B.shape[1]
    rows, cols = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    return nt.gemm_err(A, B, C, rows.ravel(), cols.ravel())


@pytest.mark.parametrize("shape", [(1, 1, 1), (64, 32, 64), (70, 45, 130), (200, 300, 100), (5, 1000, 3),
                                   (64, 128, 64), (100, 2048, 90)])
@pytest.mark.parametrize("spread", [0, 10, 20, 30])
def test_ozaki_error_bound(gpu, shape, spread):
    """Automatic slice count: the componentwise 2^-100 target on every output."""
    m, k, n = shape
    rng = np.random.default_rng(m + 7 * k + 13 * n + spread)
    A, B = mats(rng, m, k, n, spread)
    C = mf.matmul_ozaki(A, B, 0)
    used = mf.ozaki_slices_used()
    assert 7 <= used <= 16
    err, mag = all_err(A, B, C)
    ratio = err / mag
    assert ratio.max() <= BOUND, (used, mf.ozaki_fixups(), ratio.max(), np.log2(ratio.max()))


def test_ozaki_error_falls_with_slices(gpu):
    rng = np.random.default_rng(3)
    A, B = mats(rng, 96, 200, 80, 20)
    worst = []
    for S in (7, 8, 10, 12, 14, 16):
        err, mag = all_err(A, B, mf.matmul_ozaki(A, B, S))
        worst.append(np.log2((err / mag).max()))
        # the dropped levels L >= S bound the error (8 bits per slice): a
        # loose a-priori check
        assert worst[-1] <= -8 * S + 24, (S, worst[-1])
    assert all(a > b for a, b in zip(worst, worst[1:])), worst


def test_ozaki_special_rows_and_columns(gpu):
    rng = np.random.default_rng(4)
    A, B = mats(rng, 70, 40, 66, 10)
    A = A.copy()
    B = B.copy()
    A[3, :, :] = 0.0            # zero row -> zero outputs
    A[5, 7, 0] = np.inf         # row with Inf -> NaN outputs
    B[:, 9, :] = 0.0            # zero column
    B[11, 20, 0] = np.nan       # column with NaN
    C = mf.matmul_ozaki(datagen.as_colmajor(A), datagen.as_colmajor(B), 16)
    assert mf.ozaki_slices_used() == 16
    assert not np.delete(C[3], 20, axis=0).any() and np.isnan(C[3, 20, 0])
    assert np.isnan(C[5, :, 0]).all()
    assert not C[np.arange(70) != 5, 9].any()
    assert np.isnan(C[:, 20, 0]).all()
    ok = np.ones((70, 66), bool)
    ok[[3, 5], :] = False
    ok[:, [9, 20]] = False
    rows, cols = np.nonzero(ok)
    err, mag = nt.gemm_err(A, B, C, rows, cols)
    # 16 slices forced at k = 40 and spread 2^+-10, where the slice rule
    # would run the reference-order GEMM: only a loose accuracy check here.
    assert (err <= 2.0 ** -90 * mag).all()


def test_slice_rule():
    """The exponent-range rule with which an automatic call starts (it only
    sets the first residue pass; the certificate decides coverage)."""
    assert mf.ozaki_choose_slices(4096, 40, 40) == 14
    assert mf.ozaki_choose_slices(8192, 60, 60) == 14
    assert mf.ozaki_choose_slices(32, 20, 20) == 16
    assert mf.ozaki_choose_slices(32, 60, 60) > 16  # capped at 16 by the call
    assert mf.ozaki_choose_slices(1, 0, 0) == 13


def test_ozaki_long_k_chunks(gpu):
    """k beyond the exact-INT32 span of one launch (k * S * 2^14 < 2^31): the
    K range is cut into chunks whose DD results are added."""
    rng = np.random.default_rng(8)
    A, B = mats(rng, 64, 10000, 48, 10)
    for S in (0, 16):
        C = mf.matmul_ozaki(A, B, S)
        err, mag = all_err(A, B, C)
        assert (err <= BOUND * mag).all(), (S, mf.ozaki_slices_used(), np.log2((err / mag).max()))


def test_ozaki_top_digit_headroom(gpu):
    """Rows and columns whose largest entry sits just below the next power of
    two get one more bit of headroom, so the top digit never leaves int8."""
    rng = np.random.default_rng(9)
    A, B = mats(rng, 80, 300, 70, 10)
    A = A.copy()
    B = B.copy()
    for i in range(80):
        p = rng.integers(300)
        e = np.floor(np.log2(np.abs(A[i, :, 0]).max()))
        A[i, p] = [(2.0 - 2.0 ** -40) * 2.0 ** e * (1 if i % 2 else -1), 0.0]
    for j in range(0, 70, 3):
        B[rng.integers(300), j] = [127.0 / 64.0, 2.0 ** -60]
    C = mf.matmul_ozaki(datagen.as_colmajor(A), datagen.as_colmajor(B), 0)
    err, mag = all_err(A, B, C)
    assert (err <= BOUND * mag).all(), (mf.ozaki_slices_used(), np.log2((err / mag).max()))


def test_ozaki_empty_inner_dimension(gpu):
    C = mf.matmul_ozaki(np.zeros((4, 0, 2)), np.zeros((0, 3, 2)), 0)
    assert C.shape == (4, 3, 2) and not C.any()


def test_ozaki_4096_sampled(gpu):
    """4096^3 with exponent spread 2^+-20: sampled outputs within the bound."""
    import torch
    m = k = n = 4096
    rng = np.random.default_rng(42)
    A, B = mats(rng, m, k, n, 20)
    dA = torch.from_numpy(datagen.flat(A)).cuda()
    dB = torch.from_numpy(datagen.flat(B)).cuda()
    dC = torch.empty((m * n, 2), dtype=torch.float64, device="cuda")
    wb = mf.ozaki_workspace_bytes(m, k, n, 0)
    work = torch.empty(wb, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    mf.dev_matmul_ozaki(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, 0, work.data_ptr(), wb, s)
    torch.cuda.synchronize()
    assert mf.ozaki_slices_used() == -1  # chosen on the device
    used, fix = mf.ozaki_info(work.data_ptr(), m, k, n, 0, s)
    assert used in (14, 15) and fix <= 16384, (used, fix)
    # the certificate's row and column sums bound |C_int| below the worst
    # case k 2^(16S - 2): at least one modulus fewer on these inputs
    n_used = mf.ozaki_moduli_used(work.data_ptr(), m, k, n, 0, s)
    assert 0 < n_used < mf.ozaki_moduli(m, k, n, used), (n_used, used)
    C = datagen.unflat(dC.cpu().numpy(), m, n)
    rows = rng.choice(m, 24, replace=False)
    cols = rng.choice(n, 24, replace=False)
    rr, cc = np.meshgrid(rows, cols, indexing="ij")
    err, mag = nt.gemm_err(A, B, C, rr.ravel(), cc.ravel())
    assert (err <= BOUND * mag).all(), np.log2((err / mag).max())


# Shapes with m, n >= 512 run the CRT form (residue GEMMs + Chinese remainder
# reconstruction); the outputs checked are a row/column sample that always
# holds the ragged last rows and columns of the 256 x 256 tiles.
def _sample(rng, m, n, count=40):
    rows = np.unique(np.concatenate([rng.choice(m, count, replace=False), np.arange(m - 3, m)]))
    cols = np.unique(np.concatenate([rng.choice(n, count, replace=False), np.arange(n - 3, n)]))
    rr, cc = np.meshgrid(rows, cols, indexing="ij")
    return rr.ravel(), cc.ravel()


@pytest.mark.parametrize("shape", [(512, 32, 512), (520, 45, 700), (777, 300, 515), (512, 2048, 640)])
@pytest.mark.parametrize("spread", [0, 10, 20, 30])
def test_ozaki_crt_form_error_bound(gpu, shape, spread):
    m, k, n = shape
    rng = np.random.default_rng(m + 7 * k + 13 * n + spread)
    A, B = mats(rng, m, k, n, spread)
    C = mf.matmul_ozaki(A, B, 0)
    used = mf.ozaki_slices_used()
    assert 7 <= used <= 16
    rr, cc = _sample(rng, m, n)
    err, mag = nt.gemm_err(A, B, C, rr, cc, threads=4)
    assert (err <= BOUND * mag).all(), (used, np.log2((err / mag).max()))


def test_ozaki_crt_form_error_falls_with_slices(gpu):
    rng = np.random.default_rng(13)
    A, B = mats(rng, 512, 200, 520, 20)
    rr, cc = _sample(rng, 512, 520, 24)
    worst = []
    for S in (7, 8, 10, 12, 14, 16):
        err, mag = nt.gemm_err(A, B, mf.matmul_ozaki(A, B, S), rr, cc, threads=4)
        worst.append(np.log2((err / mag).max()))
        assert worst[-1] <= -8 * S + 24, (S, worst[-1])
    assert all(a > b for a, b in zip(worst, worst[1:])), worst


def test_ozaki_crt_form_special_rows_and_columns(gpu):
    rng = np.random.default_rng(14)
    m, k, n = 530, 40, 520
    A, B = mats(rng, m, k, n, 10)
    A = A.copy()
    B = B.copy()
    A[3, :, :] = 0.0
    A[300, 7, 0] = np.inf
    B[:, 9, :] = 0.0
    B[11, 400, 0] = np.nan
    C = mf.matmul_ozaki(datagen.as_colmajor(A), datagen.as_colmajor(B), 16)
    assert not np.delete(C[3], 400, axis=0).any() and np.isnan(C[3, 400, 0])
    assert np.isnan(C[300, :, 0]).all()
    assert not C[np.arange(m) != 300, 9].any()
    assert np.isnan(C[:, 400, 0]).all()
    rr, cc = _sample(rng, m, n)
    ok = ~np.isin(rr, [3, 300]) & ~np.isin(cc, [9, 400])
    err, mag = nt.gemm_err(A, B, C, rr[ok], cc[ok], threads=4)
    assert (err <= 2.0 ** -90 * mag).all()


def test_ozaki_crt_form_long_k(gpu):
    rng = np.random.default_rng(15)
    A, B = mats(rng, 512, 10000, 512, 10)
    rr, cc = _sample(rng, 512, 512, 10)
    for S in (0, 16):
        C = mf.matmul_ozaki(A, B, S)
        err, mag = nt.gemm_err(A, B, C, rr, cc, threads=4)
        assert (err <= BOUND * mag).all(), (S, mf.ozaki_slices_used(), np.log2((err / mag).max()))


def test_ozaki_crt_form_large_residue_sums(gpu):
    """Constant operands at k = 32768: every residue product is the same, so
    the INT32 sums reach k (m_t / 2)^2 ~ 2^29, beyond float's exact integers;
    the residue reduction must still be exact. C = k a b exactly."""
    from fractions import Fraction
    m, k, n = 512, 32768, 512
    a = np.array([1.0 / 3.0, (1.0 / 3.0) * 2.0 ** -54])
    b = np.array([-5.0 / 7.0, (5.0 / 7.0) * 2.0 ** -55])
    A = np.broadcast_to(a, (m, k, 2))
    B = np.broadcast_to(b, (k, n, 2))
    exact = k * (Fraction(a[0]) + Fraction(a[1])) * (Fraction(b[0]) + Fraction(b[1]))
    for S in (0, 16):
        C = mf.matmul_ozaki(datagen.as_colmajor(A), datagen.as_colmajor(B), S)
        vals = {(float(h), float(l)) for h, l in C.reshape(-1, 2)}
        assert len(vals) == 1, (S, len(vals))
        h, l = vals.pop()
        assert abs(Fraction(h) + Fraction(l) - exact) <= abs(exact) * Fraction(2) ** -100, S


def test_ozaki_products_stage(gpu):
    """mfhip_ozaki_products runs the CRT form's residue GEMMs alone: a
    multiply after it gives the full call's bits; the triangular form has no
    separate products stage and reports an error."""
    import torch
    rng = np.random.default_rng(16)
    s = torch.cuda.current_stream().cuda_stream
    for (m, k, n), crt in (((520, 70, 515), True), ((96, 70, 80), False)):
        A, B = mats(rng, m, k, n, 10)
        S = 12
        assert (mf.ozaki_moduli(m, k, n, S) > 0) == crt
        dA = torch.from_numpy(datagen.flat(A)).cuda()
        dB = torch.from_numpy(datagen.flat(B)).cuda()
        dC = torch.empty((m * n, 2), dtype=torch.float64, device="cuda")
        wb = mf.ozaki_workspace_bytes(m, k, n, S)
        work = torch.empty(wb, dtype=torch.uint8, device="cuda")
        mf.dev_matmul_ozaki(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, S, work.data_ptr(), wb, s)
        full = dC.cpu().numpy().copy()
        dC.zero_()
        mf.dev_ozaki_prepare(dA.data_ptr(), dB.data_ptr(), m, k, n, S, work.data_ptr(), wb, s)
        if crt:
            mf.dev_ozaki_products(m, k, n, S, work.data_ptr(), wb, s)
        else:
            with pytest.raises(RuntimeError):
                mf.dev_ozaki_products(m, k, n, S, work.data_ptr(), wb, s)
        mf.dev_ozaki_multiply(dC.data_ptr(), m, k, n, S, work.data_ptr(), wb, s)
        torch.cuda.synchronize()
        assert np.array_equal(dC.cpu().numpy().view(np.uint64), full.view(np.uint64))


def test_ozaki_products_stage_of_an_automatic_call(gpu):
    """mfhip_ozaki_products with slices <= 0 reruns the residue GEMMs of the
    last automatic call on the workspace (its slice and moduli counts): the
    digit planes it writes are the call's own, so a second automatic call
    still gives the first one's bits; an undersized workspace is refused."""
    import torch
    rng = np.random.default_rng(17)
    s = torch.cuda.current_stream().cuda_stream
    for m, k, n in ((520, 70, 515), (256, 40, 300)):
        A, B = mats(rng, m, k, n, 20)
        dA = torch.from_numpy(datagen.flat(A)).cuda()
        dB = torch.from_numpy(datagen.flat(B)).cuda()
        dC = torch.empty((m * n, 2), dtype=torch.float64, device="cuda")
        wb = mf.ozaki_workspace_bytes(m, k, n, 0)
        work = torch.empty(wb, dtype=torch.uint8, device="cuda")
        mf.dev_matmul_ozaki(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, 0, work.data_ptr(), wb, s)
        first = dC.cpu().numpy().copy()
        n_used = mf.ozaki_moduli_used(work.data_ptr(), m, k, n, 0, s)
        assert n_used > 0
        mf.dev_ozaki_products(m, k, n, 0, work.data_ptr(), wb, s)
        assert mf.ozaki_moduli_used(work.data_ptr(), m, k, n, 0, s) == n_used
        with pytest.raises(RuntimeError):
            mf.dev_ozaki_products(m, k, n, 0, work.data_ptr(), wb - 256, s)
        dC.zero_()
        mf.dev_matmul_ozaki(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, 0, work.data_ptr(), wb, s)
        torch.cuda.synchronize()
        assert np.array_equal(dC.cpu().numpy().view(np.uint64), first.view(np.uint64))


@pytest.mark.parametrize("shape", [(512, 1, 512), (515, 3, 600), (600, 33, 513), (1024, 64, 768)])
def test_ozaki_crt_form_tiny_k_and_uneven_tile_grids(gpu, shape):
    """k far below one 2-block stage (zero-padded K blocks), and tile grids
    that do not split evenly over the XCDs (513 columns: 3 tile columns, so
    the column-major tile order runs), all outputs checked."""
    m, k, n = shape
    rng = np.random.default_rng(m + 3 * k + 5 * n)
    A, B = mats(rng, m, k, n, 10)
    for S in (0, 9, 16):
        C = mf.matmul_ozaki(A, B, S)
        used = mf.ozaki_slices_used()
        if used == 0:
            continue  # the slice rule chose the reference-order GEMM (tested elsewhere)
        assert mf.ozaki_moduli(m, k, n, used) > 0
        rr, cc = _sample(rng, m, n, 32)
        err, mag = nt.gemm_err(A, B, C, rr, cc, threads=4)
        bound = BOUND if S == 0 else 2.0 ** (-8 * used + 24)
        assert (err <= bound * mag).all(), (S, used, np.log2((err / mag).max()))


# ---- the automatic call's guarantee on structured inputs ----
# Per-row / per-column power-of-two scaling fits iid entries; these inputs
# place each row's largest entries where the column is small (or grade the
# magnitudes along k), so (|A||B|)_ij falls far below 2^(E_i + F_j) and a
# slice count from the exponent ranges alone misses the bound (2^-79.5 at
# k = 4096 for the crossed case). The certificate must catch every such
# output: raise S or recompute it with the compensated dot.

def _dd(rng, hi):
    return datagen._with_lo(rng, hi)


def structured(rng, m, k, n, kind, frac=1.0):
    """(A, B) of the named structure; frac of the rows / columns carry it,
    the rest are iid with spread 2^+-10."""
    sgn = lambda shape: np.where(rng.random(shape) < 0.5, -1.0, 1.0) * rng.uniform(0.5, 1.0, shape)
    A = np.exp2(rng.uniform(-10, 10, (m, k))) * sgn((m, k))
    B = np.exp2(rng.uniform(-10, 10, (k, n))) * sgn((k, n))
    rows = np.nonzero(rng.random(m) < frac)[0]
    cols = np.nonzero(rng.random(n) < frac)[0]
    if kind == "crossed":
        # row: one 2^20 entry at p = 0, 2^-20 elsewhere; column: 2^-20 at
        # p = 0, 2^20 elsewhere
        A[rows] = 2.0 ** -20 * sgn((rows.size, k))
        A[rows, 0] = 2.0 ** 20 * sgn(rows.size)
        B[:, cols] = 2.0 ** 20 * sgn((k, cols.size))
        B[0, cols] = 2.0 ** -20 * sgn(cols.size)
    elif kind == "graded":
        # a_ip = 2^(-40 p / k) (from 1 down to 2^-40 along the row), columns
        # graded the other way
        g = np.exp2(-40.0 * np.arange(k) / k)
        A[rows] = g[None, :] * sgn((rows.size, k))
        B[:, cols] = g[::-1, None] * sgn((k, cols.size))
    elif kind == "apart":
        # row and column maxima at different k: row i peaks at p = i mod k,
        # column j at p = (j + k/2) mod k, both 2^30 over a 2^-10 floor
        A[rows] = 2.0 ** -10 * sgn((rows.size, k))
        A[rows, rows % k] = 2.0 ** 30 * sgn(rows.size)
        B[:, cols] = 2.0 ** -10 * sgn((k, cols.size))
        B[(cols + k // 2) % k, cols] = 2.0 ** 30 * sgn(cols.size)
    elif kind == "outlier":
        # one 2^40 entry per row / column at a random p over iid 2^+-10
        A[rows, rng.integers(0, k, rows.size)] = 2.0 ** 40 * sgn(rows.size)
        B[rng.integers(0, k, cols.size), cols] = 2.0 ** 40 * sgn(cols.size)
    return datagen.as_colmajor(_dd(rng, A)), datagen.as_colmajor(_dd(rng, B)), rows, cols


def _structured_sample(rng, m, n, rows, cols, count=24):
    r = np.unique(np.concatenate([rng.choice(m, count, replace=False), rows[:count], np.arange(m - 2, m)]))
    c = np.unique(np.concatenate([rng.choice(n, count, replace=False), cols[:count], np.arange(n - 2, n)]))
    rr, cc = np.meshgrid(r, c, indexing="ij")
    return rr.ravel(), cc.ravel()


@pytest.mark.parametrize("kind", ["crossed", "graded", "apart", "outlier"])
@pytest.mark.parametrize("frac", [1.0, 0.05])
def test_ozaki_structured_inputs_meet_the_bound(gpu, kind, frac):
    m, k, n = 520, 2048, 530
    rng = np.random.default_rng(sum(map(ord, kind)) + int(100 * frac))
    A, B, rows, cols = structured(rng, m, k, n, kind, frac)
    print("structured kind=%s frac=%g shape=%dx%dx%d" % (kind, frac, m, k, n), flush=True)
    C = mf.matmul_ozaki(A, B, 0)
    used, fix = mf.ozaki_slices_used(), mf.ozaki_fixups()
    rr, cc = _structured_sample(rng, m, n, rows, cols)
    err, mag = nt.gemm_err(A, B, C, rr, cc, threads=4)
    worst = np.log2((err / mag).max())
    assert (err <= BOUND * mag).all(), (kind, frac, used, fix, worst)
    if kind == "crossed":
        # no slice count up to 16 covers a crossed output: all of them are
        # recomputed (frac = 1: every output)
        assert fix >= rows.size * cols.size, (fix, rows.size, cols.size)


def test_ozaki_structured_small_and_ragged(gpu):
    """The automatic call runs the CRT form at every size: tiny and ragged
    shapes with crossed rows, every output checked."""
    rng = np.random.default_rng(77)
    for (m, k, n) in ((1, 1, 1), (3, 700, 5), (64, 300, 40), (130, 1000, 257)):
        A, B, rows, cols = structured(rng, m, k, n, "crossed", 0.3)
        C = mf.matmul_ozaki(A, B, 0)
        err, mag = all_err(A, B, C)
        assert (err <= BOUND * mag).all(), ((m, k, n), mf.ozaki_slices_used(), mf.ozaki_fixups())


def _dev_ozaki(A, B, m, k, n, torch):
    dA = torch.from_numpy(datagen.flat(A)).cuda()
    dB = torch.from_numpy(datagen.flat(B)).cuda()
    dC = torch.empty((m * n, 2), dtype=torch.float64, device="cuda")
    wb = mf.ozaki_workspace_bytes(m, k, n, 0)
    work = torch.empty(wb, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    mf.dev_matmul_ozaki(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, 0, work.data_ptr(), wb, s)
    used, fix = mf.ozaki_info(work.data_ptr(), m, k, n, 0, s)
    C = datagen.unflat(dC.cpu().numpy(), m, n)
    del dA, dB, dC, work
    torch.cuda.empty_cache()
    return C, used, fix


@pytest.mark.parametrize("size", [4096, 8192])
def test_ozaki_structured_at_scale(gpu, size):
    """4096^3 and 8192^3 with 2% crossed rows and columns over iid 2^+-10
    entries: the crossed outputs are recomputed, the others certified; a
    sample holding crossed and plain outputs is within the bound."""
    import torch
    m = k = n = size
    rng = np.random.default_rng(size)
    A, B, rows, cols = structured(rng, m, k, n, "crossed", 0.02)
    print("structured at scale %d^3: %d crossed rows, %d crossed columns" % (size, rows.size, cols.size), flush=True)
    C, used, fix = _dev_ozaki(A, B, m, k, n, torch)
    assert fix >= rows.size * cols.size, (fix, rows.size * cols.size)
    rr, cc = _structured_sample(rng, m, n, rows, cols, 16)
    err, mag = nt.gemm_err(A, B, C, rr, cc, threads=8)
    assert (err <= BOUND * mag).all(), (used, fix, np.log2((err / mag).max()))


def test_ozaki_automatic_call_captures_in_a_graph(gpu):
    """slices <= 0 never synchronises: the whole call (scan, plans,
    certificate, residues, products, reconstruction, fix-up) is captured in
    a HIP graph; replays with eager work and synchronisations between them
    give the eager call's bits, also when the certificate raises S and when
    outputs are fixed up."""
    import torch
    rng = np.random.default_rng(91)
    m, k, n = 600, 700, 520
    A, B, rows, cols = structured(rng, m, k, n, "crossed", 0.05)
    dA = torch.from_numpy(datagen.flat(A)).cuda()
    dB = torch.from_numpy(datagen.flat(B)).cuda()
    dC = torch.zeros((m * n, 2), dtype=torch.float64, device="cuda")
    wb = mf.ozaki_workspace_bytes(m, k, n, 0)
    work = torch.empty(wb, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    mf.dev_matmul_ozaki(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, 0, work.data_ptr(), wb, s)
    eager = dC.cpu().numpy().copy()
    used, fix = mf.ozaki_info(work.data_ptr(), m, k, n, 0, s)
    assert fix > 0
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            mf.dev_matmul_ozaki(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, 0, work.data_ptr(), wb,
                                side.cuda_stream)
    torch.cuda.current_stream().wait_stream(side)
    for rep in range(3):
        dC.fill_(float("nan"))
        junk = torch.randn(1 << 20, device="cuda", dtype=torch.float64).sum()  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(dC.cpu().numpy().view(np.uint64), eager.view(np.uint64)), rep
        assert mf.ozaki_info(work.data_ptr(), m, k, n, 0, s) == (used, fix)
        del junk
