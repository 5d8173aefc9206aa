"""GPU parity of product / maxval / minval (fsrc/multifloats.fypp:1783-1862,
1975-2090): maxval and minval bit for bit against the oracle's restatement of
the reference's scan in every form (rank 1 across many workgroups, with
masks, ties, signed zeros and NaN limbs; dim forms); product with dim bit for
bit, rank-1 product within its relative bound of the exact product."""
import mpmath
import numpy as np
import pytest

import multifloats_amd as mf
from multifloats_amd import datagen
from tests import _native as nt

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 5, 4095, 4096, 4097, (1 << 20) + 3, 5_000_000]


def _data(rng, n, style):
    if style == "random":
        return datagen.spread(rng, n, 30)
    if style == "ties":
        pool = np.array([[1.0, 0.0], [1.0, 1e-17], [1.0, -1e-17], [-0.0, 0.0], [0.0, -0.0], [0.0, 0.0],
                         [-3.0, 0.0], [2.0, 1e-16], [2.0, 1e-16]])
        return pool[rng.integers(0, len(pool), size=n)]
    # NaN limbs: NaN hi elements, and finite-hi / NaN-lo elements, sprinkled
    a = datagen.spread(rng, n, 3)
    if n:
        k = rng.integers(0, n, size=max(1, n // 500))
        a[k[: len(k) // 2], 0] = np.nan
        a[k[len(k) // 2:], 1] = np.nan
    return a


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("style", ["random", "ties", "nan"])
def test_extrema_rank1_bits(gpu, n, style):
    rng = np.random.default_rng(n * 7 + len(style))
    a = _data(rng, n, style)
    mask = rng.random(n) < 0.7
    for kind in ("maxval", "minval"):
        print("extrema", kind, style, n, flush=True)
        nt.assert_bits(getattr(mf, kind)(a), nt.or_fibre(kind, a), "%s %s n=%d" % (kind, style, n))
        nt.assert_bits(getattr(mf, kind)(a, mask=mask), nt.or_fibre(kind, a, mask),
                       "%s %s masked n=%d" % (kind, style, n))


def test_extrema_first_nan_and_special_positions(gpu):
    rng = np.random.default_rng(3)
    n = 300_000
    base = datagen.config0(rng, n)
    cases = []
    a = base.copy()
    a[0] = (np.nan, 2.0)          # first element NaN: it is the answer
    cases.append(a)
    a = base.copy()
    a[0] = (0.25, np.nan)         # first has NaN lo; larger hi later still wins
    cases.append(a)
    a = base.copy()
    k = int(np.argmax(a[:, 0]))
    a[k + 1:] = a[k]
    a[k, 1] = np.nan              # first element of the top group has NaN lo: it wins
    cases.append(a)
    a = np.zeros((n, 2))
    a[::2, 0] = -0.0              # only signed zeros: the first one
    cases.append(a)
    for i, a in enumerate(cases):
        for kind in ("maxval", "minval"):
            nt.assert_bits(getattr(mf, kind)(a), nt.or_fibre(kind, a), "case %d %s" % (i, kind))
    # fully masked / empty: 0
    for kind in ("maxval", "minval"):
        assert (getattr(mf, kind)(base, mask=np.zeros(n, bool)) == 0).all()


@pytest.mark.parametrize("kind", ["product", "maxval", "minval"])
@pytest.mark.parametrize("shape,dim", [((37, 23), 1), ((37, 23), 2), ((5, 300, 7), 2), ((300, 4, 9), 1),
                                       ((4, 5, 600), 3)])
def test_reduce_dim_bits(gpu, kind, shape, dim):
    rng = np.random.default_rng(sum(shape) + dim)
    a = datagen.config0(rng, shape) + (np.array([1.0, 0.0]) if kind == "product" else 0.0)
    if kind != "product":
        a.reshape(-1, 2)[rng.integers(0, a.size // 2, size=5)] = (np.nan, 0.0)
    mask = rng.random(shape) < 0.6
    r = len(shape)
    inner = int(np.prod(shape[:dim - 1]))
    ln = shape[dim - 1]
    outer = int(np.prod(shape[dim:]))
    buf = np.ascontiguousarray(a.transpose(list(range(r - 1, -1, -1)) + [r])).reshape(-1, 2)
    mb = np.ascontiguousarray(mask.transpose(list(range(r - 1, -1, -1)))).reshape(-1).astype(np.int32)
    for m, mbuf in ((None, None), (mask, mb)):
        got = getattr(mf, kind)(a, dim=dim, mask=m)
        gbuf = np.ascontiguousarray(got.transpose(list(range(r - 2, -1, -1)) + [r - 1])).reshape(-1, 2)
        nt.assert_bits(gbuf, nt.or_fibre_dim(kind, buf, mbuf, inner, ln, outer), "%s dim=%d" % (kind, dim))


@pytest.mark.parametrize("n", [1, 7, 4096, 20000])
def test_product_rank1_bound(gpu, n):
    rng = np.random.default_rng(n)
    # factors near 1 keep the product of 20000 of them inside the double range
    a = datagen._with_lo(rng, rng.uniform(0.97, 1.03, size=n) * np.where(rng.random(n) < 0.5, -1, 1))
    mpmath.mp.prec = 400
    exact = mpmath.mpf(1)
    for x in a:
        exact *= mpmath.mpf(float(x[0])) + mpmath.mpf(float(x[1]))
    r = mf.product(a)
    got = mpmath.mpf(float(r[0])) + mpmath.mpf(float(r[1]))
    assert abs(got - exact) <= 4 * n * mpmath.mpf(2) ** -104 * abs(exact)
    mask = rng.random(n) < 0.5
    rm = mf.product(a, mask=mask)
    ex = mpmath.mpf(1)
    for x in a[mask]:
        ex *= mpmath.mpf(float(x[0])) + mpmath.mpf(float(x[1]))
    gm = mpmath.mpf(float(rm[0])) + mpmath.mpf(float(rm[1]))
    assert abs(gm - ex) <= 4 * n * mpmath.mpf(2) ** -104 * abs(ex)


def test_product_exact_cases(gpu):
    rng = np.random.default_rng(9)
    n = (1 << 21) + 11
    e = rng.integers(-1, 2, size=n)            # powers of two: every partial product is exact
    a = np.stack([np.ldexp(1.0, e) * np.where(rng.random(n) < 0.5, -1.0, 1.0), np.zeros(n)], axis=1)
    want = np.ldexp(1.0, int(e.sum())) * np.prod(np.sign(a[:, 0]))
    r = mf.product(a)
    assert r[0] == want and r[1] == 0.0
    assert (mf.product(np.zeros((0, 2))) == np.array([1.0, 0.0])).all()
    assert (mf.product(a, mask=np.zeros(n, bool)) == np.array([1.0, 0.0])).all()


def test_device_entries(gpu):
    import torch
    torch.cuda.init()
    rng = np.random.default_rng(10)
    n = 123457
    a = _data(rng, n, "nan")
    mask = (rng.random(n) < 0.5).astype(np.int32)
    da = torch.from_numpy(a).cuda()
    dm = torch.from_numpy(mask).cuda()
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    wb = mf.reduce_workspace_bytes(n)
    work = torch.empty(wb, dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    for kind, nm in ((mf.MFHIP_MAXVAL, "maxval"), (mf.MFHIP_MINVAL, "minval")):
        mf.dev_reduce(kind, da.data_ptr(), 0, n, out.data_ptr(), work.data_ptr(), 0, st)
        torch.cuda.synchronize()
        nt.assert_bits(out.cpu().numpy(), nt.or_fibre(nm, a), "device " + nm)
        mf.dev_reduce_masked(kind, da.data_ptr(), dm.data_ptr(), n, out.data_ptr(), work.data_ptr(), st)
        torch.cuda.synchronize()
        nt.assert_bits(out.cpu().numpy(), nt.or_fibre(nm, a, mask), "device masked " + nm)
    inner, ln, outer = 7, 97, 181
    buf = datagen.config0(rng, inner * ln * outer)
    dbuf = torch.from_numpy(buf).cuda()
    dout = torch.zeros((inner * outer, 2), dtype=torch.float64, device="cuda")
    for kind, nm in ((mf.MFHIP_PRODUCT, "product"), (mf.MFHIP_MAXVAL, "maxval"), (mf.MFHIP_MINVAL, "minval")):
        mf.dev_reduce_dim(kind, dbuf.data_ptr(), 0, inner, ln, outer, dout.data_ptr(), st)
        torch.cuda.synchronize()
        nt.assert_bits(dout.cpu().numpy(), nt.or_fibre_dim(nm, buf, None, inner, ln, outer), "device dim " + nm)
