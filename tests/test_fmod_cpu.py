"""fmod, modulo (the Fortran mod and modulo) and nearest on the CPU: the
oracle against the compiled reference's bits (tests/golden/fmod.npz) and
against the mathematics, and the host build of the kernels' headers against
the oracle.

The reference's fmod loop (include/multifloats/float64x2.h:2010-2059) never
returns for some operands; the fixture records those calls (made with a time
limit when it was written) and both the oracle and the kernels give NaN there.
"""
import ctypes
import os

import mpmath
import numpy as np
import pytest

from multifloats_amd import datagen
from tests import _native as nt

G = np.load(os.path.join(nt.GOLDEN, "fmod.npz"))


def _or(nm, a, b):
    L = nt.oracle()
    getattr(L, "or_" + nm).argtypes = [nt.P, nt.P, nt.P, nt.I64]
    return nt.or_binary(nm, a, b)


def _steps(a, b):
    L = nt.oracle()
    L.or_fmod_steps.argtypes = [nt.P, nt.P, nt.P, nt.I64]
    a, b = nt.dd(a), nt.dd(b)
    st = np.zeros(a.shape[0], dtype=np.int32)
    L.or_fmod_steps(nt.ptr(a), nt.ptr(b), nt.ptr(st), a.shape[0])
    return st


def _host(nm, a, b):
    L = ctypes.CDLL(nt.HOSTMATH)
    a, b = nt.dd(a), nt.dd(b)
    c = np.empty_like(a)
    getattr(L, "h_" + nm)(nt.ptr(a), nt.ptr(b), nt.ptr(c), ctypes.c_long(a.shape[0]))
    return c


def _same_bits(x, y):
    return (x.view(np.uint64) == y.view(np.uint64)).all(1) | (np.isnan(x[:, 0]) & np.isnan(y[:, 0]))


@pytest.mark.parametrize("nm", ["fmod", "modulo"])
def test_oracle_matches_reference_bits(nm):
    a, b = G["a"], G["b"]
    ok = G[nm + "_returns"]
    got = _or(nm, a, b)
    assert ok.sum() > 3000 and (~ok).sum() > 300
    same = _same_bits(got, G[nm + "_ref"])
    assert same[ok].all(), np.nonzero(ok & ~same)[0][:10]
    # where the reference never returns: NaN
    assert np.isnan(got[~ok, 0]).all()


def test_loop_bound_and_cycles():
    """Returning calls take a few steps; every non-returning call of the
    fixture is found as a revisited remainder, never by the step cap."""
    st = _steps(G["a"], G["b"])
    ok = G["fmod_returns"]
    assert st[ok].max() <= 16
    assert (st[~ok] == -1).all()


def test_fmod_mathematics():
    """|x/y| < 2^53: within 2^-50 |y| of the exact remainder (320-bit), the
    sign of x or zero, |r| <= |y|; every finite result: |r| <= |y|.
    modulo - fmod is 0 or y, and modulo has the sign of y or is zero."""
    a, b = G["a"], G["b"]
    ex = G["fmod_exact"]
    r = _or("fmod", a, b)
    md = _or("modulo", a, b)
    mpmath.mp.prec = 320
    checked = 0
    for i in range(a.shape[0]):
        if not (np.isfinite(ex[i, 0]) and np.isfinite(r[i, 0])):
            continue
        x = mpmath.mpf(a[i, 0]) + mpmath.mpf(a[i, 1])
        y = mpmath.mpf(b[i, 0]) + mpmath.mpf(b[i, 1])
        v = mpmath.mpf(r[i, 0]) + mpmath.mpf(r[i, 1])
        assert abs(v) <= abs(y), i
        w = mpmath.mpf(md[i, 0]) + mpmath.mpf(md[i, 1])
        assert w == 0 or (w < 0) == (y < 0), i
        assert w == v or w == v + y or (v != 0 and abs(w - (v + y)) <= abs(y) * mpmath.mpf(2) ** -100), i
        if abs(x / y) < 2 ** 53 and abs(y) > 1e-290:
            e = mpmath.mpf(ex[i, 0]) + mpmath.mpf(ex[i, 1]) + mpmath.mpf(ex[i, 2])
            assert abs(v - e) <= abs(y) * mpmath.mpf(2) ** -50, i
            assert v == 0 or (v < 0) == (x < 0), i
            checked += 1
    assert checked > 2000


def test_golden_bits_pin_the_loop():
    """The fixture distinguishes the reference's loop from a correctly
    rounded remainder: rounding the exact remainder to a double-double
    disagrees with the reference's bits on many returning rows (so a
    restatement that computed the exact remainder would fail the bit test)."""
    ex = G["fmod_exact"]
    ok = G["fmod_returns"] & np.isfinite(ex[:, 0])
    rounded = np.stack([ex[:, 0], ex[:, 1] + ex[:, 2]], 1)
    diff = ~_same_bits(rounded[ok], G["fmod_ref"][ok])
    assert diff.sum() > 500


def test_nearest_oracle():
    """nearest = nextafter(x.hi, +-inf) by the sign of y.hi > 0 (NaN y and
    y.hi <= 0 step down), low limb 0."""
    sp = np.array(G["b"][-18:])  # the 18 special values (the pair grid tiles b)
    rng = np.random.default_rng(5)
    x = np.concatenate([np.repeat(sp, 18, 0), datagen._with_lo(rng, 10.0 ** rng.uniform(-300, 300, 2000)
                                                                 * rng.choice([-1, 1], 2000)),
                        np.array([[0.0, 0.0], [-0.0, 0.0], [5e-324, 0.0], [-5e-324, 0.0], [np.inf, 0.0],
                                  [-np.inf, 0.0], [1.7976931348623157e308, 0.0]] * 2)])
    y = np.concatenate([np.tile(sp, (18, 1)), datagen._with_lo(rng, rng.uniform(-1, 1, 2000)),
                        np.array([[1.0, 0.0]] * 7 + [[-1.0, 0.0]] * 7)])
    got = _or("nearest", x, y)
    with np.errstate(over="ignore"):
        want = np.nextafter(x[:, 0], np.where(y[:, 0] > 0, np.inf, -np.inf))
    assert (got[:, 1] == 0.0).all()
    assert _same_bits(got, np.stack([want, np.zeros_like(want)], 1)).all()


@pytest.mark.parametrize("nm", ["fmod", "modulo", "nearest"])
def test_hostmath_matches_oracle(nm):
    """The kernels' header built for the host: the fixture and random draws
    over the loop's regimes, bit for bit with the oracle."""
    rng = np.random.default_rng(6)
    n = 20000
    w = datagen._with_lo
    s = lambda k: rng.choice([-1.0, 1.0], k)
    a = np.concatenate([G["a"], w(rng, 2.0 ** rng.uniform(-600, 600, n) * s(n))])
    b = np.concatenate([G["b"], w(rng, 2.0 ** rng.uniform(-600, 600, n) * s(n))])
    nt.assert_bits(_host(nm, a, b), _or(nm, a, b), nm)
