#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ (run in the build container).

Inputs are seeded draws from multifloats_amd.datagen plus the special and
known-answer values the reference's own tests use; expected outputs come from
  - the reference's compiled C++ path (oracle/_ref/libmfref.so, built from the
    reference sources by oracle/Makefile), called through its C ABI: adddd,
    subdd, muldd, divdd, sqrtdd, expdd, logdd, sindd, erfdd, matmuldd_mm/mv/vm,
    and cosdd, tandd, exp2dd, expm1dd, log2dd, log10dd, log1pdd, sinhdd,
    coshdd, tanhdd, erfcdd, erfcxdd, asinhdd, acoshdd, atanhdd;
  - mpmath at 320 bits for the transcendental values (stored as three limbs so
    relative errors down to 1e-40 can be measured).
Every fixture is data only (inputs and expected outputs).

Usage: python3 tests/golden/make_golden.py [fixture ...]   (default: all)
"""
import ctypes
import os
import sys

import numpy as np
import mpmath

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.normpath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)
from multifloats_amd import datagen  # noqa: E402

mpmath.mp.prec = 320


class F2(ctypes.Structure):
    _fields_ = [("limbs", ctypes.c_double * 2)]


def load_ref():
    path = os.path.join(ROOT, "oracle", "_ref", "libmfref.so")
    if not os.path.exists(path):
        sys.exit("build the reference first: make -C oracle ref")
    L = ctypes.CDLL(path)
    for nm in ("adddd", "subdd", "muldd", "divdd"):
        getattr(L, nm).argtypes = [F2, F2]
        getattr(L, nm).restype = F2
    for nm in ("sqrtdd", "expdd", "logdd", "sindd", "erfdd", "erfcdd", "erfcxdd") + tuple(
            f + "dd" for f in TRANSC2 + TRANSC3 + TRANSC4 + TRANSC5):
        getattr(L, nm).argtypes = [F2]
        getattr(L, nm).restype = F2
    for nm in INVTRIG1 + INVTRIG3:
        getattr(L, nm + "dd").argtypes = [F2]
        getattr(L, nm + "dd").restype = F2
    for nm in BINARY2 + INVTRIG2 + ("fmod", "modulo"):
        getattr(L, nm + "dd").argtypes = [F2, F2]
        getattr(L, nm + "dd").restype = F2
    P, I64 = ctypes.c_void_p, ctypes.c_int64
    L.matmuldd_mm.argtypes = [P, P, P, I64, I64, I64, I64]
    L.matmuldd_mv.argtypes = [P, P, P, I64, I64, I64]
    L.matmuldd_vm.argtypes = [P, P, P, I64, I64, I64]
    return L


def f2(v):
    r = F2()
    r.limbs[0], r.limbs[1] = float(v[0]), float(v[1])
    return r


def call1(L, nm, a):
    fn = getattr(L, nm)
    out = np.empty_like(a)
    for i in range(a.shape[0]):
        r = fn(f2(a[i]))
        out[i] = (r.limbs[0], r.limbs[1])
    return out


def call2(L, nm, a, b):
    fn = getattr(L, nm)
    out = np.empty_like(a)
    for i in range(a.shape[0]):
        r = fn(f2(a[i]), f2(b[i]))
        out[i] = (r.limbs[0], r.limbs[1])
    return out


def limbs3(v):
    out = []
    for _ in range(3):
        d = float(v)
        out.append(d)
        v = v - mpmath.mpf(d)
    return out


def mp_of(x):
    return mpmath.mpf(float(x[0])) + mpmath.mpf(float(x[1]))


SPECIALS = [(0.0, 0.0), (-0.0, 0.0), (1.0, 0.0), (-1.0, 0.0), (2.0, 0.0), (0.5, 1e-17),
            (np.inf, 0.0), (-np.inf, 0.0), (np.nan, 0.0), (1e300, 1e283), (-1e300, 0.0),
            (1e-300, 0.0), (5e-324, 0.0), (2.2250738585072014e-308, 0.0),
            (1.7976931348623157e308, 0.0), (3.0, -1e-16), (1e16, 1.0), (-7.5, 2e-16)]


def arith(L, rng):
    n = 3000
    a = np.concatenate([datagen.config0(rng, n), datagen.spread(rng, n, 30),
                        np.array(SPECIALS * 2)])
    b = np.concatenate([datagen.config0(rng, n), datagen.spread(rng, n, 30),
                        np.array(SPECIALS[::-1] + SPECIALS)])
    # exact cancellations and equal-hi cases
    a[:50] = b[:50]
    a[50:100, 0] = -b[50:100, 0]
    out = dict(a=a, b=b)
    for nm in ("adddd", "subdd", "muldd", "divdd"):
        out[nm] = call2(L, nm, a, b)
    sq_in = np.abs(a)
    sq_in[-len(SPECIALS):] = a[-len(SPECIALS):]  # keep negatives and NaN among specials
    out["sqrt_in"] = sq_in
    out["sqrtdd"] = call1(L, "sqrtdd", sq_in)
    np.savez_compressed(os.path.join(HERE, "arith.npz"), **out)


def transc(L, rng):
    n = 1500
    cases = {
        "exp": np.concatenate([datagen.exp_inputs(rng, n), datagen.config0(rng, 200),
                               datagen._with_lo(rng, rng.uniform(-745.5, -700.0, 100)),
                               datagen._with_lo(rng, rng.uniform(708.0, 709.78, 50)),
                               np.array(SPECIALS)]),
        "log": np.concatenate([datagen.log_inputs(rng, n),
                               datagen._with_lo(rng, rng.uniform(15 / 16, 17 / 16, 300)),
                               np.abs(datagen.spread(rng, 200, 40)),
                               np.array([(2.0 ** k, 0.0) for k in range(-1074, 1024, 37)]),
                               np.array(SPECIALS)]),
        "sin": np.concatenate([datagen.sin_inputs(rng, n, near_frac=0.2),
                               datagen.config0(rng, 200),
                               np.array([(2 * np.pi * 2.0 ** k, 0.0) for k in range(0, 40)]),
                               np.array(SPECIALS)]),
        "erf": np.concatenate([datagen.erf_inputs(rng, n), datagen.config0(rng, 200),
                               datagen._with_lo(rng, rng.uniform(-0.3, 0.3, 200)),
                               datagen._with_lo(rng, 10.0 ** rng.uniform(-30, 0, 100)),
                               datagen._with_lo(rng, rng.uniform(5.5, 10.0, 100)),
                               np.array(SPECIALS)]),
    }
    fn = {"exp": mpmath.exp, "log": mpmath.log, "sin": mpmath.sin, "erf": mpmath.erf}
    out = {}
    for nm, x in cases.items():
        out[nm + "_in"] = x
        out[nm + "_ref"] = call1(L, nm + "dd", x)
        exact = np.zeros((x.shape[0], 3))
        for i in range(x.shape[0]):
            if not np.isfinite(x[i, 0]):
                exact[i] = (np.nan, 0, 0)
                continue
            v = mp_of(x[i])
            if nm == "log" and v <= 0:
                exact[i] = (np.nan, 0, 0)
                continue
            try:
                exact[i] = limbs3(fn[nm](v))
            except (OverflowError, ValueError):
                exact[i] = (np.nan, 0, 0)
        out[nm + "_exact"] = exact
    np.savez_compressed(os.path.join(HERE, "transc.npz"), **out)


TRANSC2 = ("cos", "tan", "exp2", "expm1", "log2", "log10", "log1p", "sinh", "cosh", "tanh")
TRANSC3 = ("asinh", "acosh", "atanh")
TRANSC4 = ("exp10", "exp2m1", "exp10m1", "log2p1", "log10p1")
TRANSC5 = ("sinpi", "cospi", "tanpi")
BINARY2 = ("pow", "hypot", "fdim", "fmin", "fmax", "copysign")
INVTRIG1 = ("atan", "atanpi")
INVTRIG2 = ("atan2", "atan2pi")
INVTRIG3 = ("asin", "acos", "asinpi", "acospi")


def transc2(L, rng):
    """cos, tan, exp2, expm1, log2, log10, log1p, sinh, cosh, tanh: seeded
    draws over each function's regimes (the branch points of the reference's
    algorithms included), special values, and 320-bit values for accuracy."""
    n = 1500
    w = datagen._with_lo
    U = rng.uniform
    near_half = np.array([(0.5 * s, 0.0) for s in (1, -1)] + [(np.nextafter(0.5, 0) * s, 0.0) for s in (1, -1)])
    cases = {
        "cos": np.concatenate([datagen.sin_inputs(rng, n, near_frac=0.2), datagen.config0(rng, 200),
                               np.array([(2 * np.pi * 2.0 ** k, 0.0) for k in range(0, 40)])]),
        "tan": np.concatenate([datagen.sin_inputs(rng, n, near_frac=0.2), datagen.config0(rng, 200)]),
        "exp2": np.concatenate([w(rng, U(-1100, 1100, n)), datagen.config0(rng, 200),
                                w(rng, U(1020, 1024, 50)), w(rng, U(-1080, -1070, 50)),
                                np.array([(1023.9999999999998, 0.0), (1024.0, 0.0), (-1080.0, 0.0),
                                          (-1080.0000000000002, 0.0), (0.5, 0.0), (10.0, 1e-16)])]),
        "expm1": np.concatenate([w(rng, U(-0.5, 0.5, n)), w(rng, 10.0 ** U(-30, -1, 200)),
                                 w(rng, U(-745, 709, 300)), near_half]),
        "log2": np.concatenate([datagen.log_inputs(rng, n), w(rng, U(15 / 16, 17 / 16, 300)),
                                np.array([(2.0 ** k, 0.0) for k in range(-1074, 1024, 37)])]),
        "log10": np.concatenate([datagen.log_inputs(rng, n), w(rng, U(15 / 16, 17 / 16, 300)),
                                 np.array([(10.0 ** k, 0.0) for k in range(-300, 300, 13)])]),
        "log1p": np.concatenate([w(rng, U(-0.25, 0.25, n)), w(rng, 10.0 ** U(-30, -1, 200)),
                                 w(rng, -(10.0 ** U(-30, -1, 100))), w(rng, U(-0.999, 3, 300)),
                                 w(rng, 10.0 ** U(0, 300, 100)), w(rng, -1.0 - 10.0 ** U(-10, 2, 30)),
                                 np.array([(-1.0, 0.0), (-1.0, 1e-20), (-1.0, -1e-20), (0.25, 0.0),
                                           (-0.25, 0.0)])]),
        "sinh": np.concatenate([w(rng, U(-40, 40, n)), w(rng, U(-0.5, 0.5, 300)),
                                w(rng, U(-715, 715, 100)), near_half]),
        "cosh": np.concatenate([w(rng, U(-40, 40, n)), w(rng, U(-0.5, 0.5, 300)), w(rng, U(-715, 715, 100))]),
        "tanh": np.concatenate([w(rng, U(-40, 40, n)), w(rng, U(-0.5, 0.5, 300)),
                                np.array([(38.0, 0.0), (-38.0, 0.0), (37.99999999999999, 0.0)]), near_half]),
    }
    mpfn = {"cos": mpmath.cos, "tan": mpmath.tan, "exp2": lambda v: mpmath.power(2, v), "expm1": mpmath.expm1,
            "log2": lambda v: mpmath.log(v, 2), "log10": mpmath.log10, "log1p": mpmath.log1p,
            "sinh": mpmath.sinh, "cosh": mpmath.cosh, "tanh": mpmath.tanh}
    out = {}
    for nm, x in cases.items():
        x = np.concatenate([x, np.array(SPECIALS)])
        out[nm + "_in"] = x
        out[nm + "_ref"] = call1(L, nm + "dd", x)
        exact = np.full((x.shape[0], 3), np.nan)
        for i in range(x.shape[0]):
            if not np.isfinite(x[i, 0]):
                continue
            try:
                v = mpfn[nm](mp_of(x[i]))
                if mpmath.im(v) == 0:
                    exact[i] = limbs3(mpmath.re(v))
            except (OverflowError, ValueError, ZeroDivisionError):
                pass
        out[nm + "_exact"] = exact
    np.savez_compressed(os.path.join(HERE, "transc2.npz"), **out)


def transc3(L, rng):
    """asinh, acosh, atanh: seeded draws over every branch of the reference's
    algorithms (|x| < 0.01 Taylor, below / above 1, beyond 1e150, the domain
    edges), special values, and 320-bit values for accuracy."""
    n = 1500
    w = datagen._with_lo
    U = rng.uniform
    edge = np.array([(0.01, 0.0), (-0.01, 0.0), (np.nextafter(0.01, 0), 0.0), (1.0, 0.0), (-1.0, 0.0),
                     (1.0, -1e-20), (-1.0, 1e-20), (1.0, 1e-20), (np.nextafter(1.0, 0), 0.0), (1e150, 0.0),
                     (1.0000000000000002e150, 0.0), (-1e151, 0.0), (2.0, 0.0), (0.5, 0.0)])
    cases = {
        "asinh": np.concatenate([w(rng, U(-0.01, 0.01, 300)), w(rng, U(-1, 1, n)), w(rng, U(-40, 40, 400)),
                                 w(rng, 10.0 ** U(-30, 300, 300) * rng.choice([-1.0, 1.0], 300)), edge]),
        "acosh": np.concatenate([w(rng, 1.0 + 10.0 ** U(-16, 0, 400)), w(rng, U(1, 40, n)),
                                 w(rng, 10.0 ** U(1, 308, 300)), w(rng, U(-3, 1, 100)), edge]),
        "atanh": np.concatenate([w(rng, U(-0.01, 0.01, 300)), w(rng, U(-1, 1, n)),
                                 w(rng, (1.0 - 10.0 ** U(-16, -1, 300)) * rng.choice([-1.0, 1.0], 300)),
                                 w(rng, U(-3, 3, 100)), edge]),
    }
    mpfn = {"asinh": mpmath.asinh, "acosh": mpmath.acosh, "atanh": mpmath.atanh}
    out = {}
    for nm, x in cases.items():
        x = np.concatenate([x, np.array(SPECIALS)])
        out[nm + "_in"] = x
        out[nm + "_ref"] = call1(L, nm + "dd", x)
        exact = np.full((x.shape[0], 3), np.nan)
        for i in range(x.shape[0]):
            if not np.isfinite(x[i, 0]):
                continue
            try:
                v = mpfn[nm](mp_of(x[i]))
                if mpmath.im(v) == 0 and mpmath.isfinite(v):
                    exact[i] = limbs3(mpmath.re(v))
            except (OverflowError, ValueError, ZeroDivisionError):
                pass
        out[nm + "_exact"] = exact
    np.savez_compressed(os.path.join(HERE, "transc3.npz"), **out)


def transc4(L, rng):
    """exp10, exp2m1, exp10m1, log2p1, log10p1: seeded draws over every branch
    of the reference's algorithms (the |x| = 1/2 switch of the C23 forms, the
    overflow and underflow clamps of exp10, the log1p regimes), special values,
    and 320-bit values for accuracy."""
    n = 1500
    w = datagen._with_lo
    U = rng.uniform
    near_half = np.array([(0.5 * s, 0.0) for s in (1, -1)] + [(np.nextafter(0.5, 0) * s, 0.0) for s in (1, -1)]
                         + [(0.5 * s, -1e-20 * s) for s in (1, -1)])
    hi_lim = (1024.0 - 2.0 ** -42) * float.fromhex("0x1.34413509f79ffp-2")  # exp10's clamps
    lo_lim = -1080.0 * float.fromhex("0x1.34413509f79ffp-2")
    ten_edges = np.array([(hi_lim, 0.0), (np.nextafter(hi_lim, 0), 0.0), (np.nextafter(hi_lim, 400), 0.0),
                          (hi_lim, 1e-14), (lo_lim, 0.0), (np.nextafter(lo_lim, 0), 0.0),
                          (np.nextafter(lo_lim, -400), 0.0), (lo_lim, -1e-14),
                          (308.0, 0.0), (-323.0, 0.0), (-307.5, 0.0)]
                         + [(float(k), 0.0) for k in range(-22, 23)])
    log1p_like = np.concatenate([w(rng, U(-0.25, 0.25, n)), w(rng, 10.0 ** U(-30, -1, 200)),
                                 w(rng, -(10.0 ** U(-30, -1, 100))), w(rng, U(-0.999, 3, 300)),
                                 w(rng, 10.0 ** U(0, 300, 100)), w(rng, -1.0 - 10.0 ** U(-10, 2, 30)),
                                 np.array([(-1.0, 0.0), (-1.0, 1e-20), (-1.0, -1e-20), (0.25, 0.0), (-0.25, 0.0),
                                           (1.0, 0.0), (9.0, 0.0), (1023.0, 0.0)])])
    cases = {
        "exp10": np.concatenate([w(rng, U(-330, 310, n)), datagen.config0(rng, 300), w(rng, U(300, 309, 100)),
                                 w(rng, U(-326, -320, 100)), ten_edges]),
        "exp2m1": np.concatenate([w(rng, U(-0.5, 0.5, n)), w(rng, 10.0 ** U(-30, -1, 200) * rng.choice([-1.0, 1.0], 200)),
                                  w(rng, U(-1100, 1100, 300)), w(rng, U(1020, 1024, 40)), near_half,
                                  np.array([(1024.0, 0.0), (-1080.0, 0.0), (-1100.0, 0.0), (1.0, 0.0)])]),
        "exp10m1": np.concatenate([w(rng, U(-0.5, 0.5, n)), w(rng, 10.0 ** U(-30, -1, 200) * rng.choice([-1.0, 1.0], 200)),
                                   w(rng, U(-330, 310, 300)), near_half, ten_edges]),
        "log2p1": log1p_like.copy(),
        "log10p1": log1p_like.copy(),
    }
    ln2, ln10 = mpmath.log(2), mpmath.log(10)
    mpfn = {"exp10": lambda v: mpmath.power(10, v), "exp2m1": lambda v: mpmath.expm1(v * ln2),
            "exp10m1": lambda v: mpmath.expm1(v * ln10), "log2p1": lambda v: mpmath.log1p(v) / ln2,
            "log10p1": lambda v: mpmath.log1p(v) / ln10}
    out = {}
    for nm, x in cases.items():
        x = np.concatenate([x, np.array(SPECIALS)])
        out[nm + "_in"] = x
        out[nm + "_ref"] = call1(L, nm + "dd", x)
        exact = np.full((x.shape[0], 3), np.nan)
        for i in range(x.shape[0]):
            if not np.isfinite(x[i, 0]):
                continue
            try:
                v = mpfn[nm](mp_of(x[i]))
                if mpmath.im(v) == 0 and mpmath.isfinite(v):
                    exact[i] = limbs3(mpmath.re(v))
            except (OverflowError, ValueError, ZeroDivisionError):
                pass
        out[nm + "_exact"] = exact
    np.savez_compressed(os.path.join(HERE, "transc4.npz"), **out)


def transc5(L, rng):
    """sinpi, cospi, tanpi: seeded draws over every quadrant (n = RN(2|x|)
    mod 4) and both sin/cos kernel branches (|pi r| above and below pi/8),
    exact multiples of 1/4 and 1/2 (zeros, poles, +-1), points next to them,
    large |x| up to and past the 2^52 cap, special values, and 320-bit values
    for accuracy."""
    n = 1500
    w = datagen._with_lo
    U = rng.uniform
    quarters = np.array([(k / 4.0, 0.0) for k in range(-16, 17)])
    near = np.concatenate([w(rng, np.round(U(-40, 40, 200) * 4) / 4 + 10.0 ** U(-17, -3, 200)
                            * rng.choice([-1.0, 1.0], 200))])
    big = np.array([(2.0 ** 52, 0.0), (np.nextafter(2.0 ** 52, 0), 0.0), (-(2.0 ** 52), 0.0), (2.0 ** 51 + 0.5, 0.0),
                    (2.0 ** 50 + 0.25, 0.0), (3.0e15, 0.0), (1e300, 0.0)])
    common = [w(rng, U(-2, 2, n)), w(rng, U(-1e6, 1e6, 300)), w(rng, 10.0 ** U(-30, -1, 200) * rng.choice([-1.0, 1.0], 200)),
              w(rng, U(-0.125, 0.125, 200)), quarters, near, big]
    cases = {nm: np.concatenate(common) for nm in TRANSC5}
    mpfn = {"sinpi": mpmath.sinpi, "cospi": mpmath.cospi, "tanpi": lambda v: mpmath.sinpi(v) / mpmath.cospi(v)}
    out = {}
    for nm, x in cases.items():
        x = np.concatenate([x, np.array(SPECIALS)])
        out[nm + "_in"] = x
        out[nm + "_ref"] = call1(L, nm + "dd", x)
        exact = np.full((x.shape[0], 3), np.nan)
        for i in range(x.shape[0]):
            if not np.isfinite(x[i, 0]) or abs(x[i, 0]) >= 2.0 ** 52:
                continue
            try:
                v = mpfn[nm](mp_of(x[i]))
                if mpmath.im(v) == 0 and mpmath.isfinite(v):
                    exact[i] = limbs3(mpmath.re(v))
            except (OverflowError, ValueError, ZeroDivisionError):
                pass
        out[nm + "_exact"] = exact
    np.savez_compressed(os.path.join(HERE, "transc5.npz"), **out)


def binary2(L, rng):
    """pow, hypot, fdim, fmin, fmax, copysign: every pair of special values,
    and per-function draws over the reference's branches (pow: integer y
    through powi, negative and non-finite x through exp(y log x), x near 1 on
    log2's direct path, results near overflow and underflow; hypot: wide
    exponent spreads and subnormals; the order-based functions: equal hi
    limbs with different lo, signed zeros, NaN in either place), with
    320-bit values of pow and hypot for accuracy."""
    n = 1500
    w = datagen._with_lo
    U = rng.uniform
    sp = np.array(SPECIALS)
    ga, gb = np.repeat(sp, len(sp), axis=0), np.tile(sp, (len(sp), 1))
    sgn = lambda k: rng.choice([-1.0, 1.0], k)
    cases = {
        "pow": (np.concatenate([w(rng, 10.0 ** U(-5, 3, n)), w(rng, U(0.9, 1.1, 300)), w(rng, -(10.0 ** U(-2, 2, 200))),
                                w(rng, 10.0 ** U(-300, 300, 200)), w(rng, U(0.5, 2, 200))]),
                np.concatenate([w(rng, U(-60, 60, n)), w(rng, U(-1e4, 1e4, 300)), np.round(U(-40, 40, 200))[:, None]
                                * np.array([1.0, 0.0]), w(rng, U(-1, 1, 200)), np.round(U(-2e9, 2e9, 200))[:, None]
                                * np.array([1.0, 0.0])])),
        "hypot": (w(rng, 2.0 ** U(-500, 500, 2 * n) * sgn(2 * n)), w(rng, 2.0 ** U(-500, 500, 2 * n) * sgn(2 * n))),
    }
    base = w(rng, U(-4, 4, n))
    other = base.copy()
    other[: n // 2, 1] = other[: n // 2, 1] * U(-2, 2, n // 2)  # equal hi, different lo
    other[n // 2:] = w(rng, U(-4, 4, n - n // 2))
    zeros = np.array([(0.0, 0.0), (-0.0, 0.0), (0.0, -1e-320), (-0.0, 1e-320), (0.0, 1e-300), (-0.0, -1e-300)])
    za, zb = np.repeat(zeros, len(zeros), axis=0), np.tile(zeros, (len(zeros), 1))
    for nm in ("fdim", "fmin", "fmax", "copysign"):
        cases[nm] = (np.concatenate([base, za]), np.concatenate([other, zb]))
    mpfn = {"pow": lambda a, b: mpmath.power(a, b) if a > 0 or b == int(b) else None,
            "hypot": lambda a, b: mpmath.sqrt(a * a + b * b)}
    out = {}
    for nm, (a, b) in cases.items():
        a, b = np.concatenate([a, ga]), np.concatenate([b, gb])
        out[nm + "_a"], out[nm + "_b"] = a, b
        out[nm + "_ref"] = call2(L, nm + "dd", a, b)
        if nm in mpfn:
            exact = np.full((a.shape[0], 3), np.nan)
            for i in range(a.shape[0]):
                if not (np.isfinite(a[i, 0]) and np.isfinite(b[i, 0])):
                    continue
                try:
                    v = mpfn[nm](mp_of(a[i]), mp_of(b[i]))
                    if v is not None and mpmath.isfinite(v) and abs(v) < 1e308:
                        exact[i] = limbs3(v)
                except (OverflowError, ValueError, ZeroDivisionError):
                    pass
            out[nm + "_exact"] = exact
    np.savez_compressed(os.path.join(HERE, "binary2.npz"), **out)


def call2_guarded(nm, a, b, stall=1.0):
    """The reference's nm(a_i, b_i) for every i, in a forked child that marks
    each finished call in shared memory; a call that makes no progress for
    `stall` seconds is recorded as not returning (the child is killed and a
    new one resumes after it). Returns (outputs, returned)."""
    import mmap
    import signal
    import time
    n = a.shape[0]
    out = np.full((n, 2), np.nan)
    done = np.zeros(n, dtype=bool)
    start = 0
    while start < n:
        buf = mmap.mmap(-1, n * 24)
        view = np.frombuffer(buf, dtype=np.float64).reshape(n, 3)
        pid = os.fork()
        if pid == 0:  # child: its own copy of the library
            L = load_ref()
            fn = getattr(L, nm)
            for i in range(start, n):
                r = fn(f2(a[i]), f2(b[i]))
                view[i, 0], view[i, 1], view[i, 2] = r.limbs[0], r.limbs[1], 1.0
            os._exit(0)
        last, t_last = -1, time.time()
        while True:
            fin, _ = os.waitpid(pid, os.WNOHANG)
            prog = int(np.count_nonzero(view[start:, 2]))
            if fin:
                break
            if prog != last:
                last, t_last = prog, time.time()
            elif time.time() - t_last > stall:
                os.kill(pid, signal.SIGKILL)
                os.waitpid(pid, 0)
                break
            time.sleep(0.01)
        got = view[:, 2] == 1.0
        out[got] = view[got, :2]
        done |= got
        nxt = start + int(np.count_nonzero(view[start:, 2]))
        if nxt < n and not got[nxt]:
            nxt += 1  # the call at nxt did not return
        start = nxt
        del view
        buf.close()
    return out, done


def fmod_golden(L, rng):
    """fmod and modulo: every pair of special values, and draws over the
    reference's loop regimes (x/y below 2^53, the DD-quotient steps of wide
    exponent gaps up to the double range, near-multiples, negative operands
    and the add-back path, subnormal divisors), outputs of the compiled
    reference and a flag for the calls that do not return (recorded with a
    time limit), plus the exact remainders (320-bit) for the mathematical
    checks. Inputs are canonical double-doubles."""
    w = datagen._with_lo
    U = rng.uniform
    sgn = lambda k: rng.choice([-1.0, 1.0], k)
    sp = np.array(SPECIALS)
    ga, gb = np.repeat(sp, len(sp), axis=0), np.tile(sp, (len(sp), 1))
    n = 600
    ys = w(rng, 10.0 ** U(-3, 3, n) * sgn(n))
    near = ys * np.round(U(1, 1e6, n))[:, None]  # near-multiples: q y, perturbed in lo
    near = w(rng, near[:, 0])
    parts = [
        (w(rng, 10.0 ** U(-3, 3, n) * sgn(n)), w(rng, 10.0 ** U(-3, 3, n) * sgn(n))),
        (w(rng, 2.0 ** U(0, 52, n) * sgn(n)), w(rng, U(0.5, 2, n) * sgn(n))),
        (w(rng, 2.0 ** U(53, 1000, n) * sgn(n)), w(rng, 2.0 ** U(-20, 20, n) * sgn(n))),
        (w(rng, 2.0 ** U(-500, 1000, n) * sgn(n)), w(rng, 2.0 ** U(-1000, 500, n) * sgn(n))),
        (near, ys),
        (w(rng, 2.0 ** U(-10, 60, 200) * sgn(200)), np.stack([2.0 ** U(-1074, -1022, 200) * sgn(200),
                                                                np.zeros(200)], 1)),
    ]
    a = np.concatenate([p[0] for p in parts] + [ga])
    b = np.concatenate([p[1] for p in parts] + [gb])
    out = {"a": a, "b": b}
    for nm in ("fmod", "modulo"):
        r, ok = call2_guarded(nm + "dd", a, b)
        out[nm + "_ref"], out[nm + "_returns"] = r, ok
        print(nm, "calls that do not return:", int((~ok).sum()), flush=True)
    exact = np.full((a.shape[0], 3), np.nan)
    for i in range(a.shape[0]):
        if not (np.isfinite(a[i, 0]) and np.isfinite(b[i, 0])) or b[i, 0] == 0.0:
            continue
        mpmath.mp.prec = 2400
        x, y = mp_of(a[i]), mp_of(b[i])
        r = x - mpmath.floor(abs(x / y)) * mpmath.sign(x / y) * y
        mpmath.mp.prec = 320
        exact[i] = limbs3(r)
    out["fmod_exact"] = exact
    np.savez_compressed(os.path.join(HERE, "fmod.npz"), **out)


def invtrig(L, rng):
    """atan, atanpi, atan2, atan2pi (this project's own atan, so the bar is
    accuracy): the reference's outputs and 320-bit values on draws over every
    branch of both designs (|x| below and above 1/8 and 8, table cells and
    their edges, huge and tiny |x|), all four atan2 quadrants with either
    ratio, every pair of special values, and signed zeros with and without
    a nonzero low limb."""
    n = 1500
    w = datagen._with_lo
    U = rng.uniform
    sg = lambda k: rng.choice([-1.0, 1.0], k)
    cells = np.array([((k + d) / 64.0, 0.0) for k in range(8, 513, 7) for d in (-0.5, 0.5, 0.0)])
    x1 = np.concatenate([w(rng, U(-0.2, 0.2, n)), w(rng, U(-9, 9, 2 * n)), w(rng, 10.0 ** U(-20, 300, n) * sg(n)),
                         cells, np.array([(0.125, 0.0), (np.nextafter(0.125, 0), 0.0), (8.0, 0.0), (8.0, 1e-16),
                                          (1.85e-16, 0.0), (1.8e-16, 1e-33)]), np.array(SPECIALS)])
    sp = np.array(SPECIALS)
    zeros = np.array([(0.0, 0.0), (-0.0, 0.0), (0.0, 1e-300), (-0.0, -1e-300), (0.0, -5e-324), (-0.0, 3e-310)])
    ya = np.concatenate([w(rng, U(-5, 5, 2 * n)), w(rng, 10.0 ** U(-30, 30, n) * sg(n)), np.repeat(sp, len(sp), axis=0),
                         np.repeat(zeros, len(zeros), axis=0)])
    xa = np.concatenate([w(rng, U(-5, 5, 2 * n)), w(rng, 10.0 ** U(-30, 30, n) * sg(n)), np.tile(sp, (len(sp), 1)),
                         np.tile(zeros, (len(zeros), 1))])
    out = {"x": x1, "y2": ya, "x2": xa}
    for nm in INVTRIG1:
        out[nm + "_ref"] = call1(L, nm + "dd", x1)
    for nm in INVTRIG2:
        out[nm + "_ref"] = call2(L, nm + "dd", ya, xa)
    pi = mpmath.pi
    ex = {nm: np.full((x1.shape[0], 3), np.nan) for nm in INVTRIG1}
    for i in range(x1.shape[0]):
        if np.isfinite(x1[i, 0]) and x1[i, 0] != 0.0:
            v = mpmath.atan(mp_of(x1[i]))
            ex["atan"][i], ex["atanpi"][i] = limbs3(v), limbs3(v / pi)
    ex2 = {nm: np.full((xa.shape[0], 3), np.nan) for nm in INVTRIG2}
    for i in range(xa.shape[0]):
        yv, xv = ya[i], xa[i]
        if np.isfinite(yv[0]) and np.isfinite(xv[0]) and yv[0] != 0.0 and xv[0] != 0.0:
            v = mpmath.atan2(mp_of(yv), mp_of(xv))
            ex2["atan2"][i], ex2["atan2pi"][i] = limbs3(v), limbs3(v / pi)
    for nm, e in list(ex.items()) + list(ex2.items()):
        out[nm + "_exact"] = e
    np.savez_compressed(os.path.join(HERE, "invtrig.npz"), **out)


def invtrig3(L, rng):
    """asin, acos, asinpi, acospi (own designs, so the bar is accuracy): the
    reference's outputs and 320-bit values on draws over (-1, 1), near +-1
    (1 - 10^-16 .. 1 - 10^-1), tiny |x|, the small-ratio, table and
    reflected paths of the atan reduction, the domain edges and special
    values."""
    n = 2000
    w = datagen._with_lo
    U = rng.uniform
    sg = lambda k: rng.choice([-1.0, 1.0], k)
    x = np.concatenate([w(rng, U(-1, 1, n)), w(rng, (1 - 10.0 ** U(-16, -1, 600)) * sg(600)),
                        w(rng, 10.0 ** U(-18, -1, 600) * sg(600)), w(rng, U(-0.15, 0.15, 600)),
                        np.array([(1.0, 0.0), (-1.0, 0.0), (1.0, -1e-17), (-1.0, 1e-17), (1.0, 1e-17), (0.5, 0.0),
                                  (-0.5, 0.0), (np.nextafter(1.0, 0), 0.0), (1.85e-16, 0.0), (1e-300, 0.0)]),
                        np.array(SPECIALS)])
    out = {"x": x}
    pi = mpmath.pi
    for nm in INVTRIG3:
        out[nm + "_ref"] = call1(L, nm + "dd", x)
    fn = {"asin": mpmath.asin, "acos": mpmath.acos}
    for nm in INVTRIG3:
        ex = np.full((x.shape[0], 3), np.nan)
        for i in range(x.shape[0]):
            v = mp_of(x[i]) if np.isfinite(x[i, 0]) else None
            if v is None or abs(v) > 1:
                continue
            r = fn[nm[:4]](v) / (pi if nm.endswith("pi") else 1)
            if r != 0:
                ex[i] = limbs3(r)
        out[nm + "_exact"] = ex
    np.savez_compressed(os.path.join(HERE, "invtrig3.npz"), **out)


def special2(L, rng):
    """erfc and erfcx: per-regime draws (the branch points of both the
    reference's fits and this project's rows), the reference's own outputs on
    them, and 320-bit values. Accuracy is judged against the reference's error
    on these same inputs."""
    w = datagen._with_lo
    U = rng.uniform
    n = 600
    erfc_x = np.concatenate([w(rng, U(lo, hi, n)) for lo, hi in
                             [(0, 0.25), (0.25, 1.25), (1.25, 4), (4, 9), (9, 26.5), (26.5, 27.5),
                              (-0.25, 0), (-2, -0.25), (-9, -2), (-12, -9)]] +
                            [w(rng, 10.0 ** U(-30, -1, 100)), np.array(SPECIALS)])
    erfcx_x = np.concatenate([w(rng, U(lo, hi, n)) for lo, hi in
                              [(0, 1.25), (1.25, 4), (4, 32), (32, 1000), (-2, 0), (-5, -2), (-26.5, -5),
                               (-27.5, -26.5)]] +
                             [w(rng, 10.0 ** U(3, 300, 100)), w(rng, 10.0 ** U(-30, -1, 100)), np.array(SPECIALS)])
    out = {}
    for nm, x, f in (("erfc", erfc_x, mpmath.erfc),
                     ("erfcx", erfcx_x, lambda v: mpmath.exp(v * v) * mpmath.erfc(v))):
        out[nm + "_in"] = x
        out[nm + "_ref"] = call1(L, nm + "dd", x)
        exact = np.full((x.shape[0], 3), np.nan)
        for i in range(x.shape[0]):
            if np.isfinite(x[i, 0]):
                try:
                    exact[i] = limbs3(f(mp_of(x[i])))
                except (OverflowError, ValueError):
                    pass
        out[nm + "_exact"] = exact
    np.savez_compressed(os.path.join(HERE, "special2.npz"), **out)


def matmul(L, rng):
    shapes = [(1, 1, 1), (7, 1, 9), (3, 9, 1), (2, 20, 5), (16, 16, 16), (4, 32, 6),
              (8, 8, 8), (33, 17, 31), (65, 70, 66)]
    for k in (1, 7, 8, 9, 16, 17, 64, 1000):
        shapes.append((5, k, 3))
    out = {}
    idx = 0
    for (m, k, n) in shapes:
        for ri in (0, 8, 64):
            for e in (None, 20):
                A = datagen.matrix(rng, m, k, e)
                B = datagen.matrix(rng, k, n, e)
                assert datagen.is_colmajor(A) and datagen.is_colmajor(B)
                C = datagen.as_colmajor(np.zeros((m, n, 2)))
                L.matmuldd_mm(A.ctypes.data, B.ctypes.data, C.ctypes.data, m, k, n, ri)
                x = datagen.config0(rng, k)
                y = np.zeros((m, 2))
                L.matmuldd_mv(A.ctypes.data, x.ctypes.data, y.ctypes.data, m, k, ri)
                z = np.zeros((n, 2))
                L.matmuldd_vm(x.ctypes.data, B.ctypes.data, z.ctypes.data, k, n, ri)
                key = "c%03d" % idx
                out[key + "_shape"] = np.array([m, k, n, ri])
                out[key + "_A"] = datagen.flat(A)
                out[key + "_B"] = datagen.flat(B)
                out[key + "_C"] = datagen.flat(C)
                out[key + "_x"] = x
                out[key + "_mv"] = y
                out[key + "_vm"] = z
                idx += 1
    out["count"] = np.array([idx])
    np.savez_compressed(os.path.join(HERE, "matmul.npz"), **out)


def cmatmul(L, rng):
    """Complex products through the reference's own composition
    (fsrc/multifloats.fypp:2609-2621): four matmuldd_mm calls on the split
    real/imaginary parts, then subdd / adddd elementwise."""
    out = {}
    shapes = [(1, 1, 1), (7, 5, 3), (33, 17, 31), (16, 64, 8), (1, 40, 9), (9, 40, 1)]
    idx = 0
    for (m, k, n) in shapes:
        for ri in (0, 8):
            Ar, Ai = datagen.matrix(rng, m, k, 10), datagen.matrix(rng, m, k, 10)
            Br, Bi = datagen.matrix(rng, k, n, 10), datagen.matrix(rng, k, n, 10)

            def mm(X, Y):
                Z = datagen.as_colmajor(np.zeros((m, n, 2)))
                L.matmuldd_mm(X.ctypes.data, Y.ctypes.data, Z.ctypes.data, m, k, n, ri)
                return datagen.flat(Z)
            re = call2(L, "subdd", mm(Ar, Br), mm(Ai, Bi))
            im = call2(L, "adddd", mm(Ar, Bi), mm(Ai, Br))
            key = "z%03d" % idx
            out[key + "_shape"] = np.array([m, k, n, ri])
            out[key + "_A"] = np.concatenate([datagen.flat(Ar), datagen.flat(Ai)], axis=1)
            out[key + "_B"] = np.concatenate([datagen.flat(Br), datagen.flat(Bi)], axis=1)
            out[key + "_C"] = np.concatenate([re, im], axis=1)
            idx += 1
    out["count"] = np.array([idx])
    np.savez_compressed(os.path.join(HERE, "cmatmul.npz"), **out)


def c0(L):
    """Config C0: C = A B at m = n = k = 256 with config-0 entries (seed
    20261204: A then B from datagen.matrix), renorm interval 8, by the
    compiled reference. Stored compactly: the SHA-256 of all outputs (the
    column-major float64x2 bytes) and the first column in full."""
    import hashlib
    import json
    rng = np.random.default_rng(20261204)
    m = n = k = 256
    A = datagen.as_colmajor(datagen.matrix(rng, m, k))
    B = datagen.as_colmajor(datagen.matrix(rng, k, n))
    C = datagen.as_colmajor(np.zeros((m, n, 2)))
    L.matmuldd_mm(A.ctypes.data, B.ctypes.data, C.ctypes.data, m, k, n, 8)
    flat = np.ascontiguousarray(datagen.flat(C))
    out = {"seed": 20261204, "m": m, "k": k, "n": n, "renorm_interval": 8,
           "sha256": hashlib.sha256(flat.tobytes()).hexdigest(),
           "first_column": [[float.hex(float(h)), float.hex(float(l))] for h, l in flat[:m]]}
    with open(os.path.join(HERE, "c0_256.json"), "w") as f:
        json.dump(out, f, indent=0)


def main(which):
    L = load_ref()
    rng = np.random.default_rng(20261203)
    if not which or "arith" in which or "transc" in which or "matmul" in which:
        # these three share one generator stream, so they are written together
        arith(L, rng)
        transc(L, rng)
        matmul(L, rng)
    if not which or "cmatmul" in which:
        cmatmul(L, np.random.default_rng(777))
    if not which or "transc2" in which:
        transc2(L, np.random.default_rng(4242))
    if not which or "special2" in which:
        special2(L, np.random.default_rng(777777))
    if not which or "transc3" in which:
        transc3(L, np.random.default_rng(31337))
    if not which or "transc4" in which:
        transc4(L, np.random.default_rng(5151))
    if not which or "transc5" in which:
        transc5(L, np.random.default_rng(6161))
    if not which or "binary2" in which:
        binary2(L, np.random.default_rng(7171))
    if not which or "invtrig" in which:
        invtrig(L, np.random.default_rng(8181))
    if not which or "invtrig3" in which:
        invtrig3(L, np.random.default_rng(9191))
    if not which or "c0" in which:
        c0(L)
    if not which or "fmod" in which:
        fmod_golden(L, np.random.default_rng(1212))
    for f in ("arith.npz", "transc.npz", "matmul.npz", "cmatmul.npz", "transc2.npz", "special2.npz", "transc3.npz",
              "transc4.npz", "transc5.npz", "binary2.npz", "invtrig.npz",
              "invtrig3.npz", "c0_256.json", "fmod.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main(sys.argv[1:])
