"""Input sets and exact values that pin the functions of this project's own
design (erf, erfc, erfcx, atan, asin, acos; the reference fits these with
rationals of its own) against 256-bit MPFR values (oracle/libmpfrpin.so).

Each set holds at least 100,000 double-double inputs (seeded, so the CPU and
GPU tests draw the same ones): points spread over every Taylor row or table
entry of the design, both sides of every row edge and switch point, tiny and
huge arguments, both signs. exact(nm) evaluates them in a thread pool (the
MPFR calls release the GIL) and caches the result for the session.
"""
import ctypes
import functools
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from multifloats_amd import datagen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MPFR_LIB = os.path.join(ROOT, "oracle", "libmpfrpin.so")
FUNCS = ("erf", "erfc", "erfcx", "atan", "asin", "acos")
CODE = {nm: i for i, nm in enumerate(FUNCS)}
# Largest relative error allowed on these sets where the exact value is a
# normal double-double. erf: the reference's published bar (max_rel 1.55e-32
# vs MPFR, doc/dev/benchmark/baseline/benchmark-raptor-lake.json:1636); the
# others: this build's measured maxima (erfc 5.84e-32, erfcx 5.00e-32, atan
# 1.04e-32, asin 6.2e-33, acos 7.6e-33) with a small margin. The compiled
# reference on the same points: erf 1.95e-32, erfc 8.48e-32, erfcx 8.04e-32,
# atan 3.48e-32, asin 3.26e-32, acos 4.28e-32.
CAPS = {"erf": 1.55e-32, "erfc": 6.0e-32, "erfcx": 5.5e-32, "atan": 1.2e-32, "asin": 8e-33, "acos": 9e-33}


def _w(rng, hi):
    return datagen._with_lo(rng, np.asarray(hi, dtype=np.float64))


def _around(rng, edges, per=12):
    """Points on both sides of each edge: e (1 +- 2^-j) for j in 8..52 (a
    random pick of `per`), and the edge itself."""
    edges = np.asarray(edges, dtype=np.float64)
    j = rng.integers(8, 53, size=(edges.size, per))
    s = rng.choice([-1.0, 1.0], size=(edges.size, per))
    pts = edges[:, None] * (1.0 + s * np.exp2(-j.astype(np.float64)))
    return np.concatenate([edges, pts.ravel()])


def _uniform_rows(rng, lo, hi, rows, per):
    """`per` uniform points in each of `rows` equal rows of [lo, hi)."""
    w = (hi - lo) / rows
    base = lo + w * np.arange(rows)
    return (base[:, None] + w * rng.random((rows, per))).ravel()


@functools.lru_cache(maxsize=None)
def points(nm):
    """(n, 2) double-double inputs of function nm."""
    rng = np.random.default_rng(1000 + CODE[nm])
    sgn = lambda x: x * rng.choice([-1.0, 1.0], size=np.shape(x))
    if nm == "erf":
        # row 0 = [0, 1/4) (x + x Q(x^2)), rows of 1/16 from 1/4 to 0.25 + 141/16
        edges = 0.25 + np.arange(142) / 16.0
        hi = np.concatenate([
            _uniform_rows(rng, 0.25, 0.25 + 141 / 16.0, 141, 600),
            rng.uniform(0.0, 0.25, 7000), 10.0 ** rng.uniform(-300, -0.6, 5000),
            _around(rng, edges), 10.0 ** rng.uniform(1, 300, 2000)])
        hi = sgn(hi)
    elif nm == "erfc":
        # 1 - erf below 1/4, Taylor rows of erfc on [1/4, 4), exp(-x^2) erfcx
        # beyond 4, 0 beyond 27.3, 2 - erfc(-x) below 0
        edges = np.concatenate([0.25 + np.arange(61) / 16.0, 4.0 * np.exp2(np.arange(97) / 32.0), [27.3]])
        hi = np.concatenate([
            _uniform_rows(rng, 0.25, 4.0, 60, 700), rng.uniform(4.0, 27.3, 34000),
            rng.uniform(0.0, 0.25, 4000), 10.0 ** rng.uniform(-300, -0.6, 3000), rng.uniform(27.3, 30.0, 1000),
            -rng.uniform(0.0, 10.0, 12000), -(10.0 ** rng.uniform(-300, 2.5, 2000)),
            _around(rng, edges)])
    elif nm == "erfcx":
        # rows of 1/16 on [-2, 4), 32 rows per binade on [4, 32), the
        # asymptotic series beyond 32, 2 exp(x^2) - erfcx(-x) below -2
        edges = np.concatenate([-2.0 + np.arange(97) / 16.0, 4.0 * np.exp2(np.arange(97) / 32.0)])
        hi = np.concatenate([
            _uniform_rows(rng, -2.0, 4.0, 96, 450), 4.0 * np.exp2(_uniform_rows(rng, 0.0, 3.0, 96, 300)),
            10.0 ** rng.uniform(np.log10(32.0), 300, 13000), -rng.uniform(2.0, 26.5, 10000),
            sgn(10.0 ** rng.uniform(-300, -1, 3000)), _around(rng, edges)])
    elif nm == "atan":
        # series below 1/8, table atan(k/64) on [1/8, 8], pi/2 - atan(1/x) above 8
        ks = np.arange(8, 513) / 64.0
        near = (ks[:, None] + rng.uniform(-1 / 128.0, 1 / 128.0, (ks.size, 80))).ravel()
        hi = np.concatenate([
            near, 10.0 ** rng.uniform(-300, np.log10(0.125), 23000), 10.0 ** rng.uniform(np.log10(8.0), 300, 23000),
            rng.uniform(0.0, 8.0, 12000), _around(rng, np.concatenate([[0.125, 8.0], ks]), 6)])
        hi = sgn(hi)
    else:  # asin, acos: atan of x / sqrt(1 - x^2) (or its reciprocal) on the same table
        t = np.arange(8, 513) / 64.0
        xs = t / np.sqrt(1.0 + t * t)
        near = (xs[:, None] * (1.0 + rng.uniform(-2e-3, 2e-3, (xs.size, 50)))).ravel()
        near = near[near < 1.0]
        hi = np.concatenate([
            rng.uniform(0.0, 1.0, 42000), 1.0 - 10.0 ** rng.uniform(-16, -1, 21000), near,
            10.0 ** rng.uniform(-300, -2, 10000), _around(rng, np.concatenate([[0.5, 1.0 - 2.0 ** -53], xs]), 6)])
        hi = sgn(np.minimum(hi, 1.0))
    x = _w(rng, hi)
    if nm in ("asin", "acos"):
        # keep |x| <= 1 as a double-double (drop a lo limb that would leave it)
        over = np.abs(x[:, 0]) == 1.0
        x[over, 1] = 0.0
    assert x.shape[0] >= 100000, (nm, x.shape)
    return x


@functools.lru_cache(maxsize=None)
def exact(nm):
    """(n, 3) limbs of the 256-bit MPFR value of nm at points(nm)."""
    L = ctypes.CDLL(MPFR_LIB)
    L.pin_eval.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    x = np.ascontiguousarray(points(nm))
    out = np.zeros((x.shape[0], 3))
    cuts = np.linspace(0, x.shape[0], min(16, os.cpu_count() or 1) * 4 + 1).astype(int)

    def part(a, b):
        a, b = int(a), int(b)
        if b > a:
            L.pin_eval(CODE[nm], ctypes.c_void_p(x.ctypes.data + 16 * a), ctypes.c_void_p(out.ctypes.data + 24 * a),
                       b - a)
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        list(ex.map(lambda ab: part(*ab), zip(cuts[:-1], cuts[1:])))
    return out


def checked(nm):
    """Mask of the points whose exact value is a finite normal double-double
    (|v| > 2^-960): where a relative error is meaningful."""
    v = exact(nm)
    return np.isfinite(v[:, 0]) & (np.abs(v[:, 0]) > 2.0 ** -960)


def reference_outputs(nm):
    """The compiled reference's <nm>dd at points(nm) (through the array loop of
    oracle/_ref/librefloop.so), or None without that build."""
    from tests import _native as nt
    loops = os.path.join(os.path.dirname(nt.REF), "librefloop.so")
    if not os.path.exists(loops):
        return None
    R = ctypes.CDLL(loops)
    R.refloop_unary.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    x = np.ascontiguousarray(points(nm))
    out = np.empty_like(x)
    code = {"erf": 4, "erfc": 16, "erfcx": 17, "atan": 29, "asin": 31, "acos": 32}[nm]
    assert R.refloop_unary(code, x.ctypes.data, out.ctypes.data, x.shape[0]) == 0
    return out
