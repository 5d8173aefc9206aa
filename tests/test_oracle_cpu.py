"""CPU tests pinning the oracle: golden vectors from the compiled reference,
closed forms, invariants, brute force, and planted mistakes."""
import ctypes
from fractions import Fraction

import mpmath
import numpy as np
import pytest

from multifloats_amd import datagen
from tests import _native as nt


def test_oracle_arith_matches_reference_bits():
    g = nt.load_golden("arith.npz")
    for nm, ref in (("add", "adddd"), ("sub", "subdd"), ("mul", "muldd"), ("div", "divdd")):
        nt.assert_bits(nt.or_binary(nm, g["a"], g["b"]), g[ref], nm)
    nt.assert_bits(nt.or_unary("sqrt", g["sqrt_in"]), g["sqrtdd"], "sqrt")


@pytest.mark.parametrize("fn", ["exp", "log", "sin"])
def test_oracle_transcendentals_match_reference_bits(fn):
    g = nt.load_golden("transc.npz")
    nt.assert_bits(nt.or_unary(fn, g[fn + "_in"]), g[fn + "_ref"], fn)


@pytest.mark.parametrize("fn", nt.TRANSC2)
def test_oracle_transcendentals2_match_reference_bits(fn):
    g = nt.load_golden("transc2.npz")
    nt.assert_bits(nt.or_unary(fn, g[fn + "_in"]), g[fn + "_ref"], fn)


@pytest.mark.parametrize("fn", nt.TRANSC3)
def test_oracle_inverse_hyperbolics_match_reference_bits(fn):
    """asinh / acosh / atanh restated from inv_trig.inc:226-291 against the
    compiled reference on every branch (Taylor below 0.01, log1p and log
    forms, the 1e150 cut, domain edges, specials)."""
    g = nt.load_golden("transc3.npz")
    nt.assert_bits(nt.or_unary(fn, g[fn + "_in"]), g[fn + "_ref"], fn)


def test_inverse_hyperbolics_are_accurate_where_the_reference_is():
    """A check of the fixture itself against 320-bit values: asinh holds DD
    accuracy everywhere, and acosh / atanh wherever their argument is away
    from the points where the reference's formulas cancel."""
    g = nt.load_golden("transc3.npz")
    for fn, cap in (("asinh", 1e-31), ("acosh", 1e-16), ("atanh", 1e-17)):
        ex = g[fn + "_exact"]
        fin = np.isfinite(ex[:, 0]) & (np.abs(ex[:, 0]) > 1e-290)
        assert fin.sum() > 1000
        assert nt.rel_err_vs_exact(g[fn + "_ref"][fin], ex[fin]).max() < cap, fn


@pytest.mark.parametrize("fn", nt.TRANSC4)
def test_oracle_exp10_and_c23_forms_match_reference_bits(fn):
    """exp10 (exp_log.inc:364-397 with exp2_from_triple 86-106) and the C23
    forms exp2m1, exp10m1, log2p1, log10p1 (479-519) against the compiled
    reference: both sides of the |x| = 1/2 switch, exp10's clamps, the log1p
    regimes, specials."""
    g = nt.load_golden("transc4.npz")
    nt.assert_bits(nt.or_unary(fn, g[fn + "_in"]), g[fn + "_ref"], fn)


@pytest.mark.parametrize("fn", nt.TRANSC4)
def test_oracle_exp10_and_c23_forms_are_dd_accurate(fn):
    """The oracle itself against 320-bit values: every normal result within
    2e-31 relative. exp10 without its third log2(10) limb, or with the
    triple's last limb dropped, is off by up to ~1e-30 at |x| ~ 300; a
    wrong constant or branch is off by far more."""
    g = nt.load_golden("transc4.npz")
    x, ex = g[fn + "_in"], g[fn + "_exact"]
    fin = np.isfinite(ex[:, 0]) & (np.abs(ex[:, 0]) > 1e-290) & (np.abs(ex[:, 0]) < 1e300)
    assert fin.sum() > 1500
    assert nt.rel_err_vs_exact(nt.or_unary(fn, x)[fin], ex[fin]).max() < 2e-31, fn


@pytest.mark.parametrize("fn", nt.TRANSC5)
def test_oracle_pi_scaled_trig_matches_reference_bits(fn):
    """sinpi / cospi / tanpi (trig.inc:41-104) against the compiled reference:
    every quadrant, exact quarters (zeros, poles, +-1) and their neighbours,
    the 2^52 cap, specials."""
    g = nt.load_golden("transc5.npz")
    nt.assert_bits(nt.or_unary(fn, g[fn + "_in"]), g[fn + "_ref"], fn)


@pytest.mark.parametrize("fn", nt.TRANSC5)
def test_oracle_pi_scaled_trig_is_dd_accurate(fn):
    """The oracle against 320-bit values: every normal result within 1e-31
    relative. pi as one double, a quadrant off by one or a sign slip fails."""
    g = nt.load_golden("transc5.npz")
    x, ex = g[fn + "_in"], g[fn + "_exact"]
    fin = np.isfinite(ex[:, 0]) & (np.abs(ex[:, 0]) > 1e-290) & (np.abs(ex[:, 0]) < 1e300)
    assert fin.sum() > 2000
    assert nt.rel_err_vs_exact(nt.or_unary(fn, x)[fin], ex[fin]).max() < 1e-31, fn


@pytest.mark.parametrize("fn", nt.BINARY2)
def test_oracle_binary_functions_match_reference_bits(fn):
    """pow (exp_log.inc:531-600 with powi), hypot, fdim, fmin, fmax and
    copysign (float64x2.h) against the compiled reference: every pair of
    special values and per-branch draws."""
    g = nt.load_golden("binary2.npz")
    nt.assert_bits(nt.or_binary(fn, g[fn + "_a"], g[fn + "_b"]), g[fn + "_ref"], fn)


def test_oracle_pow_and_hypot_accuracy():
    """Against 320-bit values: hypot within 1e-31 relative; pow within
    1e-29 (1 + |y ln x|) relative (its condition number is |y ln x|), on every
    normal result the reference defines (it returns NaN for x = -1 and an
    even |y| >= 2^31, outside its integer path)."""
    g = nt.load_golden("binary2.npz")
    a, b, ex = g["hypot_a"], g["hypot_b"], g["hypot_exact"]
    fin = np.isfinite(ex[:, 0]) & (np.abs(ex[:, 0]) > 1e-290) & (np.abs(ex[:, 0]) < 1e300)
    assert fin.sum() > 3000
    assert nt.rel_err_vs_exact(nt.or_binary("hypot", a, b)[fin], ex[fin]).max() < 1e-31
    a, b, ex = g["pow_a"], g["pow_b"], g["pow_exact"]
    r = nt.or_binary("pow", a, b)
    fin = (np.isfinite(ex[:, 0]) & (np.abs(ex[:, 0]) > 1e-290) & (np.abs(ex[:, 0]) < 1e300) & np.isfinite(r[:, 0])
           & (a[:, 0] != 0))
    assert fin.sum() > 2000
    z = np.abs(b[fin, 0] * np.log(np.abs(a[fin, 0])))
    assert (nt.rel_err_vs_exact(r[fin], ex[fin]) <= 1e-29 * (1 + z)).all()


def test_oracle_order_functions_properties():
    """fmin / fmax return one of their operands bit for bit, fmin <= fmax in
    the DD order, copysign keeps the magnitude of x, fdim is x - y or +0."""
    g = nt.load_golden("binary2.npz")
    a, b = g["fmin_a"], g["fmin_b"]
    ok = ~np.isnan(a[:, 0]) & ~np.isnan(b[:, 0])
    lo, hi = nt.or_binary("fmin", a, b), nt.or_binary("fmax", a, b)
    bits = lambda v: np.ascontiguousarray(v).view(np.uint64)
    for r in (lo, hi):
        assert (np.all(bits(r) == bits(a), axis=1) | np.all(bits(r) == bits(b), axis=1)).all()
    assert ((lo[ok, 0] < hi[ok, 0]) | ((lo[ok, 0] == hi[ok, 0]) & (lo[ok, 1] <= hi[ok, 1]))).all()
    cs = nt.or_binary("copysign", a, b)
    assert np.all(np.abs(cs[ok]) == np.abs(a[ok]))
    fd = nt.or_binary("fdim", a[ok], b[ok])
    diff = nt.or_binary("sub", a[ok], b[ok])
    assert (np.all(bits(fd) == bits(diff), axis=1) | np.all(fd == 0.0, axis=1)).all()


def test_oracle_matmul_matches_reference_bits():
    g = nt.load_golden("matmul.npz")
    O = nt.oracle()
    for i in range(int(g["count"][0])):
        key = "c%03d" % i
        m, k, n, ri = (int(v) for v in g[key + "_shape"])
        A = datagen.unflat(g[key + "_A"], m, k)
        B = datagen.unflat(g[key + "_B"], k, n)
        C = nt.or_matmul(A, B, ri)
        nt.assert_bits(datagen.flat(C), g[key + "_C"], "mm %s" % key)
        x = nt.dd(g[key + "_x"])
        y = np.zeros((m, 2))
        O.or_matmul_mv(nt.mptr(nt.cm(A)), nt.ptr(x), nt.ptr(y), m, k, ri)
        nt.assert_bits(y, g[key + "_mv"], "mv %s" % key)
        z = np.zeros((n, 2))
        O.or_matmul_vm(nt.ptr(x), nt.mptr(nt.cm(B)), nt.ptr(z), k, n, ri)
        nt.assert_bits(z, g[key + "_vm"], "vm %s" % key)


def test_tables_are_the_closed_forms():
    """Every DD table entry equals its closed form to < 2^-105 relative."""
    import importlib.util, os, re
    mpmath.mp.prec = 300
    path = os.path.join(nt.ROOT, "multifloats_amd", "csrc", "dd_tables.h")
    text = open(path).read()

    def arr(name):
        m = re.search(r"double %s\[\d+\] = \{(.*?)\};" % name, text, re.S)
        return [float.fromhex(t.strip()) if "0x" in t else float(t) for t in m.group(1).split(",") if t.strip()]

    ln2 = mpmath.log(2)
    checks = {
        "mf_exp2c": [ln2 ** j / mpmath.factorial(j) for j in range(10)],
        "mf_exp2t": [mpmath.mpf(2) ** (mpmath.mpf(k - 128) / 256) for k in range(257)],
        "mf_log2c": [2 / ((2 * k + 1) * ln2) for k in range(11)],
        "mf_log2v": [mpmath.log(1 + mpmath.mpf(2 * i + 1) / 64) / ln2 for i in range(32)],
        "mf_sinc": [(-1) ** k / mpmath.factorial(2 * k + 1) for k in range(13)],
        "mf_cosc": [(-1) ** k / mpmath.factorial(2 * k) for k in range(13)],
    }
    for name, vals in checks.items():
        hi, lo = arr(name + "_hi"), arr(name + "_lo")
        assert len(hi) == len(vals)
        for h, l, v in zip(hi, lo, vals):
            if v == 0:
                continue
            assert abs((mpmath.mpf(h) + mpmath.mpf(l) - v) / v) < mpmath.mpf(2) ** -105, name
            assert abs(l) <= abs(h) * 2.0 ** -52, name



def test_crt_form_tables():
    """The Ozaki CRT-form tables, recomputed with exact integers: pairwise
    coprime moduli <= 256 whose product covers the largest need
    (2 k 2^(16 S - 2) at k = 32768, S = 16); 1/m as five 36-bit limbs of
    floor(2^180 / m); the CRT inverses y = (M_N / m_t)^-1 mod m_t and M_N as
    a double-double for every N."""
    import math, os, re
    text = open(os.path.join(nt.ROOT, "multifloats_amd", "csrc", "dd_tables.h")).read()

    def vals(decl, conv):
        body = re.search(decl + r"\[\d+\] = \{(.*?)\};", text, re.S).group(1)
        return [conv(t.strip()) for t in body.split(",") if t.strip()]

    fl = lambda t: float.fromhex(t) if "0x" in t else float(t)
    mods = vals(r"MF_TABLE int mf_oz2_m", int)
    assert vals(r"static constexpr int mf_oz2_m_host", int) == mods
    assert len(mods) == 40 and all(2 <= q <= 256 for q in mods)
    assert all(math.gcd(a, b) == 1 for i, a in enumerate(mods) for b in mods[:i])
    assert math.prod(mods) > 2 * 32768 * 2 ** (16 * 16 - 2)
    rl = vals(r"MF_TABLE double mf_oz2_rlimb", fl)
    for t, q in enumerate(mods):
        limbs = [int(v) for v in rl[5 * t:5 * t + 5]]
        assert all(0 <= v < 2 ** 36 for v in limbs)
        assert sum(v << (36 * (4 - j)) for j, v in enumerate(limbs)) == (1 << 180) // q
    ys = vals(r"MF_TABLE int mf_oz2_y", int)
    mp = vals(r"MF_TABLE double mf_oz2_mp", fl)
    for N in range(1, 41):
        MN = math.prod(mods[:N])
        for t in range(40):
            y = ys[40 * (N - 1) + t]
            if t < N:
                assert (y * (MN // mods[t])) % mods[t] == 1 % mods[t] and 0 <= y < mods[t]
            else:
                assert y == 0
        hi, lo = mp[2 * (N - 1)], mp[2 * N - 1]
        err = abs(Fraction(hi) + Fraction(lo) - MN)
        assert err <= Fraction(MN, 2 ** 106) and abs(lo) <= abs(hi) * 2.0 ** -53

def _frac(x):
    return Fraction(float(x[0])) + Fraction(float(x[1]))


def test_exact_accumulator_brute_force():
    rng = np.random.default_rng(1)
    for trial in range(20):
        k = int(rng.integers(1, 40))
        e = [None, 10, 60][trial % 3]
        a = datagen.spread(rng, k, e) if e else datagen.config0(rng, k)
        b = datagen.spread(rng, k, e) if e else datagen.config0(rng, k)
        exact = sum((_frac(a[i]) * _frac(b[i]) for i in range(k)), Fraction(0))
        c = nt.exact_dot(a, b)
        assert abs(_frac(c) - exact) <= abs(exact) * Fraction(2) ** -100 + Fraction(2) ** -1000
        # error of a deliberately perturbed value is measured exactly
        cp = c.copy()
        cp[1] += abs(c[0]) * 2.0 ** -90
        err, mag = nt.exact_dot_err(a, b, cp)
        true = float(abs(_frac(cp) - exact))
        assert abs(err - true) <= true * 1e-10 + 1e-300


def test_exact_sum_negative_and_cancelling():
    x = np.array([[1.0, 1e-20], [-1.0, 0.0], [1e-30, 0.0], [-3.5, 1e-17]])
    exact = sum((_frac(v) for v in x), Fraction(0))
    c = np.array([float(exact), 0.0])
    c[1] = float(exact - Fraction(c[0]))
    err, mag = nt.exact_sum_err(x, c)
    assert err <= abs(float(exact)) * 2.0 ** -100


def test_reductions_invariants():
    O = nt.oracle()
    # integer sums are exact
    x = np.stack([np.arange(1, 1001, dtype=np.float64), np.zeros(1000)], axis=-1)
    r = nt.res(O.or_sum(nt.ptr(x), 1000))
    assert r[0] == 500500.0 and r[1] == 0.0
    # norm2 of a Pythagorean pair
    for k in (-500, 0, 100):
        v = np.array([[3.0 * 2.0 ** k, 0.0], [4.0 * 2.0 ** k, 0.0]])
        r = nt.res(O.or_norm2(nt.ptr(v), 2))
        assert r[0] == 5.0 * 2.0 ** k and r[1] == 0.0
    # empty inputs
    e = np.zeros((1, 2))
    assert nt.res(O.or_sum(nt.ptr(e), 0)).tolist() == [0.0, 0.0]
    assert nt.res(O.or_dot(nt.ptr(e), nt.ptr(e), 0, 8)).tolist() == [0.0, 0.0]


@pytest.mark.parametrize("ri", [0, 1, 8, 64])
def test_oracle_dot_sum_within_bound(ri):
    rng = np.random.default_rng(7 + ri)
    n = 20000
    a, b = datagen.cancel_pairs(rng, n)
    O = nt.oracle()
    r = nt.res(O.or_dot(nt.ptr(nt.dd(a)), nt.ptr(nt.dd(b)), n, ri))
    err, mag = nt.exact_dot_err(a, b, r)
    assert err <= 4 * n * 2.0 ** -104 * mag
    x = datagen.sum_inputs(rng, n)
    s = nt.res(O.or_sum(nt.ptr(nt.dd(x)), n))
    err, mag = nt.exact_sum_err(x, s)
    assert err <= 4 * n * 2.0 ** -104 * mag


def test_planted_mistake_is_caught():
    """A dot product that drops the cross terms (a plausible slip) must fail
    the bound the GPU reductions are held to."""
    rng = np.random.default_rng(3)
    n = 5000
    a, b = datagen.cancel_pairs(rng, n)
    sh = sl = 0.0
    for i in range(n):  # the oracle's update with the cross term dropped
        p = a[i, 0] * b[i, 0]
        e = float(Fraction(a[i, 0]) * Fraction(b[i, 0]) - Fraction(p))
        t = sh + p
        bp = t - sh
        ap = t - bp
        sl = sl + ((e + (sh - ap)) + (p - bp))
        sh = t
    bad = np.array([sh + sl, sl - ((sh + sl) - sh)])
    err, mag = nt.exact_dot_err(a, b, bad)
    assert err > 4 * n * 2.0 ** -104 * mag


def test_oracle_gemm_closed_forms():
    rng = np.random.default_rng(5)
    m = 12
    Id = np.zeros((m, m, 2))
    for i in range(m):
        Id[i, i, 0] = 1.0
    B = datagen.matrix(rng, m, 9)
    C = nt.or_matmul(Id, B, 8)
    nt.assert_bits(C, B, "I*B")
    Ai = np.zeros((6, 5, 2))
    Bi = np.zeros((5, 4, 2))
    Ai[..., 0] = rng.integers(-9, 10, (6, 5))
    Bi[..., 0] = rng.integers(-9, 10, (5, 4))
    C = nt.or_matmul(Ai, Bi, 0)
    assert np.array_equal(C[..., 0], Ai[..., 0] @ Bi[..., 0]) and not C[..., 1].any()


def test_oracle_complex_matmul_matches_reference_composition():
    """The oracle's complex GEMM equals the reference's four-matmuldd_mm +
    subdd/adddd composition (fixtures from the compiled reference)."""
    g = nt.load_golden("cmatmul.npz")
    for i in range(int(g["count"][0])):
        key = "z%03d" % i
        m, k, n, ri = (int(v) for v in g[key + "_shape"])
        A = np.ascontiguousarray(g[key + "_A"])  # (m*k, 4) column-major complex
        B = np.ascontiguousarray(g[key + "_B"])
        C = np.zeros((m * n, 4))
        nt.oracle().or_cmatmul_mm(nt.ptr(A), nt.ptr(B), nt.ptr(C), m, k, n, ri)
        nt.assert_bits(C, g[key + "_C"], "cmm " + key)


def _colsum_fast2sum(col):
    """dd_sum_1d on one column (the oracle's rank-1 sum)."""
    return nt.res(nt.oracle().or_sum(nt.ptr(nt.dd(col)), col.shape[0]))


def test_oracle_sum_dim_matches_rank1_chains():
    """Each output of sum(array, dim) is the rank-1 chain over that fibre; the
    closing TwoSum equals the rank-1 Fast2Sum whenever |s_hi| >= |s_lo|, which
    holds for these inputs, so the bits must agree fibre by fibre."""
    rng = np.random.default_rng(21)
    shape = (5, 7, 33)
    x = datagen.config0(rng, int(np.prod(shape))).reshape(shape[::-1] + (2,)).transpose(2, 1, 0, 3)
    buf = np.ascontiguousarray(x.transpose(2, 1, 0, 3)).reshape(-1, 2)  # Fortran order
    for dim in (1, 2, 3):
        inner = int(np.prod(shape[:dim - 1]))
        ln = shape[dim - 1]
        outer = int(np.prod(shape[dim:]))
        got = nt.or_reduce_dim("sum", buf, None, inner, ln, outer)
        v = buf.reshape(outer, ln, inner, 2)
        for o in range(outer):
            for i in range(inner):
                nt.assert_bits(got[i + inner * o], _colsum_fast2sum(np.ascontiguousarray(v[o, :, i])),
                               "sum dim=%d fibre (%d,%d)" % (dim, i, o))
        got2 = nt.or_reduce_dim("norm2", buf, None, inner, ln, outer)
        for o in range(0, outer, 3):
            for i in range(inner):
                col = np.ascontiguousarray(v[o, :, i])
                nt.assert_bits(got2[i + inner * o], nt.res(nt.oracle().or_norm2(nt.ptr(col), ln)),
                               "norm2 dim=%d" % dim)


def test_oracle_sum_mask_is_compaction():
    rng = np.random.default_rng(22)
    n = 4000
    x = datagen.sum_inputs(rng, n)
    mask = rng.random(n) < 0.3
    nt.assert_bits(nt.or_sum_mask(x, mask), _colsum_fast2sum(np.ascontiguousarray(x[mask])), "mask")
    nt.assert_bits(nt.or_sum_mask(x, np.ones(n, bool)), _colsum_fast2sum(x), "all true")
    assert nt.or_sum_mask(x, np.zeros(n, bool)).tolist() == [0.0, 0.0]
    # masked dim sum skips the same elements
    buf = x.reshape(n, 2)
    got = nt.or_reduce_dim("sum", buf, mask.astype(np.int32), 1, n, 1)
    nt.assert_bits(got[0], _colsum_fast2sum(np.ascontiguousarray(x[mask])), "masked dim sum")


def test_oracle_dim_closed_forms_and_planted_transpose():
    # integer columns sum exactly; Pythagorean rows give exact norms
    m, n = 6, 50
    A = np.zeros((m * n, 2))
    A[:, 0] = np.arange(m * n)  # A(i, j) = i + m*j, column-major
    s1 = nt.or_reduce_dim("sum", A, None, 1, m, n)   # dim=1: column sums
    want1 = np.array([sum(i + m * j for i in range(m)) for j in range(n)], dtype=np.float64)
    assert (s1[:, 0] == want1).all() and (s1[:, 1] == 0).all()
    s2 = nt.or_reduce_dim("sum", A, None, m, n, 1)   # dim=2: row sums
    want2 = np.array([sum(i + m * j for j in range(n)) for i in range(m)], dtype=np.float64)
    assert (s2[:, 0] == want2).all()
    # a plausible slip -- swapping inner and outer -- changes the answer
    bad = nt.or_reduce_dim("sum", A, None, n, m, 1)
    assert bad.shape[0] != s2.shape[0] or not (bad[:, 0] == want2).all()
    P = np.array([[3.0, 0.0], [4.0, 0.0], [6.0, 0.0], [8.0, 0.0]])  # columns (3,4), (6,8)
    r = nt.or_reduce_dim("norm2", P, None, 1, 2, 2)
    assert r.tolist() == [[5.0, 0.0], [10.0, 0.0]]
    # empty fibres: zero
    assert nt.or_reduce_dim("sum", np.zeros((1, 2)), None, 3, 0, 2).tolist() == [[0.0, 0.0]] * 6


def test_oracle_complex_dot():
    rng = np.random.default_rng(23)
    n = 3000
    a = np.concatenate([datagen.config0(rng, n), datagen.config0(rng, n)], axis=1)
    b = np.concatenate([datagen.config0(rng, n), datagen.config0(rng, n)], axis=1)
    O = nt.oracle()
    r = nt.or_cdot(a, b, 8)
    # real and imaginary parts against the exact accumulator:
    # re = ar.br + ai.bi, im = ar.bi - ai.br
    ar, ai, br, bi = a[:, :2], a[:, 2:], b[:, :2], b[:, 2:]
    e, m = nt.exact_dot_err(np.concatenate([ar, ai]), np.concatenate([br, bi]), r[:2])
    assert e <= 4 * n * 2.0 ** -104 * m
    e, m = nt.exact_dot_err(np.concatenate([ar, -ai]), np.concatenate([bi, br]), r[2:])
    assert e <= 4 * n * 2.0 ** -104 * m
    # conj(a).a is real: the two cross chains are the same chain commuted
    r2 = nt.or_cdot(a, a, 8)
    assert r2[2] == 0.0 and r2[3] == 0.0 and r2[0] > 0
    # imaginary parts zero: the real dot, exactly
    z = a.copy()
    z[:, 2:] = 0
    w = b.copy()
    w[:, 2:] = 0
    r3 = nt.or_cdot(z, w, 8)
    nt.assert_bits(r3[:2], nt.res(O.or_dot(nt.ptr(nt.dd(ar)), nt.ptr(nt.dd(br)), n, 8)), "real case")
    assert r3[2] == 0.0
    # a slip that drops one of the four chains breaks the bound on the imaginary part
    slip_im = nt.res(O.or_dot(nt.ptr(nt.dd(ar)), nt.ptr(nt.dd(bi)), n, 8))
    e, m = nt.exact_dot_err(np.concatenate([ar, -ai]), np.concatenate([bi, br]), slip_im)
    assert e > 4 * n * 2.0 ** -104 * m


def _lex_pick(a, idx, best):
    """First index of the lexicographic (hi, lo) max / min among idx (Python's
    max/min return the first extremal element)."""
    if not idx:
        return None
    return best(idx, key=lambda i: (a[i, 0], a[i, 1]))


def test_oracle_maxval_minval_are_lexicographic_extrema():
    rng = np.random.default_rng(31)
    for trial in range(20):
        n = int(rng.integers(0, 300))
        a = datagen.config0(rng, n)
        if trial % 2:  # many ties, signed zeros in both limbs
            a = rng.choice(np.array([[1.0, 0.0], [1.0, 1e-17], [-0.0, 0.0], [0.0, -0.0], [0.0, 0.0], [2.0, -1e-16]]),
                           size=n)
        mask = rng.random(n) < 0.6
        for kind, best in (("maxval", max), ("minval", min)):
            for m in (None, mask):
                idx = [i for i in range(n) if m is None or m[i]]
                k = _lex_pick(a, idx, best)
                want = np.zeros(2) if k is None else a[k]
                nt.assert_bits(nt.or_fibre(kind, a, m), want, kind)
    # +0 / -0 tie: the first one stays
    z = np.array([[-0.0, 0.0], [0.0, 0.0]])
    assert np.signbit(nt.or_fibre("maxval", z)[0]) and np.signbit(nt.or_fibre("minval", z)[0])


def test_oracle_extrema_nan_rules():
    """A NaN hi limb first: the result is that element; anywhere else it is
    never selected (removing it changes nothing)."""
    rng = np.random.default_rng(32)
    a = datagen.config0(rng, 50)
    b = a.copy()
    b[17] = (np.nan, 0.0)
    b[33] = (np.nan, 1.0)
    keep = [i for i in range(50) if i not in (17, 33)]
    for kind in ("maxval", "minval"):
        nt.assert_bits(nt.or_fibre(kind, b), nt.or_fibre(kind, a[keep]), kind)
        c = b.copy()
        c[0] = (np.nan, 5.0)
        r = nt.or_fibre(kind, c)
        assert np.isnan(r[0]) and r[1] == 5.0


def test_oracle_product_within_bound_of_exact():
    rng = np.random.default_rng(33)
    for n in (1, 2, 7, 40):
        a = datagen.config0(rng, n) + np.array([1.5, 0.0])  # away from 0, no cancellation
        exact = Fraction(1)
        for x in a:
            exact *= Fraction(float(x[0])) + Fraction(float(x[1]))
        r = nt.or_fibre("product", a)
        got = Fraction(float(r[0])) + Fraction(float(r[1]))
        assert abs(got - exact) <= 4 * n * Fraction(2) ** -104 * abs(exact)
    p2 = np.array([[2.0 ** k, 0.0] for k in (-3, 5, 11, -7)])
    nt.assert_bits(nt.or_fibre("product", p2), np.array([2.0 ** 6, 0.0]), "powers of two")
    assert (nt.or_fibre("product", np.zeros((0, 2))) == np.array([1.0, 0.0])).all()


def test_oracle_fibre_dim_matches_rank1():
    rng = np.random.default_rng(34)
    inner, length, outer = 3, 11, 4
    buf = datagen.config0(rng, inner * length * outer)
    mask = (rng.random(inner * length * outer) < 0.5).astype(np.int32)
    for kind in ("product", "maxval", "minval"):
        for m in (None, mask):
            out = nt.or_fibre_dim(kind, buf, m, inner, length, outer)
            for o in range(outer):
                for i in range(inner):
                    sel = [i + inner * (j + length * o) for j in range(length)]
                    want = nt.or_fibre(kind, buf[sel], None if m is None else m[sel])
                    nt.assert_bits(out[i + inner * o], want, kind)
