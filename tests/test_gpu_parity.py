"""GPU parity: the HIP path against the compiled reference's golden vectors and
the oracle, through the C ABI (host-pointer entry points and the device API)."""
import ctypes

import numpy as np
import pytest

import multifloats_amd as mf
from multifloats_amd import datagen
from tests import _native as nt

pytestmark = pytest.mark.gpu

# max relative error of the reference's erfdd against a 200-bit oracle
# (doc/dev/benchmark/baseline/benchmark-raptor-lake.json:1636)
ERF_BOUND = 1.55e-32


@pytest.fixture(scope="module")
def torch_cuda(gpu):
    import torch
    torch.cuda.init()
    return torch


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_arith_bits_golden(gpu):
    g = nt.load_golden("arith.npz")
    for nm, ref in (("add", "adddd"), ("sub", "subdd"), ("mul", "muldd"), ("div", "divdd")):
        nt.assert_bits(getattr(mf, nm)(g["a"], g["b"]), g[ref], nm)
    nt.assert_bits(mf.sqrt(g["sqrt_in"]), g["sqrtdd"], "sqrt")


def test_arith_bits_random_vs_oracle(gpu):
    rng = np.random.default_rng(11)
    n = 1 << 20
    a = np.concatenate([datagen.config0(rng, n // 2), datagen.spread(rng, n // 2, 200)])
    b = np.concatenate([datagen.spread(rng, n // 2, 200), datagen.config0(rng, n // 2)])
    for nm in ("add", "sub", "mul", "div"):
        nt.assert_bits(getattr(mf, nm)(a, b), nt.or_binary(nm, a, b), nm)
    nt.assert_bits(mf.sqrt(np.abs(a)), nt.or_unary("sqrt", np.abs(a)), "sqrt")


@pytest.mark.parametrize("fn", ["exp", "log", "sin"])
def test_transcendental_bits_golden(gpu, fn):
    g = nt.load_golden("transc.npz")
    nt.assert_bits(getattr(mf, fn)(g[fn + "_in"]), g[fn + "_ref"], fn)


@pytest.mark.parametrize("fn", ["exp", "log", "sin"])
def test_transcendental_bits_vs_oracle(gpu, fn):
    rng = np.random.default_rng(hash(fn) % 1000)
    n = 200000
    x = {"exp": datagen.exp_inputs, "log": datagen.log_inputs, "sin": datagen.sin_inputs}[fn](rng, n)
    nt.assert_bits(getattr(mf, fn)(x), nt.or_unary(fn, x), fn)


def _transc2_inputs(rng, fn, n):
    w, U = datagen._with_lo, rng.uniform
    if fn in ("cos", "tan"):
        return datagen.sin_inputs(rng, n, near_frac=0.1)
    if fn == "exp2":
        return w(rng, U(-1100, 1100, n))
    if fn == "expm1":
        return np.concatenate([w(rng, U(-0.5, 0.5, n // 2)), w(rng, U(-745, 709, n - n // 2))])
    if fn in ("log2", "log10"):
        return np.concatenate([datagen.log_inputs(rng, n // 2), w(rng, U(15 / 16, 17 / 16, n - n // 2))])
    if fn == "log1p":
        return np.concatenate([w(rng, U(-0.25, 0.25, n // 2)), w(rng, U(-1.0, 10.0, n - n // 2))])
    return np.concatenate([w(rng, U(-0.5, 0.5, n // 2)), w(rng, U(-60, 60, n - n // 2))])


@pytest.mark.parametrize("fn", nt.TRANSC2)
def test_transcendental2_bits_golden(gpu, fn):
    g = nt.load_golden("transc2.npz")
    nt.assert_bits(getattr(mf, fn)(g[fn + "_in"]), g[fn + "_ref"], fn)


@pytest.mark.parametrize("fn", nt.TRANSC2)
def test_transcendental2_bits_vs_oracle(gpu, fn):
    rng = np.random.default_rng(sum(map(ord, fn)))
    x = _transc2_inputs(rng, fn, 200000)
    nt.assert_bits(getattr(mf, fn)(x), nt.or_unary(fn, x), fn)


@pytest.mark.parametrize("fn", nt.TRANSC3)
def test_inverse_hyperbolic_bits_golden(gpu, fn):
    g = nt.load_golden("transc3.npz")
    nt.assert_bits(getattr(mf, fn)(g[fn + "_in"]), g[fn + "_ref"], fn)


def _transc3_inputs(rng, fn, n):
    w, U = datagen._with_lo, rng.uniform
    q = n // 4
    if fn == "asinh":
        return np.concatenate([w(rng, U(-0.01, 0.01, q)), w(rng, U(-1, 1, q)), w(rng, U(-50, 50, q)),
                               w(rng, 10.0 ** U(-20, 300, n - 3 * q) * rng.choice([-1.0, 1.0], n - 3 * q))])
    if fn == "acosh":
        return np.concatenate([w(rng, 1.0 + 10.0 ** U(-16, 0, q)), w(rng, U(1, 50, q)),
                               w(rng, 10.0 ** U(1, 308, q)), w(rng, U(-2, 1, n - 3 * q))])
    return np.concatenate([w(rng, U(-0.01, 0.01, q)), w(rng, U(-1, 1, q)),
                           w(rng, (1.0 - 10.0 ** U(-16, -1, q)) * rng.choice([-1.0, 1.0], q)),
                           w(rng, U(-2, 2, n - 3 * q))])


@pytest.mark.parametrize("fn", nt.TRANSC3)
def test_inverse_hyperbolic_bits_vs_oracle(gpu, fn):
    rng = np.random.default_rng(sum(map(ord, fn)) + 3)
    x = _transc3_inputs(rng, fn, 200000)
    nt.assert_bits(getattr(mf, fn)(x), nt.or_unary(fn, x), fn)
    # the device entry point with the op code reaches the same kernel
    torch = pytest.importorskip("torch")
    xs = x[:5000]
    dx = to_dev(torch, xs)
    dy = torch.empty_like(dx)
    mf.dev_unary(getattr(mf, "MFHIP_" + fn.upper()), dx.data_ptr(), dy.data_ptr(), xs.shape[0],
                 torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    nt.assert_bits(dy.cpu().numpy(), nt.or_unary(fn, xs), "device " + fn)


@pytest.mark.parametrize("fn", nt.TRANSC4)
def test_exp10_and_c23_bits_golden(gpu, fn):
    g = nt.load_golden("transc4.npz")
    nt.assert_bits(getattr(mf, fn)(g[fn + "_in"]), g[fn + "_ref"], fn)


def _transc4_inputs(rng, fn, n):
    w, U = datagen._with_lo, rng.uniform
    q = n // 4
    if fn in ("log2p1", "log10p1"):
        return np.concatenate([w(rng, U(-0.25, 0.25, q)), w(rng, 10.0 ** U(-30, -1, q) * rng.choice([-1.0, 1.0], q)),
                               w(rng, U(-0.999, 10, q)), w(rng, 10.0 ** U(0, 300, n - 3 * q))])
    big = 1100.0 if fn == "exp2m1" else 330.0
    return np.concatenate([w(rng, U(-0.5, 0.5, q)), w(rng, 10.0 ** U(-30, -1, q) * rng.choice([-1.0, 1.0], q)),
                           w(rng, U(-big, big, q)), w(rng, U(-3, 3, n - 3 * q))])


@pytest.mark.parametrize("fn", nt.TRANSC4)
def test_exp10_and_c23_bits_vs_oracle(gpu, fn):
    rng = np.random.default_rng(sum(map(ord, fn)) + 4)
    x = _transc4_inputs(rng, fn, 200000)
    nt.assert_bits(getattr(mf, fn)(x), nt.or_unary(fn, x), fn)
    torch = pytest.importorskip("torch")
    xs = x[:5000]
    dx = to_dev(torch, xs)
    dy = torch.empty_like(dx)
    mf.dev_unary(getattr(mf, "MFHIP_" + fn.upper()), dx.data_ptr(), dy.data_ptr(), xs.shape[0],
                 torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    nt.assert_bits(dy.cpu().numpy(), nt.or_unary(fn, xs), "device " + fn)


@pytest.mark.parametrize("fn", nt.TRANSC5)
def test_pi_scaled_trig_bits_golden(gpu, fn):
    g = nt.load_golden("transc5.npz")
    nt.assert_bits(getattr(mf, fn)(g[fn + "_in"]), g[fn + "_ref"], fn)


@pytest.mark.parametrize("fn", nt.TRANSC5)
def test_pi_scaled_trig_bits_vs_oracle(gpu, fn):
    rng = np.random.default_rng(sum(map(ord, fn)) + 5)
    w, U, q = datagen._with_lo, rng.uniform, 50000
    x = np.concatenate([w(rng, U(-2, 2, q)), w(rng, U(-1e8, 1e8, q)),
                        w(rng, np.round(U(-100, 100, q) * 4) / 4 + 10.0 ** U(-17, -2, q) * rng.choice([-1.0, 1.0], q)),
                        w(rng, 10.0 ** U(-30, 16, q) * rng.choice([-1.0, 1.0], q))])
    nt.assert_bits(getattr(mf, fn)(x), nt.or_unary(fn, x), fn)
    torch = pytest.importorskip("torch")
    xs = x[::40]
    dx = to_dev(torch, xs)
    dy = torch.empty_like(dx)
    mf.dev_unary(getattr(mf, "MFHIP_" + fn.upper()), dx.data_ptr(), dy.data_ptr(), xs.shape[0],
                 torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    nt.assert_bits(dy.cpu().numpy(), nt.or_unary(fn, xs), "device " + fn)


@pytest.mark.parametrize("fn", nt.BINARY2)
def test_binary_functions_bits_golden(gpu, fn):
    g = nt.load_golden("binary2.npz")
    nt.assert_bits(getattr(mf, fn)(g[fn + "_a"], g[fn + "_b"]), g[fn + "_ref"], fn)


@pytest.mark.parametrize("fn", nt.BINARY2)
def test_binary_functions_bits_vs_oracle(gpu, fn):
    rng = np.random.default_rng(sum(map(ord, fn)) + 6)
    w, U, q = datagen._with_lo, rng.uniform, 50000
    if fn == "pow":
        a = np.concatenate([w(rng, 10.0 ** U(-4, 4, 2 * q)), w(rng, U(0.5, 2, q)), w(rng, -(10.0 ** U(-1, 1, q)))])
        b = np.concatenate([w(rng, U(-40, 40, 2 * q)), w(rng, U(-200, 200, q)),
                            np.round(U(-50, 50, q))[:, None] * np.array([1.0, 0.0])])
    elif fn == "hypot":
        a = w(rng, 2.0 ** U(-600, 600, 4 * q) * rng.choice([-1.0, 1.0], 4 * q))
        b = w(rng, 2.0 ** U(-600, 600, 4 * q) * rng.choice([-1.0, 1.0], 4 * q))
    else:
        a = w(rng, U(-4, 4, 4 * q))
        b = a.copy()
        b[:q, 1] *= U(-2, 2, q)
        b[q:] = w(rng, U(-4, 4, 3 * q))
    nt.assert_bits(getattr(mf, fn)(a, b), nt.or_binary(fn, a, b), fn)
    torch = pytest.importorskip("torch")
    xa, xb = a[::40], b[::40]
    da, db = to_dev(torch, xa), to_dev(torch, xb)
    dc = torch.empty_like(da)
    mf.dev_binary(getattr(mf, "MFHIP_" + fn.upper()), da.data_ptr(), db.data_ptr(), dc.data_ptr(), xa.shape[0],
                  torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    nt.assert_bits(dc.cpu().numpy(), nt.or_binary(fn, xa, xb), "device " + fn)


@pytest.mark.parametrize("nm", nt.INVTRIG + nt.INVTRIG3)
def test_atan_family_device_matches_host_build(gpu, nm):
    """The device kernels give the host build's bits (that build is checked
    against 320-bit values and the reference on the CPU)."""
    L = ctypes.CDLL(nt.HOSTMATH)
    rng = np.random.default_rng(sum(map(ord, nm)))
    w, U, q = datagen._with_lo, rng.uniform, 60000
    g = nt.invtrig_fixture(nm)
    if nm in nt.INVTRIG3:
        x = np.concatenate([g["x"], w(rng, U(-1, 1, q)), w(rng, (1 - 10.0 ** U(-16, 0, q)) * rng.choice([-1.0, 1.0], q))])
        want = np.empty_like(x)
        getattr(L, "h_" + nm)(nt.ptr(x), nt.ptr(want), ctypes.c_long(x.shape[0]))
        nt.assert_bits(getattr(mf, nm)(x), want, nm)
    elif nm in ("atan", "atanpi"):
        x = np.concatenate([g["x"], w(rng, U(-10, 10, q)), w(rng, 10.0 ** U(-20, 20, q) * rng.choice([-1.0, 1.0], q))])
        want = np.empty_like(x)
        getattr(L, "h_" + nm)(nt.ptr(x), nt.ptr(want), ctypes.c_long(x.shape[0]))
        nt.assert_bits(getattr(mf, nm)(x), want, nm)
    else:
        y = np.concatenate([g["y2"], w(rng, U(-10, 10, q)), w(rng, 10.0 ** U(-20, 20, q) * rng.choice([-1.0, 1.0], q))])
        x = np.concatenate([g["x2"], w(rng, U(-10, 10, q)), w(rng, 10.0 ** U(-20, 20, q) * rng.choice([-1.0, 1.0], q))])
        want = np.empty_like(x)
        getattr(L, "h_" + nm)(nt.ptr(y), nt.ptr(x), nt.ptr(want), ctypes.c_long(x.shape[0]))
        nt.assert_bits(getattr(mf, nm)(y, x), want, nm)


def test_fortran_min_max_dim_follow_the_module_bodies(gpu):
    """mindd_v / maxdd_v / dimdd_v against the reference module's bodies
    (fsrc/multifloats.fypp:1539-1552, 1573-1583) written out in numpy: the
    lexicographic `<`, and x - y (the oracle's DD subtraction) for dim,
    including NaN in either operand, equal hi limbs and signed zeros."""
    g = nt.load_golden("binary2.npz")
    a, b = g["fmin_a"], g["fmin_b"]
    lt = (a[:, 0] < b[:, 0]) | ((a[:, 0] == b[:, 0]) & (a[:, 1] < b[:, 1]))
    gt = (b[:, 0] < a[:, 0]) | ((a[:, 0] == b[:, 0]) & (b[:, 1] < a[:, 1]))
    nt.assert_bits(mf.fortran_min(a, b), np.where(lt[:, None], a, b), "min")
    nt.assert_bits(mf.fortran_max(a, b), np.where(lt[:, None], b, a), "max")
    nan = np.isnan(a[:, 0]) | np.isnan(b[:, 0])
    diff = nt.or_binary("sub", a, b)
    want = np.where((nan | gt)[:, None], diff, 0.0)
    nt.assert_bits(mf.fortran_dim(a, b), want, "dim")
    assert nan.sum() > 10 and (a[:, 0] == b[:, 0]).sum() > 100


def test_transcendental2_device_entry(torch_cuda):
    """mfhip_unary with the new op codes reaches the same kernels."""
    torch = torch_cuda
    rng = np.random.default_rng(8)
    stream = torch.cuda.current_stream().cuda_stream
    for fn in nt.TRANSC2:
        x = _transc2_inputs(rng, fn, 5000)
        dx = to_dev(torch, x)
        dy = torch.empty_like(dx)
        mf.dev_unary(getattr(mf, "MFHIP_" + fn.upper()), dx.data_ptr(), dy.data_ptr(), x.shape[0], stream)
        torch.cuda.synchronize()
        nt.assert_bits(dy.cpu().numpy(), nt.or_unary(fn, x), "device " + fn)


def test_erf_accuracy_golden(gpu):
    g = nt.load_golden("transc.npz")
    x, ex = g["erf_in"], g["erf_exact"]
    got = mf.erf(x)
    fin = np.isfinite(ex[:, 0]) & (np.abs(ex[:, 0]) > 1e-290)
    rel = nt.rel_err_vs_exact(got[fin], ex[fin])
    assert rel.max() <= ERF_BOUND, rel.max()
    # non-finite / zero classes as the reference: erf(+-inf) = +-1, NaN -> NaN, +-0 kept
    nf = ~np.isfinite(x[:, 0]) | (x[:, 0] == 0)
    nt.assert_bits(got[nf], g["erf_ref"][nf], "erf specials")


def test_erf_device_matches_host_build(gpu):
    """The device compile of dd_erf returns the same bits as its host build."""
    L = ctypes.CDLL(nt.HOSTMATH)
    rng = np.random.default_rng(5)
    x = np.concatenate([datagen.erf_inputs(rng, 100000), datagen.config0(rng, 10000)])
    want = np.empty_like(x)
    L.h_erf(nt.ptr(x), nt.ptr(want), ctypes.c_long(x.shape[0]))
    nt.assert_bits(mf.erf(x), want, "erf")


@pytest.mark.parametrize("nm", ["erfc", "erfcx"])
def test_erfc_family_accuracy_golden(gpu, nm):
    """Own design (the reference fits these with rationals): no less accurate
    than the reference's erfcdd / erfcxdd on the same inputs, below the caps
    where the result is a normal double-double, the reference's special values,
    and within a few subnormal ulps where erfc underflows."""
    g = nt.load_golden("special2.npz")
    x, ex = g[nm + "_in"], g[nm + "_exact"]
    got = getattr(mf, nm)(x)
    ours, ref = nt.erfc_family_errors(got, g, nm)
    assert ours <= ref and ours <= nt.ERFC_FAMILY_CAP[nm], (ours, ref)
    nf = ~np.isfinite(x[:, 0])
    nt.assert_bits(got[nf], g[nm + "_ref"][nf], nm + " non-finite inputs")
    sub = np.isfinite(ex[:, 0]) & (np.abs(ex[:, 0]) <= 2.0 ** -960) & np.isfinite(got[:, 0])
    assert (np.abs((got[sub, 0] - ex[sub, 0]) + (got[sub, 1] - ex[sub, 1])) <= 2.0 ** -1060).all()
    if nm == "erfc":
        z = np.array([[0.0, 0.0], [-0.0, 0.0]])
        assert (mf.erfc(z) == np.array([1.0, 0.0])).all()


@pytest.mark.parametrize("nm", ["erfc", "erfcx"])
def test_erfc_family_device_matches_host_build(gpu, nm):
    L = ctypes.CDLL(nt.HOSTMATH)
    rng = np.random.default_rng(6)
    x = np.concatenate([datagen._with_lo(rng, rng.uniform(-30, 30, 100000)), datagen.config0(rng, 10000),
                        datagen._with_lo(rng, 10.0 ** rng.uniform(-5, 5, 10000))])
    want = np.empty_like(x)
    getattr(L, "h_" + nm)(nt.ptr(x), nt.ptr(want), ctypes.c_long(x.shape[0]))
    nt.assert_bits(getattr(mf, nm)(x), want, nm)


def test_matmul_bits_golden(gpu):
    g = nt.load_golden("matmul.npz")
    for i in range(int(g["count"][0])):
        key = "c%03d" % i
        m, k, n, ri = (int(v) for v in g[key + "_shape"])
        A = datagen.unflat(g[key + "_A"], m, k)
        B = datagen.unflat(g[key + "_B"], k, n)
        nt.assert_bits(datagen.flat(mf.matmul(A, B, ri)), g[key + "_C"], "mm " + key)
        nt.assert_bits(mf.matmul(A, g[key + "_x"], ri), g[key + "_mv"], "mv " + key)
        nt.assert_bits(mf.matmul(g[key + "_x"], B, ri), g[key + "_vm"], "vm " + key)


# (400, 37, 330): 525 tiles of 16 x 16, where that variant takes its direct-copy form
@pytest.mark.parametrize("shape", [(1, 1, 1), (63, 65, 67), (128, 300, 96), (200, 17, 129), (400, 37, 330)])
@pytest.mark.parametrize("ri", [0, 8, 5])
def test_matmul_every_variant_bits_vs_oracle(torch_cuda, shape, ri):
    torch = torch_cuda
    m, k, n = shape
    rng = np.random.default_rng(m * 1000 + k + ri)
    A = datagen.matrix(rng, m, k, 20)
    B = datagen.matrix(rng, k, n, 20)
    want = datagen.flat(nt.or_matmul(A, B, ri))
    dA, dB = to_dev(torch, datagen.flat(A)), to_dev(torch, datagen.flat(B))
    for v in range(mf.matmul_variants()):
        dC = torch.full((m * n, 2), float("nan"), dtype=torch.float64, device="cuda")
        mf.dev_matmul(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, ri,
                      torch.cuda.current_stream().cuda_stream, v)
        torch.cuda.synchronize()
        nt.assert_bits(dC.cpu().numpy(), want, "variant %d %s ri=%d" % (v, shape, ri))


def test_matmul_4096_sampled_bits(torch_cuda):
    """4096^3 with exponent spread 2^+-20 (the at-scale workload): every
    variant's sampled outputs equal the oracle's sequential chains."""
    torch = torch_cuda
    m = k = n = 4096
    rng = np.random.default_rng(42)
    A = datagen.matrix(rng, m, k, 20)
    B = datagen.matrix(rng, k, n, 20)
    rows = rng.choice(m, 12, replace=False)
    cols = rng.choice(n, 12, replace=False)
    want = nt.or_matmul(A[rows], B[:, cols], 8)
    dA, dB = to_dev(torch, datagen.flat(A)), to_dev(torch, datagen.flat(B))
    dC = torch.empty((m * n, 2), dtype=torch.float64, device="cuda")
    for v in range(mf.matmul_variants()):
        dC.fill_(float("nan"))
        mf.dev_matmul(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, 8,
                      torch.cuda.current_stream().cuda_stream, v)
        torch.cuda.synchronize()
        C = datagen.unflat(dC.cpu().numpy(), m, n)
        nt.assert_bits(C[np.ix_(rows, cols)], want, "variant %d" % v)


def test_matmul_nonfinite_and_empty(gpu):
    A = np.zeros((3, 4, 2))
    B = np.ones((4, 2, 2)) * [1.0, 0.0]
    A[0, 0, 0] = np.inf
    A[1, 2, 0] = np.nan
    C = mf.matmul(A, B, 8)
    want = nt.or_matmul(A, B, 8)
    nt.assert_bits(C, want, "nonfinite")
    assert np.isnan(C[0, 0, 0]) and np.isnan(C[1, 1, 0])
    Z = mf.matmul(np.zeros((3, 0, 2)), np.zeros((0, 2, 2)), 8)
    assert Z.shape == (3, 2, 2) and not Z.any()


@pytest.mark.parametrize("n", [0, 1, 2, 1000, 65537, 1 << 22])
@pytest.mark.parametrize("ri", [0, 8])
def test_dot_sum_norm2_bounds(gpu, n, ri):
    rng = np.random.default_rng(n + ri)
    a, b = datagen.cancel_pairs(rng, max(n, 1))
    a, b = a[:n], b[:n]
    r = mf.dot_product(a, b, ri)
    err, mag = nt.exact_dot_err(a, b, r)
    assert err <= 2 * max(n, 1) * 2.0 ** -104 * mag + 1e-300
    x = datagen.sum_inputs(rng, max(n, 1))[:n]
    s = mf.sum(x)
    err, mag = nt.exact_sum_err(x, s)
    assert err <= 2 * max(n, 1) * 2.0 ** -104 * mag + 1e-300
    r2 = mf.norm2(x)
    if n:
        import mpmath
        mpmath.mp.prec = 256
        ss = nt.exact_dot(x, x)  # sum of squares, exact to 2^-106 relative
        want = mpmath.sqrt(mpmath.mpf(ss[0]) + mpmath.mpf(ss[1]))
        got = mpmath.mpf(r2[0]) + mpmath.mpf(r2[1])
        assert abs(got - want) <= (n + 4) * mpmath.mpf(2) ** -104 * want
    else:
        assert r2.tolist() == [0.0, 0.0] and s.tolist() == [0.0, 0.0] and r.tolist() == [0.0, 0.0]


def test_reduction_invariants(gpu):
    x = np.stack([np.arange(1, 100001, dtype=np.float64), np.zeros(100000)], axis=-1)
    assert mf.sum(x).tolist() == [5000050000.0, 0.0]
    for k in (-500, 0, 100):
        v = np.array([[3.0 * 2.0 ** k, 0.0], [4.0 * 2.0 ** k, 0.0]])
        assert mf.norm2(v).tolist() == [5.0 * 2.0 ** k, 0.0]
    y = x.copy()
    y[77, 0] = np.nan
    assert np.isnan(mf.sum(y)[0])
    assert np.isnan(mf.dot_product(y, x)[0])


def test_reductions_deterministic(gpu):
    rng = np.random.default_rng(9)
    a, b = datagen.cancel_pairs(rng, 3_000_001)
    r1 = mf.dot_product(a, b)
    r2 = mf.dot_product(a, b)
    nt.assert_bits(r1, r2, "dot rerun")


def _cunflat(f, m, n):
    return np.asarray(f).reshape(n, m, 4).transpose(1, 0, 2)


def test_complex_matmul_bits_golden(gpu):
    g = nt.load_golden("cmatmul.npz")
    for i in range(int(g["count"][0])):
        key = "z%03d" % i
        m, k, n, ri = (int(v) for v in g[key + "_shape"])
        A, B = _cunflat(g[key + "_A"], m, k), _cunflat(g[key + "_B"], k, n)
        C = mf.cmatmul(A, B, ri)
        got = np.ascontiguousarray(C.transpose(1, 0, 2)).reshape(-1, 4)
        nt.assert_bits(got, g[key + "_C"], "cmm " + key)
        if n == 1:
            nt.assert_bits(mf.cmatmul(A, B[:, 0, :], ri), g[key + "_C"], "cmv " + key)
        if m == 1:
            nt.assert_bits(mf.cmatmul(A[0], B, ri), g[key + "_C"], "cvm " + key)


def test_complex_matmul_2048_sampled(torch_cuda):
    """complex 2048^3 (the at-scale complex workload): sampled outputs equal
    the oracle's composition bit for bit."""
    torch = torch_cuda
    m = k = n = 2048
    rng = np.random.default_rng(21)

    def cm(r, c):
        re, im = datagen.matrix(rng, r, c, 20), datagen.matrix(rng, r, c, 20)
        return np.concatenate([re, im], axis=2)  # (r, c, 4) logical
    A, B = cm(m, k), cm(k, n)
    flat = lambda X: np.ascontiguousarray(X.transpose(1, 0, 2)).reshape(-1, 4)
    dA, dB = to_dev(torch, flat(A)), to_dev(torch, flat(B))
    dC = torch.full((m * n, 4), float("nan"), dtype=torch.float64, device="cuda")
    mf.dev_cmatmul(dA.data_ptr(), dB.data_ptr(), dC.data_ptr(), m, k, n, 8,
                   torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    C = _cunflat(dC.cpu().numpy(), m, n)
    rows = rng.choice(m, 8, replace=False)
    cols = rng.choice(n, 8, replace=False)
    As, Bs = flat(A[rows]), flat(B[:, cols])
    want = np.zeros((8 * 8, 4))
    nt.oracle().or_cmatmul_mm(nt.ptr(As), nt.ptr(Bs), nt.ptr(want), 8, k, 8, 8)
    got = np.ascontiguousarray(C[np.ix_(rows, cols)].transpose(1, 0, 2)).reshape(-1, 4)
    nt.assert_bits(got, want, "complex 2048")


# ---- sum / norm2 along a dimension, masked sum, complex dot ----

DIM_SHAPES = [((1000, 37), 1), ((1000, 37), 2), ((5, 7, 300), 1), ((5, 7, 300), 2), ((5, 7, 300), 3),
              ((3, 0, 4), 2), ((1, 1), 1), ((257, 3), 1), ((64, 65, 2, 3), 3)]


@pytest.mark.parametrize("shape,dim", DIM_SHAPES)
@pytest.mark.parametrize("masked", [False, True])
def test_sum_norm2_dim_bits_vs_oracle(gpu, shape, dim, masked):
    rng = np.random.default_rng(hash((shape, dim, masked)) % (1 << 32))
    nel = int(np.prod(shape))
    x = datagen.config0(rng, max(nel, 1))[:nel].reshape(tuple(reversed(shape)) + (2,))
    x = x.transpose(list(range(len(shape) - 1, -1, -1)) + [len(shape)])  # logical shape, Fortran memory
    mask = (rng.random(shape) < 0.6) if masked else None
    inner = int(np.prod(shape[:dim - 1]))
    ln = shape[dim - 1]
    outer = int(np.prod(shape[dim:]))
    buf = mf._fortran_order(x).reshape(-1, 2)
    mbuf = None if mask is None else mf._mask_buf(mask, shape).reshape(-1)
    got = mf.sum(x, dim=dim, mask=mask) if len(shape) > 1 else None
    want = nt.or_reduce_dim("sum", buf, mbuf, inner, ln, outer)
    nt.assert_bits(mf._fortran_order(got).reshape(-1, 2), want, "sum dim=%d %s" % (dim, shape))
    if not masked:
        got = mf.norm2(x, dim=dim)
        want = nt.or_reduce_dim("norm2", buf, None, inner, ln, outer)
        nt.assert_bits(mf._fortran_order(got).reshape(-1, 2), want, "norm2 dim=%d %s" % (dim, shape))


@pytest.mark.parametrize("n", [0, 1, 1000, 65537, 1 << 21])
def test_masked_sum_bound(gpu, n):
    rng = np.random.default_rng(n + 5)
    x = datagen.sum_inputs(rng, max(n, 1))[:n]
    mask = rng.random(n) < 0.5
    s = mf.sum(x, mask=mask)
    err, mag = nt.exact_sum_err(np.ascontiguousarray(x[mask]) if n else x, s)
    assert err <= 2 * max(n, 1) * 2.0 ** -104 * mag + 1e-300
    assert mf.sum(x, mask=np.zeros(n, bool)).tolist() == [0.0, 0.0]
    if n:
        nt.assert_bits(mf.sum(x, mask=np.ones(n, bool)), mf.sum(x), "all-true mask = no mask")


@pytest.mark.parametrize("n", [0, 1, 3000, 1 << 20])
@pytest.mark.parametrize("ri", [0, 8])
def test_complex_dot_bound(gpu, n, ri):
    rng = np.random.default_rng(n + 11 + ri)
    N = max(n, 1)
    a = np.concatenate([datagen.config0(rng, N), datagen.config0(rng, N)], axis=1)[:n]
    b = np.concatenate([datagen.config0(rng, N), datagen.config0(rng, N)], axis=1)[:n]
    r = mf.dot_product(a, b, ri)
    if n == 0:
        assert r.tolist() == [0.0] * 4
        return
    ar, ai, br, bi = a[:, :2], a[:, 2:], b[:, :2], b[:, 2:]
    e, m = nt.exact_dot_err(np.concatenate([ar, ai]), np.concatenate([br, bi]), r[:2])
    assert e <= 4 * n * 2.0 ** -104 * m
    e, m = nt.exact_dot_err(np.concatenate([ar, -ai]), np.concatenate([bi, br]), r[2:])
    assert e <= 4 * n * 2.0 ** -104 * m
    r2 = mf.dot_product(a, a, ri)
    assert r2[2] == 0.0 and r2[3] == 0.0


def test_device_reduction_entry_points(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(12)
    n = 100_003
    x = datagen.sum_inputs(rng, n)
    mask = rng.random(n) < 0.5
    dx = to_dev(torch, x)
    dm = to_dev(torch, mask.astype(np.int32))
    out = torch.zeros(4, dtype=torch.float64, device="cuda")
    work = torch.empty(mf.cdot_workspace_bytes(n) // 8 + 4, dtype=torch.float64, device="cuda")
    mf.dev_sum_masked(dx.data_ptr(), dm.data_ptr(), n, out.data_ptr(), work.data_ptr())
    torch.cuda.synchronize()
    nt.assert_bits(out[:2].cpu().numpy(), mf.sum(x, mask=mask), "device masked sum = host entry")
    a = np.concatenate([datagen.config0(rng, n), datagen.config0(rng, n)], axis=1)
    b = np.concatenate([datagen.config0(rng, n), datagen.config0(rng, n)], axis=1)
    da, db = to_dev(torch, a), to_dev(torch, b)
    mf.dev_cdot(da.data_ptr(), db.data_ptr(), n, out.data_ptr(), work.data_ptr())
    torch.cuda.synchronize()
    nt.assert_bits(out.cpu().numpy(), mf.dot_product(a, b), "device cdot = host entry")
    m_, k_ = 300, 41
    A = datagen.config0(rng, m_ * k_)
    dA = to_dev(torch, A)
    dout = torch.zeros((k_, 2), dtype=torch.float64, device="cuda")
    mf.dev_reduce_dim(mf.MFHIP_SUM, dA.data_ptr(), None, 1, m_, k_, dout.data_ptr())
    torch.cuda.synchronize()
    nt.assert_bits(dout.cpu().numpy(), nt.or_reduce_dim("sum", A, None, 1, m_, k_), "device sum dim")


@pytest.mark.parametrize("nm", ["erf", "erfc", "erfcx", "atan", "asin", "acos"])
def test_own_designs_device_against_mpfr(gpu, nm):
    """The device kernels of the functions of this project's own design
    against 256-bit MPFR values directly, on the same >= 100k points as the
    host build's test: within the cap (erf: the reference's published
    1.55e-32)."""
    from tests import pin_points as pp
    x, v, ok = pp.points(nm), pp.exact(nm), pp.checked(nm)
    got = getattr(mf, nm)(x)
    rel = nt.rel_err_vs_exact(got[ok], v[ok])
    assert rel.max() <= pp.CAPS[nm], (nm, rel.max(), x[ok][np.argmax(rel)].tolist())
