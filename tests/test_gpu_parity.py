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
    L = ctypes.CDLL(nt.os.path.join(nt.ROOT, "tests", "native", "libhostmath.so"))
    rng = np.random.default_rng(5)
    x = np.concatenate([datagen.erf_inputs(rng, 100000), datagen.config0(rng, 10000)])
    want = np.empty_like(x)
    L.h_erf(nt.ptr(x), nt.ptr(want), ctypes.c_long(x.shape[0]))
    nt.assert_bits(mf.erf(x), want, "erf")


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


@pytest.mark.parametrize("shape", [(1, 1, 1), (63, 65, 67), (128, 300, 96), (200, 17, 129)])
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
