"""GPU tests at the measurement plan's own sizes (the configs the bench times):

  C0  C = A B at 256^3, config-0 entries, renorm 8: every output bit-identical
      to the compiled reference (fixture c0_256.json) and to the oracle;
  C1  N = 2^26 per op on the C1 distributions: + - * / sqrt exp log sin
      sampled bit for bit against the oracle, erf sampled against 256-bit
      MPFR values, sum / dot / norm2 within their bounds against the exact
      accumulator (at this N the fold chain runs two levels);
  C3  the Ozaki automatic form at 8192^3, exponent spreads 2^+-10 and 2^+-30,
      sampled outputs within 2^-100 (|A||B|)_ij;
  C4  10,000 GEMMs (m, n, k log-uniform in [8, 512]) + 50,000 dots (length
      log-uniform in [16, 65536]) in one grouped launch replayed from a HIP
      graph: sampled GEMMs bit-identical to the reference order, sampled dots
      within the bound.
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

import multifloats_amd as mf
from multifloats_amd import datagen
from tests import _native as nt

pytestmark = pytest.mark.gpu

N1 = 1 << 26


def test_c0_256_bit_identical_to_the_compiled_reference(gpu):
    with open(os.path.join(nt.GOLDEN, "c0_256.json")) as f:
        g = json.load(f)
    rng = np.random.default_rng(g["seed"])
    m, k, n = g["m"], g["k"], g["n"]
    A = datagen.matrix(rng, m, k)
    B = datagen.matrix(rng, k, n)
    C = mf.matmul(A, B, g["renorm_interval"])
    flat = np.ascontiguousarray(datagen.flat(C))
    first = np.array([[float.fromhex(h), float.fromhex(l)] for h, l in g["first_column"]])
    nt.assert_bits(flat[:m], first, "C0 first column")
    assert hashlib.sha256(flat.tobytes()).hexdigest() == g["sha256"]
    nt.assert_bits(C, nt.or_matmul(A, B, g["renorm_interval"]), "C0 vs oracle")


def _c1_inputs(rng, nm):
    if nm in ("add", "sub", "mul", "div"):
        return datagen.config0(rng, N1), datagen.config0(rng, N1)
    if nm == "sqrt":
        x = datagen.config0(rng, N1)
        return np.where(x[:, :1] < 0, -x, x), None
    return getattr(datagen, nm + "_inputs")(rng, N1), None


@pytest.mark.parametrize("nm", ["add", "sub", "mul", "div", "sqrt", "exp", "log", "sin", "erf"])
def test_c1_elementwise_at_2p26(gpu, nm):
    import torch
    rng = np.random.default_rng(2 ** 26 + sum(map(ord, nm)))
    a, b = _c1_inputs(rng, nm)
    print("C1 %s N=2^26" % nm, flush=True)
    da = torch.from_numpy(a).cuda()
    dc = torch.empty_like(da)
    s = torch.cuda.current_stream().cuda_stream
    if b is not None:
        db = torch.from_numpy(b).cuda()
        mf.dev_binary(getattr(mf, "MFHIP_" + nm.upper()), da.data_ptr(), db.data_ptr(), dc.data_ptr(), N1, s)
    else:
        mf.dev_unary(getattr(mf, "MFHIP_" + nm.upper()), da.data_ptr(), dc.data_ptr(), N1, s)
    c = dc.cpu().numpy()
    # a random sample plus both ends (and, for sin, every planted point near k pi/2)
    idx = np.unique(np.concatenate([rng.choice(N1, 1 << 16, replace=False), np.arange(64), np.arange(N1 - 64, N1)]))
    if nm == "sin":
        near = np.nonzero(np.abs(np.remainder(a[:, 0] + np.pi / 4, np.pi / 2) - np.pi / 4) < 1e-9)[0]
        assert near.size > N1 // 200
        idx = np.unique(np.concatenate([idx, near[:: max(1, near.size // 50000)]]))
    if nm == "erf":
        import tests.pin_points as pp
        L = ctypes.CDLL(pp.MPFR_LIB)
        L.pin_eval.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        x = np.ascontiguousarray(a[idx[:20000]])
        v = np.zeros((x.shape[0], 3))
        L.pin_eval(pp.CODE["erf"], x.ctypes.data, v.ctypes.data, x.shape[0])
        ok = np.isfinite(v[:, 0]) & (np.abs(v[:, 0]) > 2.0 ** -960)
        assert nt.rel_err_vs_exact(c[idx[:20000]][ok], v[ok]).max() <= pp.CAPS["erf"]
    else:
        want = nt.or_binary(nm, a[idx], b[idx]) if b is not None else nt.or_unary(nm, a[idx])
        nt.assert_bits(c[idx], want, "C1 " + nm)


def test_c1_reductions_at_2p26(gpu):
    """dot on 10% near-cancelling pairs, sum on near-cancelling neighbours,
    norm2 on config-0 entries, all at N = 2^26, within c n 2^-104 of the
    exact value. At this N the first pass leaves 32,768 (dot) and 16,384
    (sum, norm2) partials, more than one fold workgroup takes (2,048), so the
    fold chain runs a second level."""
    rng = np.random.default_rng(2 ** 26 + 1)
    a, b = datagen.cancel_pairs(rng, N1)
    print("C1 reductions N=2^26", flush=True)
    r = mf.dot_product(a, b, 8)
    err, mag = nt.exact_dot_err(a, b, r)
    assert err <= 2 * N1 * 2.0 ** -104 * mag
    x = datagen.sum_inputs(rng, N1)
    s = mf.sum(x)
    err, mag = nt.exact_sum_err(x, s)
    assert err <= 2 * N1 * 2.0 ** -104 * mag
    import mpmath
    mpmath.mp.prec = 256
    r2 = mf.norm2(a)
    ss = nt.exact_dot(a, a)
    want = mpmath.sqrt(mpmath.mpf(ss[0]) + mpmath.mpf(ss[1]))
    got = mpmath.mpf(r2[0]) + mpmath.mpf(r2[1])
    assert abs(got - want) <= (N1 + 4) * mpmath.mpf(2) ** -104 * want


@pytest.mark.parametrize("spread", [10, 30])
def test_c3_ozaki_8192(gpu, spread):
    import torch
    m = k = n = 8192
    rng = np.random.default_rng(8192 + spread)
    A = datagen.matrix(rng, m, k, spread)
    B = datagen.matrix(rng, k, n, spread)
    print("C3 Ozaki 8192^3 spread 2^+-%d" % spread, flush=True)
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
    rows = np.unique(np.concatenate([rng.choice(m, 40, replace=False), [0, m - 1]]))
    cols = np.unique(np.concatenate([rng.choice(n, 40, replace=False), [0, n - 1]]))
    rr, cc = np.meshgrid(rows, cols, indexing="ij")
    err, mag = nt.gemm_err(A, B, C, rr.ravel(), cc.ravel(), threads=8)
    assert (err <= 2.0 ** -100 * mag).all(), (used, fix, np.log2((err / mag).max()))


def test_c4_batched_full_mix(gpu):
    import torch
    rng = np.random.default_rng(4)
    sizes, dlen = datagen.batched_sizes(rng)
    assert len(sizes) == 10000 and len(dlen) == 50000
    print("C4 10000 GEMMs + 50000 dots", flush=True)
    tot = int(sum(m * k + k * n + m * n for m, k, n in sizes))
    gbuf = torch.empty(2 * tot, dtype=torch.float64, device="cuda")
    datagen.fill_config0_torch(torch, gbuf, 101)
    dtot = int(dlen.sum())
    dbuf = torch.empty(4 * dtot, dtype=torch.float64, device="cuda")
    datagen.fill_config0_torch(torch, dbuf, 202)
    dout = torch.full((len(dlen), 2), float("nan"), dtype=torch.float64, device="cuda")
    gemms, offs, off = [], [], 0
    base, dbase = gbuf.data_ptr(), dbuf.data_ptr()
    for m, k, n in sizes:
        m, k, n = int(m), int(k), int(n)
        offs.append(off)
        gemms.append((base + 16 * off, base + 16 * (off + m * k), base + 16 * (off + m * k + k * n), m, k, n, 8))
        off += m * k + k * n + m * n
    dots, doff, off = [], [], 0
    for i, L in enumerate(dlen):
        L = int(L)
        doff.append(off)
        dots.append((dbase + 16 * off, dbase + 16 * (off + L), dout.data_ptr() + 16 * i, L, 8))
        off += 2 * L
    plan = mf.GroupedPlan(gemms, dots)
    s = torch.cuda.current_stream().cuda_stream
    plan.run_graph(s)
    plan.run_graph(s)  # a replay after the capturing run
    torch.cuda.synchronize()
    gb, db = gbuf.view(-1, 2), dbuf.view(-1, 2)
    for i in rng.choice(len(sizes), 60, replace=False):
        m, k, n = (int(v) for v in sizes[i])
        o = offs[i]
        A = datagen.unflat(gb[o:o + m * k].cpu().numpy(), m, k)
        B = datagen.unflat(gb[o + m * k:o + m * k + k * n].cpu().numpy(), k, n)
        C = datagen.unflat(gb[o + m * k + k * n:o + m * k + k * n + m * n].cpu().numpy(), m, n)
        nt.assert_bits(C, nt.or_matmul(A, B, 8), "C4 gemm %d %s" % (i, (m, k, n)))
    res = dout.cpu().numpy()
    assert np.isfinite(res).all()
    for i in np.concatenate([rng.choice(len(dlen), 300, replace=False), [int(np.argmax(dlen))]]):
        L, o = int(dlen[i]), doff[i]
        a, b = db[o:o + L].cpu().numpy(), db[o + L:o + 2 * L].cpu().numpy()
        err, mag = nt.exact_dot_err(a, b, res[i])
        assert err <= 2 * L * 2.0 ** -104 * mag + 1e-300, (i, L)
    plan.close()
