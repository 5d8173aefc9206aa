"""GPU tests of the batching form: grouped GEMMs bit-identical to the
reference order, grouped dots within the reduction bound, graph replays equal
to eager runs."""
import numpy as np
import pytest

import multifloats_amd as mf
from multifloats_amd import datagen
from tests import _native as nt

pytestmark = pytest.mark.gpu


def build(torch, rng, ngemm, ndot, variants=None, hi=96, dhi=40000, long_dot=0):
    sizes, dlen = datagen.batched_sizes(rng, ngemm, ndot, hi=hi, dhi=dhi)
    sizes[0] = (1, 1, 1)
    sizes[1] = (70, 0, 33)          # k = 0 -> zeros
    tot = int(sum(m * k + k * n + m * n for m, k, n in sizes))
    gbuf = torch.empty(2 * tot, dtype=torch.float64, device="cuda")
    datagen.fill_config0_torch(torch, gbuf, 5)
    dlen[0] = 1
    if long_dot:
        dlen[1] = long_dot          # longer than one work item: chunks + partials
    dtot = int(dlen.sum())
    dbuf = torch.empty(4 * dtot, dtype=torch.float64, device="cuda")
    datagen.fill_config0_torch(torch, dbuf, 6)
    dout = torch.full((len(dlen), 2), float("nan"), dtype=torch.float64, device="cuda")
    gemms, offs, off = [], [], 0
    for m, k, n in sizes:
        m, k, n = int(m), int(k), int(n)
        offs.append(off)
        b = gbuf.data_ptr()
        gemms.append((b + 16 * off, b + 16 * (off + m * k), b + 16 * (off + m * k + k * n), m, k, n, 8))
        off += m * k + k * n + m * n
    dots, doff, off = [], [], 0
    for i, L in enumerate(dlen):
        L = int(L)
        doff.append(off)
        b = dbuf.data_ptr()
        dots.append((b + 16 * off, b + 16 * (off + L), dout.data_ptr() + 16 * i, L, 8))
        off += 2 * L
    return sizes, dlen, gbuf, dbuf, dout, gemms, dots, offs, doff


@pytest.mark.parametrize("variant", [None, 0, 1, 2, 3, 4, 5])
def test_grouped_gemm_and_dot_parity(gpu, variant):
    import torch
    rng = np.random.default_rng(3 if variant is None else variant)
    sizes, dlen, gbuf, dbuf, dout, gemms, dots, offs, doff = build(torch, rng, 300, 200,
                                                                   long_dot=2_500_000 if variant is None else 0)
    var = None if variant is None else [variant] * len(gemms)
    plan = mf.GroupedPlan(gemms, dots, var)
    s = torch.cuda.current_stream().cuda_stream
    plan.run(s)
    torch.cuda.synchronize()
    gb = gbuf.view(-1, 2).cpu().numpy()
    for i, (m, k, n) in enumerate(sizes):
        m, k, n = int(m), int(k), int(n)
        o = offs[i]
        A = datagen.unflat(gb[o:o + m * k], m, k)
        B = datagen.unflat(gb[o + m * k:o + m * k + k * n], k, n)
        C = datagen.unflat(gb[o + m * k + k * n:o + m * k + k * n + m * n], m, n)
        nt.assert_bits(C, nt.or_matmul(A, B, 8), "gemm %d %s" % (i, (m, k, n)))
    db = dbuf.view(-1, 2).cpu().numpy()
    res = dout.cpu().numpy()
    for i, L in enumerate(dlen):
        L, o = int(L), doff[i]
        err, mag = nt.exact_dot_err(db[o:o + L], db[o + L:o + 2 * L], res[i])
        assert err <= 2 * L * 2.0 ** -104 * mag + 1e-300, (i, L)
    plan.close()


def test_grouped_graph_replays_match_eager(gpu):
    import torch
    rng = np.random.default_rng(8)
    sizes, dlen, gbuf, dbuf, dout, gemms, dots, offs, doff = build(torch, rng, 120, 120, long_dot=3_000_001)
    plan = mf.GroupedPlan(gemms, dots)
    s = torch.cuda.current_stream().cuda_stream
    plan.run(s)
    torch.cuda.synchronize()
    eager_g = gbuf.clone()
    eager_d = dout.clone()
    for rep in range(3):
        # disturb the outputs with eager work between replays
        dout.fill_(float("nan"))
        plan.run_graph(s)
        torch.cuda.synchronize()
        assert torch.equal(gbuf, eager_g), rep
        nt.assert_bits(dout.cpu().numpy(), eager_d.cpu().numpy(), "dots replay %d" % rep)
    plan.close()


def test_grouped_dots_beside_and_alone_agree(gpu):
    """A plan with GEMMs runs its dots on a second stream beside the GEMM
    kinds (two elements per register buffer); a plan of dots only runs them
    alone on the caller's stream (eight per buffer). Each lane's chain is the
    same, so the bits are equal; a plan of the GEMMs only writes the mixed
    plan's GEMM bits."""
    import torch
    rng = np.random.default_rng(11)
    sizes, dlen, gbuf, dbuf, dout, gemms, dots, offs, doff = build(torch, rng, 150, 150, long_dot=2_100_000)
    s = torch.cuda.current_stream().cuda_stream
    mixed = mf.GroupedPlan(gemms, dots)
    mixed.run(s)
    torch.cuda.synchronize()
    g_mixed, d_mixed = gbuf.clone(), dout.clone()
    dots_only = mf.GroupedPlan([], dots)
    dout.fill_(float("nan"))
    dots_only.run(s)
    torch.cuda.synchronize()
    nt.assert_bits(dout.cpu().numpy(), d_mixed.cpu().numpy(), "dots alone vs beside")
    gemms_only = mf.GroupedPlan(gemms, [])
    for (m, k, n), o in zip(sizes, offs):
        c0 = 2 * (o + int(m) * int(k) + int(k) * int(n))
        gbuf[c0:c0 + 2 * int(m) * int(n)].fill_(float("nan"))
    gemms_only.run_graph(s)
    torch.cuda.synchronize()
    assert torch.equal(gbuf, g_mixed)
    for p in (mixed, dots_only, gemms_only):
        p.close()
