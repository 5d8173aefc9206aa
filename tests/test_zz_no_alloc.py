"""GPU test of mfhip_reduce_noalloc, the reduction entry point that takes no
workspace and allocates nothing: its results must be bit-identical to
mfhip_reduce with a caller workspace (same kernels), on two calls in a row,
the second on a smaller input, for every reduction kind; sizes above the
stated maximum are refused before anything is enqueued."""
import ctypes

import numpy as np
import pytest

import multifloats_amd as mf
from multifloats_amd import datagen

pytestmark = pytest.mark.gpu

KINDS = [mf.MFHIP_DOT, mf.MFHIP_SUM, mf.MFHIP_NORM2, mf.MFHIP_PRODUCT, mf.MFHIP_MAXVAL, mf.MFHIP_MINVAL]


def _inputs(rng, n, kind):
    if kind == mf.MFHIP_PRODUCT:
        # factors near 1 keep the product clear of overflow and underflow
        x = datagen.config0(rng, n)
        x[:, 0] = 1.0 + x[:, 0] * 2.0 ** -12
        x[:, 1] *= 2.0 ** -12
        return x, x
    return datagen.config0(rng, n), datagen.config0(rng, n)


def _run_pair(torch, kind, a, b, ri, stream):
    n = a.shape[0]
    da = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    db = torch.from_numpy(np.ascontiguousarray(b)).cuda()
    out_ref = torch.full((1, 2), float("nan"), dtype=torch.float64, device="cuda")
    out_na = torch.full((1, 2), float("nan"), dtype=torch.float64, device="cuda")
    wb = mf.reduce_workspace_bytes(n)
    work = torch.empty(max(wb, 16), dtype=torch.uint8, device="cuda")
    L = mf.lib()
    rc = L.mfhip_reduce(kind, da.data_ptr(), db.data_ptr(), n, ri, out_ref.data_ptr(), work.data_ptr(), stream)
    assert rc == 0, rc
    rc = L.mfhip_reduce_noalloc(kind, da.data_ptr(), db.data_ptr(), n, ri, out_na.data_ptr(), stream)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return out_ref.cpu().numpy(), out_na.cpu().numpy()


@pytest.mark.parametrize("kind", KINDS)
def test_noalloc_reduce_matches_workspace_form(gpu, kind):
    import torch
    stream = torch.cuda.current_stream().cuda_stream
    rng = np.random.default_rng(100 + kind)
    # the first size folds its partials twice (2^22 + 977 > 2048 * 2048
    # elements of the dot), the second is smaller and needs one level
    for n in ((1 << 22) + 977, 5000):
        print("noalloc kind=%d n=%d" % (kind, n), flush=True)
        a, b = _inputs(rng, n, kind)
        ref, got = _run_pair(torch, kind, a, b, 8, stream)
        assert np.isfinite(ref).all(), (kind, n, ref)
        assert np.array_equal(ref.view(np.uint64), got.view(np.uint64)), (kind, n, ref, got)


def test_noalloc_reduce_refuses_sizes_beyond_its_scratch(gpu):
    import torch
    x = torch.zeros((16, 2), dtype=torch.float64, device="cuda")
    out = torch.zeros((1, 2), dtype=torch.float64, device="cuda")
    L = mf.lib()
    rc = L.mfhip_reduce_noalloc(mf.MFHIP_SUM, x.data_ptr(), x.data_ptr(), mf.NOALLOC_MAX_N + 1, 0,
                                out.data_ptr(), torch.cuda.current_stream().cuda_stream)
    assert rc != 0
    torch.cuda.synchronize()
    # the refused call enqueued nothing; the entry point still works after it
    rc = L.mfhip_reduce_noalloc(mf.MFHIP_SUM, x.data_ptr(), x.data_ptr(), 16, 0, out.data_ptr(),
                                torch.cuda.current_stream().cuda_stream)
    assert rc == 0
    torch.cuda.synchronize()
    assert not out.cpu().numpy().any()
