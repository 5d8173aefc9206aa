"""fmod, modulo and nearest on the GPU (fmoddd_v, modulodd_v, nearestdd_v
and their mfhip_binary codes): the compiled reference's bits on every
returning call of tests/golden/fmod.npz, NaN where the reference's loop never
returns, and the oracle's bits on large draws over the loop's regimes."""
import os

import numpy as np
import pytest

import multifloats_amd as mf
from multifloats_amd import datagen
from tests import _native as nt

pytestmark = pytest.mark.gpu

G = np.load(os.path.join(nt.GOLDEN, "fmod.npz"))


def _oracle(nm, a, b):
    getattr(nt.oracle(), "or_" + nm).argtypes = [nt.P, nt.P, nt.P, nt.I64]
    return nt.or_binary(nm, a, b)


def _nan_rows_as_nan(got, want):
    """Rows whose high limb is NaN on both sides are the same value (NaN);
    their low limbs may differ in the sign of a zero: x86 generates NaN with
    the sign bit set, the GPU with it clear, and modulo's sign test
    (signbit of the remainder) reads that bit. Such rows are compared by
    NaN-ness, every other row bit for bit."""
    both = np.isnan(got[:, 0]) & np.isnan(want[:, 0])
    g, w = got.copy(), want.copy()
    g[both], w[both] = np.nan, np.nan
    return g, w


@pytest.mark.parametrize("nm", ["fmod", "modulo"])
def test_reference_bits_and_nonreturning_calls(gpu, nm):
    a, b = G["a"], G["b"]
    ok = G[nm + "_returns"]
    got = getattr(mf, nm)(a, b)
    g, w = _nan_rows_as_nan(got[ok], G[nm + "_ref"][ok])
    nt.assert_bits(g, w, nm + " vs the compiled reference")
    assert np.isnan(got[~ok, 0]).all()


@pytest.mark.parametrize("nm", ["fmod", "modulo", "nearest"])
def test_oracle_bits_on_wide_draws(gpu, nm):
    rng = np.random.default_rng(17 + len(nm))
    n = 1 << 20
    w = datagen._with_lo
    s = lambda k: rng.choice([-1.0, 1.0], k)
    a = np.concatenate([w(rng, 2.0 ** rng.uniform(-900, 900, n) * s(n)), w(rng, rng.uniform(-1e6, 1e6, n))])
    b = np.concatenate([w(rng, 2.0 ** rng.uniform(-900, 900, n) * s(n)), w(rng, rng.uniform(-3, 3, n))])
    print(nm, "n =", a.shape[0], flush=True)
    nt.assert_bits(*_nan_rows_as_nan(getattr(mf, nm)(a, b), _oracle(nm, a, b)), nm)


def test_device_codes(gpu):
    """The device-pointer form through mfhip_binary."""
    import torch
    a, b = G["a"], G["b"]
    da, db = torch.from_numpy(np.ascontiguousarray(a)).cuda(), torch.from_numpy(np.ascontiguousarray(b)).cuda()
    dc = torch.empty_like(da)
    s = torch.cuda.current_stream().cuda_stream
    for code, nm in ((mf.MFHIP_FMOD, "fmod"), (mf.MFHIP_MODULO, "modulo"), (mf.MFHIP_NEAREST, "nearest")):
        mf.dev_binary(code, da.data_ptr(), db.data_ptr(), dc.data_ptr(), a.shape[0], s)
        nt.assert_bits(*_nan_rows_as_nan(dc.cpu().numpy(), _oracle(nm, a, b)), nm)
