"""The Fortran bridge (multifloats_amd/fortran/multifloats_hip.F90) on the GPU:
a Fortran program calls the reference's generics -- matmul, dot_product, sum,
norm2, exp/log/sin/erf/sqrt and the array operators -- through the module and
every result must carry the same bits as the C-ABI entry points called from
Python on the same inputs (and, for matmul and the dim reductions, as the
oracle)."""
import os
import subprocess

import numpy as np
import pytest

import multifloats_amd as mf
from multifloats_amd import datagen
from tests import _native as nt

pytestmark = pytest.mark.gpu

EXE = os.path.join(nt.ROOT, "tests", "fortran", "_build", "bridge_check")


def _cmatrix(rng, r, c):
    re, im = datagen.matrix(rng, r, c, 10), datagen.matrix(rng, r, c, 10)
    return np.concatenate([re, im], axis=2)  # (r, c, 4), logical indexing


def _cflat(a):
    return np.ascontiguousarray(a.transpose(1, 0, 2)).reshape(-1, 4)


class _Reader:
    def __init__(self, buf):
        self.buf, self.pos = buf, 0

    def take(self, count, width=2):
        out = self.buf[self.pos:self.pos + count * width].reshape(count, width)
        self.pos += count * width
        return out


def test_fortran_bridge_routes_to_the_library(gpu, tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("bridge_check not built (amdflang missing at build time)")
    rng = np.random.default_rng(31)
    m, k, n, nv = 37, 45, 29, 5000
    A = datagen.matrix(rng, m, k, 10)
    B = datagen.matrix(rng, k, n, 10)
    CA, CB = _cmatrix(rng, m, k), _cmatrix(rng, k, n)
    x = datagen.config0(rng, nv)
    y = datagen.config0(rng, nv)
    cx = np.concatenate([datagen.config0(rng, nv), datagen.config0(rng, nv)], axis=1)
    cy = np.concatenate([datagen.config0(rng, nv), datagen.config0(rng, nv)], axis=1)
    mask = rng.random(nv) < 0.5
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as fh:
        np.array([m, k, n, nv], dtype=np.int64).tofile(fh)
        for arr in (datagen.flat(A), datagen.flat(B), _cflat(CA), _cflat(CB), x, y, cx, cy):
            np.ascontiguousarray(arr, dtype=np.float64).tofile(fh)
        mask.astype(np.int32).tofile(fh)
    print("bridge_check m=%d k=%d n=%d nv=%d" % (m, k, n, nv), flush=True)
    proc = subprocess.run([EXE, fin, fout], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "failed" not in proc.stderr, proc.stderr
    r = _Reader(np.fromfile(fout, dtype=np.float64))

    nt.assert_bits(r.take(m * n), datagen.flat(nt.or_matmul(A, B, 8)), "matmul mm")
    nt.assert_bits(r.take(m), datagen.flat(nt.or_matmul(A, B[:, :1], 8)), "matmul mv")
    nt.assert_bits(r.take(n), datagen.flat(nt.or_matmul(A[:1, :], B, 8)), "matmul vm (row of a)")
    nt.assert_bits(r.take(m * n, 4), _cflat(mf.cmatmul(CA, CB)), "cmatmul mm")
    nt.assert_bits(r.take(m, 4), mf.cmatmul(CA, np.ascontiguousarray(CB[:, 0])), "cmatmul mv")
    nt.assert_bits(r.take(n, 4), mf.cmatmul(np.ascontiguousarray(CA[0, :]), CB), "cmatmul vm")
    nt.assert_bits(r.take(m * n), datagen.flat(nt.or_matmul(A, B, 0)), "matmul ri=0")
    nt.assert_bits(r.take(1)[0], mf.dot_product(x, y, 0), "dot ri=0")
    nt.assert_bits(r.take(1)[0], mf.dot_product(x, y, 8), "dot ri=8")
    nt.assert_bits(r.take(1, 4)[0], mf.dot_product(cx, cy, 8), "complex dot")
    nt.assert_bits(r.take(1)[0], mf.sum(x), "sum")
    nt.assert_bits(r.take(1)[0], mf.sum(x, mask=mask), "sum mask")
    nt.assert_bits(r.take(1)[0], mf.sum(A), "sum rank 2")
    fa = datagen.flat(A)
    nt.assert_bits(r.take(k), nt.or_reduce_dim("sum", fa, None, 1, m, k), "sum dim=1")
    nt.assert_bits(r.take(m), nt.or_reduce_dim("sum", fa, None, m, k, 1), "sum dim=2")
    nt.assert_bits(r.take(1)[0], mf.norm2(x), "norm2")
    nt.assert_bits(r.take(k), nt.or_reduce_dim("norm2", fa, None, 1, m, k), "norm2 dim=1")
    nt.assert_bits(r.take(m), nt.or_reduce_dim("norm2", fa, None, m, k, 1), "norm2 dim=2")
    for nm in ("exp", "log", "sin", "erf", "sqrt"):
        nt.assert_bits(r.take(nv), getattr(mf, nm)(x), nm)
    for nm in ("add", "sub", "mul", "div"):
        nt.assert_bits(r.take(nv), getattr(mf, nm)(x, y), nm)
    nt.assert_bits(r.take(m * k), mf.exp(fa), "exp rank 2")
    C = datagen.unflat(r.take(m * n), m, n)
    rows, cols = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    err, mag = nt.gemm_err(A, B, C, rows.ravel(), cols.ravel())
    assert (err <= 2.0 ** -100 * mag).all(), "ozaki through the bridge"
    for nm in ("cos", "tan", "expm1"):
        nt.assert_bits(r.take(nv), getattr(mf, nm)(x), nm)
    for nm in ("log2", "log10"):
        nt.assert_bits(r.take(nv), getattr(mf, nm)(mf.mul(x, x)), nm)
    for nm in ("log1p", "sinh", "cosh", "tanh", "erfc", "erfcx"):
        nt.assert_bits(r.take(nv), getattr(mf, nm)(x), nm)
    nt.assert_bits(r.take(nv), mf.asinh(x), "asinh")
    nt.assert_bits(r.take(nv), mf.acosh(mf.add(mf.mul(x, x), mf.exp(mf.mul(y, y)))), "acosh")
    nt.assert_bits(r.take(nv), mf.atanh(x), "atanh")
    nt.assert_bits(r.take(nv), mf.sinpi(mf.add(x, y)), "sinpi")
    nt.assert_bits(r.take(nv), mf.cospi(mf.sub(x, y)), "cospi")
    nt.assert_bits(r.take(nv), mf.tanpi(x), "tanpi")
    nt.assert_bits(r.take(nv), mf.hypot(x, y), "hypot")
    nt.assert_bits(r.take(nv), mf.copysign(x, y), "sign")
    nt.assert_bits(r.take(nv), mf.atan(x), "atan")
    nt.assert_bits(r.take(nv), mf.atanpi(x), "atanpi")
    nt.assert_bits(r.take(nv), mf.atan2(x, y), "atan2")
    nt.assert_bits(r.take(nv), mf.atan2pi(x, y), "atan2pi")
    q = mf.div(mf.mul(x, y), mf.add(mf.mul(x, x), mf.mul(y, y)))
    nt.assert_bits(r.take(nv), mf.asin(q), "asin")
    nt.assert_bits(r.take(nv), mf.acos(q), "acos")
    for nm in ("min", "max", "dim"):
        nt.assert_bits(r.take(nv), getattr(mf, "fortran_" + nm)(x, y), nm)
    for nm, fn in (("mod", mf.fmod), ("modulo", mf.modulo), ("nearest", mf.nearest)):
        nt.assert_bits(r.take(nv), fn(x, y), nm)
    At = nt.cm(A.transpose(1, 0, 2))
    Cw = np.empty((m + 2, n, 2))
    Cw[:m] = A[:, :n]
    Cw[m:] = x[0]
    Cw = nt.cm(Cw)
    assert nt.or_gemm_blas("T", "N", m, n, k, x[0], At, nt.cm(B), y[0], Cw, 8) == 0
    nt.assert_bits(r.take((m + 2) * n), datagen.flat(Cw), "wgemm T N")
    assert r.pos == r.buf.size, "output size"
