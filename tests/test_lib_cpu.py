"""CPU checks of the product library and host logic (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import multifloats_amd as mf
from multifloats_amd import datagen
from tests import _native as nt

HEADER = os.path.join(nt.ROOT, "include", "multifloats_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"MULTIFLOATS_API\s+[\w\s\*]+?\b(\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert "matmuldd_mm" in syms and "mfhip_reduce" in syms and len(syms) >= 25
    L = ctypes.CDLL(mf.LIB_PATH)
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_reference_signatures_unchanged():
    """matmuldd_* keep the reference argument lists (float64x2.h:682-695)."""
    text = open(HEADER).read()
    assert re.search(r"void matmuldd_mm\(const float64x2 \*a, const float64x2 \*b, float64x2 \*c,\s*"
                     r"int64_t m, int64_t k, int64_t n, int64_t renorm_interval\)", text)
    assert re.search(r"void matmuldd_mv\(const float64x2 \*a, const float64x2 \*x, float64x2 \*y,\s*"
                     r"int64_t m, int64_t k, int64_t renorm_interval\)", text)
    assert re.search(r"void matmuldd_vm\(const float64x2 \*x, const float64x2 \*b, float64x2 \*y,\s*"
                     r"int64_t k, int64_t n, int64_t renorm_interval\)", text)


def test_header_compiles_as_c_and_cpp(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "multifloats_hip.h"\nint main(void){float64x2 x; (void)x; return sizeof(complex64x2) == 32 ? 0 : 1;}\n')
    import subprocess
    inc = os.path.join(nt.ROOT, "include")
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", inc, str(src), "-o", str(tmp_path / "a")], check=True)
    subprocess.run([str(tmp_path / "a")], check=True)
    srcpp = tmp_path / "t.cpp"
    srcpp.write_text('#include "multifloats_hip.h"\nusing multifloats::float64x2;\nint main(){float64x2 x{}; (void)x; '
                     'return sizeof(multifloats::complex64x2) == 32 ? 0 : 1;}\n')
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", inc, str(srcpp), "-o", str(tmp_path / "b")], check=True)
    subprocess.run([str(tmp_path / "b")], check=True)


def test_layout_helpers_roundtrip():
    rng = np.random.default_rng(0)
    A = datagen.matrix(rng, 5, 3)
    assert datagen.is_colmajor(A)
    f = datagen.flat(A)
    assert np.array_equal(f[1 + 5 * 2], A[1, 2])
    assert np.array_equal(datagen.unflat(f, 5, 3), A)
    B = datagen.as_colmajor(np.asfortranarray(A))
    assert datagen.is_colmajor(B) and np.array_equal(B, A)


def test_generators_are_canonical():
    rng = np.random.default_rng(1)
    for x in (datagen.config0(rng, 1000), datagen.spread(rng, 1000, 30), datagen.sin_inputs(rng, 1000, 0.5)):
        s = x[:, 0] + x[:, 1]
        assert np.array_equal(s, x[:, 0])  # |lo| <= ulp(hi)/2
    a, b = datagen.cancel_pairs(rng, 1000)
    assert np.isfinite(a).all() and np.isfinite(b).all()


def test_sin_generator_plants_points_near_multiples_of_half_pi():
    import mpmath
    mpmath.mp.prec = 200
    rng = np.random.default_rng(2)
    x = datagen.sin_inputs(rng, 2000, near_frac=0.5)
    near = 0
    for v in x[:400]:
        t = (mpmath.mpf(v[0]) + mpmath.mpf(v[1])) / (mpmath.pi / 2)
        if abs(t - mpmath.nint(t)) * mpmath.pi / 2 < 1.1e-12:
            near += 1
    assert near > 120


def test_hostmath_matches_oracle_bits():
    """The kernels' DD headers built for the host agree with the oracle."""
    L = ctypes.CDLL(os.path.join(nt.ROOT, "tests", "native", "libhostmath.so"))
    rng = np.random.default_rng(4)
    gens = {"exp": datagen.exp_inputs, "log": datagen.log_inputs, "sin": datagen.sin_inputs}
    for nm, gen in gens.items():
        x = gen(rng, 20000)
        got = np.empty_like(x)
        getattr(L, "h_" + nm)(nt.ptr(x), nt.ptr(got), ctypes.c_long(x.shape[0]))
        nt.assert_bits(got, nt.or_unary(nm, x), nm)


def test_hostmath_erf_accuracy():
    L = ctypes.CDLL(os.path.join(nt.ROOT, "tests", "native", "libhostmath.so"))
    g = nt.load_golden("transc.npz")
    x, ex = g["erf_in"], g["erf_exact"]
    got = np.empty_like(x)
    L.h_erf(nt.ptr(x), nt.ptr(got), ctypes.c_long(x.shape[0]))
    fin = np.isfinite(ex[:, 0]) & (np.abs(ex[:, 0]) > 1e-290)
    assert nt.rel_err_vs_exact(got[fin], ex[fin]).max() <= 1.55e-32
