"""CPU checks of the product library and host logic (no GPU calls)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import multifloats_amd as mf
from multifloats_amd import datagen
from tests import _native as nt

ROOT = nt.ROOT

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
    L = ctypes.CDLL(nt.HOSTMATH)
    rng = np.random.default_rng(4)
    gens = {"exp": datagen.exp_inputs, "log": datagen.log_inputs, "sin": datagen.sin_inputs}
    for nm, gen in gens.items():
        x = gen(rng, 20000)
        got = np.empty_like(x)
        getattr(L, "h_" + nm)(nt.ptr(x), nt.ptr(got), ctypes.c_long(x.shape[0]))
        nt.assert_bits(got, nt.or_unary(nm, x), nm)


def test_hostmath_erf_accuracy():
    L = ctypes.CDLL(nt.HOSTMATH)
    g = nt.load_golden("transc.npz")
    x, ex = g["erf_in"], g["erf_exact"]
    got = np.empty_like(x)
    L.h_erf(nt.ptr(x), nt.ptr(got), ctypes.c_long(x.shape[0]))
    fin = np.isfinite(ex[:, 0]) & (np.abs(ex[:, 0]) > 1e-290)
    assert nt.rel_err_vs_exact(got[fin], ex[fin]).max() <= 1.55e-32


@pytest.mark.parametrize("nm", ["erfc", "erfcx"])
def test_hostmath_erfc_family_accuracy(nm):
    """The host build of dd_erfc / dd_erfcx is no less accurate than the
    reference's erfcdd / erfcxdd on the same inputs (320-bit values), and
    below 4.5e-32 (erfc) / 6.5e-32 (erfcx) relative wherever the result is a
    normal double-double."""
    L = ctypes.CDLL(nt.HOSTMATH)
    g = nt.load_golden("special2.npz")
    x = g[nm + "_in"]
    got = np.empty_like(x)
    getattr(L, "h_" + nm)(nt.ptr(x), nt.ptr(got), ctypes.c_long(x.shape[0]))
    ours, ref = nt.erfc_family_errors(got, g, nm)
    assert ours <= ref and ours <= nt.ERFC_FAMILY_CAP[nm], (ours, ref)


FORTRAN_DIR = os.path.join(ROOT, "multifloats_amd", "fortran")


def test_fortran_bridge_source_is_generated():
    """multifloats_hip.F90 is exactly what gen_bridge.py writes."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("gen_bridge", os.path.join(FORTRAN_DIR, "gen_bridge.py"))
    g = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(g)
    with open(os.path.join(FORTRAN_DIR, "multifloats_hip.F90")) as fh:
        assert fh.read() == g.gen()


def test_fortran_bridge_binds_only_exported_symbols():
    """Every bind(c, name=...) in the bridge names a symbol the header declares
    and the library exports."""
    import re
    with open(os.path.join(FORTRAN_DIR, "multifloats_hip.F90")) as fh:
        names = set(re.findall(r"bind\(c, name='([a-z0-9_]+)'\)", fh.read()))
    assert names
    with open(os.path.join(ROOT, "include", "multifloats_hip.h")) as fh:
        hdr = fh.read()
    L = ctypes.CDLL(os.path.join(ROOT, "multifloats_amd", "libmultifloats_hip.so"))
    for nm in names:
        assert re.search(r"\b%s\(" % nm, hdr), nm
        assert hasattr(L, nm), nm


def test_fortran_bridge_fails_loudly_without_gpu(tmp_path):
    """With no GPU the bridge's calls report the HIP error and return NaN
    (no CPU fallback)."""
    exe = os.path.join(ROOT, "tests", "fortran", "_build", "bridge_check")
    if not os.path.exists(exe):
        pytest.skip("bridge_check not built")
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    m, k, n, nv = 2, 3, 2, 4
    fin = tmp_path / "in.bin"
    with open(fin, "wb") as fh:
        np.array([m, k, n, nv], dtype=np.int64).tofile(fh)
        np.ones(2 * (m * k + k * n) + 4 * (m * k + k * n) + 2 * 2 * nv + 4 * 2 * nv).tofile(fh)
        np.ones(nv, dtype=np.int32).tofile(fh)
    out = tmp_path / "out.bin"
    p = subprocess.run([exe, str(fin), str(out)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0
    assert "multifloats_hip: matmuldd_mm failed" in p.stderr
    vals = np.fromfile(out, dtype=np.float64)
    # the last record is wgemm's C, (m + 2) x n with two padding rows that are
    # not outputs and keep their input value
    wg = vals[-(m + 2) * n * 2:].reshape(n, m + 2, 2)
    assert "multifloats_hip: wgemmdd failed" in p.stderr
    assert vals.size > wg.size and np.isnan(vals[:-wg.size]).all()
    assert np.isnan(wg[:, :m]).all() and (wg[:, m:] == 1.0).all()


def test_grouped_tile_choice_minimises_padded_work():
    """The tuned planner gives tall, wide and small problems the tile shapes
    that waste the fewest lanes (equal rates: pure padding), and follows the
    rates of each problem's k class."""
    from multifloats_amd import autotune
    nv, nc = len(autotune.TILES), len(autotune.K_EDGES) - 1
    table = {"class_rates": [[1.0] * nv for _ in range(nc)]}
    sizes = np.array([[400, 64, 10], [10, 64, 400], [64, 64, 64], [20, 64, 20], [120, 30, 60], [8, 8, 8],
                      [96, 200, 96]])
    v = autotune.plan_variants(sizes, table)
    assert autotune.TILES[v[0]][1] == 16 and autotune.TILES[v[1]][0] == 16
    assert autotune.TILES[v[5]] == (16, 16)
    assert autotune.TILES[v[6]] == (48, 48) or autotune.padded_macs(96, 200, 96, v[6]) == 96 * 200 * 96
    for (m, k, n), q in zip(sizes, v):
        assert autotune.padded_macs(m, k, n, q) == min(autotune.padded_macs(m, k, n, w) for w in autotune.VARIANTS)
    # a faster variant wins where padding ties, per k class
    fast = [[1.0] * nv for _ in range(nc)]
    fast[autotune.k_class(64)][1] = 2.0
    assert autotune.plan_variants(np.array([[64, 64, 64], [64, 8, 64]]), {"class_rates": fast})[0] == 1
    assert autotune.plan_variants(np.array([[64, 8, 64]]), {"class_rates": fast})[0] != 1
    assert [autotune.k_class(k) for k in (8, 23, 24, 63, 64, 159, 160, 512)] == [0, 0, 1, 1, 2, 2, 3, 3]


def _host(nm):
    L = ctypes.CDLL(nt.HOSTMATH)
    f = getattr(L, "h_" + nm)

    def run(*args):
        out = np.empty_like(args[0])
        f(*[nt.ptr(np.ascontiguousarray(a)) for a in args], nt.ptr(out), ctypes.c_long(args[0].shape[0]))
        return out
    return run


@pytest.mark.parametrize("nm", nt.INVTRIG + nt.INVTRIG3)
def test_hostmath_atan_family_accuracy(nm):
    """The own atan (and the reference's atan2 / atanpi wrappers around it)
    against 320-bit values: no less accurate than the reference's atandd ...
    atan2pidd on the same inputs, below a fixed cap on every row the
    reference gets to DD accuracy, and never worse than the reference on the
    rows where the wrapper's y / x leaves the double range. asin / acos and
    their pi-scaled forms run through the same atan reduction."""
    g = nt.invtrig_fixture(nm)
    got = nt.invtrig_eval(_host(nm), nm, g)
    ex, ref = g[nm + "_exact"], g[nm + "_ref"]
    fin = np.isfinite(ex[:, 0]) & (np.abs(ex[:, 0]) > 1e-290)
    eo, er = nt.rel_err_vs_exact(got[fin], ex[fin]), nt.rel_err_vs_exact(ref[fin], ex[fin])
    dd = er < 1e-30
    assert dd.sum() > 3000
    assert eo[dd].max() <= er[dd].max() and eo[dd].max() <= nt.INVTRIG_CAP[nm], (eo[dd].max(), er[dd].max())
    assert (eo[~dd] <= er[~dd]).all()


@pytest.mark.parametrize("nm", nt.INVTRIG + nt.INVTRIG3)
def test_hostmath_atan_family_special_values_match_reference_bits(nm):
    """Non-finite arguments, zeros with and without a low limb, |x| > 1 for
    asin / acos, and every pair of the fixture's special values give the
    reference's bits."""
    g = nt.invtrig_fixture(nm)
    got = nt.invtrig_eval(_host(nm), nm, g)
    ex = g[nm + "_exact"]
    sp = ~(np.isfinite(ex[:, 0]) & (np.abs(ex[:, 0]) > 1e-290))
    assert sp.sum() > 5
    nt.assert_bits(got[sp], g[nm + "_ref"][sp], nm)


@pytest.mark.parametrize("nm", ["erf", "erfc", "erfcx", "atan", "asin", "acos"])
def test_hostmath_own_designs_against_mpfr(nm):
    """The host build of the functions of this project's own design against
    256-bit MPFR values on >= 100k points each (every row and table entry,
    both sides of each row edge and switch point, tiny and huge arguments):
    within the cap, and no less accurate than the compiled reference on the
    same points."""
    from tests import pin_points as pp
    L = ctypes.CDLL(nt.HOSTMATH)
    x, v, ok = pp.points(nm), pp.exact(nm), pp.checked(nm)
    assert x.shape[0] >= 100000 and ok.sum() >= 95000
    got = np.empty_like(x)
    getattr(L, "h_" + nm)(nt.ptr(x), nt.ptr(got), ctypes.c_long(x.shape[0]))
    ours = nt.rel_err_vs_exact(got[ok], v[ok]).max()
    assert ours <= pp.CAPS[nm], (nm, ours)
    ref = pp.reference_outputs(nm)
    if ref is not None:
        assert ours <= nt.rel_err_vs_exact(ref[ok], v[ok]).max(), nm
