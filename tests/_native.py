"""ctypes access to the oracle (test infrastructure) and helpers for tests.

Only tests, __graft_entry__.smoke() and bench.py's CPU-baseline leg load the
oracle; the product never does.
"""
import ctypes
import os

import numpy as np

from multifloats_amd import datagen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
# MF_ORACLE_LIB / MF_HOSTMATH_LIB select other builds of the same sources
# (tests/sanitize_host.sh points them at sanitizer builds)
ORACLE = os.environ.get("MF_ORACLE_LIB") or os.path.join(ROOT, "oracle", "liboracle.so")
HOSTMATH = os.environ.get("MF_HOSTMATH_LIB") or os.path.join(ROOT, "tests", "native", "libhostmath.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libmfref.so")

P, I64 = ctypes.c_void_p, ctypes.c_int64


class F2(ctypes.Structure):
    _fields_ = [("limbs", ctypes.c_double * 2)]


_or = None

# transcendentals restated bit for bit after exp/log/sin (golden file transc2.npz)
TRANSC2 = ("cos", "tan", "exp2", "expm1", "log2", "log10", "log1p", "sinh", "cosh", "tanh")
TRANSC3 = ("asinh", "acosh", "atanh")
TRANSC4 = ("exp10", "exp2m1", "exp10m1", "log2p1", "log10p1")
TRANSC5 = ("sinpi", "cospi", "tanpi")
BINARY2 = ("pow", "hypot", "fdim", "fmin", "fmax", "copysign")
INVTRIG = ("atan", "atanpi", "atan2", "atan2pi")
# worst relative error of this build on the DD-accurate rows of invtrig.npz
INVTRIG_CAP = {"atan": 1.3e-32, "atanpi": 4.0e-32, "atan2": 5.5e-32, "atan2pi": 6.0e-32, "asin": 1.3e-32,
               "acos": 2.1e-32, "asinpi": 3.3e-32, "acospi": 4.0e-32}
INVTRIG3 = ("asin", "acos", "asinpi", "acospi")


def oracle():
    global _or
    if _or is None:
        L = ctypes.CDLL(ORACLE)
        for nm in ("or_add", "or_sub", "or_mul", "or_div") + tuple("or_" + f for f in BINARY2):
            getattr(L, nm).argtypes = [P, P, P, I64]
        for nm in ("or_sqrt", "or_exp", "or_log", "or_sin") + tuple("or_" + f for f in TRANSC2 + TRANSC3 + TRANSC4 + TRANSC5):
            getattr(L, nm).argtypes = [P, P, I64]
        L.or_matmul_mm.argtypes = [P, P, P, I64, I64, I64, I64]
        L.or_cmatmul_mm.argtypes = [P, P, P, I64, I64, I64, I64]
        L.or_matmul_mv.argtypes = [P, P, P, I64, I64, I64]
        L.or_matmul_vm.argtypes = [P, P, P, I64, I64, I64]
        L.or_dot.argtypes = [P, P, I64, I64]
        L.or_dot.restype = F2
        L.or_sum.argtypes = [P, I64]
        L.or_sum.restype = F2
        L.or_norm2.argtypes = [P, I64]
        L.or_norm2.restype = F2
        L.or_sum_mask.argtypes = [P, P, I64]
        L.or_sum_mask.restype = F2
        L.or_cdot.argtypes = [P, P, I64, I64, P]
        L.or_sum_dim.argtypes = [P, P, I64, I64, I64, P]
        L.or_norm2_dim.argtypes = [P, I64, I64, I64, P]
        L.or_fibre_reduce.argtypes = [ctypes.c_int, P, P, I64, P]
        L.or_fibre_reduce_dim.argtypes = [ctypes.c_int, P, P, I64, I64, I64, P]
        L.or_gemm_blas.argtypes = [ctypes.c_char, ctypes.c_char, I64, I64, I64, P, P, I64, P, I64, P, P, I64,
                                   I64]
        L.or_gemm_blas.restype = ctypes.c_int
        c1 = ctypes.c_char
        L.or_trsm.argtypes = [c1, c1, c1, c1, I64, I64, P, P, I64, P, I64]
        L.or_trsm.restype = ctypes.c_int
        D = ctypes.POINTER(ctypes.c_double)
        L.or_exact_dot_err.argtypes = [P, I64, P, I64, I64, P, D, D]
        L.or_exact_dot.argtypes = [P, I64, P, I64, I64, P]
        L.or_exact_sum_err.argtypes = [P, I64, P, D, D]
        L.or_gemm_err.argtypes = [P, P, P, I64, I64, I64, P, P, I64, P, P]
        _or = L
    return _or


def ptr(a):
    assert a.flags.c_contiguous or a.flags.f_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def dd(a, order="C"):
    return np.require(np.asarray(a, dtype=np.float64), requirements=[order, "A", "W"])


def or_unary(nm, a):
    a = dd(a)
    c = np.empty_like(a)
    getattr(oracle(), "or_" + nm)(ptr(a), ptr(c), a.shape[0])
    return c


def or_binary(nm, a, b):
    a, b = dd(a), dd(b)
    c = np.empty_like(a)
    getattr(oracle(), "or_" + nm)(ptr(a), ptr(b), ptr(c), a.shape[0])
    return c


def cm(a):
    """Column-major float64x2 layout of an (m, n, 2) DD matrix."""
    return datagen.as_colmajor(a)


def mptr(a):
    assert datagen.is_colmajor(a)
    return ctypes.c_void_p(a.ctypes.data)


def or_matmul(A, B, ri):
    m, k = A.shape[:2]
    n = B.shape[1]
    A, B = cm(A), cm(B)
    C = cm(np.zeros((m, n, 2)))
    oracle().or_matmul_mm(mptr(A), mptr(B), mptr(C), m, k, n, ri)
    return C


def or_gemm_blas(ta, tb, m, n, k, alpha, A, B, beta, C, ri, lda=None, ldb=None, ldc=None):
    """The oracle's BLAS-style GEMM on column-major storage arrays (ld, cols, 2);
    C is updated in place. Returns wgemm.f90's info code. lda/ldb/ldc default
    to the storage arrays' first extents."""
    assert datagen.is_colmajor(A) and datagen.is_colmajor(B) and datagen.is_colmajor(C)
    al = np.array(alpha, dtype=np.float64).reshape(-1)
    be = np.array(beta, dtype=np.float64).reshape(-1)
    al = np.array([al[0], al[1] if al.size > 1 else 0.0])
    be = np.array([be[0], be[1] if be.size > 1 else 0.0])
    lda = max(A.shape[0], 1) if lda is None else lda
    ldb = max(B.shape[0], 1) if ldb is None else ldb
    ldc = max(C.shape[0], 1) if ldc is None else ldc
    return oracle().or_gemm_blas(ta.encode(), tb.encode(), m, n, k, ptr(al), mptr(A), lda, mptr(B), ldb, ptr(be),
                                 mptr(C), ldc, ri)


def or_trsm(side, uplo, ta, diag, m, n, alpha, A, B, lda=None, ldb=None):
    """The oracle's xTRSM on column-major storage arrays (ld, cols, 2); B is
    overwritten with X. Returns wgemm-style info (wtrsm.f90:24-47)."""
    assert datagen.is_colmajor(A) and datagen.is_colmajor(B)
    al = np.array(alpha, dtype=np.float64).reshape(-1)
    al = np.array([al[0], al[1] if al.size > 1 else 0.0])
    lda = max(A.shape[0], 1) if lda is None else lda
    ldb = max(B.shape[0], 1) if ldb is None else ldb
    return oracle().or_trsm(side.encode(), uplo.encode(), ta.encode(), diag.encode(), m, n, ptr(al), mptr(A), lda,
                            mptr(B), ldb)


def res(r):
    return np.array([r.limbs[0], r.limbs[1]])


def exact_sum_err(x, c):
    x = dd(x)
    c = dd(np.asarray(c))
    e, m = ctypes.c_double(), ctypes.c_double()
    oracle().or_exact_sum_err(ptr(x), x.shape[0], ptr(c), ctypes.byref(e), ctypes.byref(m))
    return e.value, m.value


def exact_dot_err(a, b, c):
    a, b, c = dd(a), dd(b), dd(np.asarray(c))
    e, m = ctypes.c_double(), ctypes.c_double()
    oracle().or_exact_dot_err(ptr(a), 1, ptr(b), 1, a.shape[0], ptr(c), ctypes.byref(e), ctypes.byref(m))
    return e.value, m.value


def exact_dot(a, b):
    a, b = dd(a), dd(b)
    out = np.zeros(2)
    oracle().or_exact_dot(ptr(a), 1, ptr(b), 1, a.shape[0], ptr(out))
    return out


def gemm_err(A, B, C, rows, cols, threads=1):
    """Exact error |C - AB|_ij and (|A||B|)_ij at the listed outputs; the
    points are split over `threads` host threads (the C routine is pure)."""
    m, k = A.shape[:2]
    n = B.shape[1]
    A, B, C = cm(A), cm(B), cm(C)
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    err = np.zeros(rows.size)
    mag = np.zeros(rows.size)
    O = oracle()

    def part(lo, hi):
        lo, hi = int(lo), int(hi)
        if hi > lo:
            O.or_gemm_err(mptr(A), mptr(B), mptr(C), m, k, n,
                          ctypes.c_void_p(rows.ctypes.data + 8 * lo), ctypes.c_void_p(cols.ctypes.data + 8 * lo),
                          hi - lo, ctypes.c_void_p(err.ctypes.data + 8 * lo),
                          ctypes.c_void_p(mag.ctypes.data + 8 * lo))
    threads = max(1, min(int(threads), rows.size))
    if threads == 1:
        part(0, rows.size)
    else:
        from concurrent.futures import ThreadPoolExecutor
        cuts = np.linspace(0, rows.size, threads + 1).astype(int)
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda t: part(cuts[t], cuts[t + 1]), range(threads)))
    return err, mag


def or_sum_mask(x, mask):
    x = dd(x)
    m = np.ascontiguousarray(mask, dtype=np.int32)
    return res(oracle().or_sum_mask(ptr(x), ptr(m), x.shape[0]))


def or_cdot(a, b, ri):
    """conj(a).b for (n, 4) complex DD vectors -> [re.hi, re.lo, im.hi, im.lo]."""
    a, b = dd(a), dd(b)
    out = np.zeros(4)
    oracle().or_cdot(ptr(a), ptr(b), a.shape[0], ri, ptr(out))
    return out


def or_reduce_dim(kind, buf, mask, inner, length, outer):
    """Oracle sum / norm2 along the middle axis of an (inner, len, outer)
    column-major DD buffer (flat (inner*len*outer, 2)); returns (inner*outer, 2)."""
    buf = dd(buf)
    out = np.zeros((max(inner * outer, 1), 2))
    if kind == "sum":
        mb = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
        oracle().or_sum_dim(ptr(buf), None if mb is None else ptr(mb), inner, length, outer, ptr(out))
    else:
        oracle().or_norm2_dim(ptr(buf), inner, length, outer, ptr(out))
    return out[:inner * outer]


FIBRE_KIND = {"product": 0, "maxval": 1, "minval": 2}


def or_fibre(kind, x, mask=None):
    """The oracle's sequential product / maxval / minval of a rank-1 array."""
    x = dd(x)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
    out = np.zeros(2)
    oracle().or_fibre_reduce(FIBRE_KIND[kind], ptr(x), None if m is None else ptr(m), x.shape[0], ptr(out))
    return out


def or_fibre_dim(kind, buf, mask, inner, length, outer):
    buf = dd(buf)
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.int32)
    out = np.zeros((inner * outer, 2))
    oracle().or_fibre_reduce_dim(FIBRE_KIND[kind], ptr(buf), None if m is None else ptr(m), inner, length, outer,
                                 ptr(out))
    return out


def bits_equal(x, y):
    """Bit-for-bit equality of DD arrays, with any NaN matching any NaN."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    nx, ny = np.isnan(x), np.isnan(y)
    same = (x.view(np.uint64) == y.view(np.uint64)) | (nx & ny)
    return same


def assert_bits(x, y, what=""):
    same = bits_equal(x, y)
    if not same.all():
        bad = np.argwhere(~same)[:5]
        raise AssertionError("%s: %d limbs differ; first %s: got %r want %r" % (
            what, int((~same).sum()), bad.tolist(),
            [np.asarray(x)[tuple(b)] for b in bad], [np.asarray(y)[tuple(b)] for b in bad]))


def rel_err_vs_exact(got, exact3):
    """|got - exact| / |exact| with exact as three limbs, for finite exact values.

    Computed in exact rational arithmetic via Python floats -> fractions would
    be slow; limb-wise subtraction is exact enough: (got.hi - e0) is exact when
    got.hi and e0 are within a factor 2 (Sterbenz), then the small terms.
    """
    g = np.asarray(got, dtype=np.float64)
    e = np.asarray(exact3, dtype=np.float64)
    d = (g[:, 0] - e[:, 0]) + (g[:, 1] - e[:, 1]) - e[:, 2]
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.abs(d) / np.abs(e[:, 0])


# relative-error caps of dd_erfc / dd_erfcx where the result is a normal
# double-double; the reference's own maxima on the same fixture are 6.2e-32 and
# 6.1e-32 (erfcx's worst case, 1/(x sqrt(pi)) for large x, is the final DD
# division in both)
ERFC_FAMILY_CAP = {"erfc": 4.5e-32, "erfcx": 6.5e-32}


def erfc_family_errors(got, g, nm):
    """(our max rel error, the reference's max rel error) on the fixture's
    inputs whose exact value is a normal double-double (|v| > 2^-960)."""
    ex = g[nm + "_exact"]
    fin = np.isfinite(ex[:, 0]) & (np.abs(ex[:, 0]) > 2.0 ** -960)
    return rel_err_vs_exact(got[fin], ex[fin]).max(), rel_err_vs_exact(g[nm + "_ref"][fin], ex[fin]).max()


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def invtrig_fixture(nm):
    return load_golden("invtrig3.npz" if nm in INVTRIG3 else "invtrig.npz")


def invtrig_eval(fn, nm, g):
    """Apply fn (one- or two-argument, numpy in/out) to the inputs of
    invtrig.npz / invtrig3.npz for function nm."""
    return fn(g["x"]) if nm in ("atan", "atanpi") + INVTRIG3 else fn(g["y2"], g["x2"])
