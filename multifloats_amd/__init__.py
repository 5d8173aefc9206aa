"""multifloats_amd -- MI355X (gfx950) double-double hot path.

Python mirror of the reference's operator surface for the hot path, over the
C ABI of libmultifloats_hip.so (include/multifloats_hip.h):

  matmul(a, b, renorm_interval=8)   Fortran matmul on real64x2 -> matmuldd_mm/mv/vm
  dot_product(a, b, renorm_interval=8), sum(a), norm2(a)
  exp, log, sin, erf, sqrt, add, sub, mul, div   (elementwise)

DD arrays are numpy float64 arrays whose last axis has length 2 (hi, lo) --
the float64x2 memory layout. Matrices are (m, n, 2) arrays in the column-major
float64x2 layout, element (i, j) at byte offset 16*(i + m*j) (datagen.matrix,
datagen.as_colmajor); other layouts are copied into it. The default renorm
interval 8 is the Fortran module's DD_FMA_RENORM_INTERVAL
(fsrc/multifloats.fypp:664); the C++ matmul_* default is 0 (float64x2.h:2500).

The device-pointer functions (dev_*) take integer device addresses (e.g.
torch tensor .data_ptr()) and a HIP stream handle; they enqueue only.

There is no CPU fallback: if the library or a GPU is missing, calls raise.
"""
import ctypes
import os

import numpy as np

from . import datagen

_HERE = os.path.dirname(os.path.abspath(__file__))
# MULTIFLOATS_HIP_LIB selects another build of the library (e.g. one with host-side sanitizers)
LIB_PATH = os.environ.get("MULTIFLOATS_HIP_LIB") or os.path.join(_HERE, "libmultifloats_hip.so")

DD_FMA_RENORM_INTERVAL = 8

MFHIP_SQRT, MFHIP_EXP, MFHIP_LOG, MFHIP_SIN, MFHIP_ERF, MFHIP_NEG = range(6)
(MFHIP_COS, MFHIP_TAN, MFHIP_EXP2, MFHIP_EXPM1, MFHIP_LOG2, MFHIP_LOG10, MFHIP_LOG1P, MFHIP_SINH, MFHIP_COSH,
 MFHIP_TANH, MFHIP_ERFC, MFHIP_ERFCX, MFHIP_ASINH, MFHIP_ACOSH, MFHIP_ATANH, MFHIP_EXP10, MFHIP_EXP2M1,
 MFHIP_EXP10M1, MFHIP_LOG2P1, MFHIP_LOG10P1, MFHIP_SINPI, MFHIP_COSPI, MFHIP_TANPI, MFHIP_ATAN, MFHIP_ATANPI,
 MFHIP_ASIN, MFHIP_ACOS, MFHIP_ASINPI, MFHIP_ACOSPI) = range(6, 35)
(MFHIP_ADD, MFHIP_SUB, MFHIP_MUL, MFHIP_DIV, MFHIP_POW, MFHIP_HYPOT, MFHIP_FDIM, MFHIP_FMIN, MFHIP_FMAX,
 MFHIP_COPYSIGN, MFHIP_ATAN2, MFHIP_ATAN2PI, MFHIP_MINF, MFHIP_MAXF, MFHIP_DIMF, MFHIP_FMOD, MFHIP_MODULO,
 MFHIP_NEAREST) = range(18)
MFHIP_DOT, MFHIP_SUM, MFHIP_NORM2, MFHIP_PRODUCT, MFHIP_MAXVAL, MFHIP_MINVAL = range(6)
NOALLOC_MAX_N = 1 << 30  # MFHIP_NOALLOC_MAX_N


class float64x2(ctypes.Structure):
    _fields_ = [("limbs", ctypes.c_double * 2)]


class complex64x2(ctypes.Structure):
    _fields_ = [("re", float64x2), ("im", float64x2)]


_lib = None
_TRANSC2 = ("cos", "tan", "exp2", "expm1", "log2", "log10", "log1p", "sinh", "cosh", "tanh")


def lib():
    """Load libmultifloats_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("multifloats_amd: %s is not built (run __graft_entry__.build())" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        P, I64, VP, I = ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int
        for nm in ("matmuldd_mm", "matmulf64x2_mm"):
            getattr(L, nm).argtypes = [P, P, P, I64, I64, I64, I64]
            getattr(L, nm).restype = None
        for nm in ("matmuldd_mv", "matmulf64x2_mv", "matmuldd_vm", "matmulf64x2_vm"):
            getattr(L, nm).argtypes = [P, P, P, I64, I64, I64]
            getattr(L, nm).restype = None
        for nm in ("adddd_v", "subdd_v", "muldd_v", "divdd_v", "powdd_v", "hypotdd_v", "fdimdd_v", "fmindd_v",
                   "fmaxdd_v", "copysigndd_v", "atan2dd_v", "atan2pidd_v", "mindd_v", "maxdd_v", "dimdd_v",
                   "fmoddd_v", "modulodd_v", "nearestdd_v"):
            getattr(L, nm).argtypes = [P, P, P, I64]
            getattr(L, nm).restype = None
        for nm in ("sqrtdd_v", "expdd_v", "logdd_v", "sindd_v", "erfdd_v", "erfcdd_v", "erfcxdd_v", "asinhdd_v",
                   "acoshdd_v", "atanhdd_v", "exp10dd_v", "exp2m1dd_v", "exp10m1dd_v", "log2p1dd_v",
                   "log10p1dd_v", "sinpidd_v", "cospidd_v", "tanpidd_v", "atandd_v", "atanpidd_v",
                   "asindd_v", "acosdd_v", "asinpidd_v", "acospidd_v") + tuple(f + "dd_v" for f in _TRANSC2):
            getattr(L, nm).argtypes = [P, P, I64]
            getattr(L, nm).restype = None
        L.sumdd.argtypes = [P, I64]
        L.sumdd.restype = float64x2
        L.norm2dd.argtypes = [P, I64]
        L.norm2dd.restype = float64x2
        L.dotdd.argtypes = [P, P, I64, I64]
        L.dotdd.restype = float64x2
        L.sumdd_mask.argtypes = [P, P, I64]
        L.sumdd_mask.restype = float64x2
        L.cdotdd.argtypes = [P, P, I64, I64]
        L.cdotdd.restype = complex64x2
        L.sumdd_dim.argtypes = [P, P, I64, I64, I64, P]
        L.sumdd_dim.restype = None
        for nm in ("productdd", "maxvaldd", "minvaldd"):
            getattr(L, nm).argtypes = [P, I64]
            getattr(L, nm).restype = float64x2
            getattr(L, nm + "_mask").argtypes = [P, P, I64]
            getattr(L, nm + "_mask").restype = float64x2
            getattr(L, nm + "_dim").argtypes = [P, P, I64, I64, I64, P]
            getattr(L, nm + "_dim").restype = None
        L.mfhip_reduce_masked.argtypes = [I, P, P, I64, P, P, VP]
        L.norm2dd_dim.argtypes = [P, I64, I64, I64, P]
        L.norm2dd_dim.restype = None
        L.mfhip_sum_masked.argtypes = [P, P, I64, P, P, VP]
        L.mfhip_cdot_workspace_bytes.argtypes = [I64]
        L.mfhip_cdot_workspace_bytes.restype = ctypes.c_size_t
        L.mfhip_cdot.argtypes = [P, P, I64, I64, P, P, VP]
        L.mfhip_reduce_dim.argtypes = [I, P, P, I64, I64, I64, P, VP]
        L.mfhip_matmul_mm.argtypes = [P, P, P, I64, I64, I64, I64, VP]
        L.mfhip_matmul_mm_variant.argtypes = [P, P, P, I64, I64, I64, I64, I, VP]
        L.mfhip_matmul_variants.argtypes = []
        L.mfhip_matmul_mv.argtypes = [P, P, P, I64, I64, I64, VP]
        L.mfhip_matmul_vm.argtypes = [P, P, P, I64, I64, I64, VP]
        L.mfhip_unary.argtypes = [I, P, P, I64, VP]
        L.mfhip_binary.argtypes = [I, P, P, P, I64, VP]
        L.mfhip_reduce_workspace_bytes.argtypes = [I64]
        L.mfhip_reduce_workspace_bytes.restype = ctypes.c_size_t
        L.mfhip_reduce.argtypes = [I, P, P, I64, I64, P, P, VP]
        L.mfhip_reduce_noalloc.argtypes = [I, P, P, I64, I64, P, VP]
        L.matmuldd_mm_ozaki.argtypes = [P, P, P, I64, I64, I64, I]
        L.matmuldd_mm_ozaki.restype = None
        L.mfhip_ozaki_workspace_bytes.argtypes = [I64, I64, I64, I]
        L.mfhip_ozaki_workspace_bytes.restype = ctypes.c_size_t
        L.mfhip_ozaki_matmul_mm.argtypes = [P, P, P, I64, I64, I64, I, P, ctypes.c_size_t, VP]
        L.cmatmuldd_mm.argtypes = [P, P, P, I64, I64, I64, I64]
        L.cmatmuldd_mv.argtypes = [P, P, P, I64, I64, I64]
        L.cmatmuldd_vm.argtypes = [P, P, P, I64, I64, I64]
        L.mfhip_cmatmul_mm.argtypes = [P, P, P, I64, I64, I64, I64, VP]
        L.mfhip_ozaki_slices_used.argtypes = []
        L.mfhip_ozaki_fixups.argtypes = []
        L.mfhip_ozaki_fixups.restype = ctypes.c_int64
        L.mfhip_ozaki_info.argtypes = [P, I64, I64, I64, I, ctypes.POINTER(I), ctypes.POINTER(I64), VP]
        L.mfhip_ozaki_moduli_used.argtypes = [P, I64, I64, I64, I, ctypes.POINTER(I), VP]
        L.mfhip_ozaki_choose_slices.argtypes = [I64, I, I]
        L.mfhip_ozaki_moduli.argtypes = [I64, I64, I64, I]
        L.mfhip_ozaki_prepare.argtypes = [P, P, I64, I64, I64, I, P, ctypes.c_size_t, VP]
        L.mfhip_ozaki_multiply.argtypes = [P, I64, I64, I64, I, P, ctypes.c_size_t, VP]
        L.mfhip_ozaki_products.argtypes = [I64, I64, I64, I, P, ctypes.c_size_t, VP]
        C = ctypes.c_char
        L.wgemmdd.argtypes = [C, C, I64, I64, I64, float64x2, P, I64, P, I64, float64x2, P, I64, I64]
        L.wgemmdd.restype = None
        L.mfhip_gemm.argtypes = [C, C, I64, I64, I64, float64x2, P, I64, P, I64, float64x2, P, I64, I64, VP]
        L.wtrsmdd.argtypes = [C, C, C, C, I64, I64, float64x2, P, I64, P, I64]
        L.wtrsmdd.restype = None
        L.mfhip_trsm_workspace_bytes.argtypes = [C, I64, I64]
        L.mfhip_trsm_workspace_bytes.restype = ctypes.c_size_t
        L.mfhip_trsm.argtypes = [C, C, C, C, I64, I64, float64x2, P, I64, P, I64, P, ctypes.c_size_t, VP]
        L.mfhip_last_error.argtypes = []
        L.mfhip_error_string.argtypes = [I]
        L.mfhip_error_string.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _check_status():
    e = lib().mfhip_last_error()
    if e:
        raise RuntimeError("multifloats_hip: %s" % lib().mfhip_error_string(e).decode())


def _dd(x):
    a = np.require(np.asarray(x, dtype=np.float64), requirements=["C", "A"])
    if a.shape[-1:] != (2,):
        raise ValueError("DD arrays need a trailing axis of length 2 (hi, lo)")
    return a


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _mat(a):
    a = datagen.as_colmajor(a)
    if a.shape[-1:] != (2,):
        raise ValueError("DD matrices are (m, n, 2) arrays")
    return a


def matmul(a, b, renorm_interval=DD_FMA_RENORM_INTERVAL):
    """Fortran matmul on real64x2 (fsrc/multifloats.fypp:2568-2593).

    a: (m, k, 2) column-major or (k, 2) vector; b: (k, n, 2) or (k, 2).
    """
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.ndim == 3 and b.ndim == 3:
        m, k = a.shape[:2]
        if b.shape[0] != k:
            raise ValueError("matmul: inner dimensions differ")
        n = b.shape[1]
        a, b = _mat(a), _mat(b)
        c = _mat(np.zeros((m, n, 2)))
        lib().matmuldd_mm(_ptr(a), _ptr(b), _ptr(c), m, k, n, renorm_interval)
    elif a.ndim == 3 and b.ndim == 2:
        m, k = a.shape[:2]
        if b.shape[0] != k:
            raise ValueError("matmul: inner dimensions differ")
        a, b = _mat(a), _dd(b)
        c = np.zeros((m, 2))
        lib().matmuldd_mv(_ptr(a), _ptr(b), _ptr(c), m, k, renorm_interval)
    elif a.ndim == 2 and b.ndim == 3:
        k, n = b.shape[:2]
        if a.shape[0] != k:
            raise ValueError("matmul: inner dimensions differ")
        a, b = _dd(a), _mat(b)
        c = np.zeros((n, 2))
        lib().matmuldd_vm(_ptr(a), _ptr(b), _ptr(c), k, n, renorm_interval)
    else:
        raise ValueError("matmul: need (matrix, matrix), (matrix, vector) or (vector, matrix)")
    _check_status()
    return c


def _f2(x):
    """A DD scalar as float64x2: a float, or a (hi, lo) pair."""
    v = np.asarray(x, dtype=np.float64).reshape(-1)
    if v.size == 1:
        return float64x2((ctypes.c_double * 2)(float(v[0]), 0.0))
    if v.size != 2:
        raise ValueError("a DD scalar is a float or a (hi, lo) pair")
    return float64x2((ctypes.c_double * 2)(float(v[0]), float(v[1])))


def _trans(t):
    """A BLAS option letter (trans, side, uplo, diag) as a C char."""
    if not isinstance(t, str) or len(t) != 1:
        raise ValueError("BLAS options are single letters, e.g. 'N', 'T', 'L', 'U'")
    return t.encode()


def gemm(transa, transb, alpha, a, b, beta=0.0, c=None, m=None, n=None, k=None,
         renorm_interval=DD_FMA_RENORM_INTERVAL):
    """BLAS xGEMM on real64x2: C = alpha op(A) op(B) + beta C (the interface of
    the reference's shim test/integration/blas/wgemm.f90:1-14).

    a, b, c are column-major storage arrays of shape (ld, cols, 2); their first
    axis is the leading dimension, so op(A) may be the leading m x k block of a
    larger array. m, n, k default to the full op(A), op(B). c is updated in
    place when it is already column-major float64 (else a column-major copy is
    made) and returned; c=None means beta C = 0 on a fresh (m, n) array.
    alpha, beta: floats or (hi, lo) pairs. The argument errors of wgemm.f90:16-51
    raise RuntimeError (C is then NaN)."""
    ta, tb = _trans(transa), _trans(transb)
    nota, notb = ta in (b"N", b"n"), tb in (b"N", b"n")
    a, b = _mat(a), _mat(b)
    if m is None:
        m = a.shape[0] if nota else a.shape[1]
    if k is None:
        k = a.shape[1] if nota else a.shape[0]
    if n is None:
        n = b.shape[1] if notb else b.shape[0]
    if (b.shape[0] if notb else b.shape[1]) < k:
        raise ValueError("gemm: op(B) has fewer than k rows")
    if c is None:
        c = _mat(np.zeros((max(m, 1), n, 2)))
        beta = 0.0
    elif not datagen.is_colmajor(c):
        c = _mat(c)
    # The library reads ld * (cols - 1) + rows elements of each operand; the
    # leading dimensions come from the arrays (and are checked by the
    # library), the column counts must be checked here.
    okt = (b"N", b"n", b"T", b"t", b"C", b"c")
    if m > 0 and n > 0:  # invalid option letters are left to the library's own check
        if k > 0 and ta in okt and a.shape[1] < (k if nota else m):
            raise ValueError("gemm: a has fewer columns than op(A) needs")
        if k > 0 and tb in okt and b.shape[1] < (n if notb else k):
            raise ValueError("gemm: b has fewer columns than op(B) needs")
        if c.shape[1] < n:
            raise ValueError("gemm: c has fewer than n columns")
    lib().wgemmdd(ta, tb, m, n, k, _f2(alpha), _ptr(a), max(a.shape[0], 1), _ptr(b), max(b.shape[0], 1),
                  _f2(beta), _ptr(c), max(c.shape[0], 1), renorm_interval)
    _check_status()
    return c


def trsm(side, uplo, transa, diag, alpha, a, b, m=None, n=None):
    """BLAS xTRSM on real64x2, bit-identical to the reference's shim
    test/integration/blas/wtrsm.f90: solves op(A) X = alpha B (side 'L') or
    X op(A) = alpha B (side 'R') for X, which overwrites b.

    a, b are column-major storage arrays (ld, cols, 2) (first axis = leading
    dimension); m, n default to b's logical extent. b is updated in place when
    it is already column-major float64 (else a column-major copy is made) and
    returned. The argument errors of wtrsm.f90:24-47 raise RuntimeError (B is
    then NaN)."""
    a = _mat(a)
    if not datagen.is_colmajor(b):
        b = _mat(b)
    m = b.shape[0] if m is None else m
    n = b.shape[1] if n is None else n
    sd = _trans(side)
    if m > 0 and n > 0:
        if b.shape[1] < n:
            raise ValueError("trsm: b has fewer than n columns")
        if sd in (b"L", b"l", b"R", b"r") and a.shape[1] < (m if sd in (b"L", b"l") else n):
            raise ValueError("trsm: a has fewer columns than the triangle needs")
    lib().wtrsmdd(_trans(side), _trans(uplo), _trans(transa), _trans(diag), m, n, _f2(alpha), _ptr(a),
                  max(a.shape[0], 1), _ptr(b), max(b.shape[0], 1))
    _check_status()
    return b


def _cmat(a):
    """(m, n, 4) complex DD matrix {re.hi, re.lo, im.hi, im.lo}, column-major."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim != 3 or a.shape[2] != 4:
        raise ValueError("complex DD matrices are (m, n, 4) arrays")
    return np.ascontiguousarray(a.transpose(1, 0, 2)).transpose(1, 0, 2)


def cmatmul(a, b, renorm_interval=DD_FMA_RENORM_INTERVAL):
    """Fortran matmul on cmplx64x2 (cdd_matmul_mm/mv/vm, fypp:2609-2649).

    Matrices are (m, n, 4) column-major; vectors (k, 4)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.ndim == 3 and b.ndim == 3:
        m, k = a.shape[:2]
        n = b.shape[1]
        if b.shape[0] != k:
            raise ValueError("cmatmul: inner dimensions differ")
        a, b = _cmat(a), _cmat(b)
        c = _cmat(np.zeros((m, n, 4)))
        lib().cmatmuldd_mm(_ptr(a), _ptr(b), _ptr(c), m, k, n, renorm_interval)
    elif a.ndim == 3 and b.ndim == 2:
        m, k = a.shape[:2]
        a, b = _cmat(a), np.ascontiguousarray(b)
        c = np.zeros((m, 4))
        lib().cmatmuldd_mv(_ptr(a), _ptr(b), _ptr(c), m, k, renorm_interval)
    elif a.ndim == 2 and b.ndim == 3:
        k, n = b.shape[:2]
        a, b = np.ascontiguousarray(a), _cmat(b)
        c = np.zeros((n, 4))
        lib().cmatmuldd_vm(_ptr(a), _ptr(b), _ptr(c), k, n, renorm_interval)
    else:
        raise ValueError("cmatmul: need (matrix, matrix), (matrix, vector) or (vector, matrix)")
    _check_status()
    return c


def dev_cmatmul(a, b, c, m, k, n, renorm_interval=DD_FMA_RENORM_INTERVAL, stream=0):
    _dev(lib().mfhip_cmatmul_mm(a, b, c, m, k, n, renorm_interval, stream))


def matmul_ozaki(a, b, slices=0):
    """Ozaki split-precision DD GEMM on INT8 matrix cores (matmuldd_mm_ozaki).

    slices=0 certifies every output to 2^-100 (|A||B|)_ij: the slice count is
    chosen on the device from the data (ozaki_slices_used() tells which) and
    outputs it does not cover are recomputed by a compensated dot
    (ozaki_fixups() tells how many). Explicit slices (7..16) skip the
    certificate."""
    a, b = _mat(a), _mat(b)
    m, k = a.shape[:2]
    if b.shape[0] != k:
        raise ValueError("matmul_ozaki: inner dimensions differ")
    n = b.shape[1]
    c = _mat(np.zeros((m, n, 2)))
    lib().matmuldd_mm_ozaki(_ptr(a), _ptr(b), _ptr(c), m, k, n, slices)
    _check_status()
    return c


def _unary(name, a):
    a = _dd(a)
    c = np.empty_like(a)
    getattr(lib(), name)(_ptr(a), _ptr(c), a.size // 2)
    _check_status()
    return c


def _binary(name, a, b):
    a, b = np.broadcast_arrays(_dd(a), _dd(b))
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    c = np.empty_like(a)
    getattr(lib(), name)(_ptr(a), _ptr(b), _ptr(c), a.size // 2)
    _check_status()
    return c


def exp(a): return _unary("expdd_v", a)
def log(a): return _unary("logdd_v", a)
def sin(a): return _unary("sindd_v", a)
def erf(a): return _unary("erfdd_v", a)
def sqrt(a): return _unary("sqrtdd_v", a)
def cos(a): return _unary("cosdd_v", a)
def tan(a): return _unary("tandd_v", a)
def exp2(a): return _unary("exp2dd_v", a)
def expm1(a): return _unary("expm1dd_v", a)
def log2(a): return _unary("log2dd_v", a)
def log10(a): return _unary("log10dd_v", a)
def log1p(a): return _unary("log1pdd_v", a)
def sinh(a): return _unary("sinhdd_v", a)
def cosh(a): return _unary("coshdd_v", a)
def tanh(a): return _unary("tanhdd_v", a)
def erfc(a): return _unary("erfcdd_v", a)
def erfcx(a): return _unary("erfcxdd_v", a)  # Fortran erfc_scaled
def asinh(a): return _unary("asinhdd_v", a)
def acosh(a): return _unary("acoshdd_v", a)
def atanh(a): return _unary("atanhdd_v", a)
def exp10(a): return _unary("exp10dd_v", a)
def exp2m1(a): return _unary("exp2m1dd_v", a)
def exp10m1(a): return _unary("exp10m1dd_v", a)
def log2p1(a): return _unary("log2p1dd_v", a)
def log10p1(a): return _unary("log10p1dd_v", a)
def sinpi(a): return _unary("sinpidd_v", a)
def cospi(a): return _unary("cospidd_v", a)
def tanpi(a): return _unary("tanpidd_v", a)
def atan(a): return _unary("atandd_v", a)
def atanpi(a): return _unary("atanpidd_v", a)
def asin(a): return _unary("asindd_v", a)
def acos(a): return _unary("acosdd_v", a)
def asinpi(a): return _unary("asinpidd_v", a)
def acospi(a): return _unary("acospidd_v", a)
def add(a, b): return _binary("adddd_v", a, b)
def sub(a, b): return _binary("subdd_v", a, b)
def mul(a, b): return _binary("muldd_v", a, b)
def div(a, b): return _binary("divdd_v", a, b)
def pow(a, b): return _binary("powdd_v", a, b)
def hypot(a, b): return _binary("hypotdd_v", a, b)
def fdim(a, b): return _binary("fdimdd_v", a, b)
def fmin(a, b): return _binary("fmindd_v", a, b)
def fmax(a, b): return _binary("fmaxdd_v", a, b)
def copysign(a, b): return _binary("copysigndd_v", a, b)
def atan2(y, x): return _binary("atan2dd_v", y, x)
def atan2pi(y, x): return _binary("atan2pidd_v", y, x)
def fortran_min(a, b): return _binary("mindd_v", a, b)  # Fortran min(a, b) on real64x2
def fortran_max(a, b): return _binary("maxdd_v", a, b)
def fortran_dim(a, b): return _binary("dimdd_v", a, b)
def fmod(a, b): return _binary("fmoddd_v", a, b)  # C fmod; Fortran mod
def modulo(a, b): return _binary("modulodd_v", a, b)  # Fortran modulo
def nearest(a, b): return _binary("nearestdd_v", a, b)  # Fortran nearest


def _res(r):
    return np.array([r.limbs[0], r.limbs[1]])


def _fortran_order(a):
    """Buffer of a (n1, ..., nr, 2) DD array in Fortran element order
    (first index fastest), limbs innermost: the layout of a real64x2 array."""
    r = a.ndim - 1
    return np.ascontiguousarray(a.transpose(list(range(r - 1, -1, -1)) + [r]))


def _from_fortran_order(buf, shape):
    """Inverse of _fortran_order for a result of logical shape `shape` + (2,)."""
    r = len(shape)
    return buf.reshape(tuple(reversed(shape)) + (2,)).transpose(list(range(r - 1, -1, -1)) + [r])


def _mask_buf(mask, shape):
    m = np.asarray(mask, dtype=bool)
    if m.shape != tuple(shape):
        raise ValueError("mask must have the array's shape")
    r = m.ndim
    return np.ascontiguousarray(m.transpose(list(range(r - 1, -1, -1)))).astype(np.int32)


def _reduce_dim(kind, a, dim, mask):
    shape = a.shape[:-1]
    r = len(shape)
    if not 1 <= dim <= r:
        raise ValueError("dim must lie in 1..%d" % r)
    inner = int(np.prod(shape[:dim - 1], dtype=np.int64))
    ln = int(shape[dim - 1])
    outer = int(np.prod(shape[dim:], dtype=np.int64))
    rest = shape[:dim - 1] + shape[dim:]
    buf = _fortran_order(a)
    out = np.empty((max(inner * outer, 1), 2))
    if kind == "norm2":
        lib().norm2dd_dim(_ptr(buf), inner, ln, outer, _ptr(out))
    else:
        mb = _mask_buf(mask, shape) if mask is not None else None
        getattr(lib(), kind + "dd_dim")(_ptr(buf), _ptr(mb) if mb is not None else None, inner, ln, outer,
                                        _ptr(out))
    _check_status()
    return _from_fortran_order(out[:inner * outer], rest)


def sum(a, dim=None, mask=None):  # noqa: A001 - mirrors the Fortran intrinsic name
    """Fortran sum(array[, dim][, mask]) on a real64x2 array of shape (..., 2)
    (fsrc/multifloats.fypp:1980-2090). Without dim: every element in Fortran
    storage order (first index fastest); with dim (1-based): the reduced
    array, bit-identical to the reference's per-output chains."""
    a = _dd(a)
    if dim is not None and a.ndim > 2:
        return _reduce_dim("sum", a, dim, mask)
    buf = _fortran_order(a).reshape(-1, 2)
    if mask is None:
        r = lib().sumdd(_ptr(buf), buf.shape[0])
    else:
        mb = _mask_buf(mask, a.shape[:-1]).reshape(-1)
        r = lib().sumdd_mask(_ptr(buf), _ptr(mb), buf.shape[0])
    _check_status()
    return _res(r)


def _masked_reduce(kind, a, dim, mask):
    a = _dd(a)
    if dim is not None and a.ndim > 2:
        return _reduce_dim(kind, a, dim, mask)
    buf = _fortran_order(a).reshape(-1, 2)
    mb = _mask_buf(mask, a.shape[:-1]).reshape(-1) if mask is not None else None
    r = getattr(lib(), kind + "dd_mask")(_ptr(buf), _ptr(mb) if mb is not None else None, buf.shape[0])
    _check_status()
    return _res(r)


def product(a, dim=None, mask=None):
    """Fortran product(array[, dim][, mask]) on real64x2 (fypp:1975-2090):
    with dim, bit-identical per-output chains; rank 1, a DD-multiply tree
    within 4 n 2^-104 relative of the exact product."""
    return _masked_reduce("product", a, dim, mask)


def maxval(a, dim=None, mask=None):
    """Fortran maxval(array[, dim][, mask]) on real64x2 (fypp:1783-1862),
    bit-identical to the reference's scan in every form (0 when empty)."""
    return _masked_reduce("maxval", a, dim, mask)


def minval(a, dim=None, mask=None):
    """Fortran minval(array[, dim][, mask]) on real64x2 (fypp:1783-1862)."""
    return _masked_reduce("minval", a, dim, mask)


def dot_product(a, b, renorm_interval=DD_FMA_RENORM_INTERVAL):
    """Fortran dot_product on rank-1 real64x2 arrays (n, 2) (fypp:2254-2308)
    or on cmplx64x2 arrays (n, 4) = {re.hi, re.lo, im.hi, im.lo}, where it is
    sum(conj(a) * b) (fypp:2321-2331)."""
    a = np.require(np.asarray(a, dtype=np.float64), requirements=["C", "A"])
    b = np.require(np.asarray(b, dtype=np.float64), requirements=["C", "A"])
    if a.shape != b.shape or a.ndim != 2 or a.shape[1] not in (2, 4):
        raise ValueError("dot_product: rank-1 arrays of equal size, (n, 2) or (n, 4)")
    if a.shape[1] == 4:
        r = lib().cdotdd(_ptr(a), _ptr(b), a.shape[0], renorm_interval)
        _check_status()
        return np.array([r.re.limbs[0], r.re.limbs[1], r.im.limbs[0], r.im.limbs[1]])
    r = lib().dotdd(_ptr(a), _ptr(b), a.shape[0], renorm_interval)
    _check_status()
    return _res(r)


def norm2(a, dim=None):
    """Fortran norm2(array[, dim]) (fypp:2338-2377); with dim (1-based) the
    results are bit-identical to the reference's per-output chains."""
    a = _dd(a)
    if dim is not None and a.ndim > 2:
        return _reduce_dim("norm2", a, dim, None)
    buf = _fortran_order(a).reshape(-1, 2)
    r = lib().norm2dd(_ptr(buf), buf.shape[0])
    _check_status()
    return _res(r)


# ---- device-pointer API (ints are device addresses) ----

def _dev(rc):
    if rc:
        raise RuntimeError("multifloats_hip: %s" % lib().mfhip_error_string(rc).decode())


def dev_matmul(a, b, c, m, k, n, renorm_interval=DD_FMA_RENORM_INTERVAL, stream=0, variant=-1):
    _dev(lib().mfhip_matmul_mm_variant(a, b, c, m, k, n, renorm_interval, variant, stream))


def dev_gemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
             renorm_interval=DD_FMA_RENORM_INTERVAL, stream=0):
    """Device form of gemm: a, b, c are device addresses (column-major DD)."""
    _dev(lib().mfhip_gemm(_trans(transa), _trans(transb), m, n, k, _f2(alpha), a, lda, b, ldb, _f2(beta),
                          c, ldc, renorm_interval, stream))


def trsm_workspace_bytes(side, m, n):
    return lib().mfhip_trsm_workspace_bytes(_trans(side), m, n)


def dev_trsm(side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb, work, work_bytes, stream=0):
    """Device form of trsm: a, b, work are device addresses."""
    _dev(lib().mfhip_trsm(_trans(side), _trans(uplo), _trans(transa), _trans(diag), m, n, _f2(alpha), a, lda, b,
                          ldb, work, work_bytes, stream))


def ozaki_slices_used():
    """Slices of the calling thread's last matmul_ozaki (host API); -1 after an
    automatic device call (see ozaki_info)."""
    return lib().mfhip_ozaki_slices_used()


def ozaki_fixups():
    """Outputs the fix-up recomputed in the calling thread's last matmul_ozaki."""
    return lib().mfhip_ozaki_fixups()


def ozaki_info(work, m, k, n, slices=0, stream=0):
    """(slices used, outputs recomputed by the fix-up) of the last Ozaki call
    on the device workspace `work`; synchronises the stream."""
    S = ctypes.c_int(0)
    F = ctypes.c_int64(0)
    _dev(lib().mfhip_ozaki_info(work, m, k, n, slices, ctypes.byref(S), ctypes.byref(F), stream))
    return S.value, F.value


def ozaki_moduli_used(work, m, k, n, slices=0, stream=0):
    """Moduli of the residue GEMMs of the last Ozaki call on the device
    workspace `work` (0 for the triangular form); synchronises the stream."""
    N = ctypes.c_int(0)
    _dev(lib().mfhip_ozaki_moduli_used(work, m, k, n, slices, ctypes.byref(N), stream))
    return N.value


def ozaki_choose_slices(k, range_a, range_b):
    return lib().mfhip_ozaki_choose_slices(k, range_a, range_b)


def ozaki_moduli(m, k, n, slices):
    """Moduli N of the CRT form for these sizes and slices (0: triangular form)."""
    return lib().mfhip_ozaki_moduli(m, k, n, slices)


def ozaki_workspace_bytes(m, k, n, slices=0):
    return lib().mfhip_ozaki_workspace_bytes(m, k, n, slices)


def dev_matmul_ozaki(a, b, c, m, k, n, slices, work, work_bytes, stream=0):
    _dev(lib().mfhip_ozaki_matmul_mm(a, b, c, m, k, n, slices, work, work_bytes, stream))


def dev_ozaki_prepare(a, b, m, k, n, slices, work, work_bytes, stream=0):
    _dev(lib().mfhip_ozaki_prepare(a, b, m, k, n, slices, work, work_bytes, stream))


def dev_ozaki_multiply(c, m, k, n, slices, work, work_bytes, stream=0):
    _dev(lib().mfhip_ozaki_multiply(c, m, k, n, slices, work, work_bytes, stream))


def dev_ozaki_products(m, k, n, slices, work, work_bytes, stream=0):
    """CRT form only: the residue GEMMs of dev_ozaki_multiply without the reconstruction."""
    _dev(lib().mfhip_ozaki_products(m, k, n, slices, work, work_bytes, stream))


def dev_unary(op, a, c, n, stream=0):
    _dev(lib().mfhip_unary(op, a, c, n, stream))


def dev_binary(op, a, b, c, n, stream=0):
    _dev(lib().mfhip_binary(op, a, b, c, n, stream))


def dev_reduce(kind, a, b, n, out, work, renorm_interval=DD_FMA_RENORM_INTERVAL, stream=0):
    _dev(lib().mfhip_reduce(kind, a, b, n, renorm_interval, out, work, stream))


def dev_reduce_noalloc(kind, a, b, n, out, renorm_interval=DD_FMA_RENORM_INTERVAL, stream=0):
    """dev_reduce without a workspace: the library's static device scratch
    (n <= NOALLOC_MAX_N; calls that overlap on one GPU race)."""
    _dev(lib().mfhip_reduce_noalloc(kind, a, b, n, renorm_interval, out, stream))


def reduce_workspace_bytes(n):
    return lib().mfhip_reduce_workspace_bytes(n)


def dev_sum_masked(a, mask, n, out, work, stream=0):
    _dev(lib().mfhip_sum_masked(a, mask, n, out, work, stream))


def cdot_workspace_bytes(n):
    return lib().mfhip_cdot_workspace_bytes(n)


def dev_cdot(a, b, n, out, work, renorm_interval=DD_FMA_RENORM_INTERVAL, stream=0):
    _dev(lib().mfhip_cdot(a, b, n, renorm_interval, out, work, stream))


def dev_reduce_masked(kind, a, mask, n, out, work, stream=0):
    _dev(lib().mfhip_reduce_masked(kind, a, mask, n, out, work, stream))


def dev_reduce_dim(kind, a, mask, inner, length, outer, out, stream=0):
    _dev(lib().mfhip_reduce_dim(kind, a, mask, inner, length, outer, out, stream))


def matmul_variants():
    return lib().mfhip_matmul_variants()


# ---- batching form (grouped persistent launch) ----

class GemmProblem(ctypes.Structure):
    _fields_ = [("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("c", ctypes.c_void_p),
                ("m", ctypes.c_int64), ("k", ctypes.c_int64), ("n", ctypes.c_int64),
                ("renorm_interval", ctypes.c_int64)]


class DotProblem(ctypes.Structure):
    _fields_ = [("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("n", ctypes.c_int64), ("renorm_interval", ctypes.c_int64)]


class GroupedPlan:
    """Many DD GEMMs and dot products in one persistent grouped launch
    (mfhip_grouped_*). gemms: list of (a, b, c, m, k, n, ri) with device
    addresses; dots: list of (a, b, out, n, ri); variants: per-GEMM tile
    variant (0 = 32x32, 1 = 64x64, 2 = 128x16, 3 = 16x128, 4 = 64x32,
    5 = 32x64, 6 = 16x16, 7 = 48x48, -1 = built-in rule). The problems must
    be independent: no output may overlap another problem's input or output
    (the dots of a plan with GEMMs run at the same time as the GEMMs)."""

    def __init__(self, gemms, dots, variants=None):
        L = lib()
        L.mfhip_grouped_create.restype = ctypes.c_void_p
        L.mfhip_grouped_create.argtypes = [ctypes.POINTER(GemmProblem), ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_int), ctypes.POINTER(DotProblem), ctypes.c_int]
        L.mfhip_grouped_items.argtypes = [ctypes.c_void_p]
        L.mfhip_grouped_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.mfhip_grouped_run_graph.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.mfhip_grouped_destroy.argtypes = [ctypes.c_void_p]
        G = (GemmProblem * max(len(gemms), 1))(*[GemmProblem(*[int(x) for x in g]) for g in gemms])
        D = (DotProblem * max(len(dots), 1))(*[DotProblem(*[int(x) for x in d]) for d in dots])
        V = None
        if variants is not None:
            V = (ctypes.c_int * max(len(variants), 1))(*[int(v) for v in variants])
        self._h = L.mfhip_grouped_create(G, len(gemms), V, D, len(dots))
        if not self._h:
            raise RuntimeError("mfhip_grouped_create failed")
        self.items = L.mfhip_grouped_items(self._h)

    def run(self, stream=0):
        _dev(lib().mfhip_grouped_run(self._h, stream))

    def run_graph(self, stream=0):
        _dev(lib().mfhip_grouped_run_graph(self._h, stream))

    def close(self):
        if self._h:
            lib().mfhip_grouped_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
