#!/usr/bin/env python3
"""Writes multifloats_hip.F90, the Fortran module that routes the reference's
array generics (matmul, dot_product, sum, norm2, the elemental math functions
exp log sin cos tan sinpi cospi tanpi expm1 log2 log10 log1p sinh cosh tanh asinh
acosh atanh erf erfc erfc_scaled sqrt, and the
array + - * /) to libmultifloats_hip.so through bind(c) interfaces.

The reference generates its module with fypp (fsrc/multifloats.fypp); fypp is
not part of this toolchain, so the rank 1..7 specifics are written out here.

    python multifloats_amd/fortran/gen_bridge.py   # rewrites multifloats_hip.F90
"""
import os

MAX_RANK = 7
HERE = os.path.dirname(os.path.abspath(__file__))
UNARY = [("exp", "expdd_v"), ("log", "logdd_v"), ("sin", "sindd_v"), ("erf", "erfdd_v"), ("sqrt", "sqrtdd_v"),
         ("cos", "cosdd_v"), ("tan", "tandd_v"), ("expm1", "expm1dd_v"), ("log2", "log2dd_v"),
         ("log10", "log10dd_v"), ("log1p", "log1pdd_v"), ("sinh", "sinhdd_v"), ("cosh", "coshdd_v"),
         ("tanh", "tanhdd_v"), ("erfc", "erfcdd_v"), ("erfc_scaled", "erfcxdd_v"), ("asinh", "asinhdd_v"),
         ("acosh", "acoshdd_v"), ("atanh", "atanhdd_v"), ("sinpi", "sinpidd_v"), ("cospi", "cospidd_v"),
         ("tanpi", "tanpidd_v"), ("atan", "atandd_v"), ("atanpi", "atanpidd_v"),
         ("asin", "asindd_v"), ("acos", "acosdd_v"), ("asinpi", "asinpidd_v"), ("acospi", "acospidd_v")]
# reductions with mask= and dim= (fypp:1783-2090): generic, C entry stem
MASKED_REDUCE = [("sum", "sumdd"), ("product", "productdd"), ("maxval", "maxvaldd"), ("minval", "minvaldd")]
BINARY = [("+", "add", "adddd_v"), ("-", "sub", "subdd_v"), ("*", "mul", "muldd_v"), ("/", "div", "divdd_v")]
# two-argument elemental generics whose reference bodies are these C functions'
# rules (fypp:1550-1607: hypot calls hypotdd; sign is copysigndd's rule; mod
# calls fmoddd; modulo is fmoddd plus y on a sign change, modulodd's rule;
# nearest steps the leading limb)
BINFN = [("hypot", "hypotdd_v"), ("sign", "copysigndd_v"), ("atan2", "atan2dd_v"), ("atan2pi", "atan2pidd_v"),
         ("min", "mindd_v"), ("max", "maxdd_v"), ("dim", "dimdd_v"), ("mod", "fmoddd_v"),
         ("modulo", "modulodd_v"), ("nearest", "nearestdd_v")]


def colons(r):
    return ",".join([":"] * r)


def shape_of(name, r):
    return ", ".join("size(%s,%d)" % (name, d) for d in range(1, r + 1))


def reduced_shape(r):
    # extents of `array` without dimension `dim`
    return ", ".join("size(array, %d + merge(1, 0, %d >= dim))" % (j, j) for j in range(1, r))


def gen():
    o = []
    w = o.append
    w("""! GENERATED by gen_bridge.py -- edit that file, not this one.
!
! multifloats_hip: Fortran bridge from the reference's array generics to the
! MI355X library libmultifloats_hip.so. Every specific here is a non-elemental
! array procedure that hands a contiguous array to one C-ABI entry point
! (include/multifloats_hip.h), which runs the hand-written HIP kernels and
! returns when the result is back in host memory.
!
! Reference interfaces replaced (fsrc/multifloats.fypp):
!   matmul          dd_matmul_mm/mv/vm 2568-2593, cdd_matmul_* 2609-2649
!   dot_product     dd_dot_product 2254-2308, cdd_dot_product 2321-2331
!   sum product     dd_sum_Nd / dd_product_Nd (+ _dim) 1975-2090 (with mask)
!   maxval minval   dd_maxval_Nd / dd_minval_Nd (+ _dim) 1783-1862 (with mask)
!   norm2           dd_norm2_Nd / dd_norm2_Nd_dim 2338-2377
!   exp log sin erf sqrt cos tan expm1 log2 log10 log1p sinh cosh tanh
!   erfc erfc_scaled asinh acosh atanh sinpi cospi tanpi atan atanpi asin acos
!   asinpi acospi
!                   elemental delegates 709-736, 1407-1412 (names fypp:57)
!   hypot sign atan2 atan2pi min max dim  elemental binary functions 1536-1592
!   + - * /         elemental native operators 2746-2837
!   DD_FMA_RENORM_INTERVAL and dd_set_fma_renorm_interval 664, 2677-2680
!
! Types: real64x2 and cmplx64x2 are SEQUENCE types with the reference's names
! and components (fypp:170-178), so values of the reference's types are values
! of these. Built with -DMULTIFLOATS_HIP_USE_REFERENCE_TYPES the module takes
! the types from the reference's module instead and only adds the specifics.
module multifloats_hip
    use, intrinsic :: iso_c_binding
#ifdef MULTIFLOATS_HIP_USE_REFERENCE_TYPES
    use multifloats, only: real64x2, cmplx64x2
#endif
    implicit none
    private

    integer, parameter :: dp = 8

#ifndef MULTIFLOATS_HIP_USE_REFERENCE_TYPES
    type, public :: real64x2
        sequence
        real(dp) :: limbs(2)
    end type real64x2

    type, public :: cmplx64x2
        sequence
        type(real64x2) :: re, im
    end type cmplx64x2
#else
    public :: real64x2, cmplx64x2
#endif

    ! by-value twins of the C structs (include/multifloats_hip.h)
    type, bind(c) :: float64x2
        real(c_double) :: limbs(2)
    end type float64x2

    type, bind(c) :: complex64x2
        type(float64x2) :: re, im
    end type complex64x2

    ! Renormalisation interval of matmul and dot_product (reference default 8).
    integer, public, protected :: DD_FMA_RENORM_INTERVAL = 8
    public :: dd_set_fma_renorm_interval
    public :: dd_matmul_ozaki
    public :: wgemm
    public :: matmul, dot_product, sum, product, maxval, minval, norm2
    public :: exp, log, sin, erf, sqrt, cos, tan, expm1, log2, log10, log1p, sinh, cosh, tanh
    public :: erfc, erfc_scaled, asinh, acosh, atanh, sinpi, cospi, tanpi
    public :: hypot, sign, atan, atanpi, atan2, atan2pi, asin, acos, asinpi, acospi, min, max, dim
    public :: mod, modulo, nearest
    public :: operator(+), operator(-), operator(*), operator(/)

    interface
        pure subroutine c_matmuldd_mm(a, b, c, m, k, n, ri) bind(c, name='matmuldd_mm')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: m, k, n, ri
        end subroutine
        pure subroutine c_matmuldd_mv(a, x, y, m, k, ri) bind(c, name='matmuldd_mv')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, x, y
            integer(c_int64_t), value :: m, k, ri
        end subroutine
        pure subroutine c_matmuldd_vm(x, b, y, k, n, ri) bind(c, name='matmuldd_vm')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: x, b, y
            integer(c_int64_t), value :: k, n, ri
        end subroutine
        pure subroutine c_cmatmuldd_mm(a, b, c, m, k, n, ri) bind(c, name='cmatmuldd_mm')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: m, k, n, ri
        end subroutine
        pure subroutine c_cmatmuldd_mv(a, x, y, m, k, ri) bind(c, name='cmatmuldd_mv')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, x, y
            integer(c_int64_t), value :: m, k, ri
        end subroutine
        pure subroutine c_cmatmuldd_vm(x, b, y, k, n, ri) bind(c, name='cmatmuldd_vm')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: x, b, y
            integer(c_int64_t), value :: k, n, ri
        end subroutine
        subroutine c_matmuldd_mm_ozaki(a, b, c, m, k, n, slices) bind(c, name='matmuldd_mm_ozaki')
            import :: c_ptr, c_int64_t, c_int
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: m, k, n
            integer(c_int), value :: slices
        end subroutine
        ! alpha and beta are float64x2 by value in C. The x86-64 System V ABI
        ! passes such a struct in two SSE registers, exactly as its two limbs;
        ! flang passes a second by-value bind(c) struct on the stack instead,
        ! so they are spelled here as their limbs.
        subroutine c_wgemmdd(ta, tb, m, n, k, alpha_hi, alpha_lo, a, lda, b, ldb, beta_hi, beta_lo, c, ldc, ri) &
                bind(c, name='wgemmdd')
            import :: c_ptr, c_int64_t, c_char, c_double
            character(kind=c_char), value :: ta, tb
            integer(c_int64_t), value :: m, n, k, lda, ldb, ldc, ri
            real(c_double), value :: alpha_hi, alpha_lo, beta_hi, beta_lo
            type(c_ptr), value :: a, b, c
        end subroutine
        pure type(float64x2) function c_dotdd(a, b, n, ri) bind(c, name='dotdd')
            import :: c_ptr, c_int64_t, float64x2
            type(c_ptr), value :: a, b
            integer(c_int64_t), value :: n, ri
        end function
        pure type(complex64x2) function c_cdotdd(a, b, n, ri) bind(c, name='cdotdd')
            import :: c_ptr, c_int64_t, complex64x2
            type(c_ptr), value :: a, b
            integer(c_int64_t), value :: n, ri
        end function
""")
    for f, c in MASKED_REDUCE:
        w("""        pure type(float64x2) function c_%(c)s_mask(a, mask, n) bind(c, name='%(c)s_mask')
            import :: c_ptr, c_int64_t, float64x2
            type(c_ptr), value :: a, mask
            integer(c_int64_t), value :: n
        end function
        pure subroutine c_%(c)s_dim(a, mask, inner, len, outer, out) bind(c, name='%(c)s_dim')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, mask, out
            integer(c_int64_t), value :: inner, len, outer
        end subroutine
""" % dict(c=c))
    w("""        pure type(float64x2) function c_norm2dd(a, n) bind(c, name='norm2dd')
            import :: c_ptr, c_int64_t, float64x2
            type(c_ptr), value :: a
            integer(c_int64_t), value :: n
        end function
        pure subroutine c_norm2dd_dim(a, inner, len, outer, out) bind(c, name='norm2dd_dim')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, out
            integer(c_int64_t), value :: inner, len, outer
        end subroutine
""")
    for f, c in UNARY:
        w("""        pure subroutine c_%s(a, c, n) bind(c, name='%s')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
""" % (c, c))
    for _, nm, c in BINARY + [(None, f, c) for f, c in BINFN]:
        w("""        pure subroutine c_%s(a, b, c, n) bind(c, name='%s')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: n
        end subroutine
""" % (c, c))
    w("    end interface\n")

    # generic interfaces
    w("""    interface matmul
        module procedure dd_matmul_mm, dd_matmul_mv, dd_matmul_vm
        module procedure cdd_matmul_mm, cdd_matmul_mv, cdd_matmul_vm
    end interface

    interface dot_product
        module procedure dd_dot_product, cdd_dot_product
    end interface
""")
    for f, _ in MASKED_REDUCE:
        w("    interface %s\n" % f)
        for r in range(1, MAX_RANK + 1):
            w("        module procedure dd_%s_%dd\n" % (f, r))
            if r > 1:
                w("        module procedure dd_%s_%dd_dim\n" % (f, r))
        w("    end interface\n\n")
    w("    interface norm2\n")
    for r in range(1, MAX_RANK + 1):
        w("        module procedure dd_norm2_%dd\n" % r)
        if r > 1:
            w("        module procedure dd_norm2_%dd_dim\n" % r)
    w("    end interface\n\n")
    for f, _ in UNARY:
        w("    interface %s\n" % f)
        for r in range(1, MAX_RANK + 1):
            w("        module procedure dd_%s_%dd\n" % (f, r))
        w("    end interface\n\n")
    for op, nm, _ in BINARY:
        w("    interface operator(%s)\n" % op)
        for r in range(1, MAX_RANK + 1):
            w("        module procedure dd_%s_%dd\n" % (nm, r))
        w("    end interface\n\n")
    for f, _ in BINFN:
        w("    interface %s\n" % f)
        for r in range(1, MAX_RANK + 1):
            w("        module procedure dd_%s_%dd\n" % (f, r))
        w("    end interface\n\n")

    w("contains\n\n")
    w("""    subroutine dd_set_fma_renorm_interval(n)
        integer, intent(in) :: n
        DD_FMA_RENORM_INTERVAL = n
    end subroutine

    pure function ri() result(r)
        integer(c_int64_t) :: r
        r = int(DD_FMA_RENORM_INTERVAL, c_int64_t)
    end function

    pure function dd_matmul_mm(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(real64x2), target :: res(size(a,1), size(b,2))
        call c_matmuldd_mm(c_loc(a), c_loc(b), c_loc(res), int(size(a,1), c_int64_t), &
            int(size(a,2), c_int64_t), int(size(b,2), c_int64_t), ri())
    end function

    pure function dd_matmul_mv(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:)
        type(real64x2), target :: res(size(a,1))
        call c_matmuldd_mv(c_loc(a), c_loc(b), c_loc(res), int(size(a,1), c_int64_t), &
            int(size(a,2), c_int64_t), ri())
    end function

    pure function dd_matmul_vm(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:,:)
        type(real64x2), target :: res(size(b,2))
        call c_matmuldd_vm(c_loc(a), c_loc(b), c_loc(res), int(size(a), c_int64_t), &
            int(size(b,2), c_int64_t), ri())
    end function

    ! Complex matmul: one fused kernel, bit-identical to the reference's
    ! four-real-matmul composition (fypp:2609-2621).
    pure function cdd_matmul_mm(a, b) result(res)
        type(cmplx64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(cmplx64x2), target :: res(size(a,1), size(b,2))
        call c_cmatmuldd_mm(c_loc(a), c_loc(b), c_loc(res), int(size(a,1), c_int64_t), &
            int(size(a,2), c_int64_t), int(size(b,2), c_int64_t), ri())
    end function

    pure function cdd_matmul_mv(a, b) result(res)
        type(cmplx64x2), intent(in), target, contiguous :: a(:,:), b(:)
        type(cmplx64x2), target :: res(size(a,1))
        call c_cmatmuldd_mv(c_loc(a), c_loc(b), c_loc(res), int(size(a,1), c_int64_t), &
            int(size(a,2), c_int64_t), ri())
    end function

    pure function cdd_matmul_vm(a, b) result(res)
        type(cmplx64x2), intent(in), target, contiguous :: a(:), b(:,:)
        type(cmplx64x2), target :: res(size(b,2))
        call c_cmatmuldd_vm(c_loc(a), c_loc(b), c_loc(res), int(size(a), c_int64_t), &
            int(size(b,2), c_int64_t), ri())
    end function

    ! Precision form (no reference counterpart): Ozaki INT8 split GEMM,
    ! |C - AB| <= 2^-100 (|A||B|); slices absent or <= 0 lets the library
    ! choose from the data.
    function dd_matmul_ozaki(a, b, slices) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        integer, intent(in), optional :: slices
        type(real64x2), target :: res(size(a,1), size(b,2))
        integer(c_int) :: s
        s = 0
        if (present(slices)) s = int(slices, c_int)
        call c_matmuldd_mm_ozaki(c_loc(a), c_loc(b), c_loc(res), int(size(a,1), c_int64_t), &
            int(size(a,2), c_int64_t), int(size(b,2), c_int64_t), s)
    end function

    ! xGEMM on real64x2 with the argument list of the reference's test shim
    ! (test/integration/blas/wgemm.f90:1-14): C = alpha op(A) op(B) + beta C,
    ! transa/transb 'N', 'T' or 'C'. op(A) op(B) is formed like matmul (same
    ! chains, DD_FMA_RENORM_INTERVAL); bad arguments leave NaN in C.
    subroutine wgemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
        character, intent(in) :: transa, transb
        integer, intent(in) :: m, n, k, lda, ldb, ldc
        type(real64x2), intent(in) :: alpha, beta
        type(real64x2), intent(in), target :: a(lda, *), b(ldb, *)
        type(real64x2), intent(inout), target :: c(ldc, *)
        call c_wgemmdd(transa, transb, int(m, c_int64_t), int(n, c_int64_t), int(k, c_int64_t), &
            alpha%limbs(1), alpha%limbs(2), c_loc(a(1, 1)), int(lda, c_int64_t), c_loc(b(1, 1)), &
            int(ldb, c_int64_t), beta%limbs(1), beta%limbs(2), c_loc(c(1, 1)), int(ldc, c_int64_t), ri())
    end subroutine

    pure function dd_dot_product(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:)
        type(real64x2) :: res
        type(float64x2) :: r
        r = c_dotdd(c_loc(a), c_loc(b), int(min(size(a), size(b)), c_int64_t), ri())
        res%limbs = r%limbs
    end function

    pure function cdd_dot_product(a, b) result(res)
        type(cmplx64x2), intent(in), target, contiguous :: a(:), b(:)
        type(cmplx64x2) :: res
        type(complex64x2) :: r
        r = c_cdotdd(c_loc(a), c_loc(b), int(min(size(a), size(b)), c_int64_t), ri())
        res%re%limbs = r%re%limbs
        res%im%limbs = r%im%limbs
    end function

""")
    for r in range(1, MAX_RANK + 1):
        C = colons(r)
        for f, c in MASKED_REDUCE:
            w("""    pure function dd_%(f)s_%(r)dd(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(%(C)s)
        logical, intent(in), target, contiguous, optional :: mask(%(C)s)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_%(c)s_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_%(c)s_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%%limbs = s%%limbs
    end function

""" % dict(r=r, C=C, f=f, c=c))
        w("""    pure function dd_norm2_%(r)dd(array) result(res)
        type(real64x2), intent(in), target, contiguous :: array(%(C)s)
        type(real64x2) :: res
        type(float64x2) :: s
        s = c_norm2dd(c_loc(array), int(size(array), c_int64_t))
        res%%limbs = s%%limbs
    end function

""" % dict(r=r, C=C))
        if r > 1:
          for f, c in MASKED_REDUCE:
            w("""    pure function dd_%(f)s_%(r)dd_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(%(C)s)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(%(C)s)
        type(real64x2), target :: res(%(RS)s)
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, %(r)d
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_%(c)s_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_%(c)s_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

""" % dict(r=r, C=C, RS=reduced_shape(r), f=f, c=c))
          w("""    pure function dd_norm2_%(r)dd_dim(array, dim) result(res)
        type(real64x2), intent(in), target, contiguous :: array(%(C)s)
        integer, intent(in) :: dim
        type(real64x2), target :: res(%(RS)s)
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, %(r)d
            outer = outer * size(array, d_)
        end do
        call c_norm2dd_dim(c_loc(array), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
    end function

""" % dict(r=r, C=C, RS=reduced_shape(r)))
    for f, c in UNARY:
        for r in range(1, MAX_RANK + 1):
            C = colons(r)
            w("""    pure function dd_%(f)s_%(r)dd(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(%(C)s)
        type(real64x2), target :: res(%(S)s)
        call c_%(c)s(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

""" % dict(f=f, r=r, C=C, c=c, S=shape_of("x", r)))
    for op, nm, c in BINARY + [(None, f, c) for f, c in BINFN]:
        for r in range(1, MAX_RANK + 1):
            C = colons(r)
            w("""    pure function dd_%(nm)s_%(r)dd(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(%(C)s), b(%(C)s)
        type(real64x2), target :: res(%(S)s)
        call c_%(c)s(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

""" % dict(nm=nm, r=r, C=C, c=c, S=shape_of("a", r)))
    w("end module multifloats_hip\n")
    return "".join(o)


if __name__ == "__main__":
    with open(os.path.join(HERE, "multifloats_hip.F90"), "w") as fh:
        fh.write(gen())
