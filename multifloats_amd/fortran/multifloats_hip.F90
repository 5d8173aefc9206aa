! GENERATED by gen_bridge.py -- edit that file, not this one.
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
        pure type(float64x2) function c_sumdd_mask(a, mask, n) bind(c, name='sumdd_mask')
            import :: c_ptr, c_int64_t, float64x2
            type(c_ptr), value :: a, mask
            integer(c_int64_t), value :: n
        end function
        pure subroutine c_sumdd_dim(a, mask, inner, len, outer, out) bind(c, name='sumdd_dim')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, mask, out
            integer(c_int64_t), value :: inner, len, outer
        end subroutine
        pure type(float64x2) function c_productdd_mask(a, mask, n) bind(c, name='productdd_mask')
            import :: c_ptr, c_int64_t, float64x2
            type(c_ptr), value :: a, mask
            integer(c_int64_t), value :: n
        end function
        pure subroutine c_productdd_dim(a, mask, inner, len, outer, out) bind(c, name='productdd_dim')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, mask, out
            integer(c_int64_t), value :: inner, len, outer
        end subroutine
        pure type(float64x2) function c_maxvaldd_mask(a, mask, n) bind(c, name='maxvaldd_mask')
            import :: c_ptr, c_int64_t, float64x2
            type(c_ptr), value :: a, mask
            integer(c_int64_t), value :: n
        end function
        pure subroutine c_maxvaldd_dim(a, mask, inner, len, outer, out) bind(c, name='maxvaldd_dim')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, mask, out
            integer(c_int64_t), value :: inner, len, outer
        end subroutine
        pure type(float64x2) function c_minvaldd_mask(a, mask, n) bind(c, name='minvaldd_mask')
            import :: c_ptr, c_int64_t, float64x2
            type(c_ptr), value :: a, mask
            integer(c_int64_t), value :: n
        end function
        pure subroutine c_minvaldd_dim(a, mask, inner, len, outer, out) bind(c, name='minvaldd_dim')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, mask, out
            integer(c_int64_t), value :: inner, len, outer
        end subroutine
        pure type(float64x2) function c_norm2dd(a, n) bind(c, name='norm2dd')
            import :: c_ptr, c_int64_t, float64x2
            type(c_ptr), value :: a
            integer(c_int64_t), value :: n
        end function
        pure subroutine c_norm2dd_dim(a, inner, len, outer, out) bind(c, name='norm2dd_dim')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, out
            integer(c_int64_t), value :: inner, len, outer
        end subroutine
        pure subroutine c_expdd_v(a, c, n) bind(c, name='expdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_logdd_v(a, c, n) bind(c, name='logdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_sindd_v(a, c, n) bind(c, name='sindd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_erfdd_v(a, c, n) bind(c, name='erfdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_sqrtdd_v(a, c, n) bind(c, name='sqrtdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_cosdd_v(a, c, n) bind(c, name='cosdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_tandd_v(a, c, n) bind(c, name='tandd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_expm1dd_v(a, c, n) bind(c, name='expm1dd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_log2dd_v(a, c, n) bind(c, name='log2dd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_log10dd_v(a, c, n) bind(c, name='log10dd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_log1pdd_v(a, c, n) bind(c, name='log1pdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_sinhdd_v(a, c, n) bind(c, name='sinhdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_coshdd_v(a, c, n) bind(c, name='coshdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_tanhdd_v(a, c, n) bind(c, name='tanhdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_erfcdd_v(a, c, n) bind(c, name='erfcdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_erfcxdd_v(a, c, n) bind(c, name='erfcxdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_asinhdd_v(a, c, n) bind(c, name='asinhdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_acoshdd_v(a, c, n) bind(c, name='acoshdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_atanhdd_v(a, c, n) bind(c, name='atanhdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_sinpidd_v(a, c, n) bind(c, name='sinpidd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_cospidd_v(a, c, n) bind(c, name='cospidd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_tanpidd_v(a, c, n) bind(c, name='tanpidd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_atandd_v(a, c, n) bind(c, name='atandd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_atanpidd_v(a, c, n) bind(c, name='atanpidd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_asindd_v(a, c, n) bind(c, name='asindd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_acosdd_v(a, c, n) bind(c, name='acosdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_asinpidd_v(a, c, n) bind(c, name='asinpidd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_acospidd_v(a, c, n) bind(c, name='acospidd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_adddd_v(a, b, c, n) bind(c, name='adddd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_subdd_v(a, b, c, n) bind(c, name='subdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_muldd_v(a, b, c, n) bind(c, name='muldd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_divdd_v(a, b, c, n) bind(c, name='divdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_hypotdd_v(a, b, c, n) bind(c, name='hypotdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_copysigndd_v(a, b, c, n) bind(c, name='copysigndd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_atan2dd_v(a, b, c, n) bind(c, name='atan2dd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_atan2pidd_v(a, b, c, n) bind(c, name='atan2pidd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_mindd_v(a, b, c, n) bind(c, name='mindd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_maxdd_v(a, b, c, n) bind(c, name='maxdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_dimdd_v(a, b, c, n) bind(c, name='dimdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_fmoddd_v(a, b, c, n) bind(c, name='fmoddd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_modulodd_v(a, b, c, n) bind(c, name='modulodd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: n
        end subroutine
        pure subroutine c_nearestdd_v(a, b, c, n) bind(c, name='nearestdd_v')
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: a, b, c
            integer(c_int64_t), value :: n
        end subroutine
    end interface
    interface matmul
        module procedure dd_matmul_mm, dd_matmul_mv, dd_matmul_vm
        module procedure cdd_matmul_mm, cdd_matmul_mv, cdd_matmul_vm
    end interface

    interface dot_product
        module procedure dd_dot_product, cdd_dot_product
    end interface
    interface sum
        module procedure dd_sum_1d
        module procedure dd_sum_2d
        module procedure dd_sum_2d_dim
        module procedure dd_sum_3d
        module procedure dd_sum_3d_dim
        module procedure dd_sum_4d
        module procedure dd_sum_4d_dim
        module procedure dd_sum_5d
        module procedure dd_sum_5d_dim
        module procedure dd_sum_6d
        module procedure dd_sum_6d_dim
        module procedure dd_sum_7d
        module procedure dd_sum_7d_dim
    end interface

    interface product
        module procedure dd_product_1d
        module procedure dd_product_2d
        module procedure dd_product_2d_dim
        module procedure dd_product_3d
        module procedure dd_product_3d_dim
        module procedure dd_product_4d
        module procedure dd_product_4d_dim
        module procedure dd_product_5d
        module procedure dd_product_5d_dim
        module procedure dd_product_6d
        module procedure dd_product_6d_dim
        module procedure dd_product_7d
        module procedure dd_product_7d_dim
    end interface

    interface maxval
        module procedure dd_maxval_1d
        module procedure dd_maxval_2d
        module procedure dd_maxval_2d_dim
        module procedure dd_maxval_3d
        module procedure dd_maxval_3d_dim
        module procedure dd_maxval_4d
        module procedure dd_maxval_4d_dim
        module procedure dd_maxval_5d
        module procedure dd_maxval_5d_dim
        module procedure dd_maxval_6d
        module procedure dd_maxval_6d_dim
        module procedure dd_maxval_7d
        module procedure dd_maxval_7d_dim
    end interface

    interface minval
        module procedure dd_minval_1d
        module procedure dd_minval_2d
        module procedure dd_minval_2d_dim
        module procedure dd_minval_3d
        module procedure dd_minval_3d_dim
        module procedure dd_minval_4d
        module procedure dd_minval_4d_dim
        module procedure dd_minval_5d
        module procedure dd_minval_5d_dim
        module procedure dd_minval_6d
        module procedure dd_minval_6d_dim
        module procedure dd_minval_7d
        module procedure dd_minval_7d_dim
    end interface

    interface norm2
        module procedure dd_norm2_1d
        module procedure dd_norm2_2d
        module procedure dd_norm2_2d_dim
        module procedure dd_norm2_3d
        module procedure dd_norm2_3d_dim
        module procedure dd_norm2_4d
        module procedure dd_norm2_4d_dim
        module procedure dd_norm2_5d
        module procedure dd_norm2_5d_dim
        module procedure dd_norm2_6d
        module procedure dd_norm2_6d_dim
        module procedure dd_norm2_7d
        module procedure dd_norm2_7d_dim
    end interface

    interface exp
        module procedure dd_exp_1d
        module procedure dd_exp_2d
        module procedure dd_exp_3d
        module procedure dd_exp_4d
        module procedure dd_exp_5d
        module procedure dd_exp_6d
        module procedure dd_exp_7d
    end interface

    interface log
        module procedure dd_log_1d
        module procedure dd_log_2d
        module procedure dd_log_3d
        module procedure dd_log_4d
        module procedure dd_log_5d
        module procedure dd_log_6d
        module procedure dd_log_7d
    end interface

    interface sin
        module procedure dd_sin_1d
        module procedure dd_sin_2d
        module procedure dd_sin_3d
        module procedure dd_sin_4d
        module procedure dd_sin_5d
        module procedure dd_sin_6d
        module procedure dd_sin_7d
    end interface

    interface erf
        module procedure dd_erf_1d
        module procedure dd_erf_2d
        module procedure dd_erf_3d
        module procedure dd_erf_4d
        module procedure dd_erf_5d
        module procedure dd_erf_6d
        module procedure dd_erf_7d
    end interface

    interface sqrt
        module procedure dd_sqrt_1d
        module procedure dd_sqrt_2d
        module procedure dd_sqrt_3d
        module procedure dd_sqrt_4d
        module procedure dd_sqrt_5d
        module procedure dd_sqrt_6d
        module procedure dd_sqrt_7d
    end interface

    interface cos
        module procedure dd_cos_1d
        module procedure dd_cos_2d
        module procedure dd_cos_3d
        module procedure dd_cos_4d
        module procedure dd_cos_5d
        module procedure dd_cos_6d
        module procedure dd_cos_7d
    end interface

    interface tan
        module procedure dd_tan_1d
        module procedure dd_tan_2d
        module procedure dd_tan_3d
        module procedure dd_tan_4d
        module procedure dd_tan_5d
        module procedure dd_tan_6d
        module procedure dd_tan_7d
    end interface

    interface expm1
        module procedure dd_expm1_1d
        module procedure dd_expm1_2d
        module procedure dd_expm1_3d
        module procedure dd_expm1_4d
        module procedure dd_expm1_5d
        module procedure dd_expm1_6d
        module procedure dd_expm1_7d
    end interface

    interface log2
        module procedure dd_log2_1d
        module procedure dd_log2_2d
        module procedure dd_log2_3d
        module procedure dd_log2_4d
        module procedure dd_log2_5d
        module procedure dd_log2_6d
        module procedure dd_log2_7d
    end interface

    interface log10
        module procedure dd_log10_1d
        module procedure dd_log10_2d
        module procedure dd_log10_3d
        module procedure dd_log10_4d
        module procedure dd_log10_5d
        module procedure dd_log10_6d
        module procedure dd_log10_7d
    end interface

    interface log1p
        module procedure dd_log1p_1d
        module procedure dd_log1p_2d
        module procedure dd_log1p_3d
        module procedure dd_log1p_4d
        module procedure dd_log1p_5d
        module procedure dd_log1p_6d
        module procedure dd_log1p_7d
    end interface

    interface sinh
        module procedure dd_sinh_1d
        module procedure dd_sinh_2d
        module procedure dd_sinh_3d
        module procedure dd_sinh_4d
        module procedure dd_sinh_5d
        module procedure dd_sinh_6d
        module procedure dd_sinh_7d
    end interface

    interface cosh
        module procedure dd_cosh_1d
        module procedure dd_cosh_2d
        module procedure dd_cosh_3d
        module procedure dd_cosh_4d
        module procedure dd_cosh_5d
        module procedure dd_cosh_6d
        module procedure dd_cosh_7d
    end interface

    interface tanh
        module procedure dd_tanh_1d
        module procedure dd_tanh_2d
        module procedure dd_tanh_3d
        module procedure dd_tanh_4d
        module procedure dd_tanh_5d
        module procedure dd_tanh_6d
        module procedure dd_tanh_7d
    end interface

    interface erfc
        module procedure dd_erfc_1d
        module procedure dd_erfc_2d
        module procedure dd_erfc_3d
        module procedure dd_erfc_4d
        module procedure dd_erfc_5d
        module procedure dd_erfc_6d
        module procedure dd_erfc_7d
    end interface

    interface erfc_scaled
        module procedure dd_erfc_scaled_1d
        module procedure dd_erfc_scaled_2d
        module procedure dd_erfc_scaled_3d
        module procedure dd_erfc_scaled_4d
        module procedure dd_erfc_scaled_5d
        module procedure dd_erfc_scaled_6d
        module procedure dd_erfc_scaled_7d
    end interface

    interface asinh
        module procedure dd_asinh_1d
        module procedure dd_asinh_2d
        module procedure dd_asinh_3d
        module procedure dd_asinh_4d
        module procedure dd_asinh_5d
        module procedure dd_asinh_6d
        module procedure dd_asinh_7d
    end interface

    interface acosh
        module procedure dd_acosh_1d
        module procedure dd_acosh_2d
        module procedure dd_acosh_3d
        module procedure dd_acosh_4d
        module procedure dd_acosh_5d
        module procedure dd_acosh_6d
        module procedure dd_acosh_7d
    end interface

    interface atanh
        module procedure dd_atanh_1d
        module procedure dd_atanh_2d
        module procedure dd_atanh_3d
        module procedure dd_atanh_4d
        module procedure dd_atanh_5d
        module procedure dd_atanh_6d
        module procedure dd_atanh_7d
    end interface

    interface sinpi
        module procedure dd_sinpi_1d
        module procedure dd_sinpi_2d
        module procedure dd_sinpi_3d
        module procedure dd_sinpi_4d
        module procedure dd_sinpi_5d
        module procedure dd_sinpi_6d
        module procedure dd_sinpi_7d
    end interface

    interface cospi
        module procedure dd_cospi_1d
        module procedure dd_cospi_2d
        module procedure dd_cospi_3d
        module procedure dd_cospi_4d
        module procedure dd_cospi_5d
        module procedure dd_cospi_6d
        module procedure dd_cospi_7d
    end interface

    interface tanpi
        module procedure dd_tanpi_1d
        module procedure dd_tanpi_2d
        module procedure dd_tanpi_3d
        module procedure dd_tanpi_4d
        module procedure dd_tanpi_5d
        module procedure dd_tanpi_6d
        module procedure dd_tanpi_7d
    end interface

    interface atan
        module procedure dd_atan_1d
        module procedure dd_atan_2d
        module procedure dd_atan_3d
        module procedure dd_atan_4d
        module procedure dd_atan_5d
        module procedure dd_atan_6d
        module procedure dd_atan_7d
    end interface

    interface atanpi
        module procedure dd_atanpi_1d
        module procedure dd_atanpi_2d
        module procedure dd_atanpi_3d
        module procedure dd_atanpi_4d
        module procedure dd_atanpi_5d
        module procedure dd_atanpi_6d
        module procedure dd_atanpi_7d
    end interface

    interface asin
        module procedure dd_asin_1d
        module procedure dd_asin_2d
        module procedure dd_asin_3d
        module procedure dd_asin_4d
        module procedure dd_asin_5d
        module procedure dd_asin_6d
        module procedure dd_asin_7d
    end interface

    interface acos
        module procedure dd_acos_1d
        module procedure dd_acos_2d
        module procedure dd_acos_3d
        module procedure dd_acos_4d
        module procedure dd_acos_5d
        module procedure dd_acos_6d
        module procedure dd_acos_7d
    end interface

    interface asinpi
        module procedure dd_asinpi_1d
        module procedure dd_asinpi_2d
        module procedure dd_asinpi_3d
        module procedure dd_asinpi_4d
        module procedure dd_asinpi_5d
        module procedure dd_asinpi_6d
        module procedure dd_asinpi_7d
    end interface

    interface acospi
        module procedure dd_acospi_1d
        module procedure dd_acospi_2d
        module procedure dd_acospi_3d
        module procedure dd_acospi_4d
        module procedure dd_acospi_5d
        module procedure dd_acospi_6d
        module procedure dd_acospi_7d
    end interface

    interface operator(+)
        module procedure dd_add_1d
        module procedure dd_add_2d
        module procedure dd_add_3d
        module procedure dd_add_4d
        module procedure dd_add_5d
        module procedure dd_add_6d
        module procedure dd_add_7d
    end interface

    interface operator(-)
        module procedure dd_sub_1d
        module procedure dd_sub_2d
        module procedure dd_sub_3d
        module procedure dd_sub_4d
        module procedure dd_sub_5d
        module procedure dd_sub_6d
        module procedure dd_sub_7d
    end interface

    interface operator(*)
        module procedure dd_mul_1d
        module procedure dd_mul_2d
        module procedure dd_mul_3d
        module procedure dd_mul_4d
        module procedure dd_mul_5d
        module procedure dd_mul_6d
        module procedure dd_mul_7d
    end interface

    interface operator(/)
        module procedure dd_div_1d
        module procedure dd_div_2d
        module procedure dd_div_3d
        module procedure dd_div_4d
        module procedure dd_div_5d
        module procedure dd_div_6d
        module procedure dd_div_7d
    end interface

    interface hypot
        module procedure dd_hypot_1d
        module procedure dd_hypot_2d
        module procedure dd_hypot_3d
        module procedure dd_hypot_4d
        module procedure dd_hypot_5d
        module procedure dd_hypot_6d
        module procedure dd_hypot_7d
    end interface

    interface sign
        module procedure dd_sign_1d
        module procedure dd_sign_2d
        module procedure dd_sign_3d
        module procedure dd_sign_4d
        module procedure dd_sign_5d
        module procedure dd_sign_6d
        module procedure dd_sign_7d
    end interface

    interface atan2
        module procedure dd_atan2_1d
        module procedure dd_atan2_2d
        module procedure dd_atan2_3d
        module procedure dd_atan2_4d
        module procedure dd_atan2_5d
        module procedure dd_atan2_6d
        module procedure dd_atan2_7d
    end interface

    interface atan2pi
        module procedure dd_atan2pi_1d
        module procedure dd_atan2pi_2d
        module procedure dd_atan2pi_3d
        module procedure dd_atan2pi_4d
        module procedure dd_atan2pi_5d
        module procedure dd_atan2pi_6d
        module procedure dd_atan2pi_7d
    end interface

    interface min
        module procedure dd_min_1d
        module procedure dd_min_2d
        module procedure dd_min_3d
        module procedure dd_min_4d
        module procedure dd_min_5d
        module procedure dd_min_6d
        module procedure dd_min_7d
    end interface

    interface max
        module procedure dd_max_1d
        module procedure dd_max_2d
        module procedure dd_max_3d
        module procedure dd_max_4d
        module procedure dd_max_5d
        module procedure dd_max_6d
        module procedure dd_max_7d
    end interface

    interface dim
        module procedure dd_dim_1d
        module procedure dd_dim_2d
        module procedure dd_dim_3d
        module procedure dd_dim_4d
        module procedure dd_dim_5d
        module procedure dd_dim_6d
        module procedure dd_dim_7d
    end interface

    interface mod
        module procedure dd_mod_1d
        module procedure dd_mod_2d
        module procedure dd_mod_3d
        module procedure dd_mod_4d
        module procedure dd_mod_5d
        module procedure dd_mod_6d
        module procedure dd_mod_7d
    end interface

    interface modulo
        module procedure dd_modulo_1d
        module procedure dd_modulo_2d
        module procedure dd_modulo_3d
        module procedure dd_modulo_4d
        module procedure dd_modulo_5d
        module procedure dd_modulo_6d
        module procedure dd_modulo_7d
    end interface

    interface nearest
        module procedure dd_nearest_1d
        module procedure dd_nearest_2d
        module procedure dd_nearest_3d
        module procedure dd_nearest_4d
        module procedure dd_nearest_5d
        module procedure dd_nearest_6d
        module procedure dd_nearest_7d
    end interface

contains

    subroutine dd_set_fma_renorm_interval(n)
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

    pure function dd_sum_1d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:)
        logical, intent(in), target, contiguous, optional :: mask(:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_sumdd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_sumdd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_product_1d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:)
        logical, intent(in), target, contiguous, optional :: mask(:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_productdd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_productdd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_maxval_1d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:)
        logical, intent(in), target, contiguous, optional :: mask(:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_maxvaldd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_maxvaldd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_minval_1d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:)
        logical, intent(in), target, contiguous, optional :: mask(:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_minvaldd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_minvaldd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_norm2_1d(array) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:)
        type(real64x2) :: res
        type(float64x2) :: s
        s = c_norm2dd(c_loc(array), int(size(array), c_int64_t))
        res%limbs = s%limbs
    end function

    pure function dd_sum_2d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_sumdd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_sumdd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_product_2d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_productdd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_productdd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_maxval_2d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_maxvaldd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_maxvaldd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_minval_2d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_minvaldd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_minvaldd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_norm2_2d(array) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        s = c_norm2dd(c_loc(array), int(size(array), c_int64_t))
        res%limbs = s%limbs
    end function

    pure function dd_sum_2d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 2
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_sumdd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_sumdd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_product_2d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 2
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_productdd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_productdd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_maxval_2d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 2
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_maxvaldd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_maxvaldd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_minval_2d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 2
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_minvaldd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_minvaldd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_norm2_2d_dim(array, dim) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:)
        integer, intent(in) :: dim
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 2
            outer = outer * size(array, d_)
        end do
        call c_norm2dd_dim(c_loc(array), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
    end function

    pure function dd_sum_3d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_sumdd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_sumdd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_product_3d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_productdd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_productdd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_maxval_3d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_maxvaldd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_maxvaldd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_minval_3d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_minvaldd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_minvaldd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_norm2_3d(array) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        s = c_norm2dd(c_loc(array), int(size(array), c_int64_t))
        res%limbs = s%limbs
    end function

    pure function dd_sum_3d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 3
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_sumdd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_sumdd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_product_3d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 3
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_productdd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_productdd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_maxval_3d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 3
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_maxvaldd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_maxvaldd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_minval_3d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 3
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_minvaldd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_minvaldd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_norm2_3d_dim(array, dim) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:)
        integer, intent(in) :: dim
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 3
            outer = outer * size(array, d_)
        end do
        call c_norm2dd_dim(c_loc(array), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
    end function

    pure function dd_sum_4d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_sumdd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_sumdd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_product_4d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_productdd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_productdd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_maxval_4d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_maxvaldd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_maxvaldd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_minval_4d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_minvaldd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_minvaldd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_norm2_4d(array) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        s = c_norm2dd(c_loc(array), int(size(array), c_int64_t))
        res%limbs = s%limbs
    end function

    pure function dd_sum_4d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 4
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_sumdd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_sumdd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_product_4d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 4
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_productdd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_productdd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_maxval_4d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 4
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_maxvaldd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_maxvaldd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_minval_4d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 4
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_minvaldd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_minvaldd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_norm2_4d_dim(array, dim) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:)
        integer, intent(in) :: dim
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 4
            outer = outer * size(array, d_)
        end do
        call c_norm2dd_dim(c_loc(array), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
    end function

    pure function dd_sum_5d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_sumdd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_sumdd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_product_5d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_productdd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_productdd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_maxval_5d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_maxvaldd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_maxvaldd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_minval_5d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_minvaldd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_minvaldd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_norm2_5d(array) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        s = c_norm2dd(c_loc(array), int(size(array), c_int64_t))
        res%limbs = s%limbs
    end function

    pure function dd_sum_5d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)), size(array, 4 + merge(1, 0, 4 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 5
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_sumdd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_sumdd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_product_5d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)), size(array, 4 + merge(1, 0, 4 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 5
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_productdd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_productdd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_maxval_5d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)), size(array, 4 + merge(1, 0, 4 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 5
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_maxvaldd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_maxvaldd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_minval_5d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)), size(array, 4 + merge(1, 0, 4 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 5
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_minvaldd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_minvaldd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_norm2_5d_dim(array, dim) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:)
        integer, intent(in) :: dim
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)), size(array, 4 + merge(1, 0, 4 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 5
            outer = outer * size(array, d_)
        end do
        call c_norm2dd_dim(c_loc(array), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
    end function

    pure function dd_sum_6d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_sumdd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_sumdd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_product_6d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_productdd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_productdd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_maxval_6d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_maxvaldd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_maxvaldd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_minval_6d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_minvaldd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_minvaldd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_norm2_6d(array) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        s = c_norm2dd(c_loc(array), int(size(array), c_int64_t))
        res%limbs = s%limbs
    end function

    pure function dd_sum_6d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)), size(array, 4 + merge(1, 0, 4 >= dim)), size(array, 5 + merge(1, 0, 5 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 6
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_sumdd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_sumdd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_product_6d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)), size(array, 4 + merge(1, 0, 4 >= dim)), size(array, 5 + merge(1, 0, 5 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 6
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_productdd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_productdd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_maxval_6d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)), size(array, 4 + merge(1, 0, 4 >= dim)), size(array, 5 + merge(1, 0, 5 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 6
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_maxvaldd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_maxvaldd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_minval_6d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)), size(array, 4 + merge(1, 0, 4 >= dim)), size(array, 5 + merge(1, 0, 5 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 6
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_minvaldd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_minvaldd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_norm2_6d_dim(array, dim) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:)
        integer, intent(in) :: dim
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)), size(array, 4 + merge(1, 0, 4 >= dim)), size(array, 5 + merge(1, 0, 5 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 6
            outer = outer * size(array, d_)
        end do
        call c_norm2dd_dim(c_loc(array), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
    end function

    pure function dd_sum_7d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_sumdd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_sumdd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_product_7d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_productdd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_productdd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_maxval_7d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_maxvaldd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_maxvaldd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_minval_7d(array, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:,:)
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        if (present(mask)) then
            s = c_minvaldd_mask(c_loc(array), c_loc(mask), int(size(array), c_int64_t))
        else
            s = c_minvaldd_mask(c_loc(array), c_null_ptr, int(size(array), c_int64_t))
        end if
        res%limbs = s%limbs
    end function

    pure function dd_norm2_7d(array) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:,:)
        type(real64x2) :: res
        type(float64x2) :: s
        s = c_norm2dd(c_loc(array), int(size(array), c_int64_t))
        res%limbs = s%limbs
    end function

    pure function dd_sum_7d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)), size(array, 4 + merge(1, 0, 4 >= dim)), size(array, 5 + merge(1, 0, 5 >= dim)), size(array, 6 + merge(1, 0, 6 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 7
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_sumdd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_sumdd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_product_7d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)), size(array, 4 + merge(1, 0, 4 >= dim)), size(array, 5 + merge(1, 0, 5 >= dim)), size(array, 6 + merge(1, 0, 6 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 7
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_productdd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_productdd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_maxval_7d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)), size(array, 4 + merge(1, 0, 4 >= dim)), size(array, 5 + merge(1, 0, 5 >= dim)), size(array, 6 + merge(1, 0, 6 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 7
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_maxvaldd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_maxvaldd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_minval_7d_dim(array, dim, mask) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:,:)
        integer, intent(in) :: dim
        logical, intent(in), target, contiguous, optional :: mask(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)), size(array, 4 + merge(1, 0, 4 >= dim)), size(array, 5 + merge(1, 0, 5 >= dim)), size(array, 6 + merge(1, 0, 6 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 7
            outer = outer * size(array, d_)
        end do
        if (present(mask)) then
            call c_minvaldd_dim(c_loc(array), c_loc(mask), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        else
            call c_minvaldd_dim(c_loc(array), c_null_ptr, inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
        end if
    end function

    pure function dd_norm2_7d_dim(array, dim) result(res)
        type(real64x2), intent(in), target, contiguous :: array(:,:,:,:,:,:,:)
        integer, intent(in) :: dim
        type(real64x2), target :: res(size(array, 1 + merge(1, 0, 1 >= dim)), size(array, 2 + merge(1, 0, 2 >= dim)), size(array, 3 + merge(1, 0, 3 >= dim)), size(array, 4 + merge(1, 0, 4 >= dim)), size(array, 5 + merge(1, 0, 5 >= dim)), size(array, 6 + merge(1, 0, 6 >= dim)))
        integer(c_int64_t) :: inner, outer
        integer :: d_
        inner = 1
        do d_ = 1, dim - 1
            inner = inner * size(array, d_)
        end do
        outer = 1
        do d_ = dim + 1, 7
            outer = outer * size(array, d_)
        end do
        call c_norm2dd_dim(c_loc(array), inner, int(size(array, dim), c_int64_t), outer, c_loc(res))
    end function

    pure function dd_exp_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_expdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_exp_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_expdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_exp_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_expdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_exp_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_expdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_exp_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_expdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_exp_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_expdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_exp_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_expdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_logdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_logdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_logdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_logdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_logdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_logdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_logdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sin_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_sindd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sin_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_sindd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sin_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_sindd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sin_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_sindd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sin_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_sindd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sin_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_sindd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sin_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_sindd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erf_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_erfdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erf_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_erfdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erf_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_erfdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erf_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_erfdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erf_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_erfdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erf_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_erfdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erf_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_erfdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sqrt_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_sqrtdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sqrt_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_sqrtdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sqrt_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_sqrtdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sqrt_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_sqrtdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sqrt_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_sqrtdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sqrt_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_sqrtdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sqrt_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_sqrtdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cos_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_cosdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cos_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_cosdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cos_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_cosdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cos_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_cosdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cos_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_cosdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cos_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_cosdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cos_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_cosdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tan_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_tandd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tan_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_tandd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tan_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_tandd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tan_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_tandd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tan_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_tandd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tan_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_tandd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tan_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_tandd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_expm1_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_expm1dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_expm1_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_expm1dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_expm1_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_expm1dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_expm1_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_expm1dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_expm1_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_expm1dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_expm1_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_expm1dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_expm1_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_expm1dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log2_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_log2dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log2_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_log2dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log2_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_log2dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log2_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_log2dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log2_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_log2dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log2_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_log2dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log2_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_log2dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log10_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_log10dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log10_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_log10dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log10_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_log10dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log10_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_log10dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log10_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_log10dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log10_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_log10dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log10_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_log10dd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log1p_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_log1pdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log1p_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_log1pdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log1p_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_log1pdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log1p_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_log1pdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log1p_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_log1pdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log1p_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_log1pdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_log1p_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_log1pdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sinh_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_sinhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sinh_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_sinhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sinh_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_sinhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sinh_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_sinhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sinh_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_sinhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sinh_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_sinhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sinh_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_sinhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cosh_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_coshdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cosh_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_coshdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cosh_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_coshdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cosh_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_coshdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cosh_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_coshdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cosh_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_coshdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cosh_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_coshdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tanh_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_tanhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tanh_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_tanhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tanh_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_tanhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tanh_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_tanhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tanh_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_tanhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tanh_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_tanhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tanh_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_tanhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erfc_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_erfcdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erfc_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_erfcdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erfc_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_erfcdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erfc_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_erfcdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erfc_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_erfcdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erfc_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_erfcdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erfc_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_erfcdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erfc_scaled_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_erfcxdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erfc_scaled_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_erfcxdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erfc_scaled_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_erfcxdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erfc_scaled_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_erfcxdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erfc_scaled_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_erfcxdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erfc_scaled_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_erfcxdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_erfc_scaled_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_erfcxdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asinh_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_asinhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asinh_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_asinhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asinh_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_asinhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asinh_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_asinhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asinh_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_asinhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asinh_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_asinhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asinh_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_asinhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acosh_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_acoshdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acosh_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_acoshdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acosh_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_acoshdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acosh_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_acoshdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acosh_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_acoshdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acosh_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_acoshdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acosh_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_acoshdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atanh_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_atanhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atanh_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_atanhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atanh_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_atanhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atanh_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_atanhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atanh_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_atanhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atanh_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_atanhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atanh_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_atanhdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sinpi_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_sinpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sinpi_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_sinpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sinpi_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_sinpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sinpi_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_sinpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sinpi_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_sinpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sinpi_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_sinpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_sinpi_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_sinpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cospi_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_cospidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cospi_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_cospidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cospi_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_cospidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cospi_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_cospidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cospi_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_cospidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cospi_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_cospidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_cospi_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_cospidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tanpi_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_tanpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tanpi_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_tanpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tanpi_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_tanpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tanpi_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_tanpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tanpi_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_tanpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tanpi_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_tanpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_tanpi_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_tanpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atan_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_atandd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atan_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_atandd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atan_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_atandd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atan_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_atandd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atan_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_atandd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atan_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_atandd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atan_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_atandd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atanpi_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_atanpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atanpi_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_atanpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atanpi_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_atanpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atanpi_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_atanpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atanpi_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_atanpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atanpi_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_atanpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_atanpi_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_atanpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asin_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_asindd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asin_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_asindd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asin_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_asindd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asin_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_asindd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asin_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_asindd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asin_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_asindd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asin_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_asindd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acos_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_acosdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acos_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_acosdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acos_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_acosdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acos_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_acosdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acos_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_acosdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acos_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_acosdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acos_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_acosdd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asinpi_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_asinpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asinpi_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_asinpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asinpi_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_asinpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asinpi_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_asinpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asinpi_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_asinpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asinpi_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_asinpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_asinpi_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_asinpidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acospi_1d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:)
        type(real64x2), target :: res(size(x,1))
        call c_acospidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acospi_2d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:)
        type(real64x2), target :: res(size(x,1), size(x,2))
        call c_acospidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acospi_3d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3))
        call c_acospidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acospi_4d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4))
        call c_acospidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acospi_5d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5))
        call c_acospidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acospi_6d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6))
        call c_acospidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_acospi_7d(x) result(res)
        type(real64x2), intent(in), target, contiguous :: x(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(x,1), size(x,2), size(x,3), size(x,4), size(x,5), size(x,6), size(x,7))
        call c_acospidd_v(c_loc(x), c_loc(res), int(size(x), c_int64_t))
    end function

    pure function dd_add_1d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:)
        type(real64x2), target :: res(size(a,1))
        call c_adddd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_add_2d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(real64x2), target :: res(size(a,1), size(a,2))
        call c_adddd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_add_3d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:), b(:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3))
        call c_adddd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_add_4d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:), b(:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4))
        call c_adddd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_add_5d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:), b(:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5))
        call c_adddd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_add_6d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:), b(:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6))
        call c_adddd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_add_7d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:,:), b(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6), size(a,7))
        call c_adddd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_sub_1d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:)
        type(real64x2), target :: res(size(a,1))
        call c_subdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_sub_2d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(real64x2), target :: res(size(a,1), size(a,2))
        call c_subdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_sub_3d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:), b(:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3))
        call c_subdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_sub_4d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:), b(:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4))
        call c_subdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_sub_5d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:), b(:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5))
        call c_subdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_sub_6d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:), b(:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6))
        call c_subdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_sub_7d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:,:), b(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6), size(a,7))
        call c_subdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_mul_1d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:)
        type(real64x2), target :: res(size(a,1))
        call c_muldd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_mul_2d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(real64x2), target :: res(size(a,1), size(a,2))
        call c_muldd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_mul_3d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:), b(:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3))
        call c_muldd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_mul_4d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:), b(:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4))
        call c_muldd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_mul_5d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:), b(:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5))
        call c_muldd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_mul_6d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:), b(:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6))
        call c_muldd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_mul_7d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:,:), b(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6), size(a,7))
        call c_muldd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_div_1d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:)
        type(real64x2), target :: res(size(a,1))
        call c_divdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_div_2d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(real64x2), target :: res(size(a,1), size(a,2))
        call c_divdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_div_3d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:), b(:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3))
        call c_divdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_div_4d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:), b(:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4))
        call c_divdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_div_5d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:), b(:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5))
        call c_divdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_div_6d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:), b(:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6))
        call c_divdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_div_7d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:,:), b(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6), size(a,7))
        call c_divdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_hypot_1d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:)
        type(real64x2), target :: res(size(a,1))
        call c_hypotdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_hypot_2d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(real64x2), target :: res(size(a,1), size(a,2))
        call c_hypotdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_hypot_3d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:), b(:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3))
        call c_hypotdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_hypot_4d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:), b(:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4))
        call c_hypotdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_hypot_5d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:), b(:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5))
        call c_hypotdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_hypot_6d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:), b(:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6))
        call c_hypotdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_hypot_7d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:,:), b(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6), size(a,7))
        call c_hypotdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_sign_1d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:)
        type(real64x2), target :: res(size(a,1))
        call c_copysigndd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_sign_2d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(real64x2), target :: res(size(a,1), size(a,2))
        call c_copysigndd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_sign_3d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:), b(:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3))
        call c_copysigndd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_sign_4d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:), b(:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4))
        call c_copysigndd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_sign_5d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:), b(:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5))
        call c_copysigndd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_sign_6d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:), b(:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6))
        call c_copysigndd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_sign_7d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:,:), b(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6), size(a,7))
        call c_copysigndd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_atan2_1d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:)
        type(real64x2), target :: res(size(a,1))
        call c_atan2dd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_atan2_2d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(real64x2), target :: res(size(a,1), size(a,2))
        call c_atan2dd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_atan2_3d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:), b(:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3))
        call c_atan2dd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_atan2_4d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:), b(:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4))
        call c_atan2dd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_atan2_5d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:), b(:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5))
        call c_atan2dd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_atan2_6d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:), b(:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6))
        call c_atan2dd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_atan2_7d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:,:), b(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6), size(a,7))
        call c_atan2dd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_atan2pi_1d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:)
        type(real64x2), target :: res(size(a,1))
        call c_atan2pidd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_atan2pi_2d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(real64x2), target :: res(size(a,1), size(a,2))
        call c_atan2pidd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_atan2pi_3d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:), b(:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3))
        call c_atan2pidd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_atan2pi_4d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:), b(:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4))
        call c_atan2pidd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_atan2pi_5d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:), b(:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5))
        call c_atan2pidd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_atan2pi_6d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:), b(:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6))
        call c_atan2pidd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_atan2pi_7d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:,:), b(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6), size(a,7))
        call c_atan2pidd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_min_1d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:)
        type(real64x2), target :: res(size(a,1))
        call c_mindd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_min_2d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(real64x2), target :: res(size(a,1), size(a,2))
        call c_mindd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_min_3d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:), b(:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3))
        call c_mindd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_min_4d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:), b(:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4))
        call c_mindd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_min_5d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:), b(:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5))
        call c_mindd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_min_6d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:), b(:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6))
        call c_mindd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_min_7d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:,:), b(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6), size(a,7))
        call c_mindd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_max_1d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:)
        type(real64x2), target :: res(size(a,1))
        call c_maxdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_max_2d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(real64x2), target :: res(size(a,1), size(a,2))
        call c_maxdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_max_3d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:), b(:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3))
        call c_maxdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_max_4d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:), b(:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4))
        call c_maxdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_max_5d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:), b(:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5))
        call c_maxdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_max_6d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:), b(:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6))
        call c_maxdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_max_7d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:,:), b(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6), size(a,7))
        call c_maxdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_dim_1d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:)
        type(real64x2), target :: res(size(a,1))
        call c_dimdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_dim_2d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(real64x2), target :: res(size(a,1), size(a,2))
        call c_dimdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_dim_3d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:), b(:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3))
        call c_dimdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_dim_4d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:), b(:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4))
        call c_dimdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_dim_5d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:), b(:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5))
        call c_dimdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_dim_6d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:), b(:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6))
        call c_dimdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_dim_7d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:,:), b(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6), size(a,7))
        call c_dimdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_mod_1d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:)
        type(real64x2), target :: res(size(a,1))
        call c_fmoddd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_mod_2d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(real64x2), target :: res(size(a,1), size(a,2))
        call c_fmoddd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_mod_3d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:), b(:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3))
        call c_fmoddd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_mod_4d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:), b(:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4))
        call c_fmoddd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_mod_5d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:), b(:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5))
        call c_fmoddd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_mod_6d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:), b(:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6))
        call c_fmoddd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_mod_7d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:,:), b(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6), size(a,7))
        call c_fmoddd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_modulo_1d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:)
        type(real64x2), target :: res(size(a,1))
        call c_modulodd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_modulo_2d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(real64x2), target :: res(size(a,1), size(a,2))
        call c_modulodd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_modulo_3d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:), b(:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3))
        call c_modulodd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_modulo_4d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:), b(:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4))
        call c_modulodd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_modulo_5d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:), b(:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5))
        call c_modulodd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_modulo_6d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:), b(:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6))
        call c_modulodd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_modulo_7d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:,:), b(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6), size(a,7))
        call c_modulodd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_nearest_1d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:), b(:)
        type(real64x2), target :: res(size(a,1))
        call c_nearestdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_nearest_2d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:), b(:,:)
        type(real64x2), target :: res(size(a,1), size(a,2))
        call c_nearestdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_nearest_3d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:), b(:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3))
        call c_nearestdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_nearest_4d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:), b(:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4))
        call c_nearestdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_nearest_5d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:), b(:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5))
        call c_nearestdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_nearest_6d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:), b(:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6))
        call c_nearestdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

    pure function dd_nearest_7d(a, b) result(res)
        type(real64x2), intent(in), target, contiguous :: a(:,:,:,:,:,:,:), b(:,:,:,:,:,:,:)
        type(real64x2), target :: res(size(a,1), size(a,2), size(a,3), size(a,4), size(a,5), size(a,6), size(a,7))
        call c_nearestdd_v(c_loc(a), c_loc(b), c_loc(res), int(min(size(a), size(b)), c_int64_t))
    end function

end module multifloats_hip
