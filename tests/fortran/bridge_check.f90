! Drives the multifloats_hip Fortran module the way a caller of the
! reference's Fortran module would (fsrc/multifloats.fypp generics): reads
! inputs written by tests/test_gpu_fortran.py, calls the generics and writes
! every result, in order, to the output file for the test to compare.
!
!   bridge_check <input file> <output file>
!
! Input (stream, native endianness): int64 m, k, n, nv; then A(m,k), B(k,n),
! CA(m,k), CB(k,n) (complex), x(nv), y(nv), cx(nv), cy(nv), mask(nv) (int32).
program bridge_check
    use multifloats_hip
    implicit none
    integer, parameter :: dp = 8
    integer(8) :: m, k, n, nv
    type(real64x2), allocatable :: a(:,:), b(:,:), x(:), y(:), at(:,:), cc(:,:)
    type(cmplx64x2), allocatable :: ca(:,:), cb(:,:), cx(:), cy(:)
    integer(4), allocatable :: imask(:)
    logical, allocatable :: mask(:)
    character(len=512) :: fin, fout
    integer :: u
    type(real64x2) :: s
    type(cmplx64x2) :: cs

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) m, k, n, nv
    allocate(a(m,k), b(k,n), ca(m,k), cb(k,n), x(nv), y(nv), cx(nv), cy(nv), imask(nv), mask(nv))
    read(u) a, b, ca, cb, x, y, cx, cy, imask
    close(u)
    mask = imask /= 0

    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    ! matmul with the default renorm interval (8), then 0
    write(u) matmul(a, b)
    write(u) matmul(a, b(:,1))
    write(u) matmul(a(1,:), b)          ! a non-contiguous row: copy-in
    write(u) matmul(ca, cb)
    write(u) matmul(ca, cb(:,1))
    write(u) matmul(ca(1,:), cb)
    call dd_set_fma_renorm_interval(0)
    write(u) matmul(a, b)
    s = dot_product(x, y)
    write(u) s
    call dd_set_fma_renorm_interval(8)
    s = dot_product(x, y)
    write(u) s
    cs = dot_product(cx, cy)
    write(u) cs
    s = sum(x)
    write(u) s
    s = sum(x, mask=mask)
    write(u) s
    s = sum(a)
    write(u) s
    write(u) sum(a, dim=1)
    write(u) sum(a, dim=2)
    s = norm2(x)
    write(u) s
    write(u) norm2(a, dim=1)
    write(u) norm2(a, dim=2)
    write(u) exp(x)
    write(u) log(x)
    write(u) sin(x)
    write(u) erf(x)
    write(u) sqrt(x)
    write(u) x + y
    write(u) x - y
    write(u) x * y
    write(u) x / y
    write(u) exp(a)                     ! rank 2
    write(u) dd_matmul_ozaki(a, b)
    write(u) cos(x)
    write(u) tan(x)
    write(u) expm1(x)
    write(u) log2(x * x)
    write(u) log10(x * x)
    write(u) log1p(x)
    write(u) sinh(x)
    write(u) cosh(x)
    write(u) tanh(x)
    write(u) erfc(x)
    write(u) erfc_scaled(x)
    write(u) asinh(x)
    write(u) acosh(x * x + exp(y * y))  ! argument >= 1
    write(u) atanh(x)
    write(u) sinpi(x + y)
    write(u) cospi(x - y)
    write(u) tanpi(x)
    write(u) hypot(x, y)
    write(u) sign(x, y)
    write(u) atan(x)
    write(u) atanpi(x)
    write(u) atan2(x, y)
    write(u) atan2pi(x, y)
    write(u) asin(x * y / (x * x + y * y))   ! |argument| <= 1/2
    write(u) acos(x * y / (x * x + y * y))
    write(u) min(x, y)
    write(u) max(x, y)
    write(u) dim(x, y)
    write(u) mod(x, y)
    write(u) modulo(x, y)
    write(u) nearest(x, y)
    ! xGEMM: C = x(1) A B + y(1) C with A passed transposed and C padded to m + 2 rows
    at = transpose(a)
    allocate(cc(m + 2, n))
    cc(1:m, :) = a(:, 1:n)
    cc(m + 1:m + 2, :) = x(1)
    call wgemm('T', 'N', int(m), int(n), int(k), x(1), at, int(k), b, int(k), y(1), cc, int(m + 2))
    write(u) cc
    close(u)
end program bridge_check
