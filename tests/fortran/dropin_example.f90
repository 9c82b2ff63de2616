! A PROPACK caller as a maintainer would write it: the reference's calling
! sequences of DLANSVD and DLANSVD_IRL (double/dlansvd.F:1-3,
! double/dlansvd_irl.F:1-3), linked against libpropack_amd instead of the
! reference.  First with the caller's own host APROD (the reference's
! callback interface, double/dlansvd.F:20-33), then with the library's
! device CSR operator (pk_dcsr_aprod + IPARM from pk_csr_to_iparm).
! Built by the test build step; run by tests/test_dropin_programs.py.
module example_mat
  implicit none
  integer, parameter :: m = 300, n = 200, per = 6
  integer :: rowptr(m + 1), colind(m * per)
  double precision :: val(m * per)
contains
  subroutine build()
    integer :: i, t, q
    q = 0
    rowptr(1) = 0
    do i = 1, m
      do t = 1, per
        q = q + 1
        colind(q) = mod(i * 7 + t * 31 + t * t, n)        ! 0-based column
        val(q) = sin(0.37d0 * i + 1.1d0 * t) + 0.1d0 * t
      end do
      rowptr(i + 1) = q
    end do
  end subroutine
end module

subroutine my_aprod(transa, mm, nn, x, y, dparm, iparm)
  use example_mat
  implicit none
  character(len=1) :: transa
  integer :: mm, nn, iparm(*)
  double precision :: x(*), y(*), dparm(*)
  integer :: i, q
  if (transa == 'n' .or. transa == 'N') then
    do i = 1, mm
      y(i) = 0d0
      do q = rowptr(i) + 1, rowptr(i + 1)
        y(i) = y(i) + val(q) * x(colind(q) + 1)
      end do
    end do
  else
    y(1:nn) = 0d0
    do i = 1, mm
      do q = rowptr(i) + 1, rowptr(i + 1)
        y(colind(q) + 1) = y(colind(q) + 1) + val(q) * x(i)
      end do
    end do
  end if
end subroutine

program dropin_example
  use iso_c_binding
  use example_mat
  implicit none
  interface
    integer(c_int) function pk_csr_create(mm, nn, nnz, rp, ci, va, prec, where, op) &
        bind(C, name='pk_csr_create')
      import
      integer(c_int), value :: mm, nn, prec, where
      integer(c_long_long), value :: nnz
      integer(c_int) :: rp(*), ci(*)
      real(c_double) :: va(*)
      type(c_ptr) :: op
    end function
    integer(c_int) function pk_csr_to_iparm(op, iparm) bind(C, name='pk_csr_to_iparm')
      import
      type(c_ptr), value :: op
      integer(c_int) :: iparm(*)
    end function
  end interface
  integer, parameter :: kmax = 200, nsv = 8
  external :: my_aprod, pk_dcsr_aprod
  double precision :: U(m, kmax + 1), V(n, kmax), sigma(kmax), bnd(kmax), doption(4), dparm(1)
  double precision :: work(1), tol
  integer :: iwork(1), ioption(2), iparm(2), info, k, dim, neig, i
  type(c_ptr) :: op

  call build()
  tol = 1d-12
  doption(1) = sqrt(epsilon(1d0))
  doption(2) = epsilon(1d0)**0.75d0
  doption(3) = 0d0
  doption(4) = 0.002d0
  ioption(1) = 1
  ioption(2) = 1
  ! start vector U(:,1)
  do i = 1, m
    U(i, 1) = cos(0.11d0 * i)
  end do

  ! 1) host APROD, non-restarted driver
  k = nsv
  call dlansvd('y', 'y', m, n, k, kmax, my_aprod, U, m, sigma, bnd, V, n, tol, work, 1, &
               iwork, 1, doption, ioption, info, dparm, iparm)
  write (*, '(a,i6,i6)') 'HOST_INFO_K ', info, k
  write (*, '(a,8es26.17)') 'HOST_SIGMA ', sigma(1:nsv)

  ! 2) device CSR operator, restarted driver
  info = pk_csr_create(m, n, int(rowptr(m + 1), c_long_long), rowptr, colind, val, 0, 0, op)
  if (info /= 0) stop 2
  info = pk_csr_to_iparm(op, iparm)
  do i = 1, m
    U(i, 1) = cos(0.11d0 * i)
  end do
  dim = 30
  neig = nsv
  doption(3) = 0d0
  call dlansvd_irl('L', 'y', 'y', m, n, dim, dim - nsv, neig, 1000, pk_dcsr_aprod, U, m, &
                   sigma, bnd, V, n, tol, work, 1, iwork, 1, doption, ioption, info, dparm, &
                   iparm)
  write (*, '(a,i6,i6,i6)') 'DEV_INFO_K ', info, neig, dim
  write (*, '(a,8es26.17)') 'DEV_SIGMA ', sigma(1:nsv)
  ! residual of the first triplet through the host operator
  call my_aprod('n', m, n, V(1, 1), U(1, kmax + 1), dparm, iparm)
  write (*, '(a,es12.4)') 'DEV_RESID1 ', norm2(U(:, kmax + 1) - sigma(1) * U(:, 1)) / sigma(1)
end program
