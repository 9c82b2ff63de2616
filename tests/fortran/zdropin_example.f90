! A complex PROPACK caller as a maintainer would write it: the reference's
! calling sequences of ZLANSVD and ZLANSVD_IRL (complex16/zlansvd.F:1-3,
! complex16/zlansvd_irl.F:1-3), linked against libpropack_amd instead of
! the reference.  First with the caller's own host APROD ('n': Y = A X,
! 'c': Y = A^H X, complex16/zlanbpro.F:28-41), then with the library's
! device CSR operator (pk_zcsr_aprod + IPARM from pk_csr_to_iparm).
! Built by the test build step; run by tests/test_dropin_programs.py.
module zexample_mat
  implicit none
  integer, parameter :: m = 260, n = 190, per = 5
  integer :: rowptr(m + 1), colind(m * per)
  complex(8) :: val(m * per)
contains
  subroutine build()
    integer :: i, t, q
    q = 0
    rowptr(1) = 0
    do i = 1, m
      do t = 1, per
        q = q + 1
        colind(q) = mod(i * 13 + t * 37 + t * t, n)       ! 0-based column
        val(q) = cmplx(sin(0.29d0 * i + 0.9d0 * t) + 0.1d0 * t, &
                       cos(0.41d0 * i - 0.6d0 * t), kind=8)
      end do
      rowptr(i + 1) = q
    end do
  end subroutine
end module

subroutine my_zaprod(transa, mm, nn, x, y, zparm, iparm)
  use zexample_mat
  implicit none
  character(len=1) :: transa
  integer :: mm, nn, iparm(*)
  complex(8) :: x(*), y(*), zparm(*)
  integer :: i, q
  if (transa == 'n' .or. transa == 'N') then
    do i = 1, mm
      y(i) = (0d0, 0d0)
      do q = rowptr(i) + 1, rowptr(i + 1)
        y(i) = y(i) + val(q) * x(colind(q) + 1)
      end do
    end do
  else
    y(1:nn) = (0d0, 0d0)
    do i = 1, mm
      do q = rowptr(i) + 1, rowptr(i + 1)
        y(colind(q) + 1) = y(colind(q) + 1) + conjg(val(q)) * x(i)
      end do
    end do
  end if
end subroutine

program zdropin_example
  use iso_c_binding
  use zexample_mat
  implicit none
  interface
    integer(c_int) function pk_csr_create(mm, nn, nnz, rp, ci, va, prec, where, op) &
        bind(C, name='pk_csr_create')
      import
      integer(c_int), value :: mm, nn, prec, where
      integer(c_long_long), value :: nnz
      integer(c_int) :: rp(*), ci(*)
      complex(c_double_complex) :: va(*)
      type(c_ptr) :: op
    end function
    integer(c_int) function pk_csr_to_iparm(op, iparm) bind(C, name='pk_csr_to_iparm')
      import
      type(c_ptr), value :: op
      integer(c_int) :: iparm(*)
    end function
  end interface
  integer, parameter :: kmax = 190, nsv = 6
  external :: my_zaprod, pk_zcsr_aprod
  complex(8) :: U(m, kmax + 1), V(n, kmax), zwork(1), zparm(1)
  double precision :: sigma(kmax), bnd(kmax), doption(4), work(1), tol
  integer :: iwork(1), ioption(2), iparm(2), info, k, dim, neig, i
  type(c_ptr) :: op

  call build()
  tol = 1d-12
  doption(1) = sqrt(epsilon(1d0))
  doption(2) = epsilon(1d0)**0.75d0
  doption(3) = 0d0
  doption(4) = 0.002d0
  ioption(1) = 0
  ioption(2) = 1
  do i = 1, m
    U(i, 1) = cmplx(cos(0.11d0 * i), sin(0.07d0 * i), kind=8)
  end do

  ! 1) host APROD, non-restarted driver
  k = nsv
  call zlansvd('y', 'y', m, n, k, kmax, my_zaprod, U, m, sigma, bnd, V, n, tol, work, 1, &
               zwork, 1, iwork, 1, doption, ioption, info, zparm, iparm)
  write (*, '(a,i6,i6)') 'HOST_INFO_K ', info, k
  write (*, '(a,6es26.17)') 'HOST_SIGMA ', sigma(1:nsv)

  ! 2) device CSR operator, restarted driver
  info = pk_csr_create(m, n, int(rowptr(m + 1), c_long_long), rowptr, colind, val, 2, 0, op)
  if (info /= 0) stop 2
  info = pk_csr_to_iparm(op, iparm)
  do i = 1, m
    U(i, 1) = cmplx(cos(0.11d0 * i), sin(0.07d0 * i), kind=8)
  end do
  dim = 24
  neig = nsv
  doption(3) = 0d0
  call zlansvd_irl('L', 'y', 'y', m, n, dim, dim - nsv, neig, 1000, pk_zcsr_aprod, U, m, &
                   sigma, bnd, V, n, tol, work, 1, zwork, 1, iwork, 1, doption, ioption, info, &
                   zparm, iparm)
  write (*, '(a,i6,i6,i6)') 'DEV_INFO_K ', info, neig, dim
  write (*, '(a,6es26.17)') 'DEV_SIGMA ', sigma(1:nsv)
  call my_zaprod('n', m, n, V(1, 1), U(1, kmax + 1), zparm, iparm)
  write (*, '(a,es12.4)') 'DEV_RESID1 ', sqrt(sum(abs(U(:, kmax + 1) - sigma(1) * U(:, 1))**2)) &
      / sigma(1)
end program
