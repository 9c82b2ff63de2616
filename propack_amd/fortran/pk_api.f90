! iso_c_binding interfaces of libpropack_amd's C ABI (include/propack_amd.h)
! as used by the host Fortran drivers, plus the host-side helpers of the
! Lanczos SVD that stay on the CPU in double precision.
module pk_api
  use iso_c_binding
  implicit none
  private
  integer(c_int), parameter, public :: PK_F64 = 0, PK_F32 = 1, PK_C128 = 2, PK_C64 = 3
  integer(c_int), parameter, public :: PK_OK = 0

  public :: pk_lbd_create, pk_lbd_destroy, pk_lbd_put, pk_lbd_get, pk_lbd_nrm2, &
            pk_lbd_getu0, pk_lbd_lanbpro, pk_lbd_update, pk_lbd_sync, pk_lbd_stats, &
            pk_lbd_reset_stats
  public :: pk_lapy2, pk_lartg, pk_bdqr, pk_bsvdstep, pk_refinebounds, lower1

  interface
    integer(c_int) function pk_lbd_create(prec, m, n, kmax, op, eng) &
        bind(C, name='pk_lbd_create')
      import :: c_int, c_ptr
      integer(c_int), value :: prec, m, n, kmax
      type(c_ptr), value :: op
      type(c_ptr) :: eng
    end function
    integer(c_int) function pk_lbd_destroy(eng) bind(C, name='pk_lbd_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: eng
    end function
    integer(c_int) function pk_lbd_put(eng, which, col, x, prec) bind(C, name='pk_lbd_put')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: eng
      character(kind=c_char), value :: which
      integer(c_int), value :: col, prec
      type(*), intent(in) :: x(*)
    end function
    integer(c_int) function pk_lbd_get(eng, which, ncols, x, ldx, prec) &
        bind(C, name='pk_lbd_get')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: eng
      character(kind=c_char), value :: which
      integer(c_int), value :: ncols, ldx, prec
      type(*) :: x(*)
    end function
    integer(c_int) function pk_lbd_nrm2(eng, which, col, nrm) bind(C, name='pk_lbd_nrm2')
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: eng
      character(kind=c_char), value :: which
      integer(c_int), value :: col
      real(c_double) :: nrm
    end function
    integer(c_int) function pk_lbd_getu0(eng, transa, j, ntry, u0norm, icgs, anormest, ierr) &
        bind(C, name='pk_lbd_getu0')
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: eng
      character(kind=c_char), value :: transa
      integer(c_int), value :: j, ntry, icgs
      real(c_double) :: u0norm, anormest
      integer(c_int) :: ierr
    end function
    integer(c_int) function pk_lbd_lanbpro(eng, k0, k, alpha, beta, rnorm, doption, ioption, &
        ierr) bind(C, name='pk_lbd_lanbpro')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: eng
      integer(c_int), value :: k0
      integer(c_int) :: k, ierr
      real(c_double) :: alpha(*), beta(*), rnorm, doption(*)
      integer(c_int) :: ioption(*)
    end function
    integer(c_int) function pk_lbd_update(eng, which, transb, kdim, ncols, b, ldb) &
        bind(C, name='pk_lbd_update')
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: eng
      character(kind=c_char), value :: which, transb
      integer(c_int), value :: kdim, ncols, ldb
      real(c_double), intent(in) :: b(*)
    end function
    integer(c_int) function pk_lbd_sync(eng) bind(C, name='pk_lbd_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: eng
    end function
    integer(c_int) function pk_lbd_stats(eng, out) bind(C, name='pk_lbd_stats')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: eng
      integer(c_long_long) :: out(*)
    end function
    integer(c_int) function pk_lbd_reset_stats(eng) bind(C, name='pk_lbd_reset_stats')
      import :: c_int, c_ptr
      type(c_ptr), value :: eng
    end function
  end interface

contains

  ! Lower-case a one-character option (the reference tests options with
  ! LSAME, i.e. case-insensitively).
  pure character(len=1) function lower1(c)
    character(len=1), intent(in) :: c
    if (c >= 'A' .and. c <= 'Z') then
      lower1 = achar(iachar(c) + 32)
    else
      lower1 = c
    end if
  end function

  ! sqrt(x**2 + y**2) without destructive overflow (LAPACK DLAPY2).
  pure double precision function pk_lapy2(x, y)
    double precision, intent(in) :: x, y
    double precision :: w, z
    w = max(abs(x), abs(y))
    z = min(abs(x), abs(y))
    if (z == 0d0) then
      pk_lapy2 = w
    else
      pk_lapy2 = w * sqrt(1d0 + (z / w)**2)
    end if
  end function

  ! Plane rotation [c s; -s c] [f; g] = [r; 0] as LAPACK 3.0's DLARTG defines
  ! it (double/Lapack_Util/dlartg.f): r = sqrt(f^2+g^2) computed with
  ! power-of-two rescaling when f, g are near the overflow/underflow limits,
  ! and c > 0 whenever |f| > |g|.
  subroutine pk_lartg(f, g, c, s, r)
    double precision, intent(in) :: f, g
    double precision, intent(out) :: c, s, r
    double precision :: f1, g1, scale, safmn2, safmx2
    integer :: count, i
    if (g == 0d0) then
      c = 1d0; s = 0d0; r = f
      return
    end if
    if (f == 0d0) then
      c = 0d0; s = 1d0; r = g
      return
    end if
    ! safmn2 = 2**int(log2(safmin/eps)/2), safmin = 2**-1022, eps = 2**-53
    safmn2 = 2d0**int(log(2d0**(-1022) / 2d0**(-53)) / log(2d0) / 2d0)
    safmx2 = 1d0 / safmn2
    f1 = f; g1 = g
    scale = max(abs(f1), abs(g1))
    count = 0
    if (scale >= safmx2) then
      do
        count = count + 1
        f1 = f1 * safmn2; g1 = g1 * safmn2
        scale = max(abs(f1), abs(g1))
        if (scale < safmx2) exit
      end do
      r = sqrt(f1**2 + g1**2)
      c = f1 / r; s = g1 / r
      do i = 1, count
        r = r * safmx2
      end do
    else if (scale <= safmn2) then
      do
        count = count + 1
        f1 = f1 * safmx2; g1 = g1 * safmx2
        scale = max(abs(f1), abs(g1))
        if (scale > safmn2) exit
      end do
      r = sqrt(f1**2 + g1**2)
      c = f1 / r; s = g1 / r
      do i = 1, count
        r = r * safmn2
      end do
    else
      r = sqrt(f1**2 + g1**2)
      c = f1 / r; s = g1 / r
    end if
    if (abs(f) > abs(g) .and. c < 0d0) then
      c = -c; s = -s; r = -r
    end if
  end subroutine

  ! QR factorisation of the (n+1) x n lower bidiagonal B (diagonal d,
  ! subdiagonal e) by n Givens rotations, as DBDQR does (double/dbsvd.F:87-157):
  ! on return d, e hold the upper bidiagonal R, [.. c1 c2]' = Q'[0 .. 0 1]',
  ! and, if jobq = 'y', qt = Q^T ((n+1) x (n+1), leading dimension ldq).
  ! With ignorelast, e(n) is treated as zero and c1, c2 are left untouched.
  subroutine pk_bdqr(ignorelast, jobq, n, d, e, c1, c2, qt, ldq)
    logical, intent(in) :: ignorelast
    character(len=1), intent(in) :: jobq
    integer, intent(in) :: n, ldq
    double precision, intent(inout) :: d(*), e(*), c1, c2, qt(ldq, *)
    integer :: i, j
    double precision :: cs, sn, r
    logical :: wantq
    if (n < 1) return
    wantq = lower1(jobq) == 'y'
    if (wantq) then
      do j = 1, n + 1
        qt(1:n + 1, j) = 0d0
        qt(j, j) = 1d0
      end do
    end if
    do i = 1, n - 1
      call pk_lartg(d(i), e(i), cs, sn, r)
      d(i) = r
      e(i) = sn * d(i + 1)
      d(i + 1) = cs * d(i + 1)
      if (wantq) then
        do j = 1, i
          qt(i + 1, j) = -sn * qt(i, j)
          qt(i, j) = cs * qt(i, j)
        end do
        qt(i, i + 1) = sn
        qt(i + 1, i + 1) = cs
      end if
    end do
    if (.not. ignorelast) then
      call pk_lartg(d(n), e(n), cs, sn, r)
      d(n) = r
      e(n) = 0d0
      c1 = sn
      c2 = cs
      if (wantq) then
        ! the last rotation acts on rows n and n+1 of Q^T
        do j = 1, n
          qt(n + 1, j) = -sn * qt(n, j)
          qt(n, j) = cs * qt(n, j)
        end do
        qt(n, n + 1) = sn
        qt(n + 1, n + 1) = cs
      end if
    end if
  end subroutine

  ! One implicit shifted LQ sweep on the lower bidiagonal (d, e) of order k
  ! (DBSVDSTEP, double/dbsvd.F:5-82), accumulating the left rotations into
  ! columns of u (m rows) and the right ones into columns of v (n rows).
  subroutine pk_bsvdstep(jobu, jobv, m, n, k, sigma, d, e, u, ldu, v, ldv)
    character(len=1), intent(in) :: jobu, jobv
    integer, intent(in) :: m, n, k, ldu, ldv
    double precision, intent(in) :: sigma
    double precision, intent(inout) :: d(*), e(*), u(ldu, *), v(ldv, *)
    integer :: i
    double precision :: c, s, x, y, r
    logical :: dou, dov
    if (k <= 1) return
    dou = lower1(jobu) == 'y'
    dov = lower1(jobv) == 'y'
    ! first rotation from the shifted first column of B B^T - sigma^2 I
    x = d(1) * d(1) - sigma * sigma
    y = e(1) * d(1)
    do i = 1, k - 1
      if (i > 1) then
        call pk_lartg(x, y, c, s, e(i - 1))
      else
        call pk_lartg(x, y, c, s, r)
      end if
      x = c * d(i) + s * e(i)
      e(i) = -s * d(i) + c * e(i)
      d(i) = x
      y = s * d(i + 1)
      d(i + 1) = c * d(i + 1)
      if (dou .and. m > 0) call rot(m, u(1, i), u(1, i + 1), c, s)
      call pk_lartg(x, y, c, s, d(i))
      x = c * e(i) + s * d(i + 1)
      d(i + 1) = -s * e(i) + c * d(i + 1)
      e(i) = x
      y = s * e(i + 1)
      e(i + 1) = c * e(i + 1)
      if (dov .and. n > 0) call rot(n, v(1, i), v(1, i + 1), c, s)
    end do
    call pk_lartg(x, y, c, s, e(k - 1))
    x = c * d(k) + s * e(k)
    e(k) = -s * d(k) + c * e(k)
    d(k) = x
    if (dou .and. m > 0) call rot(m, u(1, k), u(1, k + 1), c, s)
  contains
    subroutine rot(nn, a, b, cc, ss)
      integer, intent(in) :: nn
      double precision, intent(inout) :: a(nn), b(nn)
      double precision, intent(in) :: cc, ss
      integer :: l
      double precision :: t
      do l = 1, nn
        t = cc * a(l) + ss * b(l)
        b(l) = cc * b(l) - ss * a(l)
        a(l) = t
      end do
    end subroutine
  end subroutine

  ! Error-bound refinement (DREFINEBOUNDS, double/dbsvd.F:162-231): bounds of
  ! Ritz values closer than eps34*theta are merged, then every bound is
  ! sharpened by the gap theorem.  n is min(m, n) of A.
  subroutine pk_refinebounds(n, k, theta, bound, tol, eps34)
    integer, intent(in) :: n, k
    double precision, intent(in) :: theta(*), tol, eps34
    double precision, intent(inout) :: bound(*)
    integer :: i, l
    double precision :: gap
    if (k <= 1) return
    do i = 1, k
      do l = -1, 1, 2
        if ((l == 1 .and. i < k) .or. (l == -1 .and. i > 1)) then
          if (abs(theta(i) - theta(i + l)) < eps34 * theta(i)) then
            if (bound(i) > tol .and. bound(i + l) > tol) then
              bound(i + l) = pk_lapy2(bound(i), bound(i + l))
              bound(i) = 0d0
            end if
          end if
        end if
      end do
    end do
    do i = 1, k
      if (i < k .or. k == n) then
        if (i == 1) then
          gap = abs(theta(i) - theta(i + 1)) - max(bound(i), bound(i + 1))
        else if (i == n) then
          gap = abs(theta(i - 1) - theta(i)) - max(bound(i - 1), bound(i))
        else
          gap = abs(theta(i) - theta(i + 1)) - max(bound(i), bound(i + 1))
          gap = min(gap, abs(theta(i - 1) - theta(i)) - max(bound(i - 1), bound(i)))
        end if
        if (gap > bound(i)) bound(i) = bound(i) * (bound(i) / gap)
      end if
    end do
  end subroutine

end module pk_api
