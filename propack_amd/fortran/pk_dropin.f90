! Drop-in replacements for PROPACK's public drivers with identical calling
! sequences and IN/OUT semantics:
!   DLANSVD     (double/dlansvd.F:1-3)         SLANSVD     (single/slansvd.F:1-3)
!   DLANSVD_IRL (double/dlansvd_irl.F:1-3)     SLANSVD_IRL (single/slansvd_irl.F:1-3)
!   ZLANSVD     (complex16/zlansvd.F:1-3)      CLANSVD     (complex8/clansvd.F:1-3)
!   ZLANSVD_IRL (complex16/zlansvd_irl.F:1-3)  CLANSVD_IRL (complex8/clansvd_irl.F:1-3)
! plus CLEARSTAT / PRINTSTAT over COMMON /timing/ (double/printstat.F:4-75).
! The complex drivers run the real drivers' code (the bidiagonal B is real
! in both trees) over an engine with complex bases.
!
! APROD selection: when APROD is the library's PK_DCSR_APROD (or
! PK_SCSR_APROD, PK_ZCSR_APROD, PK_CCSR_APROD) and IPARM(1:2) holds a device
! operator handle (pk_csr_to_iparm), the whole Lanczos process stays on the
! GPU.  Any other APROD is called from the host with X/Y staged through
! pinned memory.
!
! WORK/ZWORK/IWORK are accepted with the reference's sizes and not used (the
! engine owns its device workspace).  K / NEIG / DIM and DOPTION(3) are
! written back as the reference does.  Errors of the device layer return
! INFO <= -100 instead of stopping the process.

module pk_dropin_util
  use iso_c_binding
  use pk_api
  implicit none
contains
  ! Create an engine for A of shape m x n holding kmax right Lanczos
  ! vectors.  dev = .true. when APROD is the library's own operator symbol
  ! of the matching precision with a live handle in IPARM(1:2); otherwise
  ! the caller registers APROD as a host callback.
  subroutine open_engine(prec, m, n, kmax, aprod, iparm, eng, dev, info)
    integer(c_int), intent(in) :: prec
    integer, intent(in) :: m, n, kmax
    external :: aprod
    integer :: iparm(*)
    type(c_ptr), intent(out) :: eng
    logical, intent(out) :: dev
    integer, intent(out) :: info
    integer :: kind
    integer(c_long_long) :: handle
    type(c_ptr) :: op
    integer(c_int) :: rc
    call pk_f_classify(aprod, iparm, kind, handle)
    op = c_null_ptr
    dev = (kind == 1 .and. prec == PK_F64) .or. (kind == 2 .and. prec == PK_F32) .or. &
          (kind == 3 .and. prec == PK_C128) .or. (kind == 4 .and. prec == PK_C64)
    if (dev) op = transfer(handle, op)
    eng = c_null_ptr
    rc = pk_lbd_create(prec, m, n, kmax, op, eng)
    info = rc
  end subroutine

  integer(c_long_long) function engine_id(eng)
    type(c_ptr), intent(in) :: eng
    engine_id = transfer(eng, engine_id)
  end function
end module pk_dropin_util

subroutine dlansvd(jobu, jobv, m, n, k, kmax, aprod, U, ldu, sigma, bnd, V, ldv, tolin, work, &
                   lwork, iwork, liwork, doption, ioption, info, dparm, iparm)
  use iso_c_binding
  use pk_api
  use pk_drivers
  use pk_dropin_util
  implicit none
  character(len=1) :: jobu, jobv
  integer :: m, n, k, kmax, ldu, ldv, lwork, liwork, info
  integer :: iwork(*), ioption(*), iparm(*)
  double precision :: U(ldu, *), sigma(*), bnd(*), V(ldv, *), tolin, work(*), doption(*), dparm(*)
  external :: aprod
  type(c_ptr) :: eng
  logical :: dev
  integer :: lanmax, kk
  integer(c_int) :: rc
  if (m < 1 .or. n < 1 .or. k < 1 .or. kmax < 1 .or. ldu < m .or. ldv < n) then
    info = -101
    return
  end if
  lanmax = min(n + 1, m + 1, kmax)
  call open_engine(PK_F64, m, n, lanmax, aprod, iparm, eng, dev, info)
  if (info /= 0) return
  if (.not. dev) then
    call pk_f_set_callback_d(engine_id(eng), aprod, dparm, iparm, rc)
    if (rc /= PK_OK) info = rc
  end if
  if (info /= 0) then
    rc = pk_lbd_destroy(eng)
    return
  end if
  rc = pk_lbd_put(eng, 'u', 1, U(1, 1), PK_F64)
  if (rc /= PK_OK) then
    info = rc
  else
    kk = k
    call lansvd_core(jobu, jobv, m, n, kk, lanmax, eng, 2d0**(-53), sigma, bnd, tolin, &
                     doption, ioption, info)
    k = kk
    if (info > -100 .and. lower1(jobu) == 'y' .and. k > 0) then
      rc = pk_lbd_get(eng, 'u', k, U, ldu, PK_F64)
      if (rc /= PK_OK) info = rc
    end if
    if (info > -100 .and. lower1(jobv) == 'y' .and. k > 0) then
      rc = pk_lbd_get(eng, 'v', k, V, ldv, PK_F64)
      if (rc /= PK_OK) info = rc
    end if
  end if
  rc = pk_lbd_destroy(eng)
end subroutine

subroutine dlansvd_irl(which, jobu, jobv, m, n, dim, p, neig, maxiter, aprod, U, ldu, sigma, &
                       bnd, V, ldv, tolin, work, lwork, iwork, liwork, doption, ioption, info, &
                       dparm, iparm)
  use iso_c_binding
  use pk_api
  use pk_drivers
  use pk_dropin_util
  implicit none
  character(len=1) :: which, jobu, jobv
  integer :: m, n, dim, p, neig, maxiter, ldu, ldv, lwork, liwork, info
  integer :: iwork(*), ioption(*), iparm(*)
  double precision :: U(ldu, *), sigma(*), bnd(*), V(ldv, *), tolin, work(*), doption(*), dparm(*)
  external :: aprod
  type(c_ptr) :: eng
  logical :: dev
  integer :: dd, ne
  integer(c_int) :: rc
  if (m < 1 .or. n < 1 .or. dim < 1 .or. neig < 1 .or. p < 0 .or. ldu < m .or. ldv < n) then
    info = -101
    return
  end if
  dd = min(dim, n + 1, m + 1)
  call open_engine(PK_F64, m, n, dd, aprod, iparm, eng, dev, info)
  if (info /= 0) return
  if (.not. dev) then
    call pk_f_set_callback_d(engine_id(eng), aprod, dparm, iparm, rc)
    if (rc /= PK_OK) info = rc
  end if
  if (info /= 0) then
    rc = pk_lbd_destroy(eng)
    return
  end if
  rc = pk_lbd_put(eng, 'u', 1, U(1, 1), PK_F64)
  if (rc /= PK_OK) then
    info = rc
  else
    ne = neig
    call lansvd_irl_core(which, jobu, jobv, m, n, dim, p, ne, maxiter, eng, 2d0**(-53), &
                         sigma, bnd, tolin, doption, ioption, info)
    neig = ne
    if (info > -100 .and. lower1(jobu) == 'y' .and. neig > 0) then
      rc = pk_lbd_get(eng, 'u', neig, U, ldu, PK_F64)
      if (rc /= PK_OK) info = rc
    end if
    if (info > -100 .and. lower1(jobv) == 'y' .and. neig > 0) then
      rc = pk_lbd_get(eng, 'v', neig, V, ldv, PK_F64)
      if (rc /= PK_OK) info = rc
    end if
  end if
  rc = pk_lbd_destroy(eng)
end subroutine

subroutine slansvd(jobu, jobv, m, n, k, kmax, aprod, U, ldu, sigma, bnd, V, ldv, tolin, work, &
                   lwork, iwork, liwork, soption, ioption, info, sparm, iparm)
  use iso_c_binding
  use pk_api
  use pk_drivers
  use pk_dropin_util
  implicit none
  character(len=1) :: jobu, jobv
  integer :: m, n, k, kmax, ldu, ldv, lwork, liwork, info
  integer :: iwork(*), ioption(*), iparm(*)
  real :: U(ldu, *), sigma(*), bnd(*), V(ldv, *), tolin, work(*), soption(*), sparm(*)
  external :: aprod
  type(c_ptr) :: eng
  logical :: dev
  integer :: lanmax, kk, k0
  integer(c_int) :: rc
  double precision, allocatable :: sd(:), bd(:)
  double precision :: dopt(4)
  if (m < 1 .or. n < 1 .or. k < 1 .or. kmax < 1 .or. ldu < m .or. ldv < n) then
    info = -101
    return
  end if
  lanmax = min(n + 1, m + 1, kmax)
  call open_engine(PK_F32, m, n, lanmax, aprod, iparm, eng, dev, info)
  if (info /= 0) return
  if (.not. dev) then
    call pk_f_set_callback_s(engine_id(eng), aprod, sparm, iparm, rc)
    if (rc /= PK_OK) info = rc
  end if
  if (info /= 0) then
    rc = pk_lbd_destroy(eng)
    return
  end if
  rc = pk_lbd_put(eng, 'u', 1, U(1, 1), PK_F32)
  if (rc /= PK_OK) then
    info = rc
  else
    k0 = k
    allocate(sd(k0), bd(k0))
    sd = dble(sigma(1:k0))
    bd = dble(bnd(1:k0))
    dopt = 0d0
    dopt(1:3) = dble(soption(1:3))
    kk = k
    call lansvd_core(jobu, jobv, m, n, kk, lanmax, eng, 2d0**(-24), sd, bd, dble(tolin), dopt, &
                     ioption, info)
    k = kk
    sigma(1:k0) = real(sd)
    bnd(1:k0) = real(bd)
    soption(3) = real(dopt(3))
    if (info > -100 .and. lower1(jobu) == 'y' .and. k > 0) then
      rc = pk_lbd_get(eng, 'u', k, U, ldu, PK_F32)
      if (rc /= PK_OK) info = rc
    end if
    if (info > -100 .and. lower1(jobv) == 'y' .and. k > 0) then
      rc = pk_lbd_get(eng, 'v', k, V, ldv, PK_F32)
      if (rc /= PK_OK) info = rc
    end if
  end if
  rc = pk_lbd_destroy(eng)
end subroutine

subroutine slansvd_irl(which, jobu, jobv, m, n, dim, p, neig, maxiter, aprod, U, ldu, sigma, &
                       bnd, V, ldv, tolin, work, lwork, iwork, liwork, soption, ioption, info, &
                       sparm, iparm)
  use iso_c_binding
  use pk_api
  use pk_drivers
  use pk_dropin_util
  implicit none
  character(len=1) :: which, jobu, jobv
  integer :: m, n, dim, p, neig, maxiter, ldu, ldv, lwork, liwork, info
  integer :: iwork(*), ioption(*), iparm(*)
  real :: U(ldu, *), sigma(*), bnd(*), V(ldv, *), tolin, work(*), soption(*), sparm(*)
  external :: aprod
  type(c_ptr) :: eng
  logical :: dev
  integer :: dd, ne, n0
  integer(c_int) :: rc
  double precision, allocatable :: sd(:), bd(:)
  double precision :: dopt(4)
  if (m < 1 .or. n < 1 .or. dim < 1 .or. neig < 1 .or. p < 0 .or. ldu < m .or. ldv < n) then
    info = -101
    return
  end if
  dd = min(dim, n + 1, m + 1)
  call open_engine(PK_F32, m, n, dd, aprod, iparm, eng, dev, info)
  if (info /= 0) return
  if (.not. dev) then
    call pk_f_set_callback_s(engine_id(eng), aprod, sparm, iparm, rc)
    if (rc /= PK_OK) info = rc
  end if
  if (info /= 0) then
    rc = pk_lbd_destroy(eng)
    return
  end if
  rc = pk_lbd_put(eng, 'u', 1, U(1, 1), PK_F32)
  if (rc /= PK_OK) then
    info = rc
  else
    n0 = neig
    allocate(sd(n0), bd(n0))
    sd = dble(sigma(1:n0))
    bd = dble(bnd(1:n0))
    dopt = dble(soption(1:4))
    ne = neig
    call lansvd_irl_core(which, jobu, jobv, m, n, dim, p, ne, maxiter, eng, 2d0**(-24), sd, &
                         bd, dble(tolin), dopt, ioption, info)
    neig = ne
    sigma(1:n0) = real(sd)
    bnd(1:n0) = real(bd)
    soption(3) = real(dopt(3))
    if (info > -100 .and. lower1(jobu) == 'y' .and. neig > 0) then
      rc = pk_lbd_get(eng, 'u', neig, U, ldu, PK_F32)
      if (rc /= PK_OK) info = rc
    end if
    if (info > -100 .and. lower1(jobv) == 'y' .and. neig > 0) then
      rc = pk_lbd_get(eng, 'v', neig, V, ldv, PK_F32)
      if (rc /= PK_OK) info = rc
    end if
  end if
  rc = pk_lbd_destroy(eng)
end subroutine

subroutine zlansvd(jobu, jobv, m, n, k, kmax, aprod, U, ldu, sigma, bnd, V, ldv, tolin, work, &
                   lwork, zwork, lzwrk, iwork, liwork, doption, ioption, info, zparm, iparm)
  use iso_c_binding
  use pk_api
  use pk_drivers
  use pk_dropin_util
  implicit none
  character(len=1) :: jobu, jobv
  integer :: m, n, k, kmax, ldu, ldv, lwork, lzwrk, liwork, info
  integer :: iwork(*), ioption(*), iparm(*)
  complex(8) :: U(ldu, *), V(ldv, *), zwork(*), zparm(*)
  double precision :: sigma(*), bnd(*), tolin, work(*), doption(*)
  external :: aprod
  type(c_ptr) :: eng
  logical :: dev
  integer :: lanmax, kk
  integer(c_int) :: rc
  if (m < 1 .or. n < 1 .or. k < 1 .or. kmax < 1 .or. ldu < m .or. ldv < n) then
    info = -101
    return
  end if
  lanmax = min(n + 1, m + 1, kmax)
  call open_engine(PK_C128, m, n, lanmax, aprod, iparm, eng, dev, info)
  if (info /= 0) return
  if (.not. dev) then
    call pk_f_set_callback_z(engine_id(eng), aprod, zparm, iparm, rc)
    if (rc /= PK_OK) info = rc
  end if
  if (info /= 0) then
    rc = pk_lbd_destroy(eng)
    return
  end if
  rc = pk_lbd_put(eng, 'u', 1, U(1, 1), PK_C128)
  if (rc /= PK_OK) then
    info = rc
  else
    kk = k
    call lansvd_core(jobu, jobv, m, n, kk, lanmax, eng, 2d0**(-53), sigma, bnd, tolin, &
                     doption, ioption, info)
    k = kk
    if (info > -100 .and. lower1(jobu) == 'y' .and. k > 0) then
      rc = pk_lbd_get(eng, 'u', k, U, ldu, PK_C128)
      if (rc /= PK_OK) info = rc
    end if
    if (info > -100 .and. lower1(jobv) == 'y' .and. k > 0) then
      rc = pk_lbd_get(eng, 'v', k, V, ldv, PK_C128)
      if (rc /= PK_OK) info = rc
    end if
  end if
  rc = pk_lbd_destroy(eng)
end subroutine

subroutine zlansvd_irl(which, jobu, jobv, m, n, dim, p, neig, maxiter, aprod, U, ldu, sigma, &
                       bnd, V, ldv, tolin, work, lwork, zwork, lzwrk, iwork, liwork, doption, &
                       ioption, info, zparm, iparm)
  use iso_c_binding
  use pk_api
  use pk_drivers
  use pk_dropin_util
  implicit none
  character(len=1) :: which, jobu, jobv
  integer :: m, n, dim, p, neig, maxiter, ldu, ldv, lwork, lzwrk, liwork, info
  integer :: iwork(*), ioption(*), iparm(*)
  complex(8) :: U(ldu, *), V(ldv, *), zwork(*), zparm(*)
  double precision :: sigma(*), bnd(*), tolin, work(*), doption(*)
  external :: aprod
  type(c_ptr) :: eng
  logical :: dev
  integer :: dd, ne
  integer(c_int) :: rc
  if (m < 1 .or. n < 1 .or. dim < 1 .or. neig < 1 .or. p < 0 .or. ldu < m .or. ldv < n) then
    info = -101
    return
  end if
  dd = min(dim, n + 1, m + 1)
  call open_engine(PK_C128, m, n, dd, aprod, iparm, eng, dev, info)
  if (info /= 0) return
  if (.not. dev) then
    call pk_f_set_callback_z(engine_id(eng), aprod, zparm, iparm, rc)
    if (rc /= PK_OK) info = rc
  end if
  if (info /= 0) then
    rc = pk_lbd_destroy(eng)
    return
  end if
  rc = pk_lbd_put(eng, 'u', 1, U(1, 1), PK_C128)
  if (rc /= PK_OK) then
    info = rc
  else
    ne = neig
    call lansvd_irl_core(which, jobu, jobv, m, n, dim, p, ne, maxiter, eng, 2d0**(-53), &
                         sigma, bnd, tolin, doption, ioption, info)
    neig = ne
    if (info > -100 .and. lower1(jobu) == 'y' .and. neig > 0) then
      rc = pk_lbd_get(eng, 'u', neig, U, ldu, PK_C128)
      if (rc /= PK_OK) info = rc
    end if
    if (info > -100 .and. lower1(jobv) == 'y' .and. neig > 0) then
      rc = pk_lbd_get(eng, 'v', neig, V, ldv, PK_C128)
      if (rc /= PK_OK) info = rc
    end if
  end if
  rc = pk_lbd_destroy(eng)
end subroutine

subroutine clansvd(jobu, jobv, m, n, k, kmax, aprod, U, ldu, sigma, bnd, V, ldv, tolin, work, &
                   lwork, cwork, lcwrk, iwork, liwork, soption, ioption, info, cparm, iparm)
  use iso_c_binding
  use pk_api
  use pk_drivers
  use pk_dropin_util
  implicit none
  character(len=1) :: jobu, jobv
  integer :: m, n, k, kmax, ldu, ldv, lwork, lcwrk, liwork, info
  integer :: iwork(*), ioption(*), iparm(*)
  complex(4) :: U(ldu, *), V(ldv, *), cwork(*), cparm(*)
  real :: sigma(*), bnd(*), tolin, work(*), soption(*)
  external :: aprod
  type(c_ptr) :: eng
  logical :: dev
  integer :: lanmax, kk, k0
  integer(c_int) :: rc
  double precision, allocatable :: sd(:), bd(:)
  double precision :: dopt(4)
  if (m < 1 .or. n < 1 .or. k < 1 .or. kmax < 1 .or. ldu < m .or. ldv < n) then
    info = -101
    return
  end if
  lanmax = min(n + 1, m + 1, kmax)
  call open_engine(PK_C64, m, n, lanmax, aprod, iparm, eng, dev, info)
  if (info /= 0) return
  if (.not. dev) then
    call pk_f_set_callback_c(engine_id(eng), aprod, cparm, iparm, rc)
    if (rc /= PK_OK) info = rc
  end if
  if (info /= 0) then
    rc = pk_lbd_destroy(eng)
    return
  end if
  rc = pk_lbd_put(eng, 'u', 1, U(1, 1), PK_C64)
  if (rc /= PK_OK) then
    info = rc
  else
    k0 = k
    allocate(sd(k0), bd(k0))
    sd = dble(sigma(1:k0))
    bd = dble(bnd(1:k0))
    dopt = 0d0
    dopt(1:3) = dble(soption(1:3))
    kk = k
    call lansvd_core(jobu, jobv, m, n, kk, lanmax, eng, 2d0**(-24), sd, bd, dble(tolin), dopt, &
                     ioption, info)
    k = kk
    sigma(1:k0) = real(sd)
    bnd(1:k0) = real(bd)
    soption(3) = real(dopt(3))
    if (info > -100 .and. lower1(jobu) == 'y' .and. k > 0) then
      rc = pk_lbd_get(eng, 'u', k, U, ldu, PK_C64)
      if (rc /= PK_OK) info = rc
    end if
    if (info > -100 .and. lower1(jobv) == 'y' .and. k > 0) then
      rc = pk_lbd_get(eng, 'v', k, V, ldv, PK_C64)
      if (rc /= PK_OK) info = rc
    end if
  end if
  rc = pk_lbd_destroy(eng)
end subroutine

subroutine clansvd_irl(which, jobu, jobv, m, n, dim, p, neig, maxiter, aprod, U, ldu, sigma, &
                       bnd, V, ldv, tolin, work, lwork, cwork, lcwrk, iwork, liwork, soption, &
                       ioption, info, cparm, iparm)
  use iso_c_binding
  use pk_api
  use pk_drivers
  use pk_dropin_util
  implicit none
  character(len=1) :: which, jobu, jobv
  integer :: m, n, dim, p, neig, maxiter, ldu, ldv, lwork, lcwrk, liwork, info
  integer :: iwork(*), ioption(*), iparm(*)
  complex(4) :: U(ldu, *), V(ldv, *), cwork(*), cparm(*)
  real :: sigma(*), bnd(*), tolin, work(*), soption(*)
  external :: aprod
  type(c_ptr) :: eng
  logical :: dev
  integer :: dd, ne, n0
  integer(c_int) :: rc
  double precision, allocatable :: sd(:), bd(:)
  double precision :: dopt(4)
  if (m < 1 .or. n < 1 .or. dim < 1 .or. neig < 1 .or. p < 0 .or. ldu < m .or. ldv < n) then
    info = -101
    return
  end if
  dd = min(dim, n + 1, m + 1)
  call open_engine(PK_C64, m, n, dd, aprod, iparm, eng, dev, info)
  if (info /= 0) return
  if (.not. dev) then
    call pk_f_set_callback_c(engine_id(eng), aprod, cparm, iparm, rc)
    if (rc /= PK_OK) info = rc
  end if
  if (info /= 0) then
    rc = pk_lbd_destroy(eng)
    return
  end if
  rc = pk_lbd_put(eng, 'u', 1, U(1, 1), PK_C64)
  if (rc /= PK_OK) then
    info = rc
  else
    n0 = neig
    allocate(sd(n0), bd(n0))
    sd = dble(sigma(1:n0))
    bd = dble(bnd(1:n0))
    dopt = dble(soption(1:4))
    ne = neig
    call lansvd_irl_core(which, jobu, jobv, m, n, dim, p, ne, maxiter, eng, 2d0**(-24), sd, &
                         bd, dble(tolin), dopt, ioption, info)
    neig = ne
    sigma(1:n0) = real(sd)
    bnd(1:n0) = real(bd)
    soption(3) = real(dopt(3))
    if (info > -100 .and. lower1(jobu) == 'y' .and. neig > 0) then
      rc = pk_lbd_get(eng, 'u', neig, U, ldu, PK_C64)
      if (rc /= PK_OK) info = rc
    end if
    if (info > -100 .and. lower1(jobv) == 'y' .and. neig > 0) then
      rc = pk_lbd_get(eng, 'v', neig, V, ldv, PK_C64)
      if (rc /= PK_OK) info = rc
    end if
  end if
  rc = pk_lbd_destroy(eng)
end subroutine

! Zero the counters of COMMON /timing/ (double/printstat.F:4-31).
subroutine clearstat()
  implicit none
  integer :: nopx, nreorth, ndot, nreorthu, nreorthv, nitref, nrestart, nbsvd, nlandim, nsing
  real :: tmvopx, tgetu0, tupdmu, tupdnu, tintv, tlanbpro, treorth, treorthu, treorthv, &
          telru, telrv, tbsvd, tnorm2, tlansvd, tritzvec, trestart, tdot
  common /timing/ nopx, nreorth, ndot, nreorthu, nreorthv, nitref, nrestart, nbsvd, tmvopx, &
      tgetu0, tupdmu, tupdnu, tintv, tlanbpro, treorth, treorthu, treorthv, telru, telrv, &
      tbsvd, tnorm2, tlansvd, nlandim, tritzvec, trestart, tdot, nsing
  nopx = 0; nreorth = 0; ndot = 0; nreorthu = 0; nreorthv = 0; nitref = 0
  nrestart = 0; nbsvd = 0; nlandim = 0; nsing = 0
  tmvopx = 0; tgetu0 = 0; tupdmu = 0; tupdnu = 0; tintv = 0; tlanbpro = 0; treorth = 0
  treorthu = 0; treorthv = 0; telru = 0; telrv = 0; tbsvd = 0; tnorm2 = 0; tlansvd = 0
  tritzvec = 0; trestart = 0; tdot = 0
end subroutine

! Print the counters of COMMON /timing/ (double/printstat.F:35-75).
subroutine printstat()
  implicit none
  integer :: nopx, nreorth, ndot, nreorthu, nreorthv, nitref, nrestart, nbsvd, nlandim, nsing
  real :: tmvopx, tgetu0, tupdmu, tupdnu, tintv, tlanbpro, treorth, treorthu, treorthv, &
          telru, telrv, tbsvd, tnorm2, tlansvd, tritzvec, trestart, tdot
  common /timing/ nopx, nreorth, ndot, nreorthu, nreorthv, nitref, nrestart, nbsvd, tmvopx, &
      tgetu0, tupdmu, tupdnu, tintv, tlanbpro, treorth, treorthu, treorthv, telru, telrv, &
      tbsvd, tnorm2, tlansvd, nlandim, tritzvec, trestart, tdot, nsing
  write (*, '(a)') ' -------- Lanczos SVD statistics --------'
  write (*, '(a,i12)') ' Rank of approximation           = ', nsing
  write (*, '(a,i12)') ' Dimension of Lanczos basis      = ', nlandim
  write (*, '(a,i12)') ' Number of matrix-vector products= ', nopx
  write (*, '(a,i12)') ' Number of reorthogonalizations  = ', nreorth
  write (*, '(a,i12)') ' Number of inner products in reorth = ', ndot
  write (*, '(a,i12)') ' Number of restarts              = ', nrestart
  write (*, '(a,i12)') ' Number of bidiagonal SVDs       = ', nbsvd
end subroutine

! Host bidiagonal helpers of the drivers, exported for verification against
! the reference routines (double/dbsvd.F, double/Lapack_Util/dlartg.f).
subroutine pkh_lartg(f, g, c, s, r) bind(C, name='pkh_lartg')
  use iso_c_binding
  use pk_api
  implicit none
  real(c_double), value :: f, g
  real(c_double) :: c, s, r
  call pk_lartg(f, g, c, s, r)
end subroutine

subroutine pkh_bdqr(ignorelast, jobq, n, d, e, c1, c2, qt, ldq) bind(C, name='pkh_bdqr')
  use iso_c_binding
  use pk_api
  implicit none
  integer(c_int), value :: ignorelast, n, ldq
  character(kind=c_char), value :: jobq
  real(c_double) :: d(*), e(*), c1, c2, qt(ldq, *)
  call pk_bdqr(ignorelast /= 0, jobq, n, d, e, c1, c2, qt, ldq)
end subroutine

subroutine pkh_bsvdstep(m, n, k, sigma, d, e, u, ldu, v, ldv) bind(C, name='pkh_bsvdstep')
  use iso_c_binding
  use pk_api
  implicit none
  integer(c_int), value :: m, n, k, ldu, ldv
  real(c_double), value :: sigma
  real(c_double) :: d(*), e(*), u(ldu, *), v(ldv, *)
  call pk_bsvdstep('y', 'y', m, n, k, sigma, d, e, u, ldu, v, ldv)
end subroutine

subroutine pkh_refinebounds(n, k, theta, bound, tol, eps34) bind(C, name='pkh_refinebounds')
  use iso_c_binding
  use pk_api
  implicit none
  integer(c_int), value :: n, k
  real(c_double), value :: tol, eps34
  real(c_double) :: theta(*), bound(*)
  call pk_refinebounds(n, k, theta, bound, tol, eps34)
end subroutine
