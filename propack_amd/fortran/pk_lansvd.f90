! Host drivers of the truncated SVD: the outer convergence / restart loops
! of DLANSVD (double/dlansvd.F:1-291), DLANSVD_IRL (double/dlansvd_irl.F:1-419)
! and the Ritz-vector assembly of DRITZVEC (double/dritzvec.F:1-200), written
! against the device engine of libpropack_amd.  The Lanczos steps, the
! bases and the basis-update GEMMs run on the GPU; the small bidiagonal
! problems (QR, LAPACK DBDSQR / DBDSDC, shifted sweeps, bound refinement)
! stay on the host in double precision.
!
! Entry points:
!   pk_lansvd_run / pk_lansvd_irl_run  (bind(C)) -- device-resident drivers on
!       a caller-created engine: the start vector is the engine's U(:,1), the
!       singular vectors are left in the engine's bases.
!   dlansvd, dlansvd_irl, slansvd, slansvd_irl -- the reference's Fortran
!       calling sequences (see module pk_dropin below).
module pk_drivers
  use iso_c_binding
  use pk_api
  implicit none
  private
  public :: lansvd_core, lansvd_irl_core, add_timing, irl_begin, irl_after, irl_end, &
            irl_pre_svd, irl_host_svd, irl_post_svd

  ! External LAPACK (host): bidiagonal SVD.
  interface
    subroutine dbdsqr(uplo, n, ncvt, nru, ncc, d, e, vt, ldvt, u, ldu, c, ldc, work, info)
      character(len=1) :: uplo
      integer :: n, ncvt, nru, ncc, ldvt, ldu, ldc, info
      double precision :: d(*), e(*), vt(ldvt, *), u(ldu, *), c(ldc, *), work(*)
    end subroutine
    subroutine dbdsdc(uplo, compq, n, d, e, u, ldu, vt, ldvt, q, iq, work, iwork, info)
      character(len=1) :: uplo, compq
      integer :: n, ldu, ldvt, iq(*), iwork(*), info
      double precision :: d(*), e(*), u(ldu, *), vt(ldvt, *), q(*), work(*)
    end subroutine
  end interface

  ! Per-problem state of the restarted driver between DLANBPRO calls.
  type, public :: irl_state
    type(c_ptr) :: eng = c_null_ptr
    character(len=1) :: which = 'L', jobu = 'n', jobv = 'n'
    integer :: m = 0, n = 0, dim = 0, neig = 0, maxiter = 0, k = 0
    integer :: iter = 0, nconv = 0, nmax = 0, kold = 0, ierr = 0, nbsvd = 0, nrestart = 0
    integer :: info = 0
    double precision :: eps = 0d0, eps34 = 0d0, epsn = 0d0, tol = 0d0, anorm = 0d0
    double precision :: rnorm = 0d0, dopt(4) = 0d0
    logical :: small = .false., done = .true., failed = .false.
    ! defer: the restart GEMMs are left to the caller (the batched driver
    ! runs every problem's in one launch); pending: this pass left them
    logical :: defer = .false., pending = .false.
    integer(c_long_long) :: st0(8) = 0
    double precision, allocatable :: alpha(:), beta(:), a1(:), b1(:), wb(:), shift(:), &
                                     pm(:, :), qm(:, :), wrk(:), sig(:), bd(:)
  end type

contains

  ! Effective vector length of the engine's rounding model: max(m, n) as in
  ! the reference (epsn = max(m,n)*eps/2, double/dlansvd.F:138), or the
  ! storage-rounding length of the FP32-basis form.
  double precision function eps_leff(eng, m, n)
    type(c_ptr), intent(in) :: eng
    integer, intent(in) :: m, n
    interface
      integer(c_int) function pk_lbd_eps(e, eps, leff) bind(C, name='pk_lbd_eps')
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: e
        real(c_double) :: eps, leff
      end function
    end interface
    real(c_double) :: ee, ll
    if (pk_lbd_eps(eng, ee, ll) == PK_OK) then
      eps_leff = ll
    else
      eps_leff = dble(max(m, n))
    end if
  end function

  ! Accumulate the engine's counters since st0 (taken at driver entry) into
  ! COMMON /timing/ (the reference's double/stat.h:12-15 layout) so that
  ! PRINTSTAT-style callers see them.
  subroutine add_timing(eng, st0, nbsvd_add, nrestart_add, nlandim_set, nsing_set)
    type(c_ptr), intent(in) :: eng
    integer(c_long_long), intent(in) :: st0(8)
    integer, intent(in) :: nbsvd_add, nrestart_add, nlandim_set, nsing_set
    integer :: nopx, nreorth, ndot, nreorthu, nreorthv, nitref, nrestart, nbsvd, nlandim, nsing
    real :: tmvopx, tgetu0, tupdmu, tupdnu, tintv, tlanbpro, treorth, treorthu, treorthv, &
            telru, telrv, tbsvd, tnorm2, tlansvd, tritzvec, trestart, tdot
    common /timing/ nopx, nreorth, ndot, nreorthu, nreorthv, nitref, nrestart, nbsvd, tmvopx, &
        tgetu0, tupdmu, tupdnu, tintv, tlanbpro, treorth, treorthu, treorthv, telru, telrv, &
        tbsvd, tnorm2, tlansvd, nlandim, tritzvec, trestart, tdot, nsing
    integer(c_long_long) :: st(8)
    integer(c_int) :: rc
    rc = pk_lbd_stats(eng, st)
    if (rc == PK_OK) then
      nopx = nopx + int(st(1) - st0(1))
      nreorth = nreorth + int(st(2) - st0(2))
      ndot = ndot + int(st(3) - st0(3))
    end if
    nbsvd = nbsvd + nbsvd_add
    nrestart = nrestart + nrestart_add
    if (nlandim_set >= 0) nlandim = nlandim_set
    if (nsing_set >= 0) nsing = nsing_set
  end subroutine

  ! Ritz vectors from A V_dim = U_{dim+1} B_dim (DRITZVEC): B = M [R; 0]
  ! by Givens QR, R = P S Q^T by divide and conquer (DBDSDC), then
  ! U <- U (P^T M^T)^T rows, V <- V Q on the device.
  subroutine ritzvec(which, jobu, jobv, m, n, k, dim, d, e, eng, info)
    character(len=1), intent(in) :: which, jobu, jobv
    integer, intent(in) :: m, n, k, dim
    double precision, intent(inout) :: d(*), e(*)
    type(c_ptr), intent(in) :: eng
    integer, intent(out) :: info
    double precision, allocatable :: mt(:, :), qt(:, :), p(:, :), x(:, :), wrk(:)
    integer, allocatable :: iwk(:)
    double precision :: c1, c2, dd(1)
    integer :: id(1), mstart, lapinfo
    integer(c_int) :: rc
    info = 0
    if (k < 1) return
    allocate(mt(dim + 1, dim + 1), qt(dim, dim), p(dim, dim), x(dim + 1, dim + 1))
    allocate(wrk(3 * dim * dim + 4 * dim + 8), iwk(8 * dim + 8))
    mt = 0d0; qt = 0d0; p = 0d0; x = 0d0
    c1 = 0d0; c2 = 0d0
    call pk_bdqr(dim == min(m, n), jobu, dim, d, e, c1, c2, mt, dim + 1)
    call dbdsdc('U', 'I', dim, d, e, p, dim, qt, dim, dd, id, wrk, iwk, lapinfo)
    ! X(1:dim, 1:dim+1) = P^T * M^T(1:dim, 1:dim+1)
    call dgemm('T', 'N', dim, dim + 1, dim, 1d0, p, dim, mt, dim + 1, 0d0, x, dim + 1)
    if (lower1(which) == 's') then
      mstart = dim - k + 1
    else
      mstart = 1
    end if
    if (lower1(jobu) == 'y') then
      ! U(:,1:k) <- U(:,1:dim+1) * X(mstart:mstart+k-1, :)^T
      rc = pk_lbd_update(eng, 'u', 't', dim + 1, k, x(mstart, 1), dim + 1)
      if (rc /= PK_OK) info = rc
    end if
    if (lower1(jobv) == 'y' .and. info == 0) then
      ! V(:,1:k) <- V(:,1:dim) * Q^T(mstart:mstart+k-1, :)^T
      rc = pk_lbd_update(eng, 'v', 't', dim, k, qt(mstart, 1), dim)
      if (rc /= PK_OK) info = rc
    end if
    ! the basis GEMMs return in flight: the Ritz vectors are complete on return
    if (info == 0) then
      rc = pk_lbd_sync(eng)
      if (rc /= PK_OK) info = rc
    end if
  end subroutine

  ! Non-restarted driver (DLANSVD control flow, double/dlansvd.F:130-291)
  ! on an engine whose U(:,1) holds the start vector (zero => random).
  ! eps is the unit roundoff of the engine's storage precision.
  subroutine lansvd_core(jobu, jobv, m, n, k, kmax, eng, eps, sigma, bnd, tolin, doption, &
                         ioption, info)
    character(len=1), intent(in) :: jobu, jobv
    integer, intent(in) :: m, n, kmax
    integer, intent(inout) :: k
    type(c_ptr), intent(in) :: eng
    double precision, intent(in) :: eps, tolin
    double precision, intent(inout) :: sigma(*), bnd(*), doption(*)
    integer, intent(in) :: ioption(*)
    integer, intent(out) :: info
    double precision, allocatable :: alpha(:), beta(:), a1(:), b1(:), wb(:), wrk(:)
    double precision :: eps34, epsn, tol, anorm, rnorm, dummy(1, 1)
    integer :: lanmax, neig, jold, j, jj, dj, i, ierr, lapinfo, nbsvd
    integer(c_int) :: rc, ierr_c, jj_c
    integer(c_long_long) :: st0(8)
    logical :: vectors

    rc = pk_lbd_stats(eng, st0)
    eps34 = eps**(3d0 / 4d0)
    epsn = eps_leff(eng, m, n) * eps / 2d0
    lanmax = min(n + 1, m + 1, kmax)
    tol = min(1d0, max(16d0 * eps, tolin))
    anorm = 0d0
    info = 0
    nbsvd = 0
    allocate(alpha(lanmax), beta(lanmax), a1(lanmax), b1(lanmax), wb(lanmax + 1), &
             wrk(4 * lanmax + 8))
    alpha = 0d0; beta = 0d0

    rc = pk_lbd_nrm2(eng, 'u', 1, rnorm)
    if (rc /= PK_OK) then
      info = rc
      return
    end if
    if (rnorm == 0d0) then
      rc = pk_lbd_getu0(eng, 'n', 0, 1, rnorm, ioption(1), anorm, ierr_c)
      if (rc /= PK_OK) then
        info = rc
        return
      end if
    end if

    neig = 0
    jold = 0
    j = min(k + max(8, k) + 1, lanmax)
    ierr = 0
    do while (neig < k)
      jj_c = j
      rc = pk_lbd_lanbpro(eng, jold, jj_c, alpha, beta, rnorm, doption, ioption, ierr_c)
      if (rc /= PK_OK) then
        info = rc
        return
      end if
      j = jj_c
      ierr = ierr_c
      jold = j
      ! SVD of B_j and error bounds (double/dlansvd.F:193-236)
      a1 = alpha
      b1 = beta
      wb = 0d0
      call pk_bdqr(j == min(m, n), 'N', j, a1, b1, wb(j), wb(j + 1), dummy, 1)
      call dbdsqr('U', j, 0, 1, 0, a1, b1, dummy, 1, wb, 1, dummy, 1, wrk, lapinfo)
      nbsvd = nbsvd + 1
      if (j > 5) then
        anorm = a1(1)
      else
        anorm = max(anorm, a1(1))
      end if
      wb(1:j) = abs(rnorm * wb(1:j))
      call pk_refinebounds(min(m, n), j, a1, wb, epsn * anorm, eps34)
      do i = 1, min(j, k)
        bnd(i) = wb(i)
      end do
      neig = 0
      i = 0
      do while (i < min(j, k))
        if (wb(i + 1) <= tol * a1(i + 1)) then
          neig = neig + 1
          sigma(neig) = a1(i + 1)
          i = i + 1
        else
          exit
        end if
      end do
      call pk_f_progress(j, neig, k, a1(1), rnorm)
      if (ierr < 0) then
        if (j < k) info = j
        exit
      end if
      if (j >= lanmax) then
        if (neig < k) info = -1
        exit
      end if
      ! grow the Krylov dimension (double/dlansvd.F:266-275)
      if (neig > 1) then
        dj = min(j / 2, ((k - neig) * (j - 6)) / (2 * neig + 1))
        dj = min(100, max(2, dj))
      else
        dj = j / 2
        dj = min(100, max(10, dj))
      end if
      j = min(j + dj, lanmax)
    end do

    vectors = lower1(jobu) == 'y' .or. lower1(jobv) == 'y'
    if ((neig >= k .or. info > 0) .and. vectors) then
      call ritzvec('L', jobu, jobv, m, n, neig, jold, alpha, beta, eng, jj)
      if (jj /= 0) info = jj
    end if
    call add_timing(eng, st0, nbsvd, 0, j, k)
    k = neig
  end subroutine

  ! Implicitly restarted driver (DLANSVD_IRL control flow,
  ! double/dlansvd_irl.F:164-419).  dim and neig are in/out as in the
  ! reference.  The engine must have room for dim V columns.  The body is
  ! split into irl_begin / irl_after / irl_end around the DLANBPRO call so
  ! that the batched driver runs the same code per problem.
  subroutine lansvd_irl_core(which, jobu, jobv, m, n, dim, p, neig, maxiter, eng, eps, &
                             sigma, bnd, tolin, doption, ioption, info)
    character(len=1), intent(in) :: which, jobu, jobv
    integer, intent(in) :: m, n, p, maxiter
    integer, intent(inout) :: dim, neig
    type(c_ptr), intent(in) :: eng
    double precision, intent(in) :: eps, tolin
    double precision, intent(inout) :: sigma(*), bnd(*), doption(*)
    integer, intent(in) :: ioption(*)
    integer, intent(out) :: info
    type(irl_state), target :: st
    integer(c_int) :: rc, ierr_c, kk_c

    call irl_begin(st, which, jobu, jobv, m, n, dim, p, neig, maxiter, eng, eps, tolin, &
                   doption, ioption)
    do while (.not. st%done)
      kk_c = st%dim
      rc = pk_lbd_lanbpro(eng, st%kold, kk_c, st%alpha, st%beta, st%rnorm, st%dopt, ioption, &
                          ierr_c)
      if (rc /= PK_OK) then
        st%info = rc
        st%done = .true.
        st%failed = .true.
        exit
      end if
      call irl_after(st, kk_c, ierr_c)
    end do
    call irl_end(st, sigma, bnd, dim, neig, info, doption)
  end subroutine

  ! Driver setup up to the first DLANBPRO call (double/dlansvd_irl.F:160-206).
  subroutine irl_begin(st, which, jobu, jobv, m, n, dim, p, neig, maxiter, eng, eps, tolin, &
                       doption, ioption)
    type(irl_state), intent(inout) :: st
    character(len=1), intent(in) :: which, jobu, jobv
    integer, intent(in) :: m, n, dim, p, neig, maxiter
    type(c_ptr), intent(in) :: eng
    double precision, intent(in) :: eps, tolin
    double precision, intent(in) :: doption(*)
    integer, intent(in) :: ioption(*)
    integer(c_int) :: rc, ierr_c

    st%eng = eng
    st%which = which
    st%jobu = jobu
    st%jobv = jobv
    st%m = m
    st%n = n
    st%neig = neig
    st%maxiter = maxiter
    st%eps = eps
    rc = pk_lbd_stats(eng, st%st0)
    st%eps34 = eps**(3d0 / 4d0)
    st%epsn = eps_leff(eng, m, n) * eps / 2d0
    st%dim = min(dim, n + 1, m + 1)
    st%k = st%dim - p
    st%tol = min(1d0, max(16d0 * eps, tolin))
    st%anorm = 0d0
    st%info = 0
    st%nbsvd = 0
    st%nrestart = 0
    st%small = lower1(which) == 's'
    st%dopt = doption(1:4)
    allocate(st%alpha(st%dim), st%beta(st%dim), st%a1(st%dim), st%b1(st%dim), &
             st%wb(st%dim + 1), st%shift(st%dim), st%pm(st%dim + 1, st%dim + 1), &
             st%qm(st%dim, st%dim), st%wrk(4 * st%dim + 8), st%sig(max(neig, 1)), &
             st%bd(max(neig, 1)))
    st%alpha = 0d0
    st%beta = 0d0
    st%sig = 0d0
    st%bd = 0d0
    st%iter = 0
    st%nconv = 0
    st%nmax = 0
    st%kold = 0
    st%ierr = 0
    st%done = .false.
    st%failed = .false.

    rc = pk_lbd_nrm2(eng, 'u', 1, st%rnorm)
    if (rc == PK_OK .and. st%rnorm == 0d0) &
      rc = pk_lbd_getu0(eng, 'n', 0, 1, st%rnorm, ioption(1), st%anorm, ierr_c)
    if (rc /= PK_OK) then
      st%info = rc
      st%done = .true.
      st%failed = .true.
      return
    end if
    st%done = .not. (st%nconv < st%neig .and. st%iter < st%maxiter)
  end subroutine

  ! One pass of the restart loop after its DLANBPRO call returned dim (kk)
  ! and ierr (double/dlansvd_irl.F:213-414): irl_pre_svd, the bidiagonal
  ! SVD with one row of left vectors (host DBDSQR here; the batched driver
  ! may run them on the device for every problem at once), irl_post_svd.
  subroutine irl_after(st, kk, ierr_c)
    type(irl_state), intent(inout) :: st
    integer(c_int), intent(in) :: kk, ierr_c
    call irl_pre_svd(st, kk, ierr_c)
    call irl_host_svd(st)
    call irl_post_svd(st)
  end subroutine

  ! B_dim to upper bidiagonal form in a1/b1 with the last row of the
  ! rotations in wb(dim:dim+1) (double/dlansvd_irl.F:220-226).
  subroutine irl_pre_svd(st, kk, ierr_c)
    type(irl_state), intent(inout) :: st
    integer(c_int), intent(in) :: kk, ierr_c
    double precision :: dummy(1, 1)
    integer :: dim

    st%ierr = ierr_c
    ! DLANBPRO lowers its K argument (DIM here) when it stops at an
    ! invariant subspace (double/dlanbpro.F:389,519; double/dlansvd_irl.F:213)
    st%dim = kk
    st%kold = st%k
    dim = st%dim
    st%a1 = st%alpha
    st%b1 = st%beta
    st%wb = 0d0
    call pk_bdqr(dim == min(st%m, st%n), 'N', dim, st%a1, st%b1, st%wb(dim), st%wb(dim + 1), &
                 dummy, 1)
  end subroutine

  ! Singular values of the upper bidiagonal (a1, b1) into a1 and the row
  ! wb(1:dim) times the left singular vectors into wb (double/dlansvd_irl.F:
  ! 227-229).
  subroutine irl_host_svd(st)
    type(irl_state), intent(inout) :: st
    double precision :: dummy(1, 1)
    integer :: lapinfo
    call dbdsqr('U', st%dim, 0, 1, 0, st%a1, st%b1, dummy, 1, st%wb, 1, dummy, 1, st%wrk, &
                lapinfo)
  end subroutine

  ! Error bounds, convergence count, shifts and the implicit restart
  ! (double/dlansvd_irl.F:230-394).
  subroutine irl_post_svd(st)
    type(irl_state), intent(inout) :: st
    double precision :: relgap
    integer :: i, nshft, dim, k, neig, m, n
    integer(c_int) :: rc

    dim = st%dim
    k = st%k
    neig = st%neig
    m = st%m
    n = st%n
    associate (a1 => st%a1, wb => st%wb, alpha => st%alpha, beta => st%beta, &
               shift => st%shift, pm => st%pm, qm => st%qm)
      st%nbsvd = st%nbsvd + 1
      if (dim > 5) then
        st%anorm = a1(1)
      else
        st%anorm = max(st%anorm, a1(1))
      end if
      wb(1:dim) = abs(st%rnorm * wb(1:dim))
      if (st%small) then
        call pk_refinebounds(min(m, n), dim, a1, wb, st%epsn * st%anorm, st%eps34)
      else
        call pk_refinebounds(min(m, n), min(dim, neig), a1, wb, st%epsn * st%anorm, st%eps34)
      end if
      ! converged count (double/dlansvd_irl.F:262-290)
      st%nconv = 0
      if (st%small) then
        do i = dim - neig + 1, dim
          if (wb(i) <= st%tol * a1(1)) then
            st%nconv = st%nconv + 1
            st%sig(st%nconv) = a1(i)
            st%bd(st%nconv) = wb(i)
          end if
        end do
      else
        i = 1
        do while (i <= min(dim, neig))
          if (wb(i) <= st%tol * a1(i)) then
            st%nconv = st%nconv + 1
            st%sig(st%nconv) = a1(i)
            st%bd(st%nconv) = wb(i)
            i = i + 1
          else if (i <= k) then
            ! the reference continues the count after the kept Ritz values
            ! (i = k+1, double/dlansvd_irl.F:285-286); it only matters when
            ! neig > dim - p
            i = k + 1
          else
            ! an unconverged value past k: the reference sets i = k+1 again
            ! and never leaves the loop; the count stops here instead
            exit
          end if
        end do
      end if
      st%nmax = max(st%nmax, st%nconv)
      call pk_f_progress(st%iter, st%nconv, neig, a1(1), st%rnorm)
      if (st%ierr < 0) then
        if (dim < k) st%info = dim
        st%done = .true.
        return
      end if
      if (st%nconv < neig) then
        ! exact shifts filtered by the relative gap (double/dlansvd_irl.F:318-344)
        shift = 0d0
        nshft = 0
        if (st%small) then
          do i = 1, k
            relgap = a1(i) - wb(i) - a1(dim - neig)
            nshft = nshft + 1
            if (relgap > st%dopt(4) * a1(dim - neig)) then
              shift(nshft) = a1(i)
            else
              shift(nshft) = a1(1)
            end if
          end do
        else
          do i = dim, k + 1, -1
            relgap = a1(k) - (a1(i) + wb(i))
            nshft = nshft + 1
            if (relgap > st%dopt(4) * a1(k)) then
              shift(nshft) = a1(i)
            else
              shift(nshft) = 0d0
            end if
          end do
        end if
        ! B+ = P B Q^T by dim-k shifted sweeps (double/dlansvd_irl.F:350-363)
        pm = 0d0
        qm = 0d0
        do i = 1, dim + 1
          pm(i, i) = 1d0
        end do
        do i = 1, dim
          qm(i, i) = 1d0
        end do
        do i = dim, k + 1, -1
          call pk_bsvdstep('y', 'y', dim + 1, dim, i, shift(dim - i + 1), alpha, beta, pm, &
                           size(pm, 1), qm, size(qm, 1))
        end do
        ! U(:,1:k+1) <- U(:,1:dim+1) P(:,1:k+1); V(:,1:k) <- V(:,1:dim) Q(:,1:k)
        if (st%defer) then
          st%pending = .true.
        else
          rc = pk_lbd_update(st%eng, 'u', 'n', dim + 1, k + 1, pm, size(pm, 1))
          if (rc == PK_OK) rc = pk_lbd_update(st%eng, 'v', 'n', dim, k, qm, size(qm, 1))
          if (rc /= PK_OK) then
            st%info = rc
            st%done = .true.
            st%failed = .true.
            return
          end if
        end if
        st%rnorm = beta(k)
        st%nrestart = st%nrestart + 1
      end if
    end associate
    st%iter = st%iter + 1
    st%done = .not. (st%nconv < st%neig .and. st%iter < st%maxiter)
  end subroutine

  ! Ritz vectors and outputs (double/dlansvd_irl.F:396-418).
  subroutine irl_end(st, sigma, bnd, dim, neig, info, doption)
    type(irl_state), intent(inout) :: st
    double precision, intent(inout) :: sigma(*), bnd(*), doption(*)
    integer, intent(inout) :: dim, neig
    integer, intent(out) :: info
    integer :: jj
    logical :: vectors

    doption(1:4) = st%dopt
    info = st%info
    if (st%failed) return  ! an engine error: outputs as they were
    vectors = lower1(st%jobu) == 'y' .or. lower1(st%jobv) == 'y'
    if ((st%nconv >= st%neig .or. st%info > 0) .and. vectors) then
      st%a1 = st%alpha
      st%b1 = st%beta
      call ritzvec(st%which, st%jobu, st%jobv, st%m, st%n, st%nconv, st%dim, st%a1, st%b1, &
                   st%eng, jj)
      if (jj /= 0) st%info = jj
    end if
    call add_timing(st%eng, st%st0, st%nbsvd, st%nrestart, st%dim, -1)
    if (st%nmax > 0) then
      sigma(1:st%nmax) = st%sig(1:st%nmax)
      bnd(1:st%nmax) = st%bd(1:st%nmax)
    end if
    dim = st%dim
    neig = st%nconv
    info = st%info
  end subroutine

end module pk_drivers

! Per-problem restart work of the batched driver on the library's host
! threads: each index touches only its own problem's state and engine.
module pk_batch_par
  use iso_c_binding
  use pk_drivers
  implicit none
  private
  public :: after_ctx, after_one, pre_one, post_one, pkx_parallel_for

  type :: after_ctx
    type(irl_state), pointer :: st(:) => null()
    integer(c_int), pointer :: active(:) => null(), kk(:) => null(), ierr(:) => null()
    integer(c_int), pointer :: dev(:) => null()  ! 1: the bidiagonal SVD ran on the device
    real(c_double), pointer :: rn(:) => null(), dop(:, :) => null()
  end type

  interface
    subroutine pkx_parallel_for(n, fn, ctx) bind(C, name='pkx_parallel_for')
      import :: c_int, c_funptr, c_ptr
      integer(c_int), value :: n
      type(c_funptr), value :: fn
      type(c_ptr), value :: ctx
    end subroutine
  end interface

contains

  subroutine after_one(ctx, i0) bind(C)
    type(c_ptr), value :: ctx
    integer(c_int), value :: i0
    type(after_ctx), pointer :: c
    integer :: i
    call c_f_pointer(ctx, c)
    i = i0 + 1
    if (c%active(i) == 0) return
    c%st(i)%rnorm = c%rn(i)
    c%st(i)%dopt = c%dop(:, i)
    call irl_after(c%st(i), c%kk(i), c%ierr(i))
  end subroutine

  ! The same in two halves around the device bidiagonal SVDs.
  subroutine pre_one(ctx, i0) bind(C)
    type(c_ptr), value :: ctx
    integer(c_int), value :: i0
    type(after_ctx), pointer :: c
    integer :: i
    call c_f_pointer(ctx, c)
    i = i0 + 1
    if (c%active(i) == 0) return
    c%st(i)%rnorm = c%rn(i)
    c%st(i)%dopt = c%dop(:, i)
    call irl_pre_svd(c%st(i), c%kk(i), c%ierr(i))
  end subroutine

  subroutine post_one(ctx, i0) bind(C)
    type(c_ptr), value :: ctx
    integer(c_int), value :: i0
    type(after_ctx), pointer :: c
    integer :: i
    call c_f_pointer(ctx, c)
    i = i0 + 1
    if (c%active(i) == 0) return
    if (c%dev(i) == 0) call irl_host_svd(c%st(i))
    call irl_post_svd(c%st(i))
  end subroutine

end module pk_batch_par

! ---------------------------------------------------------------------------
! Device-resident entry points (no reference counterpart): run the drivers on
! an engine created by the caller (pk_lbd_create with a device operator or a
! registered callback).  Singular values and bounds return to host arrays;
! the singular vectors stay in the engine's bases (pk_lbd_basis / pk_lbd_get).
! Option characters are passed by value; returns INFO.
integer(c_int) function pk_lansvd_run(eng, jobu, jobv, k, kmax, sigma, bnd, tolin, doption, &
                                      ioption) bind(C, name='pk_lansvd_run')
  use iso_c_binding
  use pk_api
  use pk_drivers
  implicit none
  type(c_ptr), value :: eng
  character(kind=c_char), value :: jobu, jobv
  integer(c_int) :: k
  integer(c_int), value :: kmax
  real(c_double) :: sigma(*), bnd(*), doption(*)
  real(c_double), value :: tolin
  integer(c_int) :: ioption(*)
  integer :: m, n, info, kk
  double precision :: eps
  call engine_shape(eng, m, n, eps, info)
  if (info /= 0) then
    pk_lansvd_run = info
    return
  end if
  kk = k
  call lansvd_core(jobu, jobv, m, n, kk, kmax, eng, eps, sigma, bnd, tolin, doption, ioption, &
                   info)
  k = kk
  pk_lansvd_run = info
end function

integer(c_int) function pk_lansvd_irl_run(eng, which, jobu, jobv, dim, p, neig, maxiter, sigma, &
                                          bnd, tolin, doption, ioption) &
    bind(C, name='pk_lansvd_irl_run')
  use iso_c_binding
  use pk_api
  use pk_drivers
  implicit none
  type(c_ptr), value :: eng
  character(kind=c_char), value :: which, jobu, jobv
  integer(c_int) :: dim, neig
  integer(c_int), value :: p, maxiter
  real(c_double) :: sigma(*), bnd(*), doption(*)
  real(c_double), value :: tolin
  integer(c_int) :: ioption(*)
  integer :: m, n, info, dd, ne
  double precision :: eps
  call engine_shape(eng, m, n, eps, info)
  if (info /= 0) then
    pk_lansvd_irl_run = info
    return
  end if
  dd = dim
  ne = neig
  call lansvd_irl_core(which, jobu, jobv, m, n, dd, p, ne, maxiter, eng, eps, sigma, bnd, &
                       tolin, doption, ioption, info)
  dim = dd
  neig = ne
  pk_lansvd_irl_run = info
end function

! Shape and storage unit roundoff of an engine.
! Batched restarted driver (no reference counterpart; replaces a loop of
! DLANSVD_IRL calls over independent problems, double/dlansvd_irl.F:1-3):
! nb engines of one batch (pk_batch_create) run DLANSVD_IRL's control flow
! each, their Lanczos steps advancing together through pk_batch_lanbpro.
! sigma, bnd: ldsig x nb; per problem nconv(i) and info(i).  doption(1:4)
! and ioption as for one problem (each problem gets its own doption copy).
integer(c_int) function pk_lansvd_irl_batch(nb, engs, batch, which, jobu, jobv, dim, p, neig, &
                                            maxiter, tolin, sigma, bnd, ldsig, nconv, info, &
                                            doption, ioption) &
    bind(C, name='pk_lansvd_irl_batch')
  use iso_c_binding
  use pk_api
  use pk_drivers
  use pk_batch_par
  implicit none
  integer(c_int), value :: nb, dim, p, neig, maxiter, ldsig
  type(c_ptr) :: engs(nb)
  type(c_ptr), value :: batch
  character(kind=c_char), value :: which, jobu, jobv
  real(c_double), value :: tolin
  real(c_double) :: sigma(ldsig, nb), bnd(ldsig, nb), doption(4)
  integer(c_int) :: nconv(nb), info(nb), ioption(*)
  interface
    integer(c_int) function pk_batch_lanbpro(b, active, k0, k, alpha, beta, rnorm, dopt, &
                                             iopt, ierr) bind(C, name='pk_batch_lanbpro')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b
      integer(c_int) :: active(*), k0(*), k(*), iopt(*), ierr(*)
      type(c_ptr) :: alpha(*), beta(*)
      real(c_double) :: rnorm(*), dopt(4, *)
    end function
    integer(c_int) function pk_batch_get_bsvd(b, on) bind(C, name='pk_batch_get_bsvd')
      import :: c_int, c_ptr
      type(c_ptr), value :: b
      integer(c_int) :: on
    end function
    integer(c_int) function pk_batch_bsvd(b, nd, nn, d, e, ld, sig, bot) &
        bind(C, name='pk_batch_bsvd')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: b
      integer(c_int), value :: nd, ld
      integer(c_int) :: nn(*)
      real(c_double) :: d(*), e(*), sig(*), bot(*)
    end function
    integer(c_int) function pk_batch_update(b, nsel, sel, which, transb, kdim, ncols, bm, ldb) &
        bind(C, name='pk_batch_update')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: b
      integer(c_int), value :: nsel
      character(kind=c_char), value :: which, transb
      integer(c_int) :: sel(*), kdim(*), ncols(*), ldb(*)
      type(c_ptr) :: bm(*)
    end function
  end interface
  integer, parameter :: bsvd_max = 128  ! largest order of the device bidiagonal SVD
  type(irl_state), allocatable, target :: st(:)
  integer(c_int), allocatable, target :: active(:), k0(:), kk(:), ierr(:), dev(:)
  integer(c_int), allocatable :: bn(:), idx(:), usel(:), ukd(:), unc(:), uld(:), vkd(:), &
                                 vnc(:), vld(:)
  type(c_ptr), allocatable :: ap(:), bp(:), upm(:), vqm(:)
  real(c_double), allocatable, target :: rn(:), dop(:, :)
  real(c_double), allocatable :: bd(:, :), be(:, :), bs(:, :), bb(:, :)
  type(after_ctx), target :: ctx
  integer :: i, j, m, n, rc, dd, ne, nd, ldb
  integer(c_int) :: usedev
  double precision :: eps, dummy(4), c1

  pk_lansvd_irl_batch = PK_OK
  allocate(st(nb), active(nb), k0(nb), kk(nb), ierr(nb), ap(nb), bp(nb), rn(nb), dop(4, nb), &
           dev(nb), bn(nb), idx(nb), usel(nb), ukd(nb), unc(nb), uld(nb), vkd(nb), vnc(nb), &
           vld(nb), upm(nb), vqm(nb))
  usedev = 0
  rc = pk_batch_get_bsvd(batch, usedev)
  if (rc /= PK_OK) then
    pk_lansvd_irl_batch = rc
    return
  end if
  ldb = max(dim, 1)
  if (usedev /= 0) allocate(bd(ldb, nb), be(ldb, nb), bs(ldb, nb), bb(ldb, nb))
  do i = 1, nb
    call engine_shape(engs(i), m, n, eps, rc)
    if (rc /= 0) then
      pk_lansvd_irl_batch = rc
      return
    end if
    call irl_begin(st(i), which, jobu, jobv, m, n, dim, p, neig, maxiter, engs(i), eps, tolin, &
                   doption, ioption)
    st(i)%defer = .true.
  end do
  do while (any(.not. st(:)%done))
    do i = 1, nb
      active(i) = merge(1, 0, .not. st(i)%done)
      k0(i) = st(i)%kold
      kk(i) = st(i)%dim
      ap(i) = c_loc(st(i)%alpha)
      bp(i) = c_loc(st(i)%beta)
      rn(i) = st(i)%rnorm
      dop(:, i) = st(i)%dopt
      ierr(i) = 0
    end do
    rc = pk_batch_lanbpro(batch, active, k0, kk, ap, bp, rn, dop, ioption, ierr)
    if (rc /= PK_OK) then
      pk_lansvd_irl_batch = rc
      do i = 1, nb
        if (active(i) /= 0) then
          st(i)%info = rc
          st(i)%failed = .true.
          st(i)%done = .true.
        end if
      end do
      exit
    end if
    ctx%st => st
    ctx%active => active
    ctx%kk => kk
    ctx%ierr => ierr
    ctx%dev => dev
    ctx%rn => rn
    ctx%dop => dop
    if (usedev == 0) then
      call pkx_parallel_for(nb, c_funloc(after_one), c_loc(ctx))
      call restart_gemms()
      cycle
    end if
    ! bidiagonal SVDs of every problem on the device, between the two
    ! halves of each problem's restart work
    call pkx_parallel_for(nb, c_funloc(pre_one), c_loc(ctx))
    nd = 0
    do i = 1, nb
      dev(i) = 0
      dd = st(i)%dim
      if (active(i) /= 0 .and. dd >= 1 .and. dd <= min(bsvd_max, ldb)) then
        nd = nd + 1
        idx(nd) = i
        dev(i) = 1
        bn(nd) = dd
        bd(1:dd, nd) = st(i)%a1(1:dd)
        be(:, nd) = 0d0
        if (dd > 1) be(1:dd - 1, nd) = st(i)%b1(1:dd - 1)
      end if
    end do
    rc = pk_batch_bsvd(batch, nd, bn, bd, be, ldb, bs, bb)
    if (rc /= PK_OK) then
      pk_lansvd_irl_batch = rc
      do i = 1, nb
        if (active(i) /= 0) then
          st(i)%info = rc
          st(i)%failed = .true.
          st(i)%done = .true.
        end if
      end do
      exit
    end if
    do j = 1, nd
      i = idx(j)
      dd = st(i)%dim
      c1 = st(i)%wb(dd)
      st(i)%a1(1:dd) = bs(1:dd, j)
      st(i)%wb(1:dd) = c1 * bb(1:dd, j)
    end do
    call pkx_parallel_for(nb, c_funloc(post_one), c_loc(ctx))
    call restart_gemms()
  end do
  do i = 1, nb
    dd = dim
    ne = neig
    dummy = doption
    call irl_end(st(i), sigma(:, i), bnd(:, i), dd, ne, info(i), dummy)
    nconv(i) = ne
  end do
contains
  ! The restart GEMMs this pass left (double/dlansvd_irl.F:387-395) for every
  ! problem at once: U(:,1:k+1) <- U(:,1:dim+1) P and V(:,1:k) <- V(:,1:dim) Q,
  ! one grouped launch each.
  subroutine restart_gemms()
    integer :: nu, q
    integer(c_int) :: rcu
    nu = 0
    do q = 1, nb
      if (st(q)%pending) then
        nu = nu + 1
        usel(nu) = q - 1
        ukd(nu) = st(q)%dim + 1
        unc(nu) = st(q)%k + 1
        uld(nu) = size(st(q)%pm, 1)
        upm(nu) = c_loc(st(q)%pm)
        vkd(nu) = st(q)%dim
        vnc(nu) = st(q)%k
        vld(nu) = size(st(q)%qm, 1)
        vqm(nu) = c_loc(st(q)%qm)
        st(q)%pending = .false.
      end if
    end do
    if (nu == 0) return
    rcu = pk_batch_update(batch, nu, usel, 'u', 'n', ukd, unc, upm, uld)
    if (rcu == PK_OK) rcu = pk_batch_update(batch, nu, usel, 'v', 'n', vkd, vnc, vqm, vld)
    if (rcu /= PK_OK) then
      pk_lansvd_irl_batch = rcu
      do q = 1, nu
        st(usel(q) + 1)%info = rcu
        st(usel(q) + 1)%failed = .true.
        st(usel(q) + 1)%done = .true.
      end do
    end if
  end subroutine
end function

subroutine engine_shape(eng, m, n, eps, info)
  use iso_c_binding
  implicit none
  type(c_ptr), intent(in) :: eng
  integer, intent(out) :: m, n, info
  double precision, intent(out) :: eps
  interface
    integer(c_int) function pk_lbd_shape(e, m, n, kmax, prec) bind(C, name='pk_lbd_shape')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int) :: m, n, kmax, prec
    end function
    integer(c_int) function pk_lbd_eps(e, eps, leff) bind(C, name='pk_lbd_eps')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e
      real(c_double) :: eps, leff
    end function
  end interface
  integer(c_int) :: mm, nn, kk, pp
  real(c_double) :: ee, ll
  info = pk_lbd_shape(eng, mm, nn, kk, pp)
  m = mm
  n = nn
  if (info == 0) info = pk_lbd_eps(eng, ee, ll)
  eps = ee
end subroutine

! Non-restarted FP32-basis form with FP64 polish (no reference counterpart):
! pk_lansvd_irl_mixed's scheme around the DLANSVD driver.  eng32 (FP32
! values and bases, FP64 reductions and scalars, storage-rounding
! thresholds) runs DLANSVD to max(tolin, 16*eps32); when tolin is reachable
! there, its triplets' FP64 residuals are formed on the device with eng64's
! operator and, if every one meets tolin*sigma(1), they are promoted to
! eng64 and no polish runs.  Otherwise the sum of the FP32 left Ritz vectors
! seeds an FP64 DLANSVD on eng64.  sigma32/bnd32 return the FP32 phase's
! values, info32(1:2) its INFO and converged count.
integer(c_int) function pk_lansvd_mixed(eng64, eng32, jobu, jobv, k, kmax, sigma, bnd, &
                                        sigma32, bnd32, tolin, doption, ioption, info32) &
    bind(C, name='pk_lansvd_mixed')
  use iso_c_binding
  use pk_api
  use pk_drivers
  implicit none
  type(c_ptr), value :: eng64, eng32
  character(kind=c_char), value :: jobu, jobv
  integer(c_int) :: k, info32(2)
  integer(c_int), value :: kmax
  real(c_double) :: sigma(*), bnd(*), sigma32(*), bnd32(*), doption(*)
  real(c_double), value :: tolin
  integer(c_int) :: ioption(*)
  interface
    integer(c_int) function pk_lbd_seed(dst, src, ncols) bind(C, name='pk_lbd_seed')
      import :: c_int, c_ptr
      type(c_ptr), value :: dst, src
      integer(c_int), value :: ncols
    end function
    integer(c_int) function pk_lbd_set_rounding(e, model) bind(C, name='pk_lbd_set_rounding')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: model
    end function
    integer(c_int) function pk_lbd_residuals(e, src, kk, sig, res) &
        bind(C, name='pk_lbd_residuals')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e, src
      integer(c_int), value :: kk
      real(c_double) :: sig(*), res(*)
    end function
    integer(c_int) function pk_lbd_promote(dst, src, ncols) bind(C, name='pk_lbd_promote')
      import :: c_int, c_ptr
      type(c_ptr), value :: dst, src
      integer(c_int), value :: ncols
    end function
  end interface
  integer :: m, n, info, kk, k32, m32, n32
  double precision :: eps, eps32, dopt32(4)
  double precision, allocatable :: res(:)
  integer(c_int) :: rc
  logical :: check
  call engine_shape(eng64, m, n, eps, info)
  if (info /= 0) then
    pk_lansvd_mixed = info
    return
  end if
  call engine_shape(eng32, m32, n32, eps32, info)
  if (info /= 0 .or. m32 /= m .or. n32 /= n) then
    pk_lansvd_mixed = -101
    return
  end if
  rc = pk_lbd_set_rounding(eng32, 1)
  dopt32(1) = sqrt(eps32)
  dopt32(2) = eps32**0.75d0
  dopt32(3) = 0d0
  dopt32(4) = 0d0
  check = tolin >= 16d0 * eps32
  k32 = k
  call lansvd_core('y', merge('y', 'n', check), m, n, k32, kmax, eng32, eps32, sigma32, &
                   bnd32, max(tolin, 16d0 * eps32), dopt32, ioption, info)
  info32(1) = info
  info32(2) = k32
  if (info <= -100) then
    pk_lansvd_mixed = info
    return
  end if
  if (check .and. info == 0 .and. k32 >= k .and. k32 > 0) then
    allocate(res(2 * k32))
    rc = pk_lbd_residuals(eng64, eng32, k32, sigma32, res)
    if (rc /= PK_OK) then
      pk_lansvd_mixed = rc
      return
    end if
    if (all(res(1:2 * k32:2) <= tolin * sigma32(1))) then
      rc = pk_lbd_promote(eng64, eng32, k32)
      if (rc /= PK_OK) then
        pk_lansvd_mixed = rc
        return
      end if
      sigma(1:k32) = sigma32(1:k32)
      bnd(1:k32) = bnd32(1:k32)
      k = k32
      pk_lansvd_mixed = 0
      return
    end if
  end if
  if (k32 > 0) then
    rc = pk_lbd_seed(eng64, eng32, k32)
    if (rc /= PK_OK) then
      pk_lansvd_mixed = rc
      return
    end if
  end if
  kk = k
  call lansvd_core(jobu, jobv, m, n, kk, kmax, eng64, eps, sigma, bnd, tolin, doption, &
                   ioption, info)
  k = kk
  pk_lansvd_mixed = info
end function

! FP32-basis form with FP64 polish (no reference counterpart).  eng32 holds
! FP32 values and bases (all reductions and scalars FP64); it runs the
! restarted driver to the FP32-attainable tolerance max(tolin, 16*eps32).
! Its left Ritz vectors seed an FP64 restarted solve on eng64 (same shape),
! which delivers the triplets to tolin.  sigma32/bnd32 return the FP32-phase
! values (for the per-triplet error report), sigma/bnd the polished ones.
! info32 returns the FP32 phase's INFO and converged count in info32(1:2).
integer(c_int) function pk_lansvd_irl_mixed(eng64, eng32, which, jobu, jobv, dim, p, neig, &
                                            maxiter, sigma, bnd, sigma32, bnd32, tolin, &
                                            doption, ioption, info32) &
    bind(C, name='pk_lansvd_irl_mixed')
  use iso_c_binding
  use pk_api
  use pk_drivers
  implicit none
  type(c_ptr), value :: eng64, eng32
  character(kind=c_char), value :: which, jobu, jobv
  integer(c_int) :: dim, neig, info32(2)
  integer(c_int), value :: p, maxiter
  real(c_double) :: sigma(*), bnd(*), sigma32(*), bnd32(*), doption(*)
  real(c_double), value :: tolin
  integer(c_int) :: ioption(*)
  interface
    integer(c_int) function pk_lbd_seed(dst, src, ncols) bind(C, name='pk_lbd_seed')
      import :: c_int, c_ptr
      type(c_ptr), value :: dst, src
      integer(c_int), value :: ncols
    end function
    integer(c_int) function pk_lbd_set_rounding(e, model) bind(C, name='pk_lbd_set_rounding')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: model
    end function
    integer(c_int) function pk_lbd_residuals(e, src, k, sig, res) &
        bind(C, name='pk_lbd_residuals')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e, src
      integer(c_int), value :: k
      real(c_double) :: sig(*), res(*)
    end function
    integer(c_int) function pk_lbd_promote(dst, src, ncols) bind(C, name='pk_lbd_promote')
      import :: c_int, c_ptr
      type(c_ptr), value :: dst, src
      integer(c_int), value :: ncols
    end function
  end interface
  integer :: m, n, info, dd, ne, ne32, m32, n32
  double precision :: eps, eps32, dopt32(4)
  double precision, allocatable :: res(:)
  integer(c_int) :: rc
  logical :: check
  call engine_shape(eng64, m, n, eps, info)
  if (info /= 0) then
    pk_lansvd_irl_mixed = info
    return
  end if
  call engine_shape(eng32, m32, n32, eps32, info)
  if (info /= 0 .or. m32 /= m .or. n32 /= n) then
    pk_lansvd_irl_mixed = -101
    return
  end if
  ! FP32-basis phase, left vectors kept on the device for the seed.  Its
  ! thresholds follow the storage-rounding model (FP64 accumulation).
  rc = pk_lbd_set_rounding(eng32, 1)
  dd = dim
  ne32 = neig
  ! the FP32 phase keeps its basis semi-orthogonal at its own precision:
  ! delta = sqrt(eps32), eta = eps32**(3/4) (the reference's defaults for
  ! SLANSVD_IRL callers), the caller's relative-gap filter
  dopt32(1) = sqrt(eps32)
  dopt32(2) = eps32**0.75d0
  dopt32(3) = 0d0
  dopt32(4) = doption(4)
  ! tolin reachable by the FP32 phase's own tolerance: its triplets are
  ! checked in FP64 and, when every one meets tolin, no polish runs
  check = tolin >= 16d0 * eps32
  call lansvd_irl_core(which, 'y', merge('y', 'n', check), m, n, dd, p, ne32, maxiter, eng32, &
                       eps32, sigma32, bnd32, max(tolin, 16d0 * eps32), dopt32, ioption, info)
  info32(1) = info
  info32(2) = ne32
  if (info <= -100) then
    pk_lansvd_irl_mixed = info
    return
  end if
  if (check .and. info == 0 .and. ne32 >= neig .and. ne32 > 0) then
    allocate(res(2 * ne32))
    rc = pk_lbd_residuals(eng64, eng32, ne32, sigma32, res)
    if (rc /= PK_OK) then
      pk_lansvd_irl_mixed = rc
      return
    end if
    if (all(res(1:2 * ne32:2) <= tolin * sigma32(1))) then
      rc = pk_lbd_promote(eng64, eng32, ne32)
      if (rc /= PK_OK) then
        pk_lansvd_irl_mixed = rc
        return
      end if
      sigma(1:ne32) = sigma32(1:ne32)
      bnd(1:ne32) = bnd32(1:ne32)
      dim = dd
      neig = ne32
      pk_lansvd_irl_mixed = 0
      return
    end if
  end if
  ! FP64 polish seeded with the sum of the FP32 left Ritz vectors
  if (ne32 > 0) then
    rc = pk_lbd_seed(eng64, eng32, ne32)
    if (rc /= PK_OK) then
      pk_lansvd_irl_mixed = rc
      return
    end if
  end if
  dd = dim
  ne = neig
  call lansvd_irl_core(which, jobu, jobv, m, n, dd, p, ne, maxiter, eng64, eps, sigma, bnd, &
                       tolin, doption, ioption, info)
  dim = dd
  neig = ne
  pk_lansvd_irl_mixed = info
end function
