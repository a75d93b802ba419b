! s3g_fortran_check.f90 -- the library driven from Fortran through the
! si3d_gpu module, as the reference driver would: the 100x50x20 wind-driven
! rectangle (flat 20 m bottom, 1 m layers, dx = 50 m, dt = 30 s, 0.1 Pa along
! x, FP64, multigrid PCG to 1e-10) is created from its deck, advanced nsteps
! time steps, and zeta (lm1 values) and uh (km1*lm1 values, k fastest) are
! copied into the caller's arrays.  Returns 0, or the first failing return
! code.  Built into a shared library so a test can call it in-process.
FUNCTION s3g_fortran_check(nsteps, s_out, uh_out) BIND(C, NAME="s3g_fortran_check") RESULT(rc)
   USE ISO_C_BINDING
   USE si3d_gpu
   IMPLICIT NONE
   INTEGER(C_INT), VALUE :: nsteps
   REAL(C_DOUBLE), INTENT(OUT) :: s_out(*), uh_out(*)
   INTEGER(C_INT) :: rc
   INTEGER, PARAMETER :: imm = 100, jmm = 50, km1 = 22
   TYPE(s3g_deck) :: dk
   TYPE(s3g_params) :: prm
   TYPE(C_PTR) :: ctx
   INTEGER(C_INT), ALLOCATABLE, TARGET :: bathy(:)
   REAL(C_DOUBLE), ALLOCATABLE, TARGET :: prof(:)
   INTEGER :: n

   ALLOCATE (bathy(imm * jmm), prof(km1))
   bathy = 200            ! decimetres
   prof = 20.0D0          ! uniform temperature
   dk%imm = imm; dk%jmm = jmm; dk%ibathyf = 0; dk%km1_file = 0
   dk%ddz = 1.0D0; dk%dzmin = 0.1D0; dk%datadj = 0.0D0; dk%zetainit = 0.0D0; dk%zl = 20.0D0
   dk%zlevel_file = C_NULL_PTR; dk%bathy_dm = C_LOC(bathy); dk%init_profile = C_LOC(prof)
   dk%tracer_profiles = C_NULL_PTR

   prm%dt = 30.0D0; prm%dx = 50.0D0; prm%dy = 50.0D0
   prm%itrap = 1; prm%niter = 2; prm%ismooth = 0; prm%iturb = 0; prm%iadv = 1; prm%itrmom = 1
   prm%ihd = 1; prm%ibc = 0; prm%isal = 0; prm%ifsurfbc = 0
   prm%beta = 0.05D0; prm%Av0 = 3.0D-2; prm%Dv0 = 3.0D-2; prm%Ax0 = 1.0D0; prm%Ay0 = 1.0D0
   prm%f = 0.0D0; prm%theta = 1.0D0; prm%cd = 2.5D-3; prm%cw = 1.2D-3
   prm%wa = SQRT(0.1D0 / (1.2D-3 * 1.3D0)); prm%phi = 270.0D0; prm%dtsurfbc = 600.0D0
   prm%n_surfbc = 0; prm%surfbc = C_NULL_PTR
   prm%precision = 0        ! S3G_FP64
   prm%preconditioner = 2   ! S3G_PREC_MG
   prm%pcg_tol = 1.0D-10; prm%pcg_maxit = 4000; prm%ntr = 0

   rc = s3g_create_from_deck(dk, prm, ctx)
   IF (rc < 0) RETURN
   DO n = 1, nsteps
      rc = s3g_step(ctx, n)
      IF (rc < 0) EXIT
   END DO
   IF (rc >= 0) rc = s3g_get_field_f64(ctx, cstr('s'), s_out)
   IF (rc >= 0) rc = s3g_get_field_f64(ctx, cstr('uh'), uh_out)
   CALL s3g_finalize(ctx)
END FUNCTION s3g_fortran_check

! The drop-in hand-over and time loop of the patched si3d.f90, as
! INTEGRATION.md shows it: the grid left by bathy / init (1-based index arrays
! of the reference layout, REAL zlevel / hhs / hhu / hhv) goes to s3g_init;
! every REAL state array goes over with s3g_set_field_f32 and the REAL*8
! turbulence arrays with s3g_set_field_f64; then each time step replaces the
! driver's fd / settrap / settrap2 / save calls (si3d.f90:324-393, itrap = 1:
! a leapfrog fd and settrap from step 2 on, then niter trapezoidal fd with
! settrap2 between them, then save(2)).  After nsteps steps zeta, uh and the
! temperature come back with s3g_get_field_f32 and q2 with
! s3g_get_field_f64, as before the driver's output routines.
!   f2d(lm1, 3):        s, sp, spp
!   f3d(km1, lm1, 28):  h hp hpp hu hup hupp hv hvp hvpp uh uhp uhpp u up upp
!                       vh vhp vhpp v vp vpp wp sal salp salpp rhop Av Dv
!   f8(km1, lm1, 6):    q2 q2p q2pp q2l q2lp q2lpp
FUNCTION s3g_fortran_dropin(im1, jm1, km1, lm, l2i, l2j, lEC, lWC, lNC, lSC, kmz, k1z, k1u, k1v, &
                            zlevel, hhs, hhu, hhv, prm, nsteps, f2d, f3d, f8, s_out, uh_out, sal_out, q2_out) &
                            BIND(C, NAME="s3g_fortran_dropin") RESULT(rc)
   USE ISO_C_BINDING
   USE si3d_gpu
   IMPLICIT NONE
   INTEGER(C_INT), VALUE :: im1, jm1, km1, lm, nsteps
   INTEGER(C_INT), DIMENSION(lm + 1), TARGET, INTENT(IN) :: l2i, l2j, lEC, lWC, lNC, lSC, kmz, k1z, k1u, k1v
   REAL(C_FLOAT), DIMENSION(km1), TARGET, INTENT(IN) :: zlevel
   REAL(C_FLOAT), DIMENSION(lm + 1), TARGET, INTENT(IN) :: hhs, hhu, hhv
   TYPE(s3g_params), INTENT(IN) :: prm
   REAL(C_FLOAT), INTENT(IN) :: f2d(lm + 1, 3), f3d(km1, lm + 1, 28)
   REAL(C_DOUBLE), INTENT(IN) :: f8(km1, lm + 1, 6)
   REAL(C_FLOAT), INTENT(OUT) :: s_out(lm + 1), uh_out(km1, lm + 1), sal_out(km1, lm + 1)
   REAL(C_DOUBLE), INTENT(OUT) :: q2_out(km1, lm + 1)
   INTEGER(C_INT) :: rc
   CHARACTER(LEN=5), PARAMETER :: n2d(3) = [CHARACTER(LEN=5) :: 's', 'sp', 'spp']
   CHARACTER(LEN=5), PARAMETER :: n3d(28) = [CHARACTER(LEN=5) :: 'h', 'hp', 'hpp', 'hu', 'hup', 'hupp', &
      'hv', 'hvp', 'hvpp', 'uh', 'uhp', 'uhpp', 'u', 'up', 'upp', 'vh', 'vhp', 'vhpp', 'v', 'vp', 'vpp', &
      'wp', 'sal', 'salp', 'salpp', 'rhop', 'Av', 'Dv']
   CHARACTER(LEN=5), PARAMETER :: n8(6) = [CHARACTER(LEN=5) :: 'q2', 'q2p', 'q2pp', 'q2l', 'q2lp', 'q2lpp']
   TYPE(s3g_grid) :: grid
   TYPE(C_PTR) :: ctx
   INTEGER :: i, n, iter, niter1, lastiter
   REAL(C_DOUBLE) :: thrs

   grid%im1 = im1; grid%jm1 = jm1; grid%km1 = km1; grid%lm = lm
   grid%l2i = C_LOC(l2i); grid%l2j = C_LOC(l2j)
   grid%lEC = C_LOC(lEC); grid%lWC = C_LOC(lWC); grid%lNC = C_LOC(lNC); grid%lSC = C_LOC(lSC)
   grid%kmz = C_LOC(kmz); grid%k1z = C_LOC(k1z); grid%k1u = C_LOC(k1u); grid%k1v = C_LOC(k1v)
   grid%zlevel = C_LOC(zlevel); grid%hhs = C_LOC(hhs); grid%hhu = C_LOC(hhu); grid%hhv = C_LOC(hhv)
   rc = s3g_init(grid, prm, ctx)
   IF (rc < 0) RETURN
   DO i = 1, 3
      rc = s3g_set_field_f32(ctx, cstr(n2d(i)), f2d(:, i))
      IF (rc < 0) GOTO 90
   END DO
   DO i = 1, 28
      rc = s3g_set_field_f32(ctx, cstr(n3d(i)), f3d(:, :, i))
      IF (rc < 0) GOTO 90
   END DO
   DO i = 1, 6
      rc = s3g_set_field_f64(ctx, cstr(n8(i)), f8(:, :, i))
      IF (rc < 0) GOTO 90
   END DO

   DO n = 1, nsteps
      thrs = n * prm%dt / 3600.0D0
      niter1 = prm%niter
      IF (n == 1) niter1 = MAX(prm%niter, 2)
      lastiter = 0
      IF (n > 1) THEN
         rc = s3g_fd(ctx, n, 1, 0, lastiter, thrs)
         IF (rc < 0) GOTO 90
         rc = s3g_settrap(ctx)
         IF (rc < 0) GOTO 90
      END IF
      DO iter = 1, niter1
         IF (iter == niter1) lastiter = 1
         rc = s3g_fd(ctx, n, 2, iter, lastiter, thrs)
         IF (rc < 0) GOTO 90
         IF (iter < niter1) THEN
            rc = s3g_settrap2(ctx)
            IF (rc < 0) GOTO 90
         END IF
      END DO
      rc = s3g_save(ctx, 2)
      IF (rc < 0) GOTO 90
   END DO
   rc = s3g_get_field_f32(ctx, cstr('s'), s_out)
   IF (rc >= 0) rc = s3g_get_field_f32(ctx, cstr('uh'), uh_out)
   IF (rc >= 0) rc = s3g_get_field_f32(ctx, cstr('sal'), sal_out)
   IF (rc >= 0) rc = s3g_get_field_f64(ctx, cstr('q2'), q2_out)
90 CALL s3g_finalize(ctx)
END FUNCTION s3g_fortran_dropin
