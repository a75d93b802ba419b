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
