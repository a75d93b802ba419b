!*************************************************************************
MODULE si3d_gpu
!*************************************************************************
!
!  ISO_C_BINDING interface of libsi3dgpu.so (include/si3d_gpu.h) for the
!  Fortran driver si3d.f90.  The driver keeps its input, bathy, init and
!  output routines; after init it hands the grid and state to the library
!  (s3g_init + s3g_set_field_*) and replaces the calls to fd / settrap /
!  settrap2 / save (si3d.f90:353-393) with s3g_fd / s3g_settrap /
!  s3g_settrap2 / s3g_save.  Before an output is due it pulls the fields
!  the out* routines read with s3g_get_field_*.
!
!  Return codes: 0 ok, >0 warning (1: free-surface solve hit its iteration
!  limit), <0 fatal -- the caller prints s3g_error_message() and STOPs, as
!  SolverSparse does for NSPCG errors (si3d_procedures.f90:2005-2016).
!-------------------------------------------------------------------------
   USE ISO_C_BINDING
   IMPLICIT NONE

   INTEGER(C_INT), PARAMETER :: S3G_FP64 = 0, S3G_MIXED = 1
   INTEGER(C_INT), PARAMETER :: S3G_PREC_JACOBI = 0, S3G_PREC_LINE = 1, S3G_PREC_MG = 2

   TYPE, BIND(C) :: s3g_grid
      INTEGER(C_INT) :: im1, jm1, km1, lm
      TYPE(C_PTR) :: l2i, l2j, lEC, lWC, lNC, lSC, kmz, k1z, k1u, k1v
      TYPE(C_PTR) :: zlevel, hhs, hhu, hhv
   END TYPE s3g_grid

   TYPE, BIND(C) :: s3g_params
      REAL(C_DOUBLE) :: dt, dx, dy
      INTEGER(C_INT) :: itrap, niter, ismooth, iturb, iadv, itrmom, ihd, ibc, isal, ifsurfbc
      REAL(C_DOUBLE) :: beta, Av0, Dv0, Ax0, Ay0, f, theta, cd, cw, wa, phi, dtsurfbc
      INTEGER(C_INT) :: n_surfbc
      TYPE(C_PTR)    :: surfbc
      INTEGER(C_INT) :: precision, preconditioner
      REAL(C_DOUBLE) :: pcg_tol
      INTEGER(C_INT) :: pcg_maxit
      INTEGER(C_INT) :: ntr
      ! met stations of ifsurfbc 10 / 11 (unused by the other modes)
      INTEGER(C_INT) :: icw = 0, nmetstat = 0
      TYPE(C_PTR)    :: metxy = C_NULL_PTR
      INTEGER(C_INT) :: iinterp = 1
      REAL(C_DOUBLE) :: gammaB = 0.0_C_DOUBLE, delNfactor = 0.0_C_DOUBLE
   END TYPE s3g_params

   ! input deck of the standalone driver (s3g_create_from_deck)
   TYPE, BIND(C) :: s3g_deck
      INTEGER(C_INT) :: imm, jmm, ibathyf, km1_file
      REAL(C_DOUBLE) :: ddz, dzmin, datadj, zetainit, zl
      TYPE(C_PTR)    :: zlevel_file, bathy_dm, init_profile, tracer_profiles
   END TYPE s3g_deck

   ! nopen block of si3d_inp.txt and the openbcNN.txt records (openbc0)
   TYPE, BIND(C) :: s3g_openbc
      INTEGER(C_INT) :: nopen
      REAL(C_DOUBLE) :: dtsec, dzmin
      TYPE(C_PTR)    :: iside, itype, isbc, jsbc, iebc, jebc
      INTEGER(C_INT) :: npts, nvar
      TYPE(C_PTR)    :: values
   END TYPE s3g_openbc

   INTERFACE
      FUNCTION s3g_init(grid, params, ctx) BIND(C, NAME="s3g_init")
         IMPORT :: C_INT, C_PTR, s3g_grid, s3g_params
         TYPE(s3g_grid), INTENT(IN)   :: grid
         TYPE(s3g_params), INTENT(IN) :: params
         TYPE(C_PTR), INTENT(OUT)     :: ctx
         INTEGER(C_INT) :: s3g_init
      END FUNCTION s3g_init

      FUNCTION s3g_create_from_deck(deck, params, ctx) BIND(C, NAME="s3g_create_from_deck")
         IMPORT :: C_INT, C_PTR, s3g_deck, s3g_params
         TYPE(s3g_deck), INTENT(IN)   :: deck
         TYPE(s3g_params), INTENT(IN) :: params
         TYPE(C_PTR), INTENT(OUT)     :: ctx
         INTEGER(C_INT) :: s3g_create_from_deck
      END FUNCTION s3g_create_from_deck

      FUNCTION s3g_set_field_f32(ctx, name, host) BIND(C, NAME="s3g_set_field_f32")
         IMPORT :: C_INT, C_PTR, C_CHAR, C_FLOAT
         TYPE(C_PTR), VALUE :: ctx
         CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: name
         REAL(C_FLOAT), DIMENSION(*), INTENT(IN) :: host
         INTEGER(C_INT) :: s3g_set_field_f32
      END FUNCTION s3g_set_field_f32

      FUNCTION s3g_set_field_f64(ctx, name, host) BIND(C, NAME="s3g_set_field_f64")
         IMPORT :: C_INT, C_PTR, C_CHAR, C_DOUBLE
         TYPE(C_PTR), VALUE :: ctx
         CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: name
         REAL(C_DOUBLE), DIMENSION(*), INTENT(IN) :: host
         INTEGER(C_INT) :: s3g_set_field_f64
      END FUNCTION s3g_set_field_f64

      FUNCTION s3g_set_field_i32(ctx, name, host) BIND(C, NAME="s3g_set_field_i32")
         IMPORT :: C_INT, C_PTR, C_CHAR
         TYPE(C_PTR), VALUE :: ctx
         CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: name
         INTEGER(C_INT), DIMENSION(*), INTENT(IN) :: host
         INTEGER(C_INT) :: s3g_set_field_i32
      END FUNCTION s3g_set_field_i32

      FUNCTION s3g_get_field_f32(ctx, name, host) BIND(C, NAME="s3g_get_field_f32")
         IMPORT :: C_INT, C_PTR, C_CHAR, C_FLOAT
         TYPE(C_PTR), VALUE :: ctx
         CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: name
         REAL(C_FLOAT), DIMENSION(*), INTENT(OUT) :: host
         INTEGER(C_INT) :: s3g_get_field_f32
      END FUNCTION s3g_get_field_f32

      FUNCTION s3g_get_field_f64(ctx, name, host) BIND(C, NAME="s3g_get_field_f64")
         IMPORT :: C_INT, C_PTR, C_CHAR, C_DOUBLE
         TYPE(C_PTR), VALUE :: ctx
         CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: name
         REAL(C_DOUBLE), DIMENSION(*), INTENT(OUT) :: host
         INTEGER(C_INT) :: s3g_get_field_f64
      END FUNCTION s3g_get_field_f64

      FUNCTION s3g_get_field_i32(ctx, name, host) BIND(C, NAME="s3g_get_field_i32")
         IMPORT :: C_INT, C_PTR, C_CHAR
         TYPE(C_PTR), VALUE :: ctx
         CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: name
         INTEGER(C_INT), DIMENSION(*), INTENT(OUT) :: host
         INTEGER(C_INT) :: s3g_get_field_i32
      END FUNCTION s3g_get_field_i32

      FUNCTION s3g_fd(ctx, n, istep, iter, lastiter, thrs) BIND(C, NAME="s3g_fd")
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR), VALUE :: ctx
         INTEGER(C_INT), VALUE :: n, istep, iter, lastiter
         REAL(C_DOUBLE), VALUE :: thrs
         INTEGER(C_INT) :: s3g_fd
      END FUNCTION s3g_fd

      FUNCTION s3g_set_open_boundaries(ctx, bc) BIND(C, NAME="s3g_set_open_boundaries")
         IMPORT :: C_INT, C_PTR, s3g_openbc
         TYPE(C_PTR), VALUE :: ctx
         TYPE(s3g_openbc), INTENT(IN) :: bc
         INTEGER(C_INT) :: s3g_set_open_boundaries
      END FUNCTION s3g_set_open_boundaries

      FUNCTION s3g_settrap(ctx) BIND(C, NAME="s3g_settrap")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: ctx
         INTEGER(C_INT) :: s3g_settrap
      END FUNCTION s3g_settrap

      FUNCTION s3g_settrap2(ctx) BIND(C, NAME="s3g_settrap2")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: ctx
         INTEGER(C_INT) :: s3g_settrap2
      END FUNCTION s3g_settrap2

      FUNCTION s3g_save(ctx, istep) BIND(C, NAME="s3g_save")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: ctx
         INTEGER(C_INT), VALUE :: istep
         INTEGER(C_INT) :: s3g_save
      END FUNCTION s3g_save

      FUNCTION s3g_step(ctx, n) BIND(C, NAME="s3g_step")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: ctx
         INTEGER(C_INT), VALUE :: n
         INTEGER(C_INT) :: s3g_step
      END FUNCTION s3g_step

      ! TKinE, ShearCum, BuocyCum, DisspCum for OutScalarEnergyBalance
      FUNCTION s3g_energy_balance(ctx, out, reset) BIND(C, NAME="s3g_energy_balance")
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR), VALUE :: ctx
         REAL(C_DOUBLE), INTENT(OUT) :: out(4)
         INTEGER(C_INT), VALUE :: reset
         INTEGER(C_INT) :: s3g_energy_balance
      END FUNCTION s3g_energy_balance

      FUNCTION s3g_synchronize(ctx) BIND(C, NAME="s3g_synchronize")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: ctx
         INTEGER(C_INT) :: s3g_synchronize
      END FUNCTION s3g_synchronize

      SUBROUTINE s3g_finalize(ctx) BIND(C, NAME="s3g_finalize")
         IMPORT :: C_PTR
         TYPE(C_PTR), VALUE :: ctx
      END SUBROUTINE s3g_finalize

      FUNCTION s3g_last_error() BIND(C, NAME="s3g_last_error")
         IMPORT :: C_PTR
         TYPE(C_PTR) :: s3g_last_error
      END FUNCTION s3g_last_error

      ! learned initial guess of the free-surface solve (no counterpart in the reference)
      FUNCTION s3g_learn_setup(ctx, capacity, seed) BIND(C, NAME="s3g_learn_setup")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: ctx
         INTEGER(C_INT), VALUE :: capacity, seed
         INTEGER(C_INT) :: s3g_learn_setup
      END FUNCTION s3g_learn_setup

      FUNCTION s3g_learn_record(ctx, on) BIND(C, NAME="s3g_learn_record")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: ctx
         INTEGER(C_INT), VALUE :: on
         INTEGER(C_INT) :: s3g_learn_record
      END FUNCTION s3g_learn_record

      FUNCTION s3g_learn_train(ctx, steps, batch, lr, loss) BIND(C, NAME="s3g_learn_train")
         IMPORT :: C_INT, C_PTR, C_DOUBLE
         TYPE(C_PTR), VALUE :: ctx
         INTEGER(C_INT), VALUE :: steps, batch
         REAL(C_DOUBLE), VALUE :: lr
         REAL(C_DOUBLE) :: loss
         INTEGER(C_INT) :: s3g_learn_train
      END FUNCTION s3g_learn_train

      FUNCTION s3g_learn_mode(ctx, mode) BIND(C, NAME="s3g_learn_mode")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: ctx
         INTEGER(C_INT), VALUE :: mode
         INTEGER(C_INT) :: s3g_learn_mode
      END FUNCTION s3g_learn_mode

      FUNCTION s3g_learn_flexible(ctx, period) BIND(C, NAME="s3g_learn_flexible")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: ctx
         INTEGER(C_INT), VALUE :: period
         INTEGER(C_INT) :: s3g_learn_flexible
      END FUNCTION s3g_learn_flexible
   END INTERFACE

CONTAINS

   ! Fortran string -> NUL-terminated C string
   FUNCTION cstr(s) RESULT(c)
      CHARACTER(LEN=*), INTENT(IN) :: s
      CHARACTER(KIND=C_CHAR), DIMENSION(LEN_TRIM(s)+1) :: c
      INTEGER :: i
      DO i = 1, LEN_TRIM(s)
         c(i) = s(i:i)
      END DO
      c(LEN_TRIM(s)+1) = C_NULL_CHAR
   END FUNCTION cstr

   ! Print the library's last error and stop, like the NSPCG fatal path.
   SUBROUTINE s3g_stop_on_error(rc, where)
      INTEGER(C_INT), INTENT(IN) :: rc
      CHARACTER(LEN=*), INTENT(IN) :: where
      CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
      CHARACTER(LEN=256) :: text
      INTEGER :: i
      IF (rc >= 0) RETURN
      CALL C_F_POINTER(s3g_last_error(), msg, [256])
      text = ' '
      DO i = 1, 256
         IF (msg(i) == C_NULL_CHAR) EXIT
         text(i:i) = msg(i)
      END DO
      PRINT *, ' '
      PRINT '(" **ERROR", I5, " from libsi3dgpu in ", A)', rc, where
      PRINT *, TRIM(text)
      PRINT *, ' ****STOPPING si3d due to fatal error in the GPU time step'
      STOP
   END SUBROUTINE s3g_stop_on_error

END MODULE si3d_gpu
