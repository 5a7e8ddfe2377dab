! Fortran bindings of libcumc3d_hip.so (include/cumc3d_hip.h) through
! ISO_C_BINDING.  The derived type and the interfaces mirror the C
! declarations one for one; cumc3d_check turns a non-zero status into the
! reference's error convention (STOP with a message, setup.f90:120-322).
MODULE CUMC3D_HIP
USE ISO_C_BINDING
IMPLICIT NONE

INTEGER(C_INT), PARAMETER :: CUMC3D_OK = 0, CUMC3D_EINVAL = 1, CUMC3D_EHIP = 2, &
                             CUMC3D_ENOMEM = 3, CUMC3D_ESTATE = 4, CUMC3D_ENONFINITE = 5, &
                             CUMC3D_ESTOP = 6
INTEGER(C_INT), PARAMETER :: CUMC3D_LF = 0, CUMC3D_HLL = 1, CUMC3D_HLLC = 2, CUMC3D_HLLD = 3
INTEGER(C_INT), PARAMETER :: CUMC3D_TVDMM = 0, CUMC3D_TVDVL = 1, CUMC3D_TVDMC = 2, &
                             CUMC3D_PPM = 3, CUMC3D_WENO = 4
INTEGER(C_INT), PARAMETER :: CUMC3D_MODEL_RIEMANN = 0, CUMC3D_MODEL_TORUS = 1, CUMC3D_MODEL_TORUS_SPH = 2, &
                             CUMC3D_MODEL_COLLAPSAR = 3
INTEGER(C_INT), PARAMETER :: CUMC3D_FP64 = 0, CUMC3D_FP32 = 1
INTEGER(C_INT), PARAMETER :: CUMC3D_BN_ZERO = 0, CUMC3D_BN_FIELD = 1, CUMC3D_BN_CARRY = 2

TYPE, BIND(C) :: cumc3d_grid_metrics
  TYPE(C_PTR) :: dx, dy, dz
  TYPE(C_PTR) :: x, y, z
  TYPE(C_PTR) :: dx_cb, xbar
  TYPE(C_PTR) :: dcose, dsine, sine
  TYPE(C_PTR) :: sinf
END TYPE

TYPE, BIND(C) :: cumc3d_gravity_params
  REAL(C_DOUBLE) :: prim_a(4)
  REAL(C_DOUBLE) :: eps_a, m_bh, c_code, omega, tolerance
  INTEGER(C_INT) :: relax_max, n_pot
END TYPE

TYPE, BIND(C) :: cumc3d_desc
  INTEGER(C_INT) :: nx, ny, nz
  INTEGER(C_INT) :: n_dim
  INTEGER(C_INT) :: coordinate_flag
  INTEGER(C_INT) :: boundary_flag(6)
  INTEGER(C_INT) :: riemann, recon, model, rt_gravity, hlld_bn_mode, precision
  REAL(C_DOUBLE) :: cfl, ggas, r_sh
  TYPE(C_PTR) :: xF, yF, zF
  INTEGER(C_INT) :: fixpole
  TYPE(C_PTR) :: metrics
  TYPE(C_PTR) :: gravity
END TYPE

INTERFACE
  INTEGER(C_INT) FUNCTION cumc3d_init(desc, ctx) BIND(C, NAME='cumc3d_init')
    IMPORT :: C_INT, C_PTR, cumc3d_desc
    TYPE(cumc3d_desc), INTENT(IN) :: desc
    TYPE(C_PTR), INTENT(OUT) :: ctx
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_upload(ctx, prim, eps, floors3) BIND(C, NAME='cumc3d_upload')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(IN) :: prim(*), eps(*)
    REAL(C_DOUBLE), INTENT(IN), OPTIONAL :: floors3(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_initial_update(ctx) BIND(C, NAME='cumc3d_initial_update')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: ctx
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_finddt(ctx, dt) BIND(C, NAME='cumc3d_finddt')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(OUT) :: dt
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_rk3_step(ctx, dt) BIND(C, NAME='cumc3d_rk3_step')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), VALUE :: dt
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_stage(ctx, stage, dt) BIND(C, NAME='cumc3d_stage')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: stage
    REAL(C_DOUBLE), VALUE :: dt
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_spatial(ctx, l_rk) BIND(C, NAME='cumc3d_spatial')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(INOUT) :: l_rk(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_run(ctx, max_steps, total_time, sync_each_step, t, dts, steps_done) &
      BIND(C, NAME='cumc3d_run')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: max_steps, sync_each_step
    REAL(C_DOUBLE), VALUE :: total_time
    REAL(C_DOUBLE), INTENT(INOUT) :: t
    REAL(C_DOUBLE), INTENT(OUT), OPTIONAL :: dts(*)
    INTEGER(C_INT), INTENT(OUT) :: steps_done
  END FUNCTION
  ! the same loop without device allocation (no reference counterpart: for
  ! callers that cannot allocate; dts holds at most 65536 entries)
  INTEGER(C_INT) FUNCTION cumc3d_run_noalloc(ctx, max_steps, total_time, sync_each_step, t, dts, steps_done) &
      BIND(C, NAME='cumc3d_run_noalloc')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: max_steps, sync_each_step
    REAL(C_DOUBLE), VALUE :: total_time
    REAL(C_DOUBLE), INTENT(INOUT) :: t
    REAL(C_DOUBLE), INTENT(OUT), OPTIONAL :: dts(*)
    INTEGER(C_INT), INTENT(OUT) :: steps_done
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_download(ctx, prim, eps, bcell, cs) BIND(C, NAME='cumc3d_download')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(OUT), OPTIONAL :: prim(*), eps(*), bcell(*), cs(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_profile(ctx, primitive, epsilon, bfield) BIND(C, NAME='cumc3d_profile')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(OUT), OPTIONAL :: primitive(*), epsilon(*), bfield(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_gravity(ctx, phi, m_bh, r_sh, iterations) BIND(C, NAME='cumc3d_gravity')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(OUT), OPTIONAL :: phi(*), m_bh, r_sh
    INTEGER(C_INT), INTENT(OUT), OPTIONAL :: iterations
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_divb(ctx, maxdivb) BIND(C, NAME='cumc3d_divb')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(OUT) :: maxdivb
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_finalize(ctx) BIND(C, NAME='cumc3d_finalize')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: ctx
  END FUNCTION
  TYPE(C_PTR) FUNCTION cumc3d_error_string(code) BIND(C, NAME='cumc3d_error_string')
    IMPORT :: C_INT, C_PTR
    INTEGER(C_INT), VALUE :: code
  END FUNCTION
  TYPE(C_PTR) FUNCTION cumc3d_last_error() BIND(C, NAME='cumc3d_last_error')
    IMPORT :: C_PTR
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_set_timing(ctx, enable) BIND(C, NAME='cumc3d_set_timing')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: enable
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_get_timing(ctx, ms, n) BIND(C, NAME='cumc3d_get_timing')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(OUT) :: ms(*)
    INTEGER(C_INT), VALUE :: n
  END FUNCTION
  ! learned reconstruction selector (no counterpart in the reference)
  INTEGER(C_INT) FUNCTION cumc3d_selector_init(ctx, capacity, seed) BIND(C, NAME='cumc3d_selector_init')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: capacity, seed
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_selector_set_params(ctx, w, n) BIND(C, NAME='cumc3d_selector_set_params')
    IMPORT :: C_INT, C_PTR, C_FLOAT
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_FLOAT), INTENT(IN) :: w(*)
    INTEGER(C_INT), VALUE :: n
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_selector_get_params(ctx, w, n) BIND(C, NAME='cumc3d_selector_get_params')
    IMPORT :: C_INT, C_PTR, C_FLOAT
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_FLOAT), INTENT(OUT) :: w(*)
    INTEGER(C_INT), VALUE :: n
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_selector_collect(ctx, stride, thr_ppm, thr_weno, seed, counts) &
      BIND(C, NAME='cumc3d_selector_collect')
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: stride, seed
    REAL(C_DOUBLE), VALUE :: thr_ppm, thr_weno
    INTEGER(C_INT), INTENT(OUT), OPTIONAL :: counts(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_selector_train(ctx, iters, blocks, lr, class_weight, seed, loss) &
      BIND(C, NAME='cumc3d_selector_train')
    IMPORT :: C_INT, C_PTR, C_DOUBLE, C_FLOAT
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: iters, blocks, seed
    REAL(C_DOUBLE), VALUE :: lr
    REAL(C_DOUBLE), INTENT(IN), OPTIONAL :: class_weight(*)
    REAL(C_FLOAT), INTENT(OUT), OPTIONAL :: loss(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_selector_enable(ctx, enable, every) BIND(C, NAME='cumc3d_selector_enable')
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: enable, every
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_selector_mlp(ctx, features, n, logits) BIND(C, NAME='cumc3d_selector_mlp')
    IMPORT :: C_INT, C_PTR, C_FLOAT
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_FLOAT), INTENT(IN) :: features(*)
    INTEGER(C_INT), VALUE :: n
    REAL(C_FLOAT), INTENT(OUT) :: logits(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc3d_selector_mask(ctx, dir, refresh, out) BIND(C, NAME='cumc3d_selector_mask')
    IMPORT :: C_INT, C_PTR, C_SIGNED_CHAR
    TYPE(C_PTR), VALUE :: ctx
    INTEGER(C_INT), VALUE :: dir, refresh
    INTEGER(C_SIGNED_CHAR), INTENT(OUT) :: out(*)
  END FUNCTION
  INTEGER(C_SIZE_T) FUNCTION c_strlen(s) BIND(C, NAME='strlen')
    IMPORT :: C_SIZE_T, C_PTR
    TYPE(C_PTR), VALUE :: s
  END FUNCTION
END INTERFACE

CONTAINS

! Copy of a NUL-terminated C string into a Fortran string.
FUNCTION cumc3d_fstring(p) RESULT(s)
  TYPE(C_PTR), INTENT(IN) :: p
  CHARACTER(len=:), ALLOCATABLE :: s
  CHARACTER(kind=C_CHAR), POINTER :: c(:)
  INTEGER :: i, n
  IF (.NOT. C_ASSOCIATED(p)) THEN
    s = ''
    RETURN
  END IF
  n = INT(c_strlen(p))
  CALL C_F_POINTER(p, c, [n])
  ALLOCATE(CHARACTER(len=n) :: s)
  DO i = 1, n
    s(i:i) = c(i)
  END DO
END FUNCTION

! Stop with the library's message when a call fails.
SUBROUTINE cumc3d_check(rc, where)
  INTEGER(C_INT), INTENT(IN) :: rc
  CHARACTER(len=*), INTENT(IN) :: where
  IF (rc /= CUMC3D_OK) THEN
    WRITE(*, '(A,A,A,I0,A,A,A,A)') 'cumc3d_hip: ', where, ' failed (', rc, ': ', &
        cumc3d_fstring(cumc3d_error_string(rc)), ') ', cumc3d_fstring(cumc3d_last_error())
    STOP 'cumc3d_hip error'
  END IF
END SUBROUTINE

END MODULE CUMC3D_HIP
