! Drop-in replacement of solver/mhd/rungekutta.f90: RUNGEKUTTA and FINDDT with
! the same names and the same (argument-less, module-global) interface, run on
! the MI355X through libcumc3d_hip.so.
!
! The device state is attached on the first call: the solver's module arrays
! after INITIAL_UPDATE (prim, epsilon; for the torus model the floor arrays of
! CUSTOM_DEF) are uploaded once and stay resident; each step exchanges only
! the 8-byte dt.  Before the host reads the state (the output writer,
! FIND_DIVB, a restart dump) it calls CUMC3D_HIP_SYNC_HOST, the counterpart
! of the reference's !$ACC UPDATE HOST (dataio.f90:26).  CUMC3D_HIP_DETACH
! releases the device state (CLEAR_DEVICE, openacc.f90:30-47).
!
! Build: compile with -cpp next to the solver's modules (and the model's
! param.h), with the solver's own -DFIXPOLE, -DCUMC3D_TORUS for
! model/torus-cyn, -DCUMC3D_TORUS -DCUMC3D_TORUS_SPH for model/torus-sph,
! -DCUMC3D_COLLAPSAR for model/collapsar; link libcumc3d_hip.so.
!
! Arithmetic (no counterpart in the reference, whose scheme flags live in
! parameter.h:91-115): the default is the FP64 mode, bit-exact with the
! reference.  -DCUMC3D_USE_FP32 selects the FP32 split (single-precision
! reconstruction, Riemann solver and EMFs, double-precision RK accumulators
! and energy recovery).  -DCUMC3D_USE_SELECTOR adds the learned reconstruction
! selector (with -DCUMC3D_USE_FP32, or alone in the FP64 mode): at attach it
! reads the 1699 single-precision parameters (CUMC3D_SEL_NPARAM, raw
! little-endian float32, e.g. numpy's tofile of cumc3d_selector_get_params)
! from the file named by the environment variable CUMC3D_SELECTOR_PARAMS
! (default ./selector_params.bin) and refreshes the class masks every
! CUMC3D_SELECTOR_EVERY steps (default 8).  The compile-time reconstruction
! flag of parameter.h stays the scheme of the plain path.
!
! The collapsar's CUSTOM_DEF values (prim_a, eps_a, m_bh, vel2code) and its
! param.h Poisson settings go over in cumc3d_gravity_params; the device runs
! CUSTOM_UPDATE itself (accretion, horizon check, the SOR potential), and
! CUMC3D_HIP_SYNC_HOST brings phi, m_bh and r_sh back with the state.
!
! The grid goes over as the module's GETGRID arrays (face coordinates plus
! dx, x, dx_cb, dcose, sinf, ...), so CUSTOM_GRID and FIXPOLE (which moves
! yF after dy and the sines are taken, getgrid.f90:91-98) carry over exactly.
MODULE CUMC3D_HIP_STATE
USE ISO_C_BINDING
USE CUMC3D_HIP
IMPLICIT NONE
TYPE(C_PTR), SAVE :: hip_ctx = C_NULL_PTR
LOGICAL, SAVE :: hip_attached = .false.
REAL(C_DOUBLE), ALLOCATABLE, TARGET, SAVE :: hip_grid(:)
TYPE(cumc3d_grid_metrics), TARGET, SAVE :: hip_metrics
TYPE(cumc3d_gravity_params), TARGET, SAVE :: hip_gravity

CONTAINS

! cumc3d_init + cumc3d_upload + cumc3d_initial_update from the solver's
! module state (setup.f90:340-395 allocations, parameter.h flags).
SUBROUTINE CUMC3D_HIP_ATTACH
  USE DEFINITION
#if defined(CUMC3D_TORUS) || defined(CUMC3D_COLLAPSAR)
  USE CUSTOM_DEF
#endif
  IMPLICIT NONE
  INCLUDE "param.h"
  TYPE(cumc3d_desc) :: d
  REAL(C_DOUBLE), ALLOCATABLE :: floors(:,:,:,:)
  INTEGER :: o(16)
  IF (hip_attached) RETURN
  ! prim, cons and l_rk cross the C-ABI as (8, -2:n+3, ...) blocks
  IF (imin /= 1 .OR. imax /= 8) STOP 'cumc3d drop-in: the state must hold exactly the 8 MHD variables'
  d%nx = nx
  d%ny = ny
  d%nz = nz
  d%n_dim = n_dim
  d%coordinate_flag = coordinate_flag
  d%boundary_flag = boundary_flag
  IF (LF_flag) d%riemann = CUMC3D_LF
  IF (HLL_flag) d%riemann = CUMC3D_HLL
  IF (HLLC_flag) d%riemann = CUMC3D_HLLC
  IF (HLLD_flag) d%riemann = CUMC3D_HLLD
  IF (tvdmm_flag) d%recon = CUMC3D_TVDMM
  IF (tvdvl_flag) d%recon = CUMC3D_TVDVL
  IF (tvdmc_flag) d%recon = CUMC3D_TVDMC
  IF (ppmc_flag) d%recon = CUMC3D_PPM
  IF (weno_flag) d%recon = CUMC3D_WENO
  d%hlld_bn_mode = CUMC3D_BN_CARRY
#ifdef CUMC3D_USE_FP32
  d%precision = CUMC3D_FP32
#else
  d%precision = CUMC3D_FP64
#endif
  d%cfl = cfl
  d%ggas = ggas
  d%gravity = C_NULL_PTR
#if defined(CUMC3D_COLLAPSAR)
  d%model = CUMC3D_MODEL_COLLAPSAR
  d%rt_gravity = 0
  d%r_sh = 0.0D0
  ! CUSTOM_DEF after GET_MODEL (model/collapsar/src/model.f90:44-52,238-245)
  ! and param.h:37-44; omega_weight is def.f90:31
  hip_gravity%prim_a = prim_a(irho:ivz)
  hip_gravity%eps_a = eps_a
  hip_gravity%m_bh = m_bh
  hip_gravity%c_code = clight*vel2code
  hip_gravity%omega = omega_weight
  hip_gravity%tolerance = tolerance
  hip_gravity%relax_max = relax_max
  hip_gravity%n_pot = n_pot
  d%gravity = C_LOC(hip_gravity)
#elif defined(CUMC3D_TORUS_SPH)
  d%model = CUMC3D_MODEL_TORUS_SPH
  d%rt_gravity = 0
  d%r_sh = r_sh
#elif defined(CUMC3D_TORUS)
  d%model = CUMC3D_MODEL_TORUS
  d%rt_gravity = 0
  d%r_sh = r_sh
#else
  d%model = CUMC3D_MODEL_RIEMANN
  ! model/riemann/src/custom.f90:143 adds the RT gravity for these tests
  d%rt_gravity = MERGE(1, 0, test_model == 6 .OR. test_model == 18)
  d%r_sh = 0.0D0
#endif
#ifdef FIXPOLE
  d%fixpole = 1
#else
  d%fixpole = 0
#endif
  ! GETGRID arrays (setup.f90:360-374) copied into one target buffer: faces
  ! xF(-3:n+3) (n+7 values), widths and centres (-2:n+3), the spherical
  ! dx_cb, xbar (-2:nx+3), dcose, dsine, sine (-2:ny+3) and sinf(-3:ny+3)
  o(1) = 1
  o(2) = o(1) + nx + 7
  o(3) = o(2) + ny + 7
  o(4) = o(3) + nz + 7
  o(5) = o(4) + nx + 6
  o(6) = o(5) + ny + 6
  o(7) = o(6) + nz + 6
  o(8) = o(7) + nx + 6
  o(9) = o(8) + ny + 6
  o(10) = o(9) + nz + 6
  o(11) = o(10) + nx + 6
  o(12) = o(11) + nx + 6
  o(13) = o(12) + ny + 6
  o(14) = o(13) + ny + 6
  o(15) = o(14) + ny + 6
  o(16) = o(15) + ny + 7
  ALLOCATE(hip_grid(o(16) - 1))
  hip_grid(o(1):o(2)-1) = xF(-3:nx+3)
  hip_grid(o(2):o(3)-1) = yF(-3:ny+3)
  hip_grid(o(3):o(4)-1) = zF(-3:nz+3)
  hip_grid(o(4):o(5)-1) = dx(-2:nx+3)
  hip_grid(o(5):o(6)-1) = dy(-2:ny+3)
  hip_grid(o(6):o(7)-1) = dz(-2:nz+3)
  hip_grid(o(7):o(8)-1) = x(-2:nx+3)
  hip_grid(o(8):o(9)-1) = y(-2:ny+3)
  hip_grid(o(9):o(10)-1) = z(-2:nz+3)
  hip_grid(o(10):o(11)-1) = dx_cb(-2:nx+3)
  hip_grid(o(11):o(12)-1) = xbar(-2:nx+3)
  hip_grid(o(12):o(13)-1) = dcose(-2:ny+3)
  hip_grid(o(13):o(14)-1) = dsine(-2:ny+3)
  hip_grid(o(14):o(15)-1) = sine(-2:ny+3)
  hip_grid(o(15):o(16)-1) = sinf(-3:ny+3)
  d%xF = C_LOC(hip_grid(o(1)))
  d%yF = C_LOC(hip_grid(o(2)))
  d%zF = C_LOC(hip_grid(o(3)))
  hip_metrics%dx = C_LOC(hip_grid(o(4)))
  hip_metrics%dy = C_LOC(hip_grid(o(5)))
  hip_metrics%dz = C_LOC(hip_grid(o(6)))
  hip_metrics%x = C_LOC(hip_grid(o(7)))
  hip_metrics%y = C_LOC(hip_grid(o(8)))
  hip_metrics%z = C_LOC(hip_grid(o(9)))
  hip_metrics%dx_cb = C_LOC(hip_grid(o(10)))
  hip_metrics%xbar = C_LOC(hip_grid(o(11)))
  hip_metrics%dcose = C_LOC(hip_grid(o(12)))
  hip_metrics%dsine = C_LOC(hip_grid(o(13)))
  hip_metrics%sine = C_LOC(hip_grid(o(14)))
  hip_metrics%sinf = C_LOC(hip_grid(o(15)))
  d%metrics = C_LOC(hip_metrics)
  CALL cumc3d_check(cumc3d_init(d, hip_ctx), 'cumc3d_init')
#ifdef CUMC3D_TORUS
  ALLOCATE(floors(-2:nx+3,-2:ny+3,-2:nz+3,3))
  floors(:,:,:,1) = rho_floor
  floors(:,:,:,2) = p_floor
  floors(:,:,:,3) = eps_floor
  CALL cumc3d_check(cumc3d_upload(hip_ctx, prim, epsilon, floors), 'cumc3d_upload')
  DEALLOCATE(floors)
#else
  CALL cumc3d_check(cumc3d_upload(hip_ctx, prim, epsilon), 'cumc3d_upload')
#endif
  CALL cumc3d_check(cumc3d_initial_update(hip_ctx), 'cumc3d_initial_update')
#ifdef CUMC3D_USE_SELECTOR
  CALL CUMC3D_HIP_SELECTOR
#endif
  hip_attached = .true.
END SUBROUTINE

#ifdef CUMC3D_USE_SELECTOR
! The learned selector's parameters from CUMC3D_SELECTOR_PARAMS (default
! ./selector_params.bin, CUMC3D_SEL_NPARAM float32), refreshed every
! CUMC3D_SELECTOR_EVERY steps (default 8).
SUBROUTINE CUMC3D_HIP_SELECTOR
  IMPLICIT NONE
  INTEGER, PARAMETER :: nparam = 1699
  REAL(C_FLOAT) :: w(nparam)
  CHARACTER(len=1024) :: path, val
  INTEGER :: u, ios, plen, vlen, every
  CALL GET_ENVIRONMENT_VARIABLE('CUMC3D_SELECTOR_PARAMS', path, plen)
  IF (plen == 0) path = 'selector_params.bin'
  every = 8
  CALL GET_ENVIRONMENT_VARIABLE('CUMC3D_SELECTOR_EVERY', val, vlen)
  IF (vlen > 0) READ(val, *) every
  OPEN(NEWUNIT=u, FILE=TRIM(path), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ', IOSTAT=ios)
  IF (ios /= 0) STOP 'cumc3d drop-in: cannot open the selector parameter file (CUMC3D_SELECTOR_PARAMS)'
  READ(u, IOSTAT=ios) w
  CLOSE(u)
  IF (ios /= 0) STOP 'cumc3d drop-in: the selector parameter file holds fewer than 1699 float32 values'
  CALL cumc3d_check(cumc3d_selector_init(hip_ctx, 128, 1), 'cumc3d_selector_init')
  CALL cumc3d_check(cumc3d_selector_set_params(hip_ctx, w, nparam), 'cumc3d_selector_set_params')
  CALL cumc3d_check(cumc3d_selector_enable(hip_ctx, 1, every), 'cumc3d_selector_enable')
END SUBROUTINE
#endif

! prim, epsilon, bcell and cs of the host modules from the device state
! (for the collapsar also phi, m_bh and r_sh of CUSTOM_DEF).
SUBROUTINE CUMC3D_HIP_SYNC_HOST
  USE DEFINITION
  USE MHD_MODULE, ONLY : bcell
#ifdef CUMC3D_COLLAPSAR
  USE CUSTOM_DEF, ONLY : phi, m_bh, r_sh
#endif
  IMPLICIT NONE
  IF (.NOT. hip_attached) RETURN
  CALL cumc3d_check(cumc3d_download(hip_ctx, prim, epsilon, bcell, cs), 'cumc3d_download')
#ifdef CUMC3D_COLLAPSAR
  CALL cumc3d_check(cumc3d_gravity(hip_ctx, phi, m_bh, r_sh), 'cumc3d_gravity')
#endif
END SUBROUTINE

! SPATIAL (spatial.f90:11-162) of the device state into the module's l_rk:
! l_rk(1:5) of cells 1..n and l_rk(6:8) of the faces the update advances
! (the other elements keep their values), for checking the operator against
! the reference's own SPATIAL.  The device state is not changed.
SUBROUTINE CUMC3D_HIP_SPATIAL
  USE DEFINITION, ONLY : l_rk
  IMPLICIT NONE
  CALL CUMC3D_HIP_ATTACH
  CALL cumc3d_check(cumc3d_spatial(hip_ctx, l_rk), 'SPATIAL')
END SUBROUTINE

SUBROUTINE CUMC3D_HIP_DETACH
  IMPLICIT NONE
  IF (.NOT. hip_attached) RETURN
  CALL cumc3d_check(cumc3d_finalize(hip_ctx), 'cumc3d_finalize')
  hip_ctx = C_NULL_PTR
  hip_attached = .false.
  IF (ALLOCATED(hip_grid)) DEALLOCATE(hip_grid)
END SUBROUTINE

END MODULE CUMC3D_HIP_STATE

! SSP-RK3 step with the module's dt (rungekutta.f90:12-208).
SUBROUTINE RUNGEKUTTA
USE DEFINITION, ONLY : dt
USE CUMC3D_HIP_STATE
IMPLICIT NONE
CALL CUMC3D_HIP_ATTACH
CALL cumc3d_check(cumc3d_rk3_step(hip_ctx, dt), 'RUNGEKUTTA')
END SUBROUTINE

! CFL time step into the module's dt (rungekutta.f90:221-332).
SUBROUTINE FINDDT
USE DEFINITION, ONLY : dt
USE CUMC3D_HIP_STATE
IMPLICIT NONE
CALL CUMC3D_HIP_ATTACH
CALL cumc3d_check(cumc3d_finddt(hip_ctx, dt), 'FINDDT')
END SUBROUTINE
