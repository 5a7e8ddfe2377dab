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
! param.h), -DCUMC3D_TORUS for model/torus-cyn; link libcumc3d_hip.so.
MODULE CUMC3D_HIP_STATE
USE ISO_C_BINDING
USE CUMC3D_HIP
IMPLICIT NONE
TYPE(C_PTR), SAVE :: hip_ctx = C_NULL_PTR
LOGICAL, SAVE :: hip_attached = .false.
REAL(C_DOUBLE), ALLOCATABLE, TARGET, SAVE :: hip_grid(:)

CONTAINS

! cumc3d_init + cumc3d_upload + cumc3d_initial_update from the solver's
! module state (setup.f90:340-395 allocations, parameter.h flags).
SUBROUTINE CUMC3D_HIP_ATTACH
  USE DEFINITION
#ifdef CUMC3D_TORUS
  USE CUSTOM_DEF
#endif
  IMPLICIT NONE
  INCLUDE "param.h"
  TYPE(cumc3d_desc) :: d
  REAL(C_DOUBLE), ALLOCATABLE :: floors(:,:,:,:)
  INTEGER :: o1, o2
  IF (hip_attached) RETURN
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
  d%precision = CUMC3D_FP64
  d%cfl = cfl
  d%ggas = ggas
#ifdef CUMC3D_TORUS
  d%model = CUMC3D_MODEL_TORUS
  d%rt_gravity = 0
  d%r_sh = r_sh
#else
  d%model = CUMC3D_MODEL_RIEMANN
  ! model/riemann/src/custom.f90:143 adds the RT gravity for these tests
  d%rt_gravity = MERGE(1, 0, test_model == 6 .OR. test_model == 18)
  d%r_sh = 0.0D0
#endif
  ! face coordinates xF(-3:n+3) etc. (setup.f90:363-365), n+7 values each
  o1 = nx + 7
  o2 = o1 + ny + 7
  ALLOCATE(hip_grid(o2 + nz + 7))
  hip_grid(1:o1) = xF(-3:nx+3)
  hip_grid(o1+1:o2) = yF(-3:ny+3)
  hip_grid(o2+1:) = zF(-3:nz+3)
  d%xF = C_LOC(hip_grid(1))
  d%yF = C_LOC(hip_grid(o1+1))
  d%zF = C_LOC(hip_grid(o2+1))
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
  hip_attached = .true.
END SUBROUTINE

! prim, epsilon, bcell and cs of the host modules from the device state.
SUBROUTINE CUMC3D_HIP_SYNC_HOST
  USE DEFINITION
  USE MHD_MODULE, ONLY : bcell
  IMPLICIT NONE
  IF (.NOT. hip_attached) RETURN
  CALL cumc3d_check(cumc3d_download(hip_ctx, prim, epsilon, bcell, cs), 'cumc3d_download')
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
