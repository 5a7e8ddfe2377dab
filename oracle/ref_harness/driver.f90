! Harness driver for the reference oracle (test infrastructure only).
! Runs the reference's own step sequence exactly as solver/mhd/main.f90:50-139
! does it (INITIAL_MODEL, INITIAL_UPDATE, then FINDDT / dt clip / RUNGEKUTTA),
! without the HDF5 writer, and dumps the final state as raw binary.
!
! Built with -DHIPDROP it links the drop-in RUNGEKUTTA/FINDDT of
! cumc3d_amd/fortran/rungekutta_hip.f90 instead of the reference's and
! refreshes the host state from the device before FIND_DIVB and the dump; it
! also writes ./outfile/rkiter-0-nm.hdf5 after INITIAL_UPDATE (from the host
! arrays) and ./outfile/rkiter-1-nm.hdf5 after the last step (packed on the
! device) through the print_hydroprofile of cumc3d_amd/fortran/dataio_hip.f90.
!
! It also writes OUT_FILE.grid, the GETGRID arrays the step reads
! (getgrid.f90:26-98, after CUSTOM_GRID and the FIXPOLE override), and with
! -DCUMC3D_COLLAPSAR OUT_FILE.grav: m_bh, r_sh and the potential phi of
! module CUSTOM_DEF (model/collapsar/src/def.f90:16-36) after the last step.
! It then runs SPATIAL once more on the final state (with -DHIPDROP the
! drop-in's CUMC3D_HIP_SPATIAL on the device) and writes l_rk to OUT_FILE.lrk.
!
! usage: cumc3d_ref IC_FILE OUT_FILE MAX_STEPS [THREADS]
PROGRAM REF_HARNESS
USE DEFINITION
USE MHD_MODULE
USE HARNESS_IO
#ifdef CUMC3D_COLLAPSAR
USE CUSTOM_DEF, ONLY : phi, m_bh, r_sh
#endif
#ifdef HIPDROP
USE CUMC3D_HIP_STATE, ONLY : CUMC3D_HIP_SYNC_HOST, CUMC3D_HIP_SPATIAL, CUMC3D_HIP_DETACH
#endif
IMPLICIT NONE
CHARACTER(len=1024) :: out_path, arg
INTEGER :: max_steps, nthreads, u, cnt, c0, c1, crate
REAL*8, ALLOCATABLE :: dts(:)
REAL*8 :: loop_seconds
INTEGER, EXTERNAL :: OMP_GET_NUM_PROCS

CALL GET_COMMAND_ARGUMENT(1, ic_path)
CALL GET_COMMAND_ARGUMENT(2, out_path)
CALL GET_COMMAND_ARGUMENT(3, arg)
READ(arg, *) max_steps
nthreads = OMP_GET_NUM_PROCS()
IF (COMMAND_ARGUMENT_COUNT() >= 4) THEN
  CALL GET_COMMAND_ARGUMENT(4, arg)
  READ(arg, *) nthreads
END IF
CALL OMP_SET_NUM_THREADS(nthreads)

n_iter = 0
n_step = 0
global_time = 0.0D0
CALL INITIAL_MODEL
CALL INITIAL_UPDATE
#ifdef HIPDROP
n_iter = 0
CALL print_hydroprofile
#endif

OPEN(NEWUNIT=u, FILE=TRIM(out_path)//'.grid', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE', &
     ACTION='WRITE')
WRITE(u) nx, ny, nz
WRITE(u) xF(-3:nx+3), yF(-3:ny+3), zF(-3:nz+3)
WRITE(u) dx(-2:nx+3), dy(-2:ny+3), dz(-2:nz+3), x(-2:nx+3), y(-2:ny+3), z(-2:nz+3)
WRITE(u) dx_cb(-1:nx+3), xbar(-1:nx+3), dcose(-1:ny+3), dsine(-1:ny+3), sine(-2:ny+3), sinf(-3:ny+3)
CLOSE(u)

ALLOCATE(dts(MAX(max_steps,1)))
cnt = 0
CALL SYSTEM_CLOCK(c0, crate)
DO WHILE (global_time < total_time .AND. n_step < max_steps)
  CALL FINDDT
  IF (global_time + dt > total_time) THEN
    dt = total_time - global_time
  END IF
  CALL RUNGEKUTTA
  n_step = n_step + 1
  global_time = global_time + dt
  cnt = cnt + 1
  dts(cnt) = dt
END DO
CALL SYSTEM_CLOCK(c1)
loop_seconds = DBLE(c1 - c0)/DBLE(crate)
#ifdef HIPDROP
n_iter = 1
CALL print_hydroprofile
CALL CUMC3D_HIP_SYNC_HOST
#endif

OPEN(UNIT=999, FILE=TRIM(out_path)//'.divb', STATUS='REPLACE')
CALL FIND_DIVB
CLOSE(999)

OPEN(NEWUNIT=u, FILE=TRIM(out_path), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE', ACTION='WRITE')
WRITE(u) nx, ny, nz, n_step, nthreads
WRITE(u) global_time, maxDivB, loop_seconds
WRITE(u) prim
WRITE(u) epsilon
WRITE(u) bcell
WRITE(u) cs
WRITE(u) cons
WRITE(u) dts(1:cnt)
CLOSE(u)
! SPATIAL once more, on the final state (spatial.f90:11-162), for its l_rk
! on its own: after the dump, so nothing written above depends on it.
#ifdef HIPDROP
CALL CUMC3D_HIP_SPATIAL
#else
CALL SPATIAL
#endif
OPEN(NEWUNIT=u, FILE=TRIM(out_path)//'.lrk', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE', &
     ACTION='WRITE')
WRITE(u) l_rk
CLOSE(u)
#ifdef CUMC3D_COLLAPSAR
OPEN(NEWUNIT=u, FILE=TRIM(out_path)//'.grav', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE', &
     ACTION='WRITE')
WRITE(u) m_bh, r_sh
WRITE(u) phi
CLOSE(u)
#endif
#ifdef HIPDROP
CALL CUMC3D_HIP_DETACH
#endif
END PROGRAM REF_HARNESS
