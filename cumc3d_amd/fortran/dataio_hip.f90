! Drop-in replacement of solver/mhd/dataio.f90: print_hydroprofile with the
! same name and module-global interface, writing the same HDF5 file
! (./outfile/rkiter-<n_iter>-nm.hdf5, dataio.f90:31-34) through
! libcumc3d_io.so, so the solver builds without HDF5's Fortran modules.
!
! While the state is resident on the MI355X (after the drop-in RUNGEKUTTA /
! FINDDT attached it) the three arrays are packed on the device in
! print_hydroprofile's order by cumc3d_profile and copied once; before that
! (the call after INITIAL_UPDATE, main.f90:71) the module's own arrays are
! written, as array sections exactly as dataio.f90 passes them.
SUBROUTINE print_hydroprofile
USE ISO_C_BINDING
USE DEFINITION
USE CUMC3D_HIP
USE CUMC3D_HIP_STATE, ONLY : hip_ctx, hip_attached
USE CUMC3D_IO
IMPLICIT NONE
CHARACTER(len=99) :: globalt
CHARACTER(len=256) :: filename
TYPE(cumc3d_profile_meta) :: m
REAL(C_DOUBLE), ALLOCATABLE, TARGET :: faces(:)
REAL(C_DOUBLE), ALLOCATABLE :: pp(:), pe(:), pb(:)
INTEGER :: o2, o3

! dataio.f90:30-34
write(globalt,'(I0)') n_iter
filename = './outfile/rkiter-'// trim(adjustl(globalt)) //'-nm.hdf5'

ALLOCATE(faces(nx + ny + nz + 21))
o2 = nx + 8
o3 = o2 + ny + 7
faces(1:o2-1) = xF(-3:nx+3)
faces(o2:o3-1) = yF(-3:ny+3)
faces(o3:) = zF(-3:nz+3)
m%nx = nx
m%ny = ny
m%nz = nz
m%n_dim = n_dim
m%coordinate_flag = coordinate_flag
m%time = global_time
m%xF = C_LOC(faces(1))
m%yF = C_LOC(faces(o2))
m%zF = C_LOC(faces(o3))

IF (hip_attached) THEN
  ALLOCATE(pp(5*(nx+2)*(ny+2)*(nz+2)), pe((nx+2)*(ny+2)*(nz+2)), pb(3*(nx+3)*(ny+3)*(nz+3)))
  CALL cumc3d_check(cumc3d_profile(hip_ctx, pp, pe, pb), 'cumc3d_profile')
  CALL cumc3d_io_check(cumc3d_io_write_profile(TRIM(filename)//C_NULL_CHAR, m, pp, pe, pb), filename)
  DEALLOCATE(pp, pe, pb)
ELSE
  CALL cumc3d_io_check(cumc3d_io_write_profile(TRIM(filename)//C_NULL_CHAR, m, &
       prim(imin:ibx-1,0:nx+1,0:ny+1,0:nz+1), epsilon(0:nx+1,0:ny+1,0:nz+1), &
       prim(ibx:ibz,-1:nx+1,-1:ny+1,-1:nz+1)), filename)
END IF
DEALLOCATE(faces)
END SUBROUTINE
