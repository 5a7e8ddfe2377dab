! Fortran bindings of libcumc3d_io.so (include/cumc3d_io.h) through
! ISO_C_BINDING: the HDF5 profile writer and the list-directed div B record.
MODULE CUMC3D_IO
USE ISO_C_BINDING
IMPLICIT NONE

TYPE, BIND(C) :: cumc3d_profile_meta
  INTEGER(C_INT) :: nx, ny, nz
  INTEGER(C_INT) :: n_dim
  INTEGER(C_INT) :: coordinate_flag
  REAL(C_DOUBLE) :: time
  TYPE(C_PTR) :: xF, yF, zF
END TYPE

INTERFACE
  INTEGER(C_INT) FUNCTION cumc3d_io_write_profile(path, meta, primitive, epsilon, bfield) &
      BIND(C, NAME='cumc3d_io_write_profile')
    IMPORT :: C_INT, C_CHAR, C_DOUBLE, cumc3d_profile_meta
    CHARACTER(KIND=C_CHAR), INTENT(IN) :: path(*)
    TYPE(cumc3d_profile_meta), INTENT(IN) :: meta
    REAL(C_DOUBLE), INTENT(IN) :: primitive(*), epsilon(*), bfield(*)
  END FUNCTION
  TYPE(C_PTR) FUNCTION cumc3d_io_last_error() BIND(C, NAME='cumc3d_io_last_error')
    IMPORT :: C_PTR
  END FUNCTION
END INTERFACE

CONTAINS

! STOP with the writer's message on a non-zero status.
SUBROUTINE cumc3d_io_check(rc, what)
  INTEGER(C_INT), INTENT(IN) :: rc
  CHARACTER(*), INTENT(IN) :: what
  CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
  INTEGER :: n
  IF (rc == 0) RETURN
  CALL C_F_POINTER(cumc3d_io_last_error(), msg, [512])
  n = 0
  DO WHILE (n < 512)
    IF (msg(n + 1) == C_NULL_CHAR) EXIT
    n = n + 1
  END DO
  WRITE(*,*) what, ': ', msg(1:n)
  STOP 'cumc3d_io error'
END SUBROUTINE

END MODULE CUMC3D_IO
