! File-fed initial model for the reference oracle (test infrastructure only).
! Plays the role of a model's GET_MODEL hook (model/riemann/src/model.f90:4-17):
! it fills prim(imin:imax,-2:nx+3,-2:ny+3,-2:nz+3), epsilon and ggas from a
! binary file written by cumc3d_amd.ic, so the reference and the MI355X
! library start from bit-identical states.
SUBROUTINE GET_MODEL
USE DEFINITION
USE HARNESS_IO
IMPLICIT NONE
INTEGER :: hdr(6), u
REAL*8 :: scal(2)
OPEN(NEWUNIT=u, FILE=TRIM(ic_path), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
READ(u) hdr
IF (hdr(3) /= nx .OR. hdr(4) /= ny .OR. hdr(5) /= nz) STOP 'IC grid does not match compiled grid'
READ(u) scal
ggas = scal(1)
total_time = scal(2)
READ(u) prim
READ(u) epsilon
CLOSE(u)
END SUBROUTINE GET_MODEL
