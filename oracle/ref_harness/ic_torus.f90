! File-fed initial model for the torus-cyn and torus-sph hooks (test
! infrastructure only).  Replaces model/torus-cyn/src/model.f90:6-315 and
! model/torus-sph/src/model.f90:6-324 for the oracle build: prim,
! epsilon, ggas and the floor arrays of module CUSTOM_DEF
! (model/torus-cyn/src/def.f90:6-14) are read from a binary file written by
! cumc3d_amd.ic.
SUBROUTINE GET_MODEL
USE DEFINITION
USE CUSTOM_DEF
USE HARNESS_IO
IMPLICIT NONE
INTEGER :: hdr(6), u
REAL*8 :: scal(2)
OPEN(NEWUNIT=u, FILE=TRIM(ic_path), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
READ(u) hdr
IF (hdr(3) /= nx .OR. hdr(4) /= ny .OR. hdr(5) /= nz) STOP 'IC grid does not match compiled grid'
IF (hdr(6) /= 1) STOP 'torus IC file carries no floor arrays'
READ(u) scal
ggas = scal(1)
total_time = scal(2)
READ(u) prim
READ(u) epsilon
READ(u) rho_floor
READ(u) p_floor
READ(u) eps_floor
CLOSE(u)
END SUBROUTINE GET_MODEL
