! File-fed initial model for the collapsar hooks (test infrastructure only).
! Replaces model/collapsar/src/model.f90:6-252 for the oracle build: prim,
! epsilon and ggas come from a binary file written by cumc3d_amd.ic, followed
! by the CUSTOM_DEF values GET_MODEL sets (model.f90:44-52,238-245):
! prim_a(1:8), eps_a, m_bh and vel2code.  The Poisson coefficients are formed
! by the reference's own get_poisson (model/collapsar/src/poisson.f90), as
! GET_MODEL does at model.f90:40.
SUBROUTINE GET_MODEL
USE DEFINITION
USE CUSTOM_DEF
USE HARNESS_IO
IMPLICIT NONE
INTEGER :: hdr(6), u
REAL*8 :: scal(2), grav(11)
OPEN(NEWUNIT=u, FILE=TRIM(ic_path), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
READ(u) hdr
IF (hdr(3) /= nx .OR. hdr(4) /= ny .OR. hdr(5) /= nz) STOP 'IC grid does not match compiled grid'
IF (hdr(6) /= 2) STOP 'collapsar IC file carries no gravity record'
READ(u) scal
ggas = scal(1)
total_time = scal(2)
READ(u) prim
READ(u) epsilon
READ(u) grav
CLOSE(u)
prim_a(1:8) = grav(1:8)
eps_a = grav(9)
m_bh = grav(10)
vel2code = grav(11)
m_inj = 0.0D0
CALL get_poisson
END SUBROUTINE GET_MODEL
