! Harness-side I/O for the reference oracle build (test infrastructure only).
! Holds the paths the harness driver hands to the file-fed GET_MODEL.
MODULE HARNESS_IO
IMPLICIT NONE
CHARACTER(len=1024) :: ic_path = ''
END MODULE HARNESS_IO
