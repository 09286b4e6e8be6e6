! Known-answer driver for the reference's Wan-Turek and Lennard-Jones contact
! forces.
!
! Compiled by oracle/build_ref.py together with two pieces of a reference
! program, read where they lie: its MODULE PARAM and its SUBROUTINE
! ComputeCollisionForce(DISK,WALL), either from
! 4_Oscillating_Particle/D2Q9_V7_OscillatingCylinder_Validated.f90 ("F4",
! F4:2-89 and F4:576-743) or from 1_Disk_Newtonian/D2Q9_1P_Validated_20190530.f90
! ("F1", F1:2-93 and F1:596-769), with the routine's IF(0)/IF(1) model
! switches set to the model under test.  This program is the only new code.
!
! Input (stdin, list-directed): the number of cases M, then per case the
! four wall edges WALL%EDGE(1:4) (bottom y, right x, top y, left x), N, and
! N lines "X Y DIAM".
! Output: NX NY of the module, then per case N lines "F_COL(1) F_COL(2)"
! after one call with N_PARTICLE = N (F_COL zeroed before the call).
PROGRAM CONTACT_MODEL_DRIVER
  USE PARAM
  IMPLICIT NONE
  TYPE(PARTICLE), ALLOCATABLE :: DISK(:)
  TYPE(WALL_PARTICLE), SAVE :: WALL
  INTEGER :: M, C, N, IP

  READ(*, *) M
  WRITE(*, *) NX, NY
  DO C = 1, M
    READ(*, *) WALL%EDGE(1), WALL%EDGE(2), WALL%EDGE(3), WALL%EDGE(4)
    READ(*, *) N
    ALLOCATE(DISK(N))
    DO IP = 1, N
      READ(*, *) DISK(IP)%CENTER(1), DISK(IP)%CENTER(2), DISK(IP)%DIAM
      DISK(IP)%F_COL = 0D0
    END DO
    N_PARTICLE = N
    CALL ComputeCollisionForce(DISK, WALL)
    DO IP = 1, N
      WRITE(*, *) DISK(IP)%F_COL(1), DISK(IP)%F_COL(2)
    END DO
    DEALLOCATE(DISK)
  END DO
END PROGRAM CONTACT_MODEL_DRIVER
