! Known-answer driver for the reference's N-body contact force.
!
! Compiled by oracle/build_ref.py together with two pieces of the reference
! program 3_MultiParticle_Sedimentation/D2Q9_RTParticle_20191003.f90 ("F3"),
! read where they lie: MODULE PARAM (F3:2-83) and SUBROUTINE
! ComputeCollisionForce (F3:632-806).  This program is the only new code.
!
! Input (stdin, list-directed): N, then N lines "X Y DIAM".
! It calls ComputeCollisionForce three times:
!   1. with N_PARTICLE = 1 (no pair term), which leaves the routine's local
!      D_PW at particle 1's left-wall distance;
!   2. with N_PARTICLE = N: the pair branch reads that stale D_PW (F3:760);
!   3. again with N_PARTICLE = N: D_PW is now particle N's left-wall distance.
! The routine is compiled with -fno-automatic, so its local D_PW keeps its
! value from one call to the next.
! Output: A_G RHO_S RHO_F PI NX NY, then 2*N lines "F_COL(1) F_COL(2)" (call
! 2 for particles 1..N, then call 3).
PROGRAM F3_CONTACT_DRIVER
  USE PARAM
  IMPLICIT NONE
  TYPE(PARTICLE), ALLOCATABLE :: DISK(:)
  DOUBLE PRECISION, ALLOCATABLE :: XIN(:), YIN(:), DIN(:), FC(:, :, :)
  INTEGER :: N, IP, K

  READ(*, *) N
  ALLOCATE(DISK(N), XIN(N), YIN(N), DIN(N), FC(2, N, 2))
  DO IP = 1, N
    READ(*, *) XIN(IP), YIN(IP), DIN(IP)
  END DO
  DO IP = 1, N
    DISK(IP)%CENTER(1) = XIN(IP)
    DISK(IP)%CENTER(2) = YIN(IP)
    DISK(IP)%DIAM = DIN(IP)
    DISK(IP)%F_COL = 0D0
  END DO

  N_PARTICLE = 1
  CALL ComputeCollisionForce(DISK)
  N_PARTICLE = N
  CALL ComputeCollisionForce(DISK)
  DO IP = 1, N
    FC(1, IP, 1) = DISK(IP)%F_COL(1)
    FC(2, IP, 1) = DISK(IP)%F_COL(2)
  END DO
  CALL ComputeCollisionForce(DISK)
  DO IP = 1, N
    FC(1, IP, 2) = DISK(IP)%F_COL(1)
    FC(2, IP, 2) = DISK(IP)%F_COL(2)
  END DO

  WRITE(*, *) A_G, RHO_S, RHO_F, PI, NX, NY
  DO K = 1, 2
    DO IP = 1, N
      WRITE(*, *) FC(1, IP, K), FC(2, IP, K)
    END DO
  END DO
END PROGRAM F3_CONTACT_DRIVER
