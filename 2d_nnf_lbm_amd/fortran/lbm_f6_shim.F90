! Drop-in replacements for the time-step subroutines of the reference's
! settling-disk program (6_FreeFall_YieldStress/D2Q9_BinghamSediment_
! Validated_20191127.f90, "F6"), with the reference's names and argument
! lists.  Compiled together with the program's own MODULE PARAM and MAIN (and
! without the original subroutines of the same names), the unchanged MAIN
! runs its time step on the GPU through liblbm_amd:
!
!   Collision_MRT (F6:867), STREAMING (F6:1154), BOUNDARYCONDITION (F6:1195),
!   MACROSCOPICVARIABLES (F6:1394), ComputeConcenterationFunction (F6:430),
!   ComputeForceField_SPM (F6:477), ComputeForce_HYDRO (F6:534),
!   VelocityCorrector (F6:1111), ComputeCollisionForce (F6:584).
!
! The arrays stay the caller's host arrays: each routine uploads what it
! reads, runs the GPU kernel of the same routine and downloads what it
! writes, so every other part of MAIN (the inline particle update, SUM(RHO),
! the output writers) sees the reference's data at every point of the loop.
! That costs host-device copies per call; the fused lbm_step is the fast path.
!
! The handle is created on the first call from MODULE PARAM: grid, OMEGA, NU,
! MU_0, TAU_Y, M_PAP, RHO0, A_G, the contact constants of F6:691-693, and the
! relaxation rates read back from the program's own ComputeRelaxation_MRT
! (so a program built with another rate set gets that set).  Bingham runs
! keep the reference's lagged module-global rate (F6:958) unless compiled
! with -DLBM_SHIM_LOCAL_RATE; the top-wall velocity UXN, local to F6's
! BOUNDARYCONDITION, is LBM_SHIM_UXN (default 0).
#ifndef LBM_SHIM_UXN
#define LBM_SHIM_UXN 0D0
#endif

MODULE LBM_F6_SHIM
  USE, INTRINSIC :: ISO_C_BINDING
  USE LBM_AMD
  USE PARAM
  IMPLICIT NONE
  TYPE(C_PTR), SAVE :: SH = C_NULL_PTR
  DOUBLE PRECISION, ALLOCATABLE, SAVE :: SPART(:, :)

CONTAINS

  SUBROUTINE SHIM_START(DISK)
    TYPE(PARTICLE), INTENT(IN) :: DISK(1:N_PARTICLE)
    TYPE(LBM_PARAMS) :: P
    DOUBLE PRECISION :: SAVED(0:8), S1(0:8), S2(0:8)
    INTEGER(C_INT) :: RC
    IF (C_ASSOCIATED(SH)) RETURN
    ! the program's rate set: s(4) follows omega when it is 8(2-w)/(8-w)
    SAVED = OMEGA_MRT
    CALL ComputeRelaxation_MRT(OMEGA)
    S1 = OMEGA_MRT
    CALL ComputeRelaxation_MRT(0.5D0 * OMEGA)
    S2 = OMEGA_MRT
    OMEGA_MRT = SAVED
    P%NX = NX
    P%NY = NY
    P%RHEOLOGY = MERGE(LBM_BINGHAM, LBM_NEWTONIAN, NON_NEWTONIAN)
#ifdef LBM_SHIM_LOCAL_RATE
    P%RATE_MODE = LBM_RATE_LOCAL
#else
    P%RATE_MODE = LBM_RATE_LAGGED
#endif
    IF (S1(4) /= S2(4)) THEN
      P%Q_MODE = 1
      P%SQ = 0D0
    ELSE
      P%Q_MODE = 0
      P%SQ = S1(4)
    END IF
    P%S0 = S1(0)
    P%S1 = S1(1)
    P%S2 = S1(2)
    P%S3 = S1(3)
    P%S5 = S1(5)
    P%UPDATE_ANGLE = 1
    P%STORAGE = LBM_FP64
    P%MAX_PARTICLES = N_PARTICLE
    P%MOTION = LBM_MOTION_FREE
    P%BOUNDARY = LBM_BOUNCE_BACK
    P%OMEGA = OMEGA
    P%NU = NU
    P%MU0 = MU_0
    P%TAU_Y = TAU_Y
    P%M_PAP = M_PAP
    P%GAMMA_MIN = 1.0E-15                                          ! F6:939
    P%RHO0 = RHO0
    P%GY = A_G
    P%WALL = 0D0
    P%WALL(5) = LBM_SHIM_UXN
    P%CORNER = 0D0
    P%ZETA_C = 1.0                                                 ! F6:691-693
    P%CIJ = -(RHO_S - RHO_F) * A_G * PI * 0.5 * DISK(1)%DIAM * 0.5 * DISK(1)%DIAM
    P%EPS_P = 0.5
    P%EPS_W = P%EPS_P
    P%E_P = 10.0
    P%E_W = P%E_P
    P%OSC_U = 0D0
    P%OSC_PERIOD = 0D0
    P%CONTACT = LBM_CONTACT_GLOWINSKI
    P%RESERVED = 0
    P%WALL_EDGE = 0D0
    RC = lbm_create(P, SH)
    CALL LBM_CHECK(SH, RC, 'lbm_create')
    RC = lbm_initialize(SH)
    CALL LBM_CHECK(SH, RC, 'lbm_initialize')
    ALLOCATE(SPART(N_PARTICLE, LBM_PF_COUNT))
    SPART = 0D0
  END SUBROUTINE SHIM_START

  ! TYPE(PARTICLE) records -> the library's particle block (field-major)
  SUBROUTINE SHIM_PUT_PARTICLES(DISK)
    TYPE(PARTICLE), INTENT(IN) :: DISK(1:N_PARTICLE)
    INTEGER :: IP
    INTEGER(C_INT) :: RC
    DO IP = 1, N_PARTICLE
      SPART(IP, PF_X) = DISK(IP)%CENTER(1)
      SPART(IP, PF_Y) = DISK(IP)%CENTER(2)
      SPART(IP, PF_ANG) = DISK(IP)%ANGLE
      SPART(IP, PF_VX) = DISK(IP)%VELOCITY(1)
      SPART(IP, PF_VY) = DISK(IP)%VELOCITY(2)
      SPART(IP, PF_W) = DISK(IP)%VELOCITY(3)
      SPART(IP, PF_DIAM) = DISK(IP)%DIAM
      SPART(IP, PF_FX) = DISK(IP)%FORCE(1)
      SPART(IP, PF_FY) = DISK(IP)%FORCE(2)
      SPART(IP, PF_TQ) = DISK(IP)%FORCE(3)
      SPART(IP, PF_FCX) = DISK(IP)%F_COL(1)
      SPART(IP, PF_FCY) = DISK(IP)%F_COL(2)
      SPART(IP, PF_RHOP) = DISK(IP)%RHO_P
      SPART(IP, PF_MP) = DISK(IP)%M_P
      SPART(IP, PF_IP) = DISK(IP)%I_P
      SPART(IP, PF_ZETA) = DISK(IP)%ZITA
    END DO
    RC = lbm_set_particles(SH, INT(N_PARTICLE, C_INT32_T), SPART)
    CALL LBM_CHECK(SH, RC, 'lbm_set_particles')
  END SUBROUTINE SHIM_PUT_PARTICLES

  SUBROUTINE SHIM_GET_PARTICLES()
    INTEGER(C_INT) :: RC
    RC = lbm_get_particles(SH, SPART)
    CALL LBM_CHECK(SH, RC, 'lbm_get_particles')
  END SUBROUTINE SHIM_GET_PARTICLES

END MODULE LBM_F6_SHIM

SUBROUTINE ComputeCollisionForce(DISK)
  USE LBM_F6_SHIM
  IMPLICIT NONE
  TYPE(PARTICLE) :: DISK(1:N_PARTICLE)
  INTEGER :: IP
  INTEGER(C_INT) :: RC
  CALL SHIM_START(DISK)
  CALL SHIM_PUT_PARTICLES(DISK)
  RC = lbm_contact_force(SH, C_NULL_PTR, 0_C_INT32_T, C_NULL_PTR)
  CALL LBM_CHECK(SH, RC, 'ComputeCollisionForce')
  CALL SHIM_GET_PARTICLES()
  DO IP = 1, N_PARTICLE
    DISK(IP)%F_COL(1) = SPART(IP, PF_FCX)
    DISK(IP)%F_COL(2) = SPART(IP, PF_FCY)
  END DO
END SUBROUTINE ComputeCollisionForce

SUBROUTINE ComputeConcenterationFunction(DISK, DOMAIN_IMAGE, IS_SOLID_NODE, IS_PARTICLE_NODE, &
                                         IS_WALL_NODE)
  USE LBM_F6_SHIM
  IMPLICIT NONE
  TYPE(PARTICLE) :: DISK(1:N_PARTICLE)
  INTEGER :: DOMAIN_IMAGE(0:NX, 0:NY)
  LOGICAL :: IS_SOLID_NODE(0:NX, 0:NY), IS_WALL_NODE(0:NX, 0:NY), IS_PARTICLE_NODE(0:NX, 0:NY)
  INTEGER :: IP
  INTEGER(C_INT) :: RC
  CALL SHIM_START(DISK)
  CALL SHIM_PUT_PARTICLES(DISK)
  RC = lbm_spm_phi(SH)
  CALL LBM_CHECK(SH, RC, 'ComputeConcenterationFunction')
  IS_SOLID_NODE = .FALSE.
  IS_WALL_NODE = .FALSE.
  IS_PARTICLE_NODE = .FALSE.
  DO IP = 1, N_PARTICLE
    RC = lbm_get_phi(SH, INT(IP - 1, C_INT32_T), DISK(IP)%PHI)
    CALL LBM_CHECK(SH, RC, 'lbm_get_phi')
    WHERE (DISK(IP)%PHI >= 0.5) IS_PARTICLE_NODE = .TRUE.
  END DO
  RC = lbm_get_image(SH, DOMAIN_IMAGE)
  CALL LBM_CHECK(SH, RC, 'lbm_get_image')
END SUBROUTINE ComputeConcenterationFunction

SUBROUTINE Collision_MRT(UX, UY, RHO, F, FEQ, G, Omega_BN, Domain_Image, STRESS)
  USE LBM_F6_SHIM
  IMPLICIT NONE
  DOUBLE PRECISION, TARGET :: RHO(0:NX, 0:NY), UX(0:NX, 0:NY), UY(0:NX, 0:NY)
  DOUBLE PRECISION, TARGET :: F(0:8, 0:NX, 0:NY), FEQ(0:8, 0:NX, 0:NY), G(1:2, 0:NX, 0:NY)
  DOUBLE PRECISION, TARGET :: Omega_BN(0:NX, 0:NY), STRESS(1:3, 0:NX, 0:NY)
  INTEGER :: Domain_Image(0:NX, 0:NY)
  INTEGER(C_INT) :: RC
  IF (.NOT. C_ASSOCIATED(SH)) STOP 'Collision_MRT before the first particle call'
  RC = lbm_set_f(SH, F)
  CALL LBM_CHECK(SH, RC, 'lbm_set_f')
  RC = lbm_set_fields(SH, C_LOC(RHO), C_LOC(UX), C_LOC(UY), C_LOC(G), C_LOC(Omega_BN))
  CALL LBM_CHECK(SH, RC, 'lbm_set_fields')
  RC = lbm_collide(SH)
  CALL LBM_CHECK(SH, RC, 'Collision_MRT')
  RC = lbm_get_f(SH, F)
  CALL LBM_CHECK(SH, RC, 'lbm_get_f')
  RC = lbm_get_stress(SH, STRESS)
  CALL LBM_CHECK(SH, RC, 'lbm_get_stress')
  IF (NON_NEWTONIAN) THEN
    RC = lbm_get_fields(SH, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_LOC(Omega_BN))
    CALL LBM_CHECK(SH, RC, 'lbm_get_fields')
    RC = lbm_get_image(SH, Domain_Image)
    CALL LBM_CHECK(SH, RC, 'lbm_get_image')
    ! the module-global rate left by the last node (F6:958)
    CALL ComputeRelaxation_MRT(Omega_BN(NX, NY))
  END IF
END SUBROUTINE Collision_MRT

SUBROUTINE STREAMING(F, F_POST, IS_SOLID_NODE)
  USE LBM_F6_SHIM
  IMPLICIT NONE
  DOUBLE PRECISION :: F_POST(0:8, 0:NX, 0:NY), F(0:8, 0:NX, 0:NY)
  LOGICAL :: IS_SOLID_NODE(0:NX, 0:NY)
  INTEGER(C_INT) :: RC
  RC = lbm_set_f(SH, F)
  CALL LBM_CHECK(SH, RC, 'lbm_set_f')
  RC = lbm_set_fpost(SH, F_POST)
  CALL LBM_CHECK(SH, RC, 'lbm_set_fpost')
  RC = lbm_stream(SH)
  CALL LBM_CHECK(SH, RC, 'STREAMING')
  RC = lbm_get_f(SH, F)
  CALL LBM_CHECK(SH, RC, 'lbm_get_f')
END SUBROUTINE STREAMING

SUBROUTINE BOUNDARYCONDITION(F, F_POST, RHO, UX, UY, KK)
  USE LBM_F6_SHIM
  IMPLICIT NONE
  DOUBLE PRECISION, TARGET :: F(0:8, 0:NX, 0:NY), F_POST(0:8, 0:NX, 0:NY)
  DOUBLE PRECISION, TARGET :: RHO(0:NX, 0:NY), UX(0:NX, 0:NY), UY(0:NX, 0:NY)
  INTEGER :: KK
  INTEGER(C_INT) :: RC
  RC = lbm_set_f(SH, F)
  CALL LBM_CHECK(SH, RC, 'lbm_set_f')
  RC = lbm_set_fpost(SH, F_POST)
  CALL LBM_CHECK(SH, RC, 'lbm_set_fpost')
  RC = lbm_set_fields(SH, C_LOC(RHO), C_LOC(UX), C_LOC(UY), C_NULL_PTR, C_NULL_PTR)
  CALL LBM_CHECK(SH, RC, 'lbm_set_fields')
  RC = lbm_boundary(SH, INT(KK, C_INT32_T))
  CALL LBM_CHECK(SH, RC, 'BOUNDARYCONDITION')
  RC = lbm_get_f(SH, F)
  CALL LBM_CHECK(SH, RC, 'lbm_get_f')
END SUBROUTINE BOUNDARYCONDITION

SUBROUTINE MACROSCOPICVARIABLES(F, RHO, UX, UY)
  USE LBM_F6_SHIM
  IMPLICIT NONE
  DOUBLE PRECISION, TARGET :: F(0:8, 0:NX, 0:NY)
  DOUBLE PRECISION, TARGET :: RHO(0:NX, 0:NY), UX(0:NX, 0:NY), UY(0:NX, 0:NY)
  INTEGER(C_INT) :: RC
  RC = lbm_set_f(SH, F)
  CALL LBM_CHECK(SH, RC, 'lbm_set_f')
  RC = lbm_macros(SH)
  CALL LBM_CHECK(SH, RC, 'MACROSCOPICVARIABLES')
  RC = lbm_get_fields(SH, C_LOC(RHO), C_LOC(UX), C_LOC(UY), C_NULL_PTR, C_NULL_PTR)
  CALL LBM_CHECK(SH, RC, 'lbm_get_fields')
END SUBROUTINE MACROSCOPICVARIABLES

SUBROUTINE ComputeForceField_SPM(G, DISK, UX, UY)
  USE LBM_F6_SHIM
  IMPLICIT NONE
  DOUBLE PRECISION, TARGET :: G(1:2, 0:NX, 0:NY), UX(0:NX, 0:NY), UY(0:NX, 0:NY)
  TYPE(PARTICLE) :: DISK(1:N_PARTICLE)
  INTEGER(C_INT) :: RC
  CALL SHIM_PUT_PARTICLES(DISK)
  RC = lbm_set_fields(SH, C_NULL_PTR, C_LOC(UX), C_LOC(UY), C_NULL_PTR, C_NULL_PTR)
  CALL LBM_CHECK(SH, RC, 'lbm_set_fields')
  RC = lbm_spm_forcing(SH)
  CALL LBM_CHECK(SH, RC, 'ComputeForceField_SPM')
  RC = lbm_get_fields(SH, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_LOC(G), C_NULL_PTR)
  CALL LBM_CHECK(SH, RC, 'lbm_get_fields')
END SUBROUTINE ComputeForceField_SPM

SUBROUTINE ComputeForce_HYDRO(DISK, RHO, UX, UY, G)
  USE LBM_F6_SHIM
  IMPLICIT NONE
  TYPE(PARTICLE) :: DISK(1:N_PARTICLE)
  DOUBLE PRECISION, TARGET :: G(1:2, 0:NX, 0:NY), RHO(0:NX, 0:NY), UX(0:NX, 0:NY), UY(0:NX, 0:NY)
  INTEGER :: IP
  INTEGER(C_INT) :: RC
  CALL SHIM_PUT_PARTICLES(DISK)
  RC = lbm_set_fields(SH, C_LOC(RHO), C_LOC(UX), C_LOC(UY), C_LOC(G), C_NULL_PTR)
  CALL LBM_CHECK(SH, RC, 'lbm_set_fields')
  RC = lbm_hydro_force(SH)
  CALL LBM_CHECK(SH, RC, 'ComputeForce_HYDRO')
  CALL SHIM_GET_PARTICLES()
  DO IP = 1, N_PARTICLE
    DISK(IP)%FORCE(1) = SPART(IP, PF_FX)
    DISK(IP)%FORCE(2) = SPART(IP, PF_FY)
    DISK(IP)%FORCE(3) = SPART(IP, PF_TQ)
  END DO
END SUBROUTINE ComputeForce_HYDRO

SUBROUTINE VelocityCorrector(UX, UY, G)
  USE LBM_F6_SHIM
  IMPLICIT NONE
  DOUBLE PRECISION, TARGET :: UX(0:NX, 0:NY), UY(0:NX, 0:NY), G(1:2, 0:NX, 0:NY)
  INTEGER(C_INT) :: RC
  RC = lbm_set_fields(SH, C_NULL_PTR, C_LOC(UX), C_LOC(UY), C_LOC(G), C_NULL_PTR)
  CALL LBM_CHECK(SH, RC, 'lbm_set_fields')
  RC = lbm_correct(SH)
  CALL LBM_CHECK(SH, RC, 'VelocityCorrector')
  RC = lbm_get_fields(SH, C_NULL_PTR, C_LOC(UX), C_LOC(UY), C_NULL_PTR, C_NULL_PTR)
  CALL LBM_CHECK(SH, RC, 'lbm_get_fields')
END SUBROUTINE VelocityCorrector
