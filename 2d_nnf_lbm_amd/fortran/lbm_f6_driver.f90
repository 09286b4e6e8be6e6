! Fortran driver of the settling-disk Bingham case (the MAIN program of
! 6_FreeFall_YieldStress/D2Q9_BinghamSediment_Validated_20191127.f90, "F6")
! running its time step on the GPU through the LBM_AMD binding.
!
! The run constants are evaluated from the same float32 literals as F6's
! MODULE PARAM (F6:3-66); the loop either calls the per-subroutine entries in
! F6's order (FUSED=.FALSE., F6:293-349) or the fused device step
! (FUSED=.TRUE.).  Output files and their list-directed formats follow F6:
! 01_ParticleTimehistory.plt (F6:315-318), 01_AveragedDensity.PLT
! (F6:348-349) and, with WRITE_OUTPUT, the state dump of F6:389-403 plus the
! field file of WriteResults (F6:1416-1463).
!
! With READ_INPUT the run starts from the state files of F6:174-198 in the
! working directory (as written by WRITE_OUTPUT).
!
! Usage: lbm_f6_driver [namelist-file]   (namelist /RUN/ below)
PROGRAM LBM_F6_DRIVER
  USE LBM_AMD
  IMPLICIT NONE
  INTEGER :: NX = 80, NY = 300, TIME_MAX = 1000, STORAGE = 0
  LOGICAL :: NON_NEWTONIAN = .TRUE., WRITE_OUTPUT = .TRUE., FUSED = .TRUE., READ_INPUT = .FALSE.
  DOUBLE PRECISION :: UXN = 0D0
  ! variants: RHEOLOGY (-1: from NON_NEWTONIAN; LBM_CASSON for the
  ! Papanastasiou-Casson lines F6:941, F6:944), BOUNDARY (LBM_BOUNCE_BACK,
  ! LBM_ZOU_HE of F1, LBM_EQUILIBRIUM of F4), CONTACT (LBM_CONTACT_*, with the
  ! constants of F4's blocks and the domain as the Wan-Turek wall box)
  INTEGER :: RHEOLOGY = -1, BOUNDARY = 0, CONTACT = 0
  NAMELIST /RUN/ NX, NY, TIME_MAX, NON_NEWTONIAN, WRITE_OUTPUT, FUSED, STORAGE, UXN, READ_INPUT, &
                 RHEOLOGY, BOUNDARY, CONTACT

  DOUBLE PRECISION :: PI, RE_S, BN_S, Y_G, RHO_S, RHO_F, NU, L_C, A_G, U_C, TAU_Y
  DOUBLE PRECISION :: MU_0, M_PAP, RHO0, OMEGA, AVGDEN, MASS, M_P, I_P, D
  DOUBLE PRECISION :: START, FINISH, FX_PREV, FY_PREV, VY, OMEGA_MRT(0:8), OML
  DOUBLE PRECISION, ALLOCATABLE, TARGET :: HIST(:, :)
  DOUBLE PRECISION, ALLOCATABLE, TARGET :: F(:, :, :), RHO(:, :), UX(:, :), UY(:, :)
  DOUBLE PRECISION, ALLOCATABLE, TARGET :: G(:, :, :), OMEGA_BN(:, :), STRESS(:, :, :)
  INTEGER(C_INT32_T), ALLOCATABLE :: IMAGE(:, :)
  DOUBLE PRECISION :: PART(LBM_PF_COUNT, 1)
  TYPE(LBM_PARAMS) :: P
  TYPE(C_PTR) :: H
  INTEGER(C_INT) :: RC
  INTEGER :: KK, I, J, NARG
  CHARACTER(LEN=512) :: ARG, FNAME

  NARG = COMMAND_ARGUMENT_COUNT()
  IF (NARG >= 1) THEN
    CALL GET_COMMAND_ARGUMENT(1, ARG)
    OPEN(30, FILE=TRIM(ARG), STATUS='OLD')
    READ(30, NML=RUN)
    CLOSE(30)
  END IF

  ! MODULE PARAM, F6:3-66 (default-REAL literals are float32, widened)
  PI = 3.14159265359
  RE_S = 0.78
  BN_S = 0.033
  Y_G = BN_S
  RHO_S = 2.0
  RHO_F = 1.0
  NU = 0.025
  L_C = DBLE(NX / 4)
  A_G = -RE_S * NU * NU / ((L_C**3) * (RHO_S - RHO_F) * RHO_F)
  U_C = (L_C**2) * (RHO_S - RHO_F) * (-A_G) / NU
  TAU_Y = Y_G * (RHO_S - RHO_F) * (-A_G) * L_C
  MU_0 = NU
  M_PAP = 5.5E6
  RHO0 = 1.0
  OMEGA = 1.0 / (3. * NU + 0.5)                                  ! F6:101
  IF (1.0 / OMEGA < 0.5) THEN                                    ! F6:115-118
    PRINT *, 'RELAXATION TIME EXCEEDED 0.5 -> STABILITY PROBLEM'
    STOP
  END IF

  OPEN(12, FILE='00_Simulation_Parameters.txt')                  ! F6:104-107
  WRITE(12, *) '**********************'
  WRITE(12, *) 'Lattice Parameter'
  WRITE(12, *) 'TAU_LB=', 1.0 / OMEGA
  WRITE(12, *) 'U_LB_max=', U_C
  WRITE(12, *) 'Tau_Y_LB=', TAU_Y
  WRITE(12, *) 'A_G=', A_G
  WRITE(12, *) 'MAX_ITR =', TIME_MAX
  CLOSE(12)

  P%NX = NX
  P%NY = NY
  P%RHEOLOGY = MERGE(LBM_BINGHAM, LBM_NEWTONIAN, NON_NEWTONIAN)
  IF (RHEOLOGY >= 0) P%RHEOLOGY = RHEOLOGY
  P%RATE_MODE = LBM_RATE_LOCAL
  P%Q_MODE = 1
  P%UPDATE_ANGLE = 1
  P%STORAGE = STORAGE
  P%MAX_PARTICLES = 1
  P%MOTION = LBM_MOTION_FREE
  P%BOUNDARY = BOUNDARY
  P%OSC_U = 0D0
  P%OSC_PERIOD = 0D0
  P%CONTACT = CONTACT
  P%RESERVED = 0
  P%WALL_EDGE = (/ 0D0, DBLE(NX), DBLE(NY), 0D0 /)
  P%OMEGA = OMEGA
  P%S0 = 0.0                                                     ! F6:1009-1011
  P%S1 = 1.1
  P%S2 = 1.25
  P%S3 = 0.0
  P%S5 = 0.0
  P%SQ = 0.0
  P%NU = NU
  P%MU0 = MU_0
  P%TAU_Y = TAU_Y
  P%M_PAP = M_PAP
  P%GAMMA_MIN = 1.0E-15
  P%RHO0 = RHO0
  P%GY = A_G
  P%WALL = 0D0
  P%WALL(5) = UXN
  P%CORNER = 0D0
  P%ZETA_C = 1.0                                                 ! F6:691-693
  P%CIJ = -(RHO_S - RHO_F) * A_G * PI * 0.5 * L_C * 0.5 * L_C
  P%EPS_P = 0.5
  P%EPS_W = P%EPS_P
  P%E_P = 10.0
  P%E_W = P%E_P
  IF (CONTACT == LBM_CONTACT_WAN_TUREK) THEN                    ! F4:625
    P%ZETA_C = 1.0
    P%EPS_P = 1.0
    P%EPS_W = 2.0 * P%EPS_P
  ELSE IF (CONTACT == LBM_CONTACT_LJ) THEN                       ! F4:588
    P%ZETA_C = 1.0
  END IF

  RC = lbm_create(P, H)
  CALL LBM_CHECK(H, RC, 'lbm_create')
  CALL CPU_TIME(START)
  RC = lbm_initialize(H)
  CALL LBM_CHECK(H, RC, 'lbm_initialize')

  ! the disk of F6:161-169
  PART = 0D0
  D = L_C
  PART(PF_X, 1) = 0.5 * DBLE(NX)
  PART(PF_Y, 1) = DBLE(NY) - 7.0 * L_C
  PART(PF_DIAM, 1) = D
  PART(PF_ZETA, 1) = 1.0
  PART(PF_RHOP, 1) = RHO_S
  M_P = RHO_S * PI * (0.5 * D) * (0.5 * D)
  I_P = 0.5 * M_P * (0.5 * D) * (0.5 * D)
  PART(PF_MP, 1) = M_P
  PART(PF_IP, 1) = I_P
  RC = lbm_set_particles(H, 1, PART)
  CALL LBM_CHECK(H, RC, 'lbm_set_particles')
  ! READ_INPUT (F6:174-198): state files of an earlier run in the working
  ! directory replace the initial state
  IF (READ_INPUT) THEN
    RC = lbm_restart_read(H, '.' // C_NULL_CHAR)
    CALL LBM_CHECK(H, RC, 'READ_INPUT')
    PRINT *, 'Reading Initial Data Was Succsessful...'
  END IF

  OPEN(8, FILE='01_AveragedDensity.PLT')
  WRITE(8, *) 'VARIABLES=Iteration,Density_Averaged'
  OPEN(11, FILE='01_ParticleTimehistory.plt')
  WRITE(11, *) 'VARIABLES=T,Y,X,ANG,Ux,Uy,OMG,Fx,BnT,Cd,ReT'

  IF (FUSED) THEN
    ALLOCATE(HIST(LBM_H_PCOUNT + 2, TIME_MAX))
    RC = lbm_step(H, 0, TIME_MAX, C_LOC(HIST))
    CALL LBM_CHECK(H, RC, 'lbm_step')
    FX_PREV = 0D0
    FY_PREV = 0D0
    DO KK = 1, TIME_MAX
      VY = HIST(H_VY, KK)
      WRITE(11, *) KK * NU / L_C**2, HIST(H_Y, KK), (HIST(H_X, KK)), HIST(H_ANG, KK), &
        & HIST(H_VX, KK), VY / U_C, HIST(H_W, KK), &
        & (FX_PREV + HIST(H_FHX, KK)), TAU_Y * L_C / (NU * (-VY)), (FY_PREV + HIST(H_FHY, KK)) / (NU * (-VY)), &
        & -VY * L_C / NU
      FX_PREV = HIST(H_FX, KK)
      FY_PREV = HIST(H_FY, KK)
      AVGDEN = HIST(LBM_H_PCOUNT + 1, KK) / DBLE((NX + 1) * (NY + 1))
      WRITE(8, *) KK, AVGDEN
    END DO
  ELSE
    DO KK = 1, TIME_MAX
      RC = lbm_contact_force(H, C_NULL_PTR, 0, C_NULL_PTR)
      CALL LBM_CHECK(H, RC, 'ComputeCollisionForce')
      RC = lbm_particle_update(H, KK)
      CALL LBM_CHECK(H, RC, 'particle update')
      RC = lbm_get_particles(H, PART)
      CALL LBM_CHECK(H, RC, 'lbm_get_particles')
      VY = PART(PF_VY, 1)
      WRITE(11, *) KK * NU / L_C**2, PART(PF_Y, 1), (PART(PF_X, 1)), PART(PF_ANG, 1), &
        & PART(PF_VX, 1), VY / U_C, PART(PF_W, 1), &
        & (PART(PF_FX, 1) + PART(PF_FHX, 1)), TAU_Y * L_C / (NU * (-VY)), &
        & (PART(PF_FY, 1) + PART(PF_FHY, 1)) / (NU * (-VY)), -VY * L_C / NU
      RC = lbm_spm_phi(H)
      CALL LBM_CHECK(H, RC, 'ComputeConcenterationFunction')
      RC = lbm_collide(H)
      CALL LBM_CHECK(H, RC, 'Collision_MRT')
      RC = lbm_copy_post(H)
      CALL LBM_CHECK(H, RC, 'F_POST=F')
      RC = lbm_stream(H)
      CALL LBM_CHECK(H, RC, 'STREAMING')
      RC = lbm_boundary(H, KK)
      CALL LBM_CHECK(H, RC, 'BOUNDARYCONDITION')
      RC = lbm_macros(H)
      CALL LBM_CHECK(H, RC, 'MACROSCOPICVARIABLES')
      RC = lbm_spm_forcing(H)
      CALL LBM_CHECK(H, RC, 'ComputeForceField_SPM')
      RC = lbm_hydro_force(H)
      CALL LBM_CHECK(H, RC, 'ComputeForce_HYDRO')
      RC = lbm_correct(H)
      CALL LBM_CHECK(H, RC, 'VelocityCorrector')
      RC = lbm_total_mass(H, MASS)
      CALL LBM_CHECK(H, RC, 'SUM(RHO)')
      AVGDEN = MASS / DBLE((NX + 1) * (NY + 1))
      WRITE(8, *) KK, AVGDEN
    END DO
  END IF
  CLOSE(8)
  CLOSE(11)

  IF (WRITE_OUTPUT) THEN
    ALLOCATE(F(0:8, 0:NX, 0:NY), RHO(0:NX, 0:NY), UX(0:NX, 0:NY), UY(0:NX, 0:NY))
    ALLOCATE(G(1:2, 0:NX, 0:NY), OMEGA_BN(0:NX, 0:NY), STRESS(1:3, 0:NX, 0:NY))
    ALLOCATE(IMAGE(0:NX, 0:NY))
    RC = lbm_get_f(H, F)
    CALL LBM_CHECK(H, RC, 'lbm_get_f')
    RC = lbm_get_fields(H, C_LOC(RHO), C_LOC(UX), C_LOC(UY), C_LOC(G), C_LOC(OMEGA_BN))
    CALL LBM_CHECK(H, RC, 'lbm_get_fields')
    RC = lbm_get_stress(H, STRESS)
    CALL LBM_CHECK(H, RC, 'lbm_get_stress')
    RC = lbm_get_image(H, IMAGE)
    CALL LBM_CHECK(H, RC, 'lbm_get_image')
    RC = lbm_get_particles(H, PART)
    CALL LBM_CHECK(H, RC, 'lbm_get_particles')
    ! the rate vector the last node of the sweep leaves behind (F6:958)
    OML = MERGE(OMEGA_BN(NX, NY), OMEGA, NON_NEWTONIAN)
    OMEGA_MRT = (/ 0D0, DBLE(1.1), 1.25D0, 0D0, 8. * (2. - OML) / (8. - OML), 0D0, &
      & 8. * (2. - OML) / (8. - OML), OML, OML /)
    OPEN(51, FILE='0_F_output.txt')
    OPEN(52, FILE='0_UX_output.txt')
    OPEN(53, FILE='0_UY_output.txt')
    OPEN(54, FILE='0_RHO_output.txt')
    OPEN(55, FILE='0_Gspm_output.txt')
    OPEN(56, FILE='0_OmegaBN_output.txt')
    OPEN(57, FILE='0_Particle_output.txt')
    OPEN(59, FILE='0_OmegaMRT_output.txt')
    OPEN(61, FILE='0_DomainImage_output.txt')
    DO I = 0, NX
      DO J = 0, NY
        WRITE(51, *) F(:, I, J)
        WRITE(52, *) UX(I, J)
        WRITE(53, *) UY(I, J)
        WRITE(54, *) RHO(I, J)
        WRITE(55, *) G(:, I, J)
        WRITE(56, *) OMEGA_BN(I, J)
        WRITE(61, *) IMAGE(I, J)
      END DO
    END DO
    WRITE(57, *) PART(PF_X, 1), PART(PF_Y, 1), PART(PF_VX, 1), PART(PF_VY, 1), PART(PF_W, 1), &
      & PART(PF_FX, 1), PART(PF_FY, 1), PART(PF_TQ, 1)
    WRITE(59, *) OMEGA_MRT
    CLOSE(51); CLOSE(52); CLOSE(53); CLOSE(54); CLOSE(55); CLOSE(56); CLOSE(57)
    CLOSE(59); CLOSE(61)
    ! WriteResults (F6:1416-1463), field file of the last step
    WRITE(FNAME, *) TIME_MAX
    FNAME = ADJUSTL(FNAME)
    OPEN(201, FILE='02_VelocityField_' // TRIM(FNAME) // '.plt')
    WRITE(201, *) 'VARIABLES =X/D, Y/D, UX, UY, RHO,Solid_Nodes,TAU_11,TAU_22,TAU_12'
    WRITE(201, *) 'ZONE ', 'I=', NX + 1, 'J=', NY + 1, ',', 'F=BLOCK'
    DO J = 0, NY
      WRITE(201, *) (DBLE(I), I = 0, NX)
    END DO
    DO J = 0, NY
      WRITE(201, *) (DBLE(J), I = 0, NX)
    END DO
    DO J = 0, NY
      WRITE(201, *) (UX(I, J), I = 0, NX)
    END DO
    DO J = 0, NY
      WRITE(201, *) (UY(I, J), I = 0, NX)
    END DO
    DO J = 0, NY
      WRITE(201, *) (RHO(I, J), I = 0, NX)
    END DO
    DO J = 0, NY
      WRITE(201, *) (DBLE(IMAGE(I, J)), I = 0, NX)
    END DO
    DO J = 0, NY
      WRITE(201, *) (STRESS(1, I, J), I = 0, NX)
    END DO
    DO J = 0, NY
      WRITE(201, *) (STRESS(2, I, J), I = 0, NX)
    END DO
    DO J = 0, NY
      WRITE(201, *) (STRESS(3, I, J), I = 0, NX)
    END DO
    CLOSE(201)
  END IF
  CALL CPU_TIME(FINISH)
  PRINT *, 'RUNTIME=', FINISH - START
  RC = lbm_destroy(H)
END PROGRAM LBM_F6_DRIVER
