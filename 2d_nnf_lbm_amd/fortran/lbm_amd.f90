! iso_c_binding interface of liblbm_amd.so (C ABI: include/lbm_amd.h).
!
! A Fortran driver of taha-re/2D_NNF_LBM calls the GPU time step through
! this module: one BIND(C) interface per C entry point, the lbm_params record
! as a BIND(C) derived type, and LBM_CHECK, which turns a nonzero status into
! STOP as the reference does for its only error (F6:115-118).
MODULE LBM_AMD
  USE, INTRINSIC :: ISO_C_BINDING
  IMPLICIT NONE

  INTEGER, PARAMETER :: LBM_PF_COUNT = 22, LBM_H_PCOUNT = 12
  INTEGER, PARAMETER :: LBM_NEWTONIAN = 0, LBM_BINGHAM = 1, LBM_CASSON = 2
  INTEGER, PARAMETER :: LBM_RATE_LAGGED = 0, LBM_RATE_LOCAL = 1
  INTEGER, PARAMETER :: LBM_MOTION_FREE = 0, LBM_MOTION_OSC_X = 1
  INTEGER, PARAMETER :: LBM_BOUNCE_BACK = 0, LBM_ZOU_HE = 1, LBM_EQUILIBRIUM = 2
  INTEGER, PARAMETER :: LBM_CONTACT_GLOWINSKI = 0, LBM_CONTACT_WAN_TUREK = 1, LBM_CONTACT_LJ = 2
  INTEGER, PARAMETER :: LBM_FP64 = 0, LBM_FP32 = 1, LBM_FP16 = 2, LBM_BF16 = 3
  ! particle record fields, 1-based (LBM_PF_* + 1)
  INTEGER, PARAMETER :: PF_X = 1, PF_Y = 2, PF_ANG = 3, PF_VX = 4, PF_VY = 5, PF_W = 6
  INTEGER, PARAMETER :: PF_DIAM = 7, PF_FX = 8, PF_FY = 9, PF_TQ = 10, PF_FCX = 11
  INTEGER, PARAMETER :: PF_FCY = 12, PF_RHOP = 13, PF_MP = 14, PF_IP = 15, PF_ZETA = 16
  INTEGER, PARAMETER :: PF_VOX = 17, PF_VOY = 18, PF_VOW = 19, PF_FHX = 20, PF_FHY = 21
  INTEGER, PARAMETER :: PF_FHW = 22
  ! history record fields, 1-based (LBM_H_* + 1)
  INTEGER, PARAMETER :: H_X = 1, H_Y = 2, H_ANG = 3, H_VX = 4, H_VY = 5, H_W = 6
  INTEGER, PARAMETER :: H_FX = 7, H_FY = 8, H_TQ = 9, H_FHX = 10, H_FHY = 11, H_FHW = 12

  TYPE, BIND(C) :: LBM_PARAMS
    INTEGER(C_INT32_T) :: NX, NY, RHEOLOGY, RATE_MODE, Q_MODE, UPDATE_ANGLE
    INTEGER(C_INT32_T) :: STORAGE, MAX_PARTICLES, MOTION, BOUNDARY
    REAL(C_DOUBLE) :: OMEGA, S0, S1, S2, S3, S5, SQ, NU, MU0, TAU_Y, M_PAP, GAMMA_MIN
    REAL(C_DOUBLE) :: RHO0, GY
    REAL(C_DOUBLE) :: WALL(8), CORNER(2)
    REAL(C_DOUBLE) :: ZETA_C, CIJ, EPS_P, EPS_W, E_P, E_W
    REAL(C_DOUBLE) :: OSC_U, OSC_PERIOD
    INTEGER(C_INT32_T) :: CONTACT, RESERVED
    REAL(C_DOUBLE) :: WALL_EDGE(4)
  END TYPE LBM_PARAMS

  INTERFACE
    FUNCTION lbm_create(p, h) BIND(C, NAME='lbm_create') RESULT(rc)
      IMPORT :: LBM_PARAMS, C_PTR, C_INT
      TYPE(LBM_PARAMS), INTENT(IN) :: p
      TYPE(C_PTR), INTENT(OUT) :: h
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_last_error(h) BIND(C, NAME='lbm_last_error') RESULT(s)
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: h
      TYPE(C_PTR) :: s
    END FUNCTION
    FUNCTION lbm_destroy(h) BIND(C, NAME='lbm_destroy') RESULT(rc)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_sync(h) BIND(C, NAME='lbm_sync') RESULT(rc)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_initialize(h) BIND(C, NAME='lbm_initialize') RESULT(rc)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_set_f(h, f) BIND(C, NAME='lbm_set_f') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(IN) :: f(*)
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_get_f(h, f) BIND(C, NAME='lbm_get_f') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(OUT) :: f(*)
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_set_fpost(h, f) BIND(C, NAME='lbm_set_fpost') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(IN) :: f(*)
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_get_fpost(h, f) BIND(C, NAME='lbm_get_fpost') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(OUT) :: f(*)
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_set_fields(h, rho, ux, uy, g, omega_bn) BIND(C, NAME='lbm_set_fields') RESULT(rc)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h, rho, ux, uy, g, omega_bn
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_get_fields(h, rho, ux, uy, g, omega_bn) BIND(C, NAME='lbm_get_fields') RESULT(rc)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h, rho, ux, uy, g, omega_bn
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_get_stress(h, stress) BIND(C, NAME='lbm_get_stress') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(OUT) :: stress(*)
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_get_image(h, image) BIND(C, NAME='lbm_get_image') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_INT32_T
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), INTENT(OUT) :: image(*)
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_set_particles(h, n, part) BIND(C, NAME='lbm_set_particles') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_INT32_T, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: n
      REAL(C_DOUBLE), INTENT(IN) :: part(*)
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_get_particles(h, part) BIND(C, NAME='lbm_get_particles') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(OUT) :: part(*)
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_get_phi(h, ip, phi) BIND(C, NAME='lbm_get_phi') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_INT32_T, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: ip
      REAL(C_DOUBLE), INTENT(OUT) :: phi(*)
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_collide(h) BIND(C, NAME='lbm_collide') RESULT(rc)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_copy_post(h) BIND(C, NAME='lbm_copy_post') RESULT(rc)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_stream(h) BIND(C, NAME='lbm_stream') RESULT(rc)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_boundary(h, kk) BIND(C, NAME='lbm_boundary') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_INT32_T
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: kk
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_macros(h) BIND(C, NAME='lbm_macros') RESULT(rc)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_spm_phi(h) BIND(C, NAME='lbm_spm_phi') RESULT(rc)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_spm_forcing(h) BIND(C, NAME='lbm_spm_forcing') RESULT(rc)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_hydro_force(h) BIND(C, NAME='lbm_hydro_force') RESULT(rc)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_correct(h) BIND(C, NAME='lbm_correct') RESULT(rc)
      IMPORT :: C_PTR, C_INT
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_contact_force(h, pairs, max_pairs, npairs) BIND(C, NAME='lbm_contact_force') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_INT32_T
      TYPE(C_PTR), VALUE :: h, pairs
      INTEGER(C_INT32_T), VALUE :: max_pairs
      TYPE(C_PTR), VALUE :: npairs
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_particle_update(h, kk) BIND(C, NAME='lbm_particle_update') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_INT32_T
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: kk
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_total_mass(h, mass) BIND(C, NAME='lbm_total_mass') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      REAL(C_DOUBLE), INTENT(OUT) :: mass
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_step(h, kk0, nsteps, hist) BIND(C, NAME='lbm_step') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_INT32_T
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: kk0, nsteps
      TYPE(C_PTR), VALUE :: hist
      INTEGER(C_INT) :: rc
    END FUNCTION
    ! lbm_step with the history staged in the library's static device
    ! array: allocates nothing (for callers that cannot allocate)
    FUNCTION lbm_step_hist_static(h, kk0, nsteps, hist) BIND(C, NAME='lbm_step_hist_static') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_INT32_T
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: kk0, nsteps
      TYPE(C_PTR), VALUE :: hist
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_last_pairs(h, pairs, max_pairs, npairs) BIND(C, NAME='lbm_last_pairs') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_INT32_T
      TYPE(C_PTR), VALUE :: h, pairs
      INTEGER(C_INT32_T), VALUE :: max_pairs
      TYPE(C_PTR), VALUE :: npairs
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_step_timed(h, kk0, nsteps, ms_total, ms_main) BIND(C, NAME='lbm_step_timed') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_INT32_T, C_DOUBLE
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT32_T), VALUE :: kk0, nsteps
      REAL(C_DOUBLE), INTENT(OUT) :: ms_total, ms_main
      INTEGER(C_INT) :: rc
    END FUNCTION
    ! restart files of WRITE_OUTPUT / READ_INPUT (F6:377-405, F6:174-198);
    ! dir is a NUL-terminated path: TRIM(DIR)//C_NULL_CHAR
    FUNCTION lbm_restart_write(h, dir) BIND(C, NAME='lbm_restart_write') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_CHAR
      TYPE(C_PTR), VALUE :: h
      CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: dir
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_restart_read(h, dir) BIND(C, NAME='lbm_restart_read') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_CHAR
      TYPE(C_PTR), VALUE :: h
      CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: dir
      INTEGER(C_INT) :: rc
    END FUNCTION
    ! binary snapshot of the run state in a caller buffer of `bytes` bytes
    FUNCTION lbm_snapshot_size(h, bytes) BIND(C, NAME='lbm_snapshot_size') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_INT64_T
      TYPE(C_PTR), VALUE :: h
      INTEGER(C_INT64_T), INTENT(OUT) :: bytes
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_snapshot_save(h, buf, bytes) BIND(C, NAME='lbm_snapshot_save') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_INT64_T
      TYPE(C_PTR), VALUE :: h, buf
      INTEGER(C_INT64_T), VALUE :: bytes
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION lbm_snapshot_load(h, buf, bytes) BIND(C, NAME='lbm_snapshot_load') RESULT(rc)
      IMPORT :: C_PTR, C_INT, C_INT64_T
      TYPE(C_PTR), VALUE :: h, buf
      INTEGER(C_INT64_T), VALUE :: bytes
      INTEGER(C_INT) :: rc
    END FUNCTION
    FUNCTION c_strlen(s) BIND(C, NAME='strlen') RESULT(n)
      IMPORT :: C_PTR, C_SIZE_T
      TYPE(C_PTR), VALUE :: s
      INTEGER(C_SIZE_T) :: n
    END FUNCTION
  END INTERFACE

CONTAINS

  ! Nonzero status -> print the library's message and STOP (the reference's
  ! only error handling, F6:115-118).
  SUBROUTINE LBM_CHECK(h, rc, what)
    TYPE(C_PTR), INTENT(IN) :: h
    INTEGER(C_INT), INTENT(IN) :: rc
    CHARACTER(LEN=*), INTENT(IN) :: what
    CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
    TYPE(C_PTR) :: s
    INTEGER :: n
    IF (rc == 0) RETURN
    PRINT *, 'lbm_amd: ', what, ' failed with status ', rc
    IF (C_ASSOCIATED(h)) THEN
      s = lbm_last_error(h)
      n = INT(c_strlen(s))
      CALL C_F_POINTER(s, msg, [n])
      PRINT *, 'lbm_amd: ', msg(1:n)
    END IF
    STOP 1
  END SUBROUTINE LBM_CHECK

END MODULE LBM_AMD
