! fwgpu_bind.F90 -- ISO_C_BINDING interface to libfwgpu.so (include/fwgpu.h).
!
! Thin, reference-independent layer: one bind(C) interface per C entry
! point, and the fw_config derived type with the exact layout of the C
! struct.  Host arrays are passed as C pointers (c_loc of the Fortran
! arrays); they keep the reference layout (column-major, ghosts included).
MODULE FWGPU_BIND
  USE, INTRINSIC :: ISO_C_BINDING
  IMPLICIT NONE

  INTEGER(C_INT), PARAMETER :: FW_OK = 0, FW_E_INVALID = -1, FW_E_HIP = -2, &
       FW_E_NOMEM = -3, FW_E_BLOWUP = -4, FW_E_NOTFOUND = -5, FW_E_BREAK_ITER = -6
  INTEGER(C_INT), PARAMETER :: FW_WK_NONE = 0, FW_WK_REG = 1, FW_WK_DATA2D = 2
  INTEGER(C_INT), PARAMETER :: FW_OPT_FULL_STATE = 1

  TYPE, BIND(C) :: fw_config
    INTEGER(C_INT) :: real_bytes
    INTEGER(C_INT) :: Mloc, Nloc, Mglob, Nglob, Ibeg, Iend, Jbeg, Jend
    INTEGER(C_INT) :: DISPERSION, PERIODIC, VISCOSITY_BREAKING, SHOW_BREAKING
    INTEGER(C_INT) :: DIFFUSION_SPONGE, FRICTION_SPONGE, DIRECT_SPONGE, ETA_LIMITER, FIXED_DT
    INTEGER(C_INT) :: wavemaker, NumFreq, ilo, ihi, jlo, jhi
    REAL(C_DOUBLE) :: DX, DY, DXg, DYg, CFL, MinDepth, MinDepthFrc
    REAL(C_DOUBLE) :: Gamma1, Gamma2, Gamma3, Beta_1, Beta_2, b1, b2
    REAL(C_DOUBLE) :: SWE_ETA_DEP, FroudeCap, Cbrk1, Cbrk2, T_brk, WAVEMAKER_Cbrk, nu_bkg
    REAL(C_DOUBLE) :: Xc_WK, Yc_WK, Width_WK, Ywidth_WK, EtaBlowVal, DT_fixed
    REAL(C_DOUBLE) :: CrestLimit, TroughLimit
    REAL(C_DOUBLE) :: D_gen, Beta_gen, rlamda, Tperiod, Time_ramp, FreqPeak
  END TYPE fw_config

  INTERFACE
    INTEGER(C_INT) FUNCTION fw_create(cfg, ctx) BIND(C, NAME='fw_create')
      IMPORT :: C_INT, C_PTR, fw_config
      TYPE(fw_config), INTENT(IN) :: cfg
      TYPE(C_PTR), INTENT(OUT) :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_destroy(ctx) BIND(C, NAME='fw_destroy')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    TYPE(C_PTR) FUNCTION fw_strerror(ctx) BIND(C, NAME='fw_strerror')
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_field_id(name) BIND(C, NAME='fw_field_id')
      IMPORT :: C_INT, C_CHAR
      CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: name
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_upload(ctx, field, host) BIND(C, NAME='fw_upload')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, host
      INTEGER(C_INT), VALUE :: field
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_download(ctx, field, host) BIND(C, NAME='fw_download')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, host
      INTEGER(C_INT), VALUE :: field
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_set_clock(ctx, dt, time) BIND(C, NAME='fw_set_clock')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), VALUE :: dt, time
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_get_clock(ctx, dt, time) BIND(C, NAME='fw_get_clock')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: dt, time
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_exchange(ctx) BIND(C, NAME='fw_exchange')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_estimate_dt(ctx, dt, time) BIND(C, NAME='fw_estimate_dt')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: dt, time
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_save_stage0(ctx, which) BIND(C, NAME='fw_save_stage0')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: which
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_cal_dispersion(ctx) BIND(C, NAME='fw_cal_dispersion')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_fluxes(ctx) BIND(C, NAME='fw_fluxes')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_source_terms(ctx) BIND(C, NAME='fw_source_terms')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_estimate_huv(ctx, istage) BIND(C, NAME='fw_estimate_huv')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: istage
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_update_mask(ctx) BIND(C, NAME='fw_update_mask')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_wave_breaking(ctx) BIND(C, NAME='fw_wave_breaking')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_sponge_damping(ctx) BIND(C, NAME='fw_sponge_damping')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_check_blowup(ctx, maxeta) BIND(C, NAME='fw_check_blowup')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      REAL(C_DOUBLE), INTENT(OUT) :: maxeta
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_step(ctx, nsteps, dt, time) BIND(C, NAME='fw_step')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: nsteps
      REAL(C_DOUBLE), INTENT(OUT) :: dt, time
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_synchronize(ctx) BIND(C, NAME='fw_synchronize')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
    END FUNCTION
    INTEGER(C_INT) FUNCTION fw_set_option(ctx, option, value) BIND(C, NAME='fw_set_option')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT), VALUE :: option, value
    END FUNCTION
  END INTERFACE

CONTAINS

  ! field id of a reference array name
  INTEGER FUNCTION fwgpu_id(name)
    CHARACTER(LEN=*), INTENT(IN) :: name
    fwgpu_id = fw_field_id(TRIM(name)//C_NULL_CHAR)
  END FUNCTION fwgpu_id

  ! message of the last error of a context
  FUNCTION fwgpu_message(ctx) RESULT(msg)
    TYPE(C_PTR), INTENT(IN) :: ctx
    CHARACTER(LEN=256) :: msg
    CHARACTER(KIND=C_CHAR), POINTER :: p(:)
    INTEGER :: k
    msg = ' '
    CALL C_F_POINTER(fw_strerror(ctx), p, [256])
    DO k = 1, 256
      IF (p(k) == C_NULL_CHAR) EXIT
      msg(k:k) = p(k)
    ENDDO
  END FUNCTION fwgpu_message

END MODULE FWGPU_BIND
