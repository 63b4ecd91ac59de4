! fwgpu_global.F90 -- drop-in GPU versions of the reference's per-step
! routines, for a reference build (compile against its MODULE PARAM and
! MODULE GLOBAL, src/mod_param.F / src/mod_global.F).
!
! The reference main loop (src/main.F:284-535) calls zero-argument routines
! that work on MODULE GLOBAL.  Each FWGPU_<name> below has the same name
! (prefixed), the same arguments and the same effect, with the state held in
! HBM: FWGPU_INIT uploads GLOBAL once after INITIALIZATION, the per-step
! routines run on the device, and FWGPU_SYNC_HOST copies the device state
! back into GLOBAL before host code (OUTPUT, STATISTICS, PREVIEW) reads it.
MODULE FWGPU_GLOBAL
  USE, INTRINSIC :: ISO_C_BINDING
  USE PARAM
  USE GLOBAL
  USE FWGPU_BIND
  IMPLICIT NONE
  TYPE(C_PTR), SAVE :: fwctx = C_NULL_PTR

CONTAINS

  SUBROUTINE fwgpu_check(rc, where)
    INTEGER(C_INT), INTENT(IN) :: rc
    CHARACTER(LEN=*), INTENT(IN) :: where
    IF (rc /= FW_OK) THEN
      WRITE(*,*) 'FWGPU error in ', where, ': ', TRIM(fwgpu_message(fwctx)), ' (code', rc, ')'
      IF (rc == FW_E_BLOWUP) THEN
        ICOUNT = 99998
        CALL FWGPU_SYNC_HOST
        CALL PREVIEW
      ENDIF
      STOP
    ENDIF
  END SUBROUTINE fwgpu_check

  ! --- array transfer helpers (explicit-shape TARGET dummies give c_loc) ---
  SUBROUTINE put_r(name, n, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    INTEGER, INTENT(IN) :: n
    REAL(SP), TARGET, INTENT(IN) :: a(n)
    CALL fwgpu_check(fw_upload(fwctx, fwgpu_id(name), C_LOC(a)), 'upload '//name)
  END SUBROUTINE put_r
  SUBROUTINE put_i(name, n, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    INTEGER, INTENT(IN) :: n
    INTEGER(C_INT), TARGET, INTENT(IN) :: a(n)
    CALL fwgpu_check(fw_upload(fwctx, fwgpu_id(name), C_LOC(a)), 'upload '//name)
  END SUBROUTINE put_i
  SUBROUTINE get_r(name, n, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    INTEGER, INTENT(IN) :: n
    REAL(SP), TARGET, INTENT(INOUT) :: a(n)
    CALL fwgpu_check(fw_download(fwctx, fwgpu_id(name), C_LOC(a)), 'download '//name)
  END SUBROUTINE get_r
  SUBROUTINE get_i(name, n, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    INTEGER, INTENT(IN) :: n
    INTEGER(C_INT), TARGET, INTENT(INOUT) :: a(n)
    CALL fwgpu_check(fw_download(fwctx, fwgpu_id(name), C_LOC(a)), 'download '//name)
  END SUBROUTINE get_i

  ! Inputs the reference reads at run time on this path that have no device
  ! counterpart: each stops the program here instead of letting the device
  ! run different physics.  (Time_Scheme is read by READ_INPUT but no
  ! routine of the step reads it, src/io.F:2806; the step is always SSP-RK3.)
  SUBROUTINE FWGPU_REFUSE_UNSUPPORTED
    IF (WAVEMAKER_VIS) CALL refuse('WAVEMAKER_VIS = T (breaker.F VISCOSITY_WMAKER, sources.F nu terms)')
    IF (BREAKWATER) CALL refuse('BREAKWATER = T (sponge-like breakwater friction, init.F/sources.F)')
    IF (WaveMakerCurrentBalance) CALL refuse('WaveMakerCurrentBalance = T (sources.F wavemaker current balance)')
    IF (TRIM(HIGH_ORDER) /= 'FOURTH') CALL refuse('HIGH_ORDER = '//TRIM(HIGH_ORDER)//' (only FOURTH, fluxes.F:1324)')
    IF (CONSTR(1:3) /= 'HLL') CALL refuse('CONSTRUCTION = '//TRIM(CONSTR)//' (only HLL, fluxes.F:109)')
  END SUBROUTINE FWGPU_REFUSE_UNSUPPORTED

  SUBROUTINE refuse(what)
    CHARACTER(LEN=*), INTENT(IN) :: what
    WRITE(*,*) 'FWGPU: unsupported on the device path: ', what
    STOP 2
  END SUBROUTINE refuse

  ! Create the device domain from MODULE GLOBAL (after INITIALIZATION,
  ! src/main.F:214) and upload the whole per-step state.
  SUBROUTINE FWGPU_INIT
    TYPE(fw_config) :: c
    CALL FWGPU_REFUSE_UNSUPPORTED
    c%real_bytes = STORAGE_SIZE(DX) / 8
    c%Mloc = Mloc; c%Nloc = Nloc; c%Mglob = Mglob; c%Nglob = Nglob
    c%Ibeg = Ibeg; c%Iend = Iend; c%Jbeg = Jbeg; c%Jend = Jend
    c%DISPERSION = MERGE(1, 0, DISPERSION); c%PERIODIC = MERGE(1, 0, PERIODIC)
    c%VISCOSITY_BREAKING = MERGE(1, 0, VISCOSITY_BREAKING)
    c%SHOW_BREAKING = MERGE(1, 0, SHOW_BREAKING)
    c%DIFFUSION_SPONGE = MERGE(1, 0, DIFFUSION_SPONGE)
    c%FRICTION_SPONGE = MERGE(1, 0, FRICTION_SPONGE)
    c%DIRECT_SPONGE = MERGE(1, 0, DIRECT_SPONGE)
    c%ETA_LIMITER = MERGE(1, 0, ETA_LIMITER); c%FIXED_DT = MERGE(1, 0, FIXED_DT)
    IF (WaveMaker(1:6) == 'WK_REG') THEN
      c%wavemaker = FW_WK_REG
    ELSEIF (WaveMaker(1:9) == 'WK_DATA2D') THEN
      c%wavemaker = FW_WK_DATA2D
    ELSEIF (WaveMaker(1:4) == 'INI_' .OR. WaveMaker(1:6) == 'N_WAVE') THEN
      c%wavemaker = FW_WK_NONE
    ELSE
      CALL refuse('WAVEMAKER = '//TRIM(WaveMaker))
    ENDIF
    c%NumFreq = NumFreq; c%ilo = ilo; c%ihi = ihi; c%jlo = jlo; c%jhi = jhi
    c%DX = DX; c%DY = DY; c%DXg = DXg; c%DYg = DYg; c%CFL = CFL
    c%MinDepth = MinDepth; c%MinDepthFrc = MinDepthFrc
    c%Gamma1 = Gamma1; c%Gamma2 = Gamma2; c%Gamma3 = Gamma3
    c%Beta_1 = Beta_1; c%Beta_2 = Beta_2; c%b1 = b1; c%b2 = b2
    c%SWE_ETA_DEP = SWE_ETA_DEP; c%FroudeCap = FroudeCap
    c%Cbrk1 = Cbrk1; c%Cbrk2 = Cbrk2; c%T_brk = T_brk
    c%WAVEMAKER_Cbrk = WAVEMAKER_Cbrk; c%nu_bkg = nu_bkg
    c%Xc_WK = Xc_WK; c%Yc_WK = Yc_WK; c%Width_WK = Width_WK; c%Ywidth_WK = Ywidth_WK
    c%EtaBlowVal = EtaBlowVal; c%DT_fixed = DT_fixed
    c%CrestLimit = CrestLimit; c%TroughLimit = TroughLimit
    c%D_gen = D_gen; c%Beta_gen = Beta_gen; c%rlamda = rlamda
    c%Tperiod = Tperiod; c%Time_ramp = Time_ramp; c%FreqPeak = FreqPeak
    CALL fwgpu_check(fw_create(c, fwctx), 'fw_create')
    ! the roller flux and undertow are kept up to date only with every
    ! intermediate array (output diagnostics, read by no routine of the step)
    IF (OUT_ROLLER .OR. OUT_UNDERTOW) &
      CALL fwgpu_check(fw_set_option(fwctx, FW_OPT_FULL_STATE, 1_C_INT), 'fw_set_option')
    CALL FWGPU_SYNC_DEVICE
  END SUBROUTINE FWGPU_INIT

  ! host -> device: every array of the per-step state
  SUBROUTINE FWGPU_SYNC_DEVICE
    INTEGER :: nc, nx, ny
    nc = Mloc * Nloc; nx = (Mloc + 1) * Nloc; ny = Mloc * (Nloc + 1)
    CALL put_i('MASK', nc, MASK); CALL put_i('MASK9', nc, MASK9); CALL put_i('MASK_STRUC', nc, MASK_STRUC)
    CALL put_r('Eta', nc, Eta); CALL put_r('Eta0', nc, Eta0); CALL put_r('U', nc, U); CALL put_r('V', nc, V)
    CALL put_r('U0', nc, U0); CALL put_r('V0', nc, V0); CALL put_r('HU', nc, HU); CALL put_r('HV', nc, HV)
    CALL put_r('Ubar', nc, Ubar); CALL put_r('Vbar', nc, Vbar)
    CALL put_r('Ubar0', nc, Ubar0); CALL put_r('Vbar0', nc, Vbar0)
    CALL put_r('H', nc, H); CALL put_r('Depth', nc, Depth); CALL put_r('Cd', nc, Cd)
    CALL put_r('SPONGE', nc, SPONGE); CALL put_r('nu_break', nc, nu_break)
    CALL put_r('AGE_BREAKING', nc, AGE_BREAKING); CALL put_r('ETAT', nc, ETAT)
    CALL put_r('ETAx', nc, ETAx); CALL put_r('ETAy', nc, ETAy)
    CALL put_r('U4', nc, U4); CALL put_r('V4', nc, V4); CALL put_r('U1p', nc, U1p); CALL put_r('V1p', nc, V1p)
    CALL put_r('WaveMaker_Mass', nc, WaveMaker_Mass)
    IF (ALLOCATED(CD_4_SPONGE)) CALL put_r('CD_4_SPONGE', nc, CD_4_SPONGE)
    IF (ALLOCATED(nu_sponge)) CALL put_r('nu_sponge', nc, nu_sponge)
    CALL put_r('ROLLER_FLUX', nc, ROLLER_FLUX); CALL put_r('UNDERTOW_U', nc, UNDERTOW_U)
    CALL put_r('UNDERTOW_V', nc, UNDERTOW_V); CALL put_r('ETAmean', nc, ETAmean)
    CALL put_r('DepthX', nx, DepthX); CALL put_r('P', nx, P); CALL put_r('Fx', nx, Fx); CALL put_r('Gx', nx, Gx)
    CALL put_r('DepthY', ny, DepthY); CALL put_r('Q', ny, Q); CALL put_r('Fy', ny, Fy); CALL put_r('Gy', ny, Gy)
    IF (ALLOCATED(xmk_wk)) CALL put_r('xmk_wk', Mloc, xmk_wk)
    IF (ALLOCATED(ymk_wk)) CALL put_r('ymk_wk', Nloc, ymk_wk)
    IF (WaveMaker(1:9) == 'WK_DATA2D') THEN
      CALL put_r('OMGN2D', NumFreq, OMGN2D)
      CALL put_r('Cm', nc * NumFreq, Cm); CALL put_r('Sm', nc * NumFreq, Sm)
    ENDIF
    CALL fwgpu_check(fw_set_clock(fwctx, REAL(DT, C_DOUBLE), REAL(TIME, C_DOUBLE)), 'set_clock')
  END SUBROUTINE FWGPU_SYNC_DEVICE

  ! device -> host: the arrays host-side output and diagnostics read
  SUBROUTINE FWGPU_SYNC_HOST
    INTEGER :: nc, nx, ny
    REAL(C_DOUBLE) :: d, t
    nc = Mloc * Nloc; nx = (Mloc + 1) * Nloc; ny = Mloc * (Nloc + 1)
    CALL get_r('Eta', nc, Eta); CALL get_r('U', nc, U); CALL get_r('V', nc, V)
    CALL get_r('HU', nc, HU); CALL get_r('HV', nc, HV); CALL get_r('Ubar', nc, Ubar); CALL get_r('Vbar', nc, Vbar)
    CALL get_r('H', nc, H); CALL get_i('MASK', nc, MASK); CALL get_i('MASK9', nc, MASK9)
    CALL get_r('AGE_BREAKING', nc, AGE_BREAKING); CALL get_r('nu_break', nc, nu_break)
    CALL get_r('ETAT', nc, ETAT); CALL get_r('ETAx', nc, ETAx); CALL get_r('ETAy', nc, ETAy)
    CALL get_r('SourceX', nc, SourceX); CALL get_r('SourceY', nc, SourceY)
    CALL get_r('ROLLER_FLUX', nc, ROLLER_FLUX); CALL get_r('UNDERTOW_U', nc, UNDERTOW_U)
    CALL get_r('UNDERTOW_V', nc, UNDERTOW_V)
    CALL get_r('DepthX', nx, DepthX); CALL get_r('P', nx, P); CALL get_r('Fx', nx, Fx); CALL get_r('Gx', nx, Gx)
    CALL get_r('DepthY', ny, DepthY); CALL get_r('Q', ny, Q); CALL get_r('Fy', ny, Fy); CALL get_r('Gy', ny, Gy)
    CALL get_r('SxL', nx, SxL); CALL get_r('SxR', nx, SxR); CALL get_r('SyL', ny, SyL); CALL get_r('SyR', ny, SyR)
    CALL fwgpu_check(fw_get_clock(fwctx, d, t), 'get_clock')
    DT = REAL(d, SP)
    TIME = REAL(t, SP)
  END SUBROUTINE FWGPU_SYNC_HOST

  ! ---- the per-step routines (src/main.F:295-517) ----
  SUBROUTINE FWGPU_SAVE_STAGE0(which)      ! main.F:295-297 (1) and :331-332 (2)
    INTEGER, INTENT(IN) :: which
    CALL fwgpu_check(fw_save_stage0(fwctx, INT(which, C_INT)), 'save_stage0')
  END SUBROUTINE FWGPU_SAVE_STAGE0
  SUBROUTINE FWGPU_EXCHANGE                ! bc.F:426
    CALL fwgpu_check(fw_exchange(fwctx), 'EXCHANGE')
  END SUBROUTINE FWGPU_EXCHANGE
  SUBROUTINE FWGPU_ESTIMATE_DT             ! misc.F:216 (sets DT, TIME)
    REAL(C_DOUBLE) :: d, t
    CALL fwgpu_check(fw_estimate_dt(fwctx, d, t), 'ESTIMATE_DT')
    DT = REAL(d, SP)
    TIME = REAL(t, SP)
  END SUBROUTINE FWGPU_ESTIMATE_DT
  SUBROUTINE FWGPU_CAL_DISPERSION          ! dispersion.F:62
    CALL fwgpu_check(fw_cal_dispersion(fwctx), 'CAL_DISPERSION')
  END SUBROUTINE FWGPU_CAL_DISPERSION
  SUBROUTINE FWGPU_FLUXES                  ! fluxes.F:52
    CALL fwgpu_check(fw_fluxes(fwctx), 'FLUXES')
  END SUBROUTINE FWGPU_FLUXES
  SUBROUTINE FWGPU_SOURCETERMS             ! sources.F:56
    CALL fwgpu_check(fw_source_terms(fwctx), 'SourceTerms')
  END SUBROUTINE FWGPU_SOURCETERMS
  SUBROUTINE FWGPU_ESTIMATE_HUV(ISTEP)     ! etauv_solver.F:53
    INTEGER, INTENT(IN) :: ISTEP
    CALL fwgpu_check(fw_estimate_huv(fwctx, INT(ISTEP, C_INT)), 'ESTIMATE_HUV')
  END SUBROUTINE FWGPU_ESTIMATE_HUV
  SUBROUTINE FWGPU_UPDATE_MASK             ! masks.F:51
    CALL fwgpu_check(fw_update_mask(fwctx), 'UPDATE_MASK')
  END SUBROUTINE FWGPU_UPDATE_MASK
  SUBROUTINE FWGPU_WAVE_BREAKING           ! breaker.F:60
    CALL fwgpu_check(fw_wave_breaking(fwctx), 'WAVE_BREAKING')
  END SUBROUTINE FWGPU_WAVE_BREAKING
  SUBROUTINE FWGPU_SPONGE_DAMPING          ! sponge.F:52
    CALL fwgpu_check(fw_sponge_damping(fwctx), 'SPONGE_DAMPING')
  END SUBROUTINE FWGPU_SPONGE_DAMPING
  SUBROUTINE FWGPU_CHECK_BLOWUP            ! misc.F:394
    REAL(C_DOUBLE) :: m
    CALL fwgpu_check(fw_check_blowup(fwctx, m), 'CHECK_BLOWUP')
  END SUBROUTINE FWGPU_CHECK_BLOWUP
  ! one whole time step, main.F:295-517, in a single call
  SUBROUTINE FWGPU_STEP
    REAL(C_DOUBLE) :: d, t
    CALL fwgpu_check(fw_step(fwctx, 1_C_INT, d, t), 'step')
    DT = REAL(d, SP)
    TIME = REAL(t, SP)
  END SUBROUTINE FWGPU_STEP
  SUBROUTINE FWGPU_FINALIZE
    INTEGER(C_INT) :: rc
    rc = fw_destroy(fwctx)
    fwctx = C_NULL_PTR
  END SUBROUTINE FWGPU_FINALIZE

END MODULE FWGPU_GLOBAL
