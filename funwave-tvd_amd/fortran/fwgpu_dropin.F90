! fwgpu_dropin.F90 -- the per-step routines of the reference under their own
! names and argument lists, running on the GPU.
!
! A reference main program (src/main.F) that adds
!     USE FWGPU_DROPIN
! to its USE lines binds every call of its time loop (src/main.F:284-535) to
! the module procedures below instead of the CPU routines, by use
! association: CALL EXCHANGE, CALL ESTIMATE_DT(Mloc,Nloc,DX,DY,U,V,H,
! MinDepthFrc,DT,CFL,TIME), CALL Cal_Dispersion, CALL FLUXES,
! CALL SourceTerms, CALL ESTIMATE_HUV(ISTAGE), CALL UPDATE_MASK,
! CALL WAVE_BREAKING, CALL SPONGE_DAMPING and CALL CHECK_BLOWUP keep their
! call sites unchanged.  Other files (init.F calls EXCHANGE during
! INITIALIZATION) do not use the module and keep the CPU routines.
!
! Besides the USE line the main program gains three calls:
!     CALL FWGPU_BEGIN        after CALL INITIALIZATION (upload to HBM)
!     CALL FWGPU_SYNC_OUTPUT  before CALL OUTPUT (device state back into
!                             MODULE GLOBAL when OUTPUT will write)
!     CALL FWGPU_END          after the time loop (state back, context freed)
!
! The main loop's own array copies Eta0=Eta, Ubar0=Ubar, Vbar0=Vbar
! (src/main.F:295-297) and U0=U, V0=V (:331-332) act on the host arrays,
! which the device state supersedes; their device counterparts run here, at
! the same points of the step: the first EXCHANGE of a step (the one after
! CHECK_BLOWUP or FWGPU_BEGIN) copies Eta0/Ubar0/Vbar0 first, and
! ESTIMATE_DT copies U0/V0 after the new time step is known.
MODULE FWGPU_DROPIN
  USE, INTRINSIC :: ISO_C_BINDING
  USE PARAM
  USE GLOBAL, ONLY : Mloc, Nloc, DT, TIME, TOTAL_TIME, STEADY_TIME, &
                     OUT_Hmax, OUT_Hmin, OUT_Umax, OUT_MFmax, OUT_VORmax, OUT_Time, &
                     SCREEN_COUNT, SCREEN_INTV, PLOT_COUNT, PLOT_INTV, PLOT_START_TIME, &
                     NumberStations, PLOT_COUNT_STATION, PLOT_INTV_STATION
  USE FWGPU_BIND
  USE FWGPU_GLOBAL, ONLY : fwctx, fwgpu_check, FWGPU_INIT, FWGPU_SYNC_HOST, FWGPU_FINALIZE
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: FWGPU_BEGIN, FWGPU_SYNC_OUTPUT, FWGPU_END
  PUBLIC :: EXCHANGE, ESTIMATE_DT, Cal_Dispersion, FLUXES, SourceTerms, ESTIMATE_HUV, &
            UPDATE_MASK, WAVE_BREAKING, SPONGE_DAMPING, CHECK_BLOWUP
  ! .TRUE. between CHECK_BLOWUP (or FWGPU_BEGIN) and the next EXCHANGE: that
  ! EXCHANGE opens a new time step (src/main.F:295-313)
  LOGICAL, SAVE :: step_open = .FALSE.

CONTAINS

  ! After INITIALIZATION (src/main.F:214): refuse host-side diagnostics that
  ! read MODULE GLOBAL every step (they would see stale arrays), then create
  ! the device domain and upload the state.
  SUBROUTINE FWGPU_BEGIN
    IF (STEADY_TIME < TOTAL_TIME) THEN
      WRITE(*,*) 'FWGPU: STEADY_TIME < TOTAL_TIME: MIXING_STUFF would average host arrays every step'
      STOP 2
    ENDIF
    IF (OUT_Hmax .OR. OUT_Hmin .OR. OUT_Umax .OR. OUT_MFmax .OR. OUT_VORmax .OR. OUT_Time) THEN
      WRITE(*,*) 'FWGPU: OUT_Hmax/Hmin/Umax/MFmax/VORmax/Time: MAX_MIN_PROPERTY would read host arrays every step'
      STOP 2
    ENDIF
    CALL FWGPU_INIT
    step_open = .TRUE.
  END SUBROUTINE FWGPU_BEGIN

  ! Before OUTPUT (src/main.F:288): OUTPUT adds DT to its counters and
  ! writes statistics, station records or field files once a counter
  ! reaches its interval (src/io.F:72-100).  The device state is copied
  ! back only when one of them will.
  SUBROUTINE FWGPU_SYNC_OUTPUT
    LOGICAL :: due
    due = SCREEN_COUNT + DT >= SCREEN_INTV
    IF (NumberStations > 0) due = due .OR. (PLOT_COUNT_STATION + DT >= PLOT_INTV_STATION)
    IF (TIME >= PLOT_START_TIME) due = due .OR. (PLOT_COUNT + DT >= PLOT_INTV)
    IF (due) CALL FWGPU_SYNC_HOST
  END SUBROUTINE FWGPU_SYNC_OUTPUT

  ! After the time loop: the final state back into MODULE GLOBAL (the
  ! reference writes its last station records from it), then free HBM.
  SUBROUTINE FWGPU_END
    CALL FWGPU_SYNC_HOST
    CALL FWGPU_FINALIZE
  END SUBROUTINE FWGPU_END

  ! ---- src/main.F's per-step calls ----------------------------------------

  SUBROUTINE EXCHANGE                                   ! bc.F:426
    IF (step_open) THEN                                  ! main.F:295-297
      CALL fwgpu_check(fw_save_stage0(fwctx, 1_C_INT), 'Eta0/Ubar0/Vbar0')
      step_open = .FALSE.
    ENDIF
    CALL fwgpu_check(fw_exchange(fwctx), 'EXCHANGE')
  END SUBROUTINE EXCHANGE

  ! misc.F:216: the device reduces U, V, H of its own state; the host
  ! arrays passed in are not read.  DT and TIME are set as in the reference,
  ! then U0=U, V0=V (main.F:331-332) runs on the device.
  SUBROUTINE ESTIMATE_DT(M, N, DX, DY, U, V, H, MinDepthFrc, DT, CFL, TIME)
    INTEGER, INTENT(IN) :: M, N
    REAL(SP), INTENT(IN) :: DX, DY
    REAL(SP), INTENT(IN), DIMENSION(M, N) :: U, V, H
    REAL(SP), INTENT(IN) :: CFL, MinDepthFrc
    REAL(SP), INTENT(OUT) :: DT
    REAL(SP), INTENT(INOUT) :: TIME
    REAL(C_DOUBLE) :: d, t
    CALL fwgpu_check(fw_estimate_dt(fwctx, d, t), 'ESTIMATE_DT')
    DT = REAL(d, SP)
    TIME = REAL(t, SP)
    CALL fwgpu_check(fw_save_stage0(fwctx, 2_C_INT), 'U0/V0')
  END SUBROUTINE ESTIMATE_DT

  SUBROUTINE Cal_Dispersion                             ! dispersion.F:62
    CALL fwgpu_check(fw_cal_dispersion(fwctx), 'CAL_DISPERSION')
  END SUBROUTINE Cal_Dispersion

  SUBROUTINE FLUXES                                     ! fluxes.F:52
    CALL fwgpu_check(fw_fluxes(fwctx), 'FLUXES')
  END SUBROUTINE FLUXES

  SUBROUTINE SourceTerms                                ! sources.F:56
    CALL fwgpu_check(fw_source_terms(fwctx), 'SourceTerms')
  END SUBROUTINE SourceTerms

  SUBROUTINE ESTIMATE_HUV(ISTEP)                        ! etauv_solver.F:53
    INTEGER, INTENT(IN) :: ISTEP
    CALL fwgpu_check(fw_estimate_huv(fwctx, INT(ISTEP, C_INT)), 'ESTIMATE_HUV')
  END SUBROUTINE ESTIMATE_HUV

  SUBROUTINE UPDATE_MASK                                ! masks.F:51
    CALL fwgpu_check(fw_update_mask(fwctx), 'UPDATE_MASK')
  END SUBROUTINE UPDATE_MASK

  SUBROUTINE WAVE_BREAKING                              ! breaker.F:60
    CALL fwgpu_check(fw_wave_breaking(fwctx), 'WAVE_BREAKING')
  END SUBROUTINE WAVE_BREAKING

  SUBROUTINE SPONGE_DAMPING                             ! sponge.F:52
    CALL fwgpu_check(fw_sponge_damping(fwctx), 'SPONGE_DAMPING')
  END SUBROUTINE SPONGE_DAMPING

  ! misc.F:394: on blow-up fwgpu_check writes the ICOUNT=99998 preview from
  ! the device state and stops, as the reference does
  SUBROUTINE CHECK_BLOWUP
    REAL(C_DOUBLE) :: m
    CALL fwgpu_check(fw_check_blowup(fwctx, m), 'CHECK_BLOWUP')
    step_open = .TRUE.
  END SUBROUTINE CHECK_BLOWUP

END MODULE FWGPU_DROPIN
