! ref_driver.F90 -- test harness (oracle side) that drives the reference
! FUNWAVE-TVD routines, compiled from the untouched reference sources, through
! the same per-step call sequence as the reference main program and writes
! full-precision binary state dumps.
!
! TEST INFRASTRUCTURE ONLY: never linked into, or called by, the GPU library.
!
! Call order restated from src/main.F:177-214 (init) and src/main.F:284-535
! (time loop: Eta0/Ubar0/Vbar0 copy :295-297, EXCHANGE :313, ESTIMATE_DT :320,
! U0/V0 :331-332, the three RK stages :380-487, CHECK_BLOWUP :517).
!
! Controlled by environment variables (READ_INPUT takes argv(1) as input file):
!   FWREF_PREFIX      dump file prefix (default 'fwref')
!   FWREF_NSTEPS      number of time steps to run (default 0)
!   FWREF_DUMP_INIT   1 -> write <prefix>_init.bin after INITIALIZATION
!   FWREF_DUMP_EVERY  k -> write <prefix>_step<NNNNNN>.bin every k steps
!   FWREF_DUMP_LIST   comma-separated steps (at most 32) -> write
!                     <prefix>_step<NNNNNN>.bin after each of them
!   FWREF_UNTIL_TOTAL 1 -> also stop once TIME >= TOTAL_TIME (the loop
!                     condition of src/main.F:284) and write
!                     <prefix>_final.bin after the last step, and print
!                     FWREF_STEPS=<steps run>
!   FWREF_TRACE_STEP  s -> in step s write <prefix>_trace_<stage>_<routine>.bin
!                     after every routine of every stage
!   FWREF_OUTPUT      1 -> call the reference OUTPUT at the top of each step
!                     (field files in RESULT_FOLDER, as src/main.F:288)
!   FWREF_WARMUP      w -> with FWREF_NSTEPS=n, time steps w+1..n with a
!                     monotonic wall clock and print one FWREF_TIMING line

PROGRAM FWREF_DRIVER
  USE PARAM
  USE GLOBAL
  USE FWREF_DUMP
  IMPLICIT NONE
  INTEGER :: nsteps, dump_init, dump_every, trace_step, do_output, warmup
  INTEGER :: istep, istg, until_total, ndl, kd, ndone
  INTEGER :: dump_list(32)
  LOGICAL :: dump_now
  INTEGER(8) :: c0, c1, crate
  CHARACTER(LEN=256) :: prefix, fname
  LOGICAL :: tr

  CALL env_str('FWREF_PREFIX', 'fwref', prefix)
  nsteps     = env_int('FWREF_NSTEPS', 0)
  dump_init  = env_int('FWREF_DUMP_INIT', 0)
  dump_every = env_int('FWREF_DUMP_EVERY', 0)
  trace_step = env_int('FWREF_TRACE_STEP', -1)
  do_output  = env_int('FWREF_OUTPUT', 0)
  warmup     = env_int('FWREF_WARMUP', -1)
  until_total = env_int('FWREF_UNTIL_TOTAL', 0)
  CALL env_list('FWREF_DUMP_LIST', dump_list, ndl)

  ! initialisation, src/main.F:177-214
  CALL READ_INPUT
  IF (INI_UVZ) TIME = HotStartTime
  CALL INDEX
  CALL ALLOCATE_VARIABLES
  CALL INITIALIZATION

  IF (dump_init == 1) CALL dump_state(TRIM(prefix)//'_init.bin')

  ndone = 0
  DO istep = 1, nsteps
    IF (until_total == 1 .AND. TIME >= TOTAL_TIME) EXIT
    ndone = istep
    IF (istep == warmup + 1) CALL SYSTEM_CLOCK(c0, crate)
    tr = (istep == trace_step)
    IF (do_output == 1) CALL OUTPUT
    ! src/main.F:295-297
    Eta0 = Eta
    Ubar0 = Ubar
    Vbar0 = Vbar
    CALL EXCHANGE                                          ! src/main.F:313
    CALL ESTIMATE_DT(Mloc,Nloc,DX,DY,U,V,H,MinDepthFrc,DT,CFL,TIME)  ! :320
    U0 = U                                                 ! :331-332
    V0 = V
    IF (tr) CALL dump_state(TRIM(prefix)//'_trace_0_dt.bin')
    DO istg = 1, 3                                         ! :380-487
      IF (DISPERSION) THEN
        CALL CAL_DISPERSION
        IF (tr) CALL trace(istg, 'disp')
      ENDIF
      CALL FLUXES
      IF (tr) CALL trace(istg, 'flux')
      CALL SourceTerms
      IF (tr) CALL trace(istg, 'src')
      CALL ESTIMATE_HUV(istg)
      IF (tr) CALL trace(istg, 'huv')
      CALL UPDATE_MASK
      IF (tr) CALL trace(istg, 'mask')
      CALL WAVE_BREAKING
      IF (tr) CALL trace(istg, 'brk')
      CALL EXCHANGE
      IF (tr) CALL trace(istg, 'exch')
      IF (DIRECT_SPONGE) THEN
        CALL SPONGE_DAMPING
        IF (tr) CALL trace(istg, 'spg')
      ENDIF
    ENDDO
    dump_now = .FALSE.
    IF (dump_every > 0) dump_now = MOD(istep, dump_every) == 0
    DO kd = 1, ndl
      IF (dump_list(kd) == istep) dump_now = .TRUE.
    ENDDO
    IF (dump_now) THEN
      WRITE(fname, '(A,A,I6.6,A)') TRIM(prefix), '_step', istep, '.bin'
      CALL dump_state(TRIM(fname))
    ENDIF
  ENDDO
  IF (until_total == 1) THEN
    CALL dump_state(TRIM(prefix)//'_final.bin')
    WRITE(*,'(A,I0)') 'FWREF_STEPS=', ndone
  ENDIF

  IF (warmup >= 0 .AND. nsteps > warmup) THEN
    CALL SYSTEM_CLOCK(c1)
    WRITE(*,'(A,I0,A,ES16.8)') 'FWREF_TIMING steps=', nsteps - warmup, ' seconds=', &
         REAL(c1 - c0, 8) / REAL(crate, 8)
  ENDIF

CONTAINS

  SUBROUTINE trace(istg, tag)
    INTEGER, INTENT(IN) :: istg
    CHARACTER(LEN=*), INTENT(IN) :: tag
    CHARACTER(LEN=300) :: f
    WRITE(f, '(A,A,I1,A,A,A)') TRIM(prefix), '_trace_', istg, '_', tag, '.bin'
    CALL dump_state(TRIM(f))
  END SUBROUTINE trace

  INTEGER FUNCTION env_int(name, dflt)
    CHARACTER(LEN=*), INTENT(IN) :: name
    INTEGER, INTENT(IN) :: dflt
    CHARACTER(LEN=64) :: v
    INTEGER :: ln, st
    CALL GET_ENVIRONMENT_VARIABLE(name, v, ln, st)
    IF (st /= 0 .OR. ln == 0) THEN
      env_int = dflt
    ELSE
      READ(v, *) env_int
    ENDIF
  END FUNCTION env_int

  SUBROUTINE env_list(name, vals, n)
    CHARACTER(LEN=*), INTENT(IN) :: name
    INTEGER, INTENT(OUT) :: vals(:), n
    CHARACTER(LEN=512) :: v
    INTEGER :: ln, st, q, a
    n = 0
    CALL GET_ENVIRONMENT_VARIABLE(name, v, ln, st)
    IF (st /= 0 .OR. ln == 0) RETURN
    a = 1
    DO q = 1, ln + 1
      IF (q > ln) THEN
        IF (q > a) THEN
          n = n + 1
          READ(v(a:q-1), *) vals(n)
        ENDIF
      ELSE IF (v(q:q) == ',') THEN
        n = n + 1
        READ(v(a:q-1), *) vals(n)
        a = q + 1
      ENDIF
      IF (n == SIZE(vals)) EXIT
    ENDDO
  END SUBROUTINE env_list

  SUBROUTINE env_str(name, dflt, out)
    CHARACTER(LEN=*), INTENT(IN) :: name, dflt
    CHARACTER(LEN=*), INTENT(OUT) :: out
    INTEGER :: ln, st
    CALL GET_ENVIRONMENT_VARIABLE(name, out, ln, st)
    IF (st /= 0 .OR. ln == 0) out = dflt
  END SUBROUTINE env_str

END PROGRAM FWREF_DRIVER
