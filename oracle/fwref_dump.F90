! fwref_dump.F90 -- binary state dump of MODULE GLOBAL (test infrastructure;
! read by oracle/fwdump.py).  Used by the oracle harnesses ref_driver.F90
! and gpu_driver.F90.

MODULE FWREF_DUMP
  USE PARAM
  USE GLOBAL
  IMPLICIT NONE
  INTEGER, PARAMETER :: DU = 77
CONTAINS

  SUBROUTINE rec_head(name, kind, ndim, d1, d2, d3)
    CHARACTER(LEN=*), INTENT(IN) :: name
    INTEGER, INTENT(IN) :: kind, ndim, d1, d2, d3
    CHARACTER(LEN=32) :: nm
    INTEGER(4) :: hdr(5)
    nm = name
    hdr = (/ kind, ndim, d1, d2, d3 /)
    WRITE(DU) nm, hdr
  END SUBROUTINE rec_head

  SUBROUTINE dr2(name, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(SP), INTENT(IN) :: a(:,:)
    CALL rec_head(name, STORAGE_SIZE(a)/8, 2, SIZE(a,1), SIZE(a,2), 1)
    WRITE(DU) a
  END SUBROUTINE dr2

  SUBROUTINE di2(name, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    INTEGER, INTENT(IN) :: a(:,:)
    INTEGER(4), ALLOCATABLE :: b(:,:)
    ALLOCATE(b(SIZE(a,1),SIZE(a,2)))
    b = a
    CALL rec_head(name, 104, 2, SIZE(a,1), SIZE(a,2), 1)
    WRITE(DU) b
    DEALLOCATE(b)
  END SUBROUTINE di2

  SUBROUTINE dr1(name, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(SP), INTENT(IN) :: a(:)
    CALL rec_head(name, STORAGE_SIZE(a)/8, 1, SIZE(a,1), 1, 1)
    WRITE(DU) a
  END SUBROUTINE dr1

  SUBROUTINE dr3(name, a)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(SP), INTENT(IN) :: a(:,:,:)
    CALL rec_head(name, STORAGE_SIZE(a)/8, 3, SIZE(a,1), SIZE(a,2), SIZE(a,3))
    WRITE(DU) a
  END SUBROUTINE dr3

  SUBROUTINE drs(name, x)
    CHARACTER(LEN=*), INTENT(IN) :: name
    REAL(SP), INTENT(IN) :: x
    CALL rec_head(name, STORAGE_SIZE(x)/8, 0, 1, 1, 1)
    WRITE(DU) x
  END SUBROUTINE drs

  SUBROUTINE dis(name, n)
    CHARACTER(LEN=*), INTENT(IN) :: name
    INTEGER, INTENT(IN) :: n
    INTEGER(4) :: m
    m = n
    CALL rec_head(name, 104, 0, 1, 1, 1)
    WRITE(DU) m
  END SUBROUTINE dis

  SUBROUTINE dls(name, l)
    CHARACTER(LEN=*), INTENT(IN) :: name
    LOGICAL, INTENT(IN) :: l
    IF (l) THEN
      CALL dis(name, 1)
    ELSE
      CALL dis(name, 0)
    ENDIF
  END SUBROUTINE dls

  SUBROUTINE dstr(name, s)
    CHARACTER(LEN=*), INTENT(IN) :: name
    CHARACTER(LEN=*), INTENT(IN) :: s
    CHARACTER(LEN=80) :: t
    t = s
    CALL rec_head(name, 200, 1, 80, 1, 1)
    WRITE(DU) t
  END SUBROUTINE dstr

  ! Every scalar parameter and field the per-step path reads or writes.
  SUBROUTINE dump_state(fname)
    CHARACTER(LEN=*), INTENT(IN) :: fname
    OPEN(DU, FILE=TRIM(fname), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE(DU) 'FWDUMP01'
    CALL dis('SP_BYTES', STORAGE_SIZE(DX)/8)
    CALL dis('Mglob', Mglob); CALL dis('Nglob', Nglob)
    CALL dis('Mloc', Mloc);   CALL dis('Nloc', Nloc)
    CALL dis('Ibeg', Ibeg);   CALL dis('Iend', Iend)
    CALL dis('Jbeg', Jbeg);   CALL dis('Jend', Jend)
    CALL dis('Nghost', Nghost)
    CALL drs('DX', DX); CALL drs('DY', DY); CALL drs('DXg', DXg); CALL drs('DYg', DYg)
    CALL drs('DT', DT); CALL drs('TIME', TIME); CALL drs('TOTAL_TIME', TOTAL_TIME)
    CALL drs('CFL', CFL)
    CALL drs('MinDepth', MinDepth); CALL drs('MinDepthFrc', MinDepthFrc)
    CALL drs('Gamma1', Gamma1); CALL drs('Gamma2', Gamma2); CALL drs('Gamma3', Gamma3)
    CALL drs('Beta_1', Beta_1); CALL drs('Beta_2', Beta_2)
    CALL drs('b1', b1); CALL drs('b2', b2)
    CALL drs('SWE_ETA_DEP', SWE_ETA_DEP); CALL drs('FroudeCap', FroudeCap)
    CALL drs('Cbrk1', Cbrk1); CALL drs('Cbrk2', Cbrk2); CALL drs('T_brk', T_brk)
    CALL drs('WAVEMAKER_Cbrk', WAVEMAKER_Cbrk); CALL drs('nu_bkg', nu_bkg)
    CALL drs('Xc_WK', Xc_WK); CALL drs('Yc_WK', Yc_WK)
    CALL drs('Width_WK', Width_WK); CALL drs('Ywidth_WK', Ywidth_WK)
    CALL drs('EtaBlowVal', EtaBlowVal); CALL drs('DT_fixed', DT_fixed)
    CALL drs('CrestLimit', CrestLimit); CALL drs('TroughLimit', TroughLimit)
    CALL drs('D_gen', D_gen); CALL drs('Beta_gen', Beta_gen); CALL drs('rlamda', rlamda)
    CALL drs('Tperiod', Tperiod); CALL drs('Time_ramp', Time_ramp)
    CALL drs('FreqPeak', FreqPeak)
    CALL dls('DISPERSION', DISPERSION); CALL dls('PERIODIC', PERIODIC)
    CALL dls('VISCOSITY_BREAKING', VISCOSITY_BREAKING)
    CALL dls('SHOW_BREAKING', SHOW_BREAKING); CALL dls('WAVEMAKER_VIS', WAVEMAKER_VIS)
    CALL dls('DIFFUSION_SPONGE', DIFFUSION_SPONGE)
    CALL dls('FRICTION_SPONGE', FRICTION_SPONGE)
    CALL dls('DIRECT_SPONGE', DIRECT_SPONGE); CALL dls('ETA_LIMITER', ETA_LIMITER)
    CALL dls('WaveMakerCurrentBalance', WaveMakerCurrentBalance)
    CALL dls('BREAKWATER', BREAKWATER); CALL dls('DISP_TIME_LEFT', DISP_TIME_LEFT)
    CALL dls('FIXED_DT', FIXED_DT); CALL dls('OUT_VORmax', OUT_VORmax)
    CALL dis('NumFreq', NumFreq)
    CALL dis('ilo', ilo); CALL dis('ihi', ihi); CALL dis('jlo', jlo); CALL dis('jhi', jhi)
    CALL dstr('WaveMaker', WaveMaker)
    IF (ALLOCATED(xmk_wk)) CALL dr1('xmk_wk', xmk_wk)
    IF (ALLOCATED(ymk_wk)) CALL dr1('ymk_wk', ymk_wk)
    IF (ALLOCATED(OMGN2D)) CALL dr1('OMGN2D', OMGN2D)
    IF (ALLOCATED(Cm)) CALL dr3('Cm', Cm)
    IF (ALLOCATED(Sm)) CALL dr3('Sm', Sm)
    ! cell-centred integer masks
    CALL di2('MASK', MASK); CALL di2('MASK9', MASK9); CALL di2('MASK_STRUC', MASK_STRUC)
    ! cell-centred reals
    CALL dr2('Eta', Eta); CALL dr2('Eta0', Eta0); CALL dr2('U', U); CALL dr2('V', V)
    CALL dr2('U0', U0); CALL dr2('V0', V0); CALL dr2('HU', HU); CALL dr2('HV', HV)
    CALL dr2('Ubar', Ubar); CALL dr2('Vbar', Vbar)
    CALL dr2('Ubar0', Ubar0); CALL dr2('Vbar0', Vbar0)
    CALL dr2('H', H); CALL dr2('Depth', Depth); CALL dr2('Cd', Cd); CALL dr2('SPONGE', SPONGE)
    CALL dr2('nu_break', nu_break); CALL dr2('AGE_BREAKING', AGE_BREAKING)
    CALL dr2('ETAT', ETAT); CALL dr2('ETAx', ETAx); CALL dr2('ETAy', ETAy)
    CALL dr2('Ux', Ux); CALL dr2('Uy', Uy); CALL dr2('Vx', Vx); CALL dr2('Vy', Vy)
    CALL dr2('DUx', DUx); CALL dr2('DUy', DUy); CALL dr2('DVx', DVx); CALL dr2('DVy', DVy)
    CALL dr2('Uxx', Uxx); CALL dr2('Uxy', Uxy); CALL dr2('Vxy', Vxy); CALL dr2('Vyy', Vyy)
    CALL dr2('DUxx', DUxx); CALL dr2('DUxy', DUxy); CALL dr2('DVxy', DVxy); CALL dr2('DVyy', DVyy)
    CALL dr2('Ut', Ut); CALL dr2('Vt', Vt); CALL dr2('Utx', Utx); CALL dr2('Vty', Vty)
    CALL dr2('Utxx', Utxx); CALL dr2('Utxy', Utxy); CALL dr2('Vtxy', Vtxy); CALL dr2('Vtyy', Vtyy)
    CALL dr2('DUtx', DUtx); CALL dr2('DVty', DVty)
    CALL dr2('DUtxx', DUtxx); CALL dr2('DUtxy', DUtxy); CALL dr2('DVtxy', DVtxy); CALL dr2('DVtyy', DVtyy)
    CALL dr2('U4', U4); CALL dr2('V4', V4); CALL dr2('U1p', U1p); CALL dr2('V1p', V1p)
    CALL dr2('U1pp', U1pp); CALL dr2('V1pp', V1pp)
    CALL dr2('U2', U2); CALL dr2('V2', V2); CALL dr2('U3', U3); CALL dr2('V3', V3)
    CALL dr2('SourceX', SourceX); CALL dr2('SourceY', SourceY)
    CALL dr2('WaveMaker_Mass', WaveMaker_Mass)
    CALL dr2('P_center', P_center); CALL dr2('Q_center', Q_center)
    CALL dr2('U_davg', U_davg); CALL dr2('V_davg', V_davg); CALL dr2('Wsurf', Wsurf)
    CALL dr2('BreakSourceX', BreakSourceX); CALL dr2('BreakSourceY', BreakSourceY)
    CALL dr2('FrcInsX', FrcInsX); CALL dr2('FrcInsY', FrcInsY)
    IF (ALLOCATED(CD_4_SPONGE)) CALL dr2('CD_4_SPONGE', CD_4_SPONGE)
    IF (ALLOCATED(nu_sponge)) CALL dr2('nu_sponge', nu_sponge)
    CALL dr2('ROLLER_FLUX', ROLLER_FLUX); CALL dr2('UNDERTOW_U', UNDERTOW_U)
    CALL dr2('UNDERTOW_V', UNDERTOW_V); CALL dr2('ETAmean', ETAmean)
    ! x-faces (Mloc+1, Nloc)
    CALL dr2('DepthX', DepthX); CALL dr2('P', P); CALL dr2('Fx', Fx); CALL dr2('Gx', Gx)
    CALL dr2('UxL', UxL); CALL dr2('UxR', UxR); CALL dr2('VxL', VxL); CALL dr2('VxR', VxR)
    CALL dr2('HUxL', HUxL); CALL dr2('HUxR', HUxR); CALL dr2('HVxL', HVxL); CALL dr2('HVxR', HVxR)
    CALL dr2('EtaRxL', EtaRxL); CALL dr2('EtaRxR', EtaRxR); CALL dr2('HxL', HxL); CALL dr2('HxR', HxR)
    CALL dr2('U4xL', U4xL); CALL dr2('U4xR', U4xR); CALL dr2('V4xL', V4xL); CALL dr2('V4xR', V4xR)
    CALL dr2('PL', PL); CALL dr2('PR', PR); CALL dr2('FxL', FxL); CALL dr2('FxR', FxR)
    CALL dr2('GxL', GxL); CALL dr2('GxR', GxR); CALL dr2('SxL', SxL); CALL dr2('SxR', SxR)
    ! y-faces (Mloc, Nloc+1)
    CALL dr2('DepthY', DepthY); CALL dr2('Q', Q); CALL dr2('Fy', Fy); CALL dr2('Gy', Gy)
    CALL dr2('UyL', UyL); CALL dr2('UyR', UyR); CALL dr2('VyL', VyL); CALL dr2('VyR', VyR)
    CALL dr2('HUyL', HUyL); CALL dr2('HUyR', HUyR); CALL dr2('HVyL', HVyL); CALL dr2('HVyR', HVyR)
    CALL dr2('EtaRyL', EtaRyL); CALL dr2('EtaRyR', EtaRyR); CALL dr2('HyL', HyL); CALL dr2('HyR', HyR)
    CALL dr2('U4yL', U4yL); CALL dr2('U4yR', U4yR); CALL dr2('V4yL', V4yL); CALL dr2('V4yR', V4yR)
    CALL dr2('QL', QL); CALL dr2('QR', QR); CALL dr2('FyL', FyL); CALL dr2('FyR', FyR)
    CALL dr2('GyL', GyL); CALL dr2('GyR', GyR); CALL dr2('SyL', SyL); CALL dr2('SyR', SyR)
    WRITE(DU) 'FWDUMPEN'
    CLOSE(DU)
  END SUBROUTINE dump_state

END MODULE FWREF_DUMP
