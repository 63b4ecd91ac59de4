/* fw_oracle.h -- CPU restatement of FUNWAVE-TVD's per-time-step solver.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ and by bench.py's cpu_baseline leg
 * as the checker; never linked into the GPU library.
 *
 * Compiled twice (FWO_REAL=float / double) to match the reference's single-
 * and double-precision builds (src/mod_param.F:58-68).  Arrays use the
 * reference layout: Fortran column-major, 1-based, 3 ghost cells; cell
 * arrays are (Mloc, Nloc), x-face arrays (Mloc+1, Nloc), y-face arrays
 * (Mloc, Nloc+1) (src/mod_global.F:184-317).
 */
#ifndef FW_ORACLE_H
#define FW_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Wavemaker kinds handled on the per-step path (src/sources.F:109-184). */
enum { FWO_WK_NONE = 0, FWO_WK_REG = 1, FWO_WK_DATA2D = 2 };

typedef struct {
  int Mloc, Nloc, Mglob, Nglob, Ibeg, Iend, Jbeg, Jend;
  int DISPERSION, PERIODIC, VISCOSITY_BREAKING, SHOW_BREAKING;
  int DIFFUSION_SPONGE, FRICTION_SPONGE, DIRECT_SPONGE, ETA_LIMITER, FIXED_DT;
  int wavemaker, NumFreq, ilo, ihi, jlo, jhi;
  double DX, DY, DXg, DYg, CFL, MinDepth, MinDepthFrc;
  double Gamma1, Gamma2, Gamma3, Beta_1, Beta_2, b1, b2;
  double SWE_ETA_DEP, FroudeCap, Cbrk1, Cbrk2, T_brk, WAVEMAKER_Cbrk, nu_bkg;
  double Xc_WK, Yc_WK, Width_WK, Ywidth_WK, EtaBlowVal, DT_fixed;
  double CrestLimit, TroughLimit;
  double D_gen, Beta_gen, rlamda, Tperiod, Time_ramp, FreqPeak;
} fwo_params;

/* Field table: one pointer per array, in this order (names match the
 * reference GLOBAL arrays).  Integer masks are int32. */
#define FWO_CELL_FIELDS(X) \
  X(Eta) X(Eta0) X(U) X(V) X(U0) X(V0) X(HU) X(HV) X(Ubar) X(Vbar) \
  X(Ubar0) X(Vbar0) X(H) X(Depth) X(Cd) X(SPONGE) X(nu_break) \
  X(AGE_BREAKING) X(ETAT) X(ETAx) X(ETAy) X(Ux) X(Uy) X(Vx) X(Vy) \
  X(DUx) X(DUy) X(DVx) X(DVy) X(Uxx) X(Uxy) X(Vxy) X(Vyy) X(DUxx) \
  X(DUxy) X(DVxy) X(DVyy) X(Ut) X(Vt) X(Utx) X(Vty) X(Utxx) X(Utxy) \
  X(Vtxy) X(Vtyy) X(DUtx) X(DVty) X(DUtxx) X(DUtxy) X(DVtxy) X(DVtyy) \
  X(U4) X(V4) X(U1p) X(V1p) X(U1pp) X(V1pp) X(U2) X(V2) X(U3) X(V3) \
  X(SourceX) X(SourceY) X(WaveMaker_Mass) X(P_center) X(Q_center) \
  X(U_davg) X(V_davg) X(Wsurf) X(BreakSourceX) X(BreakSourceY) \
  X(FrcInsX) X(FrcInsY) X(CD_4_SPONGE) X(nu_sponge) \
  X(ROLLER_FLUX) X(UNDERTOW_U) X(UNDERTOW_V) X(ETAmean)
#define FWO_XFACE_FIELDS(X) \
  X(DepthX) X(P) X(Fx) X(Gx) X(UxL) X(UxR) X(VxL) X(VxR) X(HUxL) X(HUxR) \
  X(HVxL) X(HVxR) X(EtaRxL) X(EtaRxR) X(HxL) X(HxR) X(U4xL) X(U4xR) \
  X(V4xL) X(V4xR) X(PL) X(PR) X(FxL) X(FxR) X(GxL) X(GxR) X(SxL) X(SxR)
#define FWO_YFACE_FIELDS(X) \
  X(DepthY) X(Q) X(Fy) X(Gy) X(UyL) X(UyR) X(VyL) X(VyR) X(HUyL) X(HUyR) \
  X(HVyL) X(HVyR) X(EtaRyL) X(EtaRyR) X(HyL) X(HyR) X(U4yL) X(U4yR) \
  X(V4yL) X(V4yR) X(QL) X(QR) X(FyL) X(FyR) X(GyL) X(GyR) X(SyL) X(SyR)
#define FWO_INT_FIELDS(X) X(MASK) X(MASK9) X(MASK_STRUC)
/* band tables: xmk_wk(Mloc), ymk_wk(Nloc), OMGN2D(NumFreq),
 * Cm/Sm (Mloc, Nloc, NumFreq) */
#define FWO_AUX_FIELDS(X) X(xmk_wk) X(ymk_wk) X(OMGN2D) X(Cm) X(Sm)

#define FWO_PTR(n) void *n;
typedef struct {
  FWO_CELL_FIELDS(FWO_PTR)
  FWO_XFACE_FIELDS(FWO_PTR)
  FWO_YFACE_FIELDS(FWO_PTR)
  FWO_INT_FIELDS(FWO_PTR)
  FWO_AUX_FIELDS(FWO_PTR)
} fwo_fields;
#undef FWO_PTR

/* Routines (argument state is the field table plus scalars DT/TIME). */
int  fwo_real_bytes(void);
void fwo_exchange(const fwo_params *p, fwo_fields *f);
void fwo_estimate_dt(const fwo_params *p, fwo_fields *f, double *dt, double *time);
void fwo_dispersion(const fwo_params *p, fwo_fields *f, double dt);
void fwo_fluxes(const fwo_params *p, fwo_fields *f);
void fwo_sources(const fwo_params *p, fwo_fields *f, double time);
void fwo_estimate_huv(const fwo_params *p, fwo_fields *f, int istage, double dt);
void fwo_update_mask(const fwo_params *p, fwo_fields *f);
void fwo_breaking(const fwo_params *p, fwo_fields *f, double dt);
void fwo_sponge(const fwo_params *p, fwo_fields *f);
double fwo_max_abs_eta(const fwo_params *p, const fwo_fields *f);
/* nsteps full time steps (src/main.F:295-487); returns 0, or 1 on blow-up */
int  fwo_run(const fwo_params *p, fwo_fields *f, int nsteps, double *dt, double *time);

#ifdef __cplusplus
}
#endif
#ifdef __cplusplus
extern "C" {
#endif
/* Unit-level hooks on single building blocks (tests only). */
/* Thomas sweep of tridiagnal.F:338-398 on one line of length n (a, c, d
 * normalised rows; a[0] and c[n-1] unused); solution in z. a/c/d modified. */
void fwo_thomas_line(int n, void *a, void *c, void *d, void *z);
/* HLL flux of fluxes.F:171-194 */
double fwo_hll(double sl, double sr, double fl, double fr, double ul, double ur);
/* CONSTRUCT_HO_X (fluxes.F:1531-1639) on one row of n cells (n >= 7, three
 * ghost cells each side): face states OutL/OutR at faces 4..n-2 (1-based,
 * face i between cells i-1 and i); mask entries int32. */
void fwo_muscl_row(int n, const void *vin, const int *mask, void *outl, void *outr);
#ifdef __cplusplus
}
#endif

#endif
