// fw_fields.h -- the per-step state of FUNWAVE-TVD as it lives in HBM.
//
// Field names are the reference MODULE GLOBAL array names
// (src/mod_global.F:184-317).  Every 2-D array is stored with one common
// row pitch (in elements, a multiple of 64 so rows start 256-byte aligned)
// and Nloc+1 rows: x (the reference's first, fastest index I) is contiguous,
// so a wavefront reading 64 consecutive I of one row is one coalesced access.
// Cell arrays use (Mloc, Nloc) of that box, x-face arrays (Mloc+1, Nloc),
// y-face arrays (Mloc, Nloc+1).
#pragma once

#define FW_CELL_FIELDS(X) \
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
#define FW_XFACE_FIELDS(X) \
  X(DepthX) X(P) X(Fx) X(Gx) X(UxL) X(UxR) X(VxL) X(VxR) X(HUxL) X(HUxR) \
  X(HVxL) X(HVxR) X(EtaRxL) X(EtaRxR) X(HxL) X(HxR) X(U4xL) X(U4xR) \
  X(V4xL) X(V4xR) X(PL) X(PR) X(FxL) X(FxR) X(GxL) X(GxR) X(SxL) X(SxR)
#define FW_YFACE_FIELDS(X) \
  X(DepthY) X(Q) X(Fy) X(Gy) X(UyL) X(UyR) X(VyL) X(VyR) X(HUyL) X(HUyR) \
  X(HVyL) X(HVyR) X(EtaRyL) X(EtaRyR) X(HyL) X(HyR) X(U4yL) X(U4yR) \
  X(V4yL) X(V4yR) X(QL) X(QR) X(FyL) X(FyR) X(GyL) X(GyR) X(SyL) X(SyR)
#define FW_INT_FIELDS(X) X(MASK) X(MASK9) X(MASK_STRUC)
#define FW_AUX_FIELDS(X) X(xmk_wk) X(ymk_wk) X(OMGN2D) X(Cm) X(Sm)

namespace fw {

enum FieldKind { K_CELL = 0, K_XFACE = 1, K_YFACE = 2, K_INT = 3, K_AUX = 4 };

enum FieldId {
#define FW_ENUM(n) F_##n,
  FW_CELL_FIELDS(FW_ENUM) FW_XFACE_FIELDS(FW_ENUM) FW_YFACE_FIELDS(FW_ENUM)
  FW_INT_FIELDS(FW_ENUM) FW_AUX_FIELDS(FW_ENUM)
#undef FW_ENUM
  F_COUNT
};

// Device pointers of every field plus solver scratch.
template <class R>
struct Dev {
#define FW_PTR(n) R *n;
  FW_CELL_FIELDS(FW_PTR) FW_XFACE_FIELDS(FW_PTR) FW_YFACE_FIELDS(FW_PTR)
#undef FW_PTR
  int *MASK, *MASK9, *MASK_STRUC;
  int *MASKn;      // second mask buffer (UPDATE_MASK reads old, writes new)
  R *AGEn;         // second breaking-age buffer
  // wavemaker band tables
  R *xmk_wk, *ymk_wk, *OMGN2D, *Cm, *Sm;  // Cm/Sm: band only, [kf][j-jlo][i-ilo]
  R *wk_exp;       // WK_REG: exp(-Beta_gen (x_i - Xc_WK)^2) per band column
  R *wk_stage;     // per-stage host-evaluated factors (sin per row / cos,sin per freq)
  // line-solve coefficients
  R *xa, *xc, *xd, *ya, *yc, *yd, *c1, *y1, *y2;
  // (U1pp + U2) + U3 and (V1pp + V2) + V3, the only form in which the fused
  // RK update reads the nonlinear dispersive terms
  R *sumU, *sumV;
  // reductions and breaking candidates
  R *red;          // block partials
  double *red64;   // FP64 partials (stats)
  int *cand, *cand_n, *flags;
  // per breaking candidate k: quadrant, px, py, pxy, ETAt, t2, t3 at
  // cinf[8k..8k+6] (formed once by the first pass)
  R *cinf;
  // Froude-cap cells left to the host (host_math): count, cell indices,
  // (u, v, Fr, max(H, MinDepthFrc)) per cell
  int *fc_n, *fc_idx;
  R *fc;
};

struct Geo {
  int M, N, P, Ib, Ie, Jb, Je;
};

// Scalars in the working precision R (converted once from fw_config).
template <class R>
struct Prm {
  R DX, DY, DXg, DYg, CFL, MinDepth, MDF, G1, G2, G3, B1, B2, b1, b2;
  R SWE, FCAP, C1, C2, TBRK, WKC, NUB, XC, YC, WW, YW, BLOW, DTF, CREST, TROUGH;
  R Dgen, Bgen, rlam, Tper, Tramp, FreqPeak;
  int DISPERSION, PERIODIC, VISC_BRK, SHOW_BRK, DIFF_SPONGE, FRIC_SPONGE;
  int DIRECT_SPONGE, ETA_LIMITER, FIXED_DT, wavemaker, NumFreq;
  int ilo, ihi, jlo, jhi;
  int full_state;  // write every intermediate array (routine-level parity)
  int host_math;   // Froude cap's ATAN2/COS/SIN evaluated on the host
  int xd_in_huv;   // the RK update forms the cross derivatives itself (the
                   // one-pass dispersion kernel does not store them)
  // 1/DX, 1/DY and whether they are exact (DX, DY powers of two): then
  // x/DX == x*(1/DX) bit for bit and the division is replaced
  R rDX, rDY;
  int dx_pow2, dy_pow2;
  // constant factors of the tridiagonal rows, evaluated once in the
  // reference's operation order (etauv_solver.F:402-411, :520-530)
  R cxa, cxb, cxc, cxd, cya, cyb, cyc, cyd, b1h;
  R h5x, h5y;       // 0.5/DX, 0.5/DY (sources.F:441-449)
  R rDXg, rDYg;     // 1/DXg, 1/DYg (etauv_solver.F:124)
};

}  // namespace fw
