/* fw_oracle.c -- CPU restatement (plain C) of FUNWAVE-TVD's per-time-step
 * Boussinesq solver, serial Cartesian build.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): used by tests/ and by bench.py's
 * cpu_baseline leg, never by the GPU library.
 *
 * Each routine restates one reference routine loop nest by loop nest, with
 * the same index ranges, the same order of floating-point operations (no
 * contraction: build with -ffp-contract=off) and the same write order where
 * later writes overwrite earlier ones.  Literal typing follows the
 * reference: `_SP` literals are REAL(SP); untyped literals are default REAL
 * (single precision), which matters in the double build only for pi
 * (src/mod_param.F:71).  Reference citations are file:line under src/.
 *
 * Pinned by: the compiled reference (oracle/_ref) on every routine and on
 * whole runs, bit for bit, see tests/test_oracle_vs_ref.py; plus closed-form
 * checks in tests/test_oracle_math.py.
 */
#include "fw_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

typedef FWO_REAL R;

#define FWO_IS_DOUBLE (sizeof(R) == 8)
/* REAL(SP) literal x_SP: exact decimal rounded once to R */
#define SPC(x) ((R)(FWO_IS_DOUBLE ? (R)(x) : (R)(x##f)))
/* default-REAL literal: rounded to single, then widened */
#define FLC(x) ((R)(x##f))

#define GRAV SPC(9.81)                 /* mod_param.F:74 */
#define SMALL SPC(0.000001)            /* mod_param.F:72 */
#define LARGE SPC(999999.0)            /* mod_param.F:73 */
#define PI FLC(3.141592653)            /* mod_param.F:71 (default REAL) */
#define ZERO ((R)0)

int fwo_real_bytes(void) { return (int)sizeof(R); }

/* ------------------------------------------------------------------ */
/* indexing: Fortran (I,J), 1-based, column-major                      */
#define DIMS                                                            \
  const int M = p->Mloc, N = p->Nloc, M1 = p->Mloc + 1, N1 = p->Nloc + 1; \
  const int Ib = p->Ibeg, Ie = p->Iend, Jb = p->Jbeg, Je = p->Jend;     \
  (void)M; (void)N; (void)M1; (void)N1; (void)Ib; (void)Ie; (void)Jb; (void)Je
#define C_(a, i, j) (a)[((size_t)(j)-1) * (size_t)M + (size_t)(i)-1]
#define X_(a, i, j) (a)[((size_t)(j)-1) * (size_t)M1 + (size_t)(i)-1]
#define Y_(a, i, j) (a)[((size_t)(j)-1) * (size_t)M + (size_t)(i)-1]
#define F(n) ((R *)f->n)
#define FI(n) ((int *)f->n)

static R rmax(R a, R b) { return a > b ? a : b; } /* Fortran MAX on non-NaN */
static R rmin(R a, R b) { return a < b ? a : b; }
static R rsign1(R b) { return signbit(b) ? (R)-1 : (R)1; } /* SIGN(1,b) */

/* ------------------------------------------------------------------ */
/* PHI_COLL, bc.F:631-1272 (serial, Cartesian).  vtype 1: even in x and y
 * (eta), 2: odd in x (u), 3: odd in y (v).  Periodic-y copy first over all
 * columns (bc.F:740-747), then x-mirror on interior rows (bc.F:753-791 ...),
 * then y-mirror over all columns when not periodic. */
static void phi_coll(const fwo_params *p, R *phi, int vtype) {
  DIMS;
  if (p->PERIODIC) {
    for (int i = 1; i <= M; i++)
      for (int k = 1; k <= 3; k++) {
        C_(phi, i, k) = C_(phi, i, Je - 3 + k);
        C_(phi, i, Je + k) = C_(phi, i, 3 + k);
      }
  }
  const int sx = (vtype == 2), sy = (vtype == 3);
  for (int j = Jb; j <= Je; j++)
    for (int k = 1; k <= 3; k++) {
      R w = C_(phi, Ib + 3 - k, j);
      C_(phi, k, j) = sx ? -w : w;
    }
  for (int j = Jb; j <= Je; j++)
    for (int k = 1; k <= 3; k++) {
      R e = C_(phi, Ie - k + 1, j);
      C_(phi, Ie + k, j) = sx ? -e : e;
    }
  if (!p->PERIODIC) {
    for (int i = 1; i <= M; i++)
      for (int k = 1; k <= 3; k++) {
        R s = C_(phi, i, Jb + 3 - k);
        C_(phi, i, k) = sy ? -s : s;
      }
    for (int i = 1; i <= M; i++)
      for (int k = 1; k <= 3; k++) {
        R n = C_(phi, i, Je - k + 1);
        C_(phi, i, Je + k) = sy ? -n : n;
      }
  }
}

/* PHI_COLL_VARIABLE_LENGTH, bc.F:539-617: even mirror on a face array of
 * leading dimension m (no periodic branch). */
static void phi_coll_vl(R *phi, int m, int ib, int ie, int jb, int je) {
#define V_(i, j) phi[((size_t)(j)-1) * (size_t)m + (size_t)(i)-1]
  for (int j = jb; j <= je; j++)
    for (int k = 1; k <= 3; k++) V_(k, j) = V_(ib + 3 - k, j);
  for (int j = jb; j <= je; j++)
    for (int k = 1; k <= 3; k++) V_(ie + k, j) = V_(ie - k + 1, j);
  for (int i = 1; i <= m; i++)
    for (int k = 1; k <= 3; k++) V_(i, k) = V_(i, jb + 3 - k);
  for (int i = 1; i <= m; i++)
    for (int k = 1; k <= 3; k++) V_(i, je + k) = V_(i, je - k + 1);
#undef V_
}

/* EXCHANGE, bc.F:426-529 */
void fwo_exchange(const fwo_params *p, fwo_fields *f) {
  DIMS;
  phi_coll(p, F(Eta), 1);
  phi_coll(p, F(Wsurf), 1);
  phi_coll(p, F(P_center), 1);
  phi_coll(p, F(Q_center), 1);
  phi_coll(p, F(U_davg), 1);
  phi_coll(p, F(V_davg), 1);
  if (p->VISCOSITY_BREAKING) phi_coll(p, F(AGE_BREAKING), 1);
  if (p->VISCOSITY_BREAKING) phi_coll(p, F(nu_break), 1);
  if (p->VISCOSITY_BREAKING || p->DIFFUSION_SPONGE) {
    phi_coll_vl(F(P), M1, Ib, Ie + 1, Jb, Je);
    phi_coll_vl(F(Q), M, Ib, Ie, Jb, Je + 1);
  }
  {
    size_t n = (size_t)M * N;
    R *rm = (R *)malloc(n * sizeof(R));
    int *mk = FI(MASK);
    for (size_t q = 0; q < n; q++) rm[q] = (R)mk[q];
    phi_coll(p, rm, 1);
    for (size_t q = 0; q < n; q++) mk[q] = (int)rm[q];
    free(rm);
  }
  phi_coll(p, F(U), 2);
  phi_coll(p, F(HU), 2);
  phi_coll(p, F(V), 3);
  phi_coll(p, F(HV), 3);
  {
    size_t n = (size_t)M * N;
    int *mk = FI(MASK);
    R *u = F(U), *v = F(V), *hu = F(HU), *hv = F(HV);
    for (size_t q = 0; q < n; q++) {
      R m = (R)mk[q];
      u[q] = u[q] * m;
      v[q] = v[q] * m;
      hu[q] = hu[q] * m;
      hv[q] = hv[q] * m;
    }
  }
}

/* sqrt / transcendental in the working precision */
#define RSQRT(x) ((R)(FWO_IS_DOUBLE ? sqrt((double)(x)) : (double)sqrtf((float)(x))))
#define RTANH(x) ((R)(FWO_IS_DOUBLE ? tanh((double)(x)) : (double)tanhf((float)(x))))
#define REXP(x) ((R)(FWO_IS_DOUBLE ? exp((double)(x)) : (double)expf((float)(x))))
#define RSIN(x) ((R)(FWO_IS_DOUBLE ? sin((double)(x)) : (double)sinf((float)(x))))
#define RCOS(x) ((R)(FWO_IS_DOUBLE ? cos((double)(x)) : (double)cosf((float)(x))))
#define RATAN2(y, x) ((R)(FWO_IS_DOUBLE ? atan2((double)(y), (double)(x)) : (double)atan2f((float)(y), (float)(x))))
#define RABS(x) ((R)(FWO_IS_DOUBLE ? fabs((double)(x)) : (double)fabsf((float)(x))))

/* ESTIMATE_DT, misc.F:216-296: min over ALL cells including ghosts */
void fwo_estimate_dt(const fwo_params *p, fwo_fields *f, double *dt, double *time) {
  DIMS;
  const R *u = F(U), *v = F(V), *h = F(H);
  const R dx = (R)p->DX, dy = (R)p->DY, mdf = (R)p->MinDepthFrc;
  R t3 = LARGE;
  for (int j = 1; j <= N; j++)
    for (int i = 1; i <= M; i++) {
      R c = RSQRT(GRAV * rmax(C_(h, i, j), mdf));
      R t1 = RABS(C_(u, i, j)) + c;
      R t2 = (t1 < SMALL) ? dx / SMALL : dx / t1;
      if (t2 < t3) t3 = t2;
      t1 = RABS(C_(v, i, j)) + c;
      t2 = (t1 < SMALL) ? dy / SMALL : dy / t1;
      if (t2 < t3) t3 = t2;
    }
  R dtt = (R)p->CFL * t3;
  R d;
  if (p->FIXED_DT) {
    d = (R)p->DT_fixed;
    while (d > dtt) d = d / SPC(2.0);
  } else {
    d = dtt;
  }
  *dt = (double)d;
  *time = (double)((R)*time + d);
}

/* ------------------------------------------------------------------ */
/* DERIVATIVE_*, derivatives.F:123-289: interior only, times MASK(=MASK9) */
static void deriv_x(const fwo_params *p, const int *mk, R dx, const R *a, R *o) {
  DIMS;
  for (int j = Jb; j <= Je; j++)
    for (int i = Ib; i <= Ie; i++)
      C_(o, i, j) = (C_(a, i + 1, j) - C_(a, i - 1, j)) / dx / SPC(2.0) * (R)C_(mk, i, j);
}
static void deriv_y(const fwo_params *p, const int *mk, R dy, const R *a, R *o) {
  DIMS;
  for (int j = Jb; j <= Je; j++)
    for (int i = Ib; i <= Ie; i++)
      C_(o, i, j) = (C_(a, i, j + 1) - C_(a, i, j - 1)) / dy / SPC(2.0) * (R)C_(mk, i, j);
}
static void deriv_xx(const fwo_params *p, const int *mk, R dx, const R *a, R *o) {
  DIMS;
  for (int j = Jb; j <= Je; j++)
    for (int i = Ib; i <= Ie; i++)
      C_(o, i, j) = (C_(a, i + 1, j) - SPC(2.0) * C_(a, i, j) + C_(a, i - 1, j)) / dx / dx * (R)C_(mk, i, j);
}
static void deriv_yy(const fwo_params *p, const int *mk, R dy, const R *a, R *o) {
  DIMS;
  for (int j = Jb; j <= Je; j++)
    for (int i = Ib; i <= Ie; i++)
      C_(o, i, j) = (C_(a, i, j + 1) - SPC(2.0) * C_(a, i, j) + C_(a, i, j - 1)) / dy / dy * (R)C_(mk, i, j);
}
static void deriv_xy(const fwo_params *p, const int *mk, R dx, R dy, const R *a, R *o) {
  DIMS;
  for (int j = Jb; j <= Je; j++)
    for (int i = Ib; i <= Ie; i++) {
      R t1 = (C_(a, i + 1, j + 1) - C_(a, i + 1, j - 1)) / SPC(2.0) / dy;
      R t2 = (C_(a, i - 1, j + 1) - C_(a, i - 1, j - 1)) / SPC(2.0) / dy;
      C_(o, i, j) = (t1 - t2) / SPC(2.0) / dx * (R)C_(mk, i, j);
    }
}

/* EXCHANGE_DISPERSION, bc.F:306-414 (Cartesian, DISP_TIME_LEFT=F) */
static void exchange_dispersion(const fwo_params *p, fwo_fields *f) {
  phi_coll(p, F(Uxx), 2);
  phi_coll(p, F(DUxx), 2);
  phi_coll(p, F(Vyy), 3);
  phi_coll(p, F(DVyy), 3);
  phi_coll(p, F(Uxy), 1);
  phi_coll(p, F(DUxy), 1);
  phi_coll(p, F(Vxy), 1);
  phi_coll(p, F(DVxy), 1);
  if (p->Gamma2 > 0) {
    phi_coll(p, F(Ut), 2);
    phi_coll(p, F(Vt), 3);
    phi_coll(p, F(Utx), 1);
    phi_coll(p, F(Vty), 1);
    phi_coll(p, F(Utxx), 2);
    phi_coll(p, F(Vtyy), 3);
    phi_coll(p, F(Utxy), 1);
    phi_coll(p, F(Vtxy), 1);
    phi_coll(p, F(DUtxx), 2);
    phi_coll(p, F(DVtyy), 3);
    phi_coll(p, F(DUtxy), 1);
    phi_coll(p, F(DVtxy), 1);
    phi_coll(p, F(Ux), 1);
    phi_coll(p, F(DUx), 1);
    phi_coll(p, F(Vy), 1);
    phi_coll(p, F(DVy), 1);
    phi_coll(p, F(Uy), 3);
    phi_coll(p, F(DUy), 3);
    phi_coll(p, F(Vx), 2);
    phi_coll(p, F(DVx), 2);
    phi_coll(p, F(ETAx), 2);
    phi_coll(p, F(ETAy), 3);
  }
}

/* CAL_DISPERSION, dispersion.F:62-506 (Cartesian) */
void fwo_dispersion(const fwo_params *p, fwo_fields *f, double dt_) {
  DIMS;
  const int *m9 = FI(MASK9);
  const R dx = (R)p->DX, dy = (R)p->DY, DT = (R)dt_;
  const R mdf = (R)p->MinDepthFrc, g2 = (R)p->Gamma2;
  const R B1 = (R)p->Beta_1, B2 = (R)p->Beta_2, b2 = (R)p->b2;
  R *U = F(U), *V = F(V), *Eta = F(Eta), *Depth = F(Depth);
  size_t ncell = (size_t)M * N;
  R *DU = (R *)calloc(ncell, sizeof(R)), *DV = (R *)calloc(ncell, sizeof(R));
  R *DUt = (R *)calloc(ncell, sizeof(R)), *DVt = (R *)calloc(ncell, sizeof(R));

  deriv_xx(p, m9, dx, U, F(Uxx));                       /* :77 */
  deriv_xy(p, m9, dx, dy, U, F(Uxy));                   /* :79 */
  deriv_xy(p, m9, dx, dy, V, F(Vxy));                   /* :81 */
  deriv_yy(p, m9, dy, V, F(Vyy));                       /* :83 */
  if (g2 > ZERO) {                                      /* :87-94 */
    deriv_x(p, m9, dx, U, F(Ux));
    deriv_x(p, m9, dx, V, F(Vx));
    deriv_y(p, m9, dy, U, F(Uy));
    deriv_y(p, m9, dy, V, F(Vy));
    deriv_x(p, m9, dx, Eta, F(ETAx));
    deriv_y(p, m9, dy, Eta, F(ETAy));
  } else if (p->SHOW_BREAKING) {                        /* :96-99 */
    deriv_x(p, m9, dx, Eta, F(ETAx));
    deriv_y(p, m9, dy, Eta, F(ETAy));
  }
  /* DU, DV, ETAT on 1..Mloc-1 x 1..Nloc-1, :103-119 */
  for (int j = 1; j <= N - 1; j++)
    for (int i = 1; i <= M - 1; i++) {
      R d = rmax(C_(Depth, i, j), mdf);
      C_(DU, i, j) = d * C_(U, i, j);
      C_(DV, i, j) = d * C_(V, i, j);
      C_(F(ETAT), i, j) = -((X_(F(P), i + 1, j) - X_(F(P), i, j)) / dx)
                          - (Y_(F(Q), i, j + 1) - Y_(F(Q), i, j)) / dy;
    }
  if (g2 > ZERO) {                                      /* :133-142 */
    for (int j = 1; j <= N; j++)
      for (int i = 1; i <= M; i++) {
        C_(F(Ut), i, j) = (C_(U, i, j) - C_(F(U0), i, j)) / DT;
        C_(F(Vt), i, j) = (C_(V, i, j) - C_(F(V0), i, j)) / DT;
        R d = rmax(C_(Depth, i, j), mdf);
        C_(DUt, i, j) = d * C_(F(Ut), i, j);
        C_(DVt, i, j) = d * C_(F(Vt), i, j);
      }
  }
  deriv_xx(p, m9, dx, DU, F(DUxx));                     /* :152-158 */
  deriv_xy(p, m9, dx, dy, DU, F(DUxy));
  deriv_xy(p, m9, dx, dy, DV, F(DVxy));
  deriv_yy(p, m9, dy, DV, F(DVyy));
  if (g2 > ZERO) {                                      /* :161-182 */
    deriv_x(p, m9, dx, DU, F(DUx));
    deriv_x(p, m9, dx, DV, F(DVx));
    deriv_y(p, m9, dy, DU, F(DUy));
    deriv_y(p, m9, dy, DV, F(DVy));
    deriv_x(p, m9, dx, F(Ut), F(Utx));
    deriv_y(p, m9, dy, F(Vt), F(Vty));
    deriv_xx(p, m9, dx, F(Ut), F(Utxx));
    deriv_yy(p, m9, dy, F(Vt), F(Vtyy));
    deriv_xy(p, m9, dx, dy, F(Ut), F(Utxy));
    deriv_xy(p, m9, dx, dy, F(Vt), F(Vtxy));
    deriv_x(p, m9, dx, DUt, F(DUtx));
    deriv_y(p, m9, dy, DVt, F(DVty));
    deriv_xx(p, m9, dx, DUt, F(DUtxx));
    deriv_yy(p, m9, dy, DVt, F(DVtyy));
    deriv_xy(p, m9, dx, dy, DUt, F(DUtxy));
    deriv_xy(p, m9, dx, dy, DVt, F(DVtxy));
  }
  /* zero cross derivatives on the physical boundaries, :189-287 */
  {
    R *z[8] = {F(Uxy), F(DUxy), F(Vxy), F(DVxy), F(Utxy), F(DUtxy), F(Vtxy), F(DVtxy)};
    for (int q = 0; q < 8; q++) {
      for (int j = 1; j <= N; j++) C_(z[q], Ib, j) = ZERO;
      for (int j = 1; j <= N; j++) C_(z[q], Ie, j) = ZERO;
      for (int i = 1; i <= M; i++) C_(z[q], i, Jb) = ZERO;
      for (int i = 1; i <= M; i++) C_(z[q], i, Je) = ZERO;
    }
  }
  exchange_dispersion(p, f);                            /* :289 */

  /* U4, V4, U1p, V1p on all cells, :293-330 */
  const R c13 = SPC(1.0) / SPC(3.0), c12 = SPC(1.0) / SPC(2.0), c16 = SPC(1.0) / SPC(6.0);
  for (int j = 1; j <= N; j++)
    for (int i = 1; i <= M; i++) {
      R dep = C_(Depth, i, j);
      R uxx = C_(F(Uxx), i, j), vxy = C_(F(Vxy), i, j), uxy = C_(F(Uxy), i, j), vyy = C_(F(Vyy), i, j);
      R duxx = C_(F(DUxx), i, j), dvxy = C_(F(DVxy), i, j), duxy = C_(F(DUxy), i, j), dvyy = C_(F(DVyy), i, j);
      R k0 = c13 - B1 + SPC(0.5) * B1 * B1;
      R u4 = k0 * dep * dep * (uxx + vxy) + (B1 - c12) * dep * (duxx + dvxy);
      R v4 = k0 * dep * dep * (uxy + vyy) + (B1 - c12) * dep * (duxy + dvyy);
      if (g2 > ZERO) {
        R UxxVxy = uxx + vxy, UxyVyy = uxy + vyy;
        R HUxxHVxy = duxx + dvxy, HUxyHVyy = duxy + dvyy;
        R rh = dep, reta = C_(Eta, i, j);
        R ken1 = (c16 - B1 + B1 * B1) * rh * reta * B2
               + (c12 * B1 * B1 - c16) * reta * reta * B2 * B2;
        R ken2 = (B1 - c12) * reta * B2;
        u4 = u4 + g2 * (R)C_(m9, i, j) * (ken1 * UxxVxy + ken2 * HUxxHVxy);
        v4 = v4 + g2 * (R)C_(m9, i, j) * (ken1 * UxyVyy + ken2 * HUxyHVyy);
      }
      C_(F(U4), i, j) = u4;
      C_(F(V4), i, j) = v4;
      C_(F(U1p), i, j) = SPC(0.5) * (SPC(1.0) - B1) * (SPC(1.0) - B1) * dep * dep * (uxx + vxy)
                       + (B1 - SPC(1.0)) * dep * (duxx + dvxy);
      C_(F(V1p), i, j) = SPC(0.5) * (SPC(1.0) - B1) * (SPC(1.0) - B1) * dep * dep * (uxy + vyy)
                       + (B1 - SPC(1.0)) * dep * (duxy + dvyy);
    }

  /* nonlinear dispersive terms, :398-503 */
  if (g2 > ZERO) {
    for (int j = Jb; j <= Je; j++)
      for (int i = Ib; i <= Ie; i++) {
#define A(n, ii, jj) C_(F(n), ii, jj)
        R UxxVxy = A(Uxx, i, j) + A(Vxy, i, j);
        R UxyVyy = A(Uxy, i, j) + A(Vyy, i, j);
        R UxxVxy_x = (A(Uxx, i + 1, j) + A(Vxy, i + 1, j) - A(Uxx, i - 1, j) - A(Vxy, i - 1, j)) / SPC(2.0) / dx;
        R UxxVxy_y = (A(Uxx, i, j + 1) + A(Vxy, i, j + 1) - A(Uxx, i, j - 1) - A(Vxy, i, j - 1)) / SPC(2.0) / dy;
        R UxyVyy_x = (A(Uxy, i + 1, j) + A(Vyy, i + 1, j) - A(Uxy, i - 1, j) - A(Vyy, i - 1, j)) / SPC(2.0) / dx;
        R UxyVyy_y = (A(Uxy, i, j + 1) + A(Vyy, i, j + 1) - A(Uxy, i, j - 1) - A(Vyy, i, j - 1)) / SPC(2.0) / dy;
        R HUxxHVxy = A(DUxx, i, j) + A(DVxy, i, j);
        R HUxyHVyy = A(DUxy, i, j) + A(DVyy, i, j);
        R HUxxHVxy_x = (A(DUxx, i + 1, j) + A(DVxy, i + 1, j) - A(DUxx, i - 1, j) - A(DVxy, i - 1, j)) / SPC(2.0) / dx;
        R HUxxHVxy_y = (A(DUxx, i, j + 1) + A(DVxy, i, j + 1) - A(DUxx, i, j - 1) - A(DVxy, i, j - 1)) / SPC(2.0) / dy;
        R HUxyHVyy_x = (A(DUxy, i + 1, j) + A(DVyy, i + 1, j) - A(DUxy, i - 1, j) - A(DVyy, i - 1, j)) / SPC(2.0) / dx;
        R HUxyHVyy_y = (A(DUxy, i, j + 1) + A(DVyy, i, j + 1) - A(DUxy, i, j - 1) - A(DVyy, i, j - 1)) / SPC(2.0) / dy;
        R rh = A(Depth, i, j);
        R rhx = (A(Depth, i + 1, j) - A(Depth, i - 1, j)) / SPC(2.0) / dx;
        R rhy = (A(Depth, i, j + 1) - A(Depth, i, j - 1)) / SPC(2.0) / dy;
        R reta = A(Eta, i, j);
        R ex = A(ETAx, i, j), ey = A(ETAy, i, j), et = A(ETAT, i, j);
        R Utx_ = A(Utx, i, j), Vty_ = A(Vty, i, j);
        R u1pp = -reta * B2 * ex * B2 * (Utx_ + Vty_) - SPC(0.5) * reta * reta * B2 * B2 * (A(Utxx, i, j) + A(Vtxy, i, j))
                 - ex * B2 * (A(DUtx, i, j) + A(DVty, i, j)) - reta * B2 * (A(DUtxx, i, j) + A(DVtxy, i, j));
        R v1pp = -reta * B2 * ey * B2 * (Utx_ + Vty_) - SPC(0.5) * reta * reta * B2 * B2 * (A(Utxy, i, j) + A(Vtyy, i, j))
                 - ey * B2 * (A(DUtx, i, j) + A(DVty, i, j)) - reta * B2 * (A(DUtxy, i, j) + A(DVtyy, i, j));
        R ken1 = B1 * (SPC(1.0) - B1) * rh * et * B2 - B1 * B1 * reta * B2 * et * B2;
        R ken2 = B1 * (SPC(1.0) - B1) * rh * reta * B2 - FLC(0.5) * B1 * B1 * reta * reta * B2 * B2;
        R ken3 = B1 * et * B2;
        R ken4 = B1 * reta * B2;
        u1pp = u1pp - ken1 * UxxVxy - ken2 * (A(Utxx, i, j) + A(Vtxy, i, j)) + ken3 * HUxxHVxy
               + ken4 * (A(DUtxx, i, j) + A(DVtxy, i, j));
        v1pp = v1pp - ken1 * UxyVyy - ken2 * (A(Utxy, i, j) + A(Vtyy, i, j)) + ken3 * HUxyHVyy
               + ken4 * (A(DUtxy, i, j) + A(DVtyy, i, j));
        A(U1pp, i, j) = u1pp;
        A(V1pp, i, j) = v1pp;
        R omB = SPC(1.0) - B1;
        ken1 = (B1 - SPC(1.0)) * (rhx + ex) * B2;
        ken2 = (B1 - SPC(1.0)) * (rh + reta) * B2;
        ken3 = (omB * omB * rh * rhx * B2 * B2 - B1 * omB * (rhx * reta * B2 + rh * ex * B2)
                + (B1 * B1 - SPC(1.0)) * reta * ex * B2 * B2);
        ken4 = (SPC(0.5) * omB * omB * rh * rh * B2 * B2 - B1 * omB * rh * reta * B2
                + SPC(0.5) * (B1 * B1 - SPC(1.0)) * reta * reta * B2 * B2);
        R ken5 = (omB * omB * rh * rhy * B2 * B2 - B1 * omB * (rhy * reta * B2 + rh * ey * B2)
                  + (B1 * B1 - SPC(1.0)) * reta * ey * B2 * B2);
        R u = A(U, i, j), v = A(V, i, j);
        R ux = A(Ux, i, j), vx = A(Vx, i, j), uy = A(Uy, i, j), vy = A(Vy, i, j);
        R dux = A(DUx, i, j), dvy = A(DVy, i, j);
        A(U2, i, j) = ken1 * (u * HUxxHVxy + v * HUxyHVyy)
                      + ken2 * (ux * HUxxHVxy + u * HUxxHVxy_x + vx * HUxyHVyy + v * HUxyHVyy_x)
                      + ken3 * (u * UxxVxy + v * UxyVyy)
                      + ken4 * (ux * UxxVxy + u * UxxVxy_x + vx * UxyVyy + v * UxyVyy_x)
                      + B2 * B2 * (dux + dvy + reta * B2 * (ux + vy))
                            * (HUxxHVxy + ex * B2 * (ux + vy) + reta * B2 * UxxVxy);
        ken1 = (B1 - SPC(1.0)) * (rhy + ey) * B2;
        A(V2, i, j) = ken1 * (u * HUxxHVxy + v * HUxyHVyy)
                      + ken2 * (uy * HUxxHVxy + u * HUxxHVxy_y + vy * HUxyHVyy + v * HUxyHVyy_y)
                      + ken5 * (u * UxxVxy + v * UxyVyy)
                      + ken4 * (uy * UxxVxy + u * UxxVxy_y + vy * UxyVyy + v * UxyVyy_y)
                      + B2 * B2 * (dux + dvy + reta * B2 * (ux + vy))
                            * (HUxyHVyy + ey * B2 * (ux + vy) + reta * B2 * UxyVyy);
        R omega_0 = vx - uy;
        R omega_1 = (b2 * rhx + B1 * ex) * B2 * (HUxyHVyy + (b2 * rh + B1 * reta) * B2 * UxyVyy)
                    - (b2 * rhy + B1 * ey) * B2 * (HUxxHVxy + (b2 * rh + B1 * reta) * B2 * UxxVxy);
        ken1 = ((B1 - c12) * (reta + rh) * B2);
        ken2 = (c13 - B1 + SPC(0.5) * B1 * B1) * rh * rh * B2 * B2
               + (c16 - B1 + B1 * B1) * rh * reta * B2
               + (c12 * B1 * B1 - c16) * reta * reta * B2 * B2;
        A(U3, i, j) = -v * omega_1 - omega_0 * (ken1 * HUxyHVyy + ken2 * UxyVyy);
        A(V3, i, j) = u * omega_1 + omega_0 * (ken1 * HUxxHVxy + ken2 * UxxVxy);
#undef A
      }
  }
  free(DU); free(DV); free(DUt); free(DVt);
}

/* ------------------------------------------------------------------ */
/* minmod pieces of CONSTRUCT_HO_X/Y, fluxes.F:1550-1598 */
static R mm3(R a, R b, R c) {
  /* a>=0: MAX(0, MIN(a, 2b, 2c)); else MIN(0, MAX(a, 2b, 2c)) */
  R tb = FLC(2.0) * b, tc = FLC(2.0) * c;
  if (a >= ZERO) return rmax(ZERO, rmin(rmin(a, tb), tc));
  return rmin(ZERO, rmax(rmax(a, tb), tc));
}
static R din_val(R vm2, R vm1, R v0, R vp1, int mm2, int mm1, int m0, int mp1) {
  R t1 = vm1 - vm2, t2 = v0 - vm1, t3 = vp1 - v0;
  R d1 = mm3(t1, t2, t3), d2 = mm3(t2, t3, t1), d3 = mm3(t3, t1, t2);
  if (mm2 == 0 || mp1 == 0) {
    t2 = v0 - vm1;
    t1 = t2;
    t3 = t2;
    d1 = mm3(t1, t2, t3);
    d2 = mm3(t2, t3, t1);
    d3 = mm3(t3, t1, t2);
  }
  if (mm1 == 0 || m0 == 0) {
    d1 = ZERO; d2 = ZERO; d3 = ZERO;
  }
  return t2 - SPC(1.0) / SPC(6.0) * (d3 - SPC(2.0) * d2 + d1);
}
/* van Leer weights, fluxes.F:1605-1635 */
static void vanleer(R a, R b, R *van1, R *van2, R *ta, R *tb) {
  if (RABS(a) <= SMALL) a = SMALL * rsign1(a);
  if (RABS(b) <= SMALL) b = SMALL * rsign1(b);
  R rat = b / a;
  R v1 = ZERO;
  if (RABS(FLC(1.0) + rat) > SMALL) v1 = (rat + RABS(rat)) / (SPC(1.0) + rat);
  rat = a / b;
  R v2 = ZERO;
  if (RABS(FLC(1.0) + rat) > SMALL) v2 = (rat + RABS(rat)) / (SPC(1.0) + rat);
  *van1 = v1; *van2 = v2; *ta = a; *tb = b;
}

/* CONSTRUCT_HO_X, fluxes.F:1531-1639 */
static void construct_x(const fwo_params *p, const int *mk, const R *vin, R *outl, R *outr) {
  DIMS;
  R *din = (R *)calloc((size_t)M * N, sizeof(R));
  for (int j = Jb; j <= Je; j++) {
    for (int i = Ib - 1; i <= Ie + 2; i++)
      C_(din, i, j) = din_val(C_(vin, i - 2, j), C_(vin, i - 1, j), C_(vin, i, j), C_(vin, i + 1, j),
                              C_(mk, i - 2, j), C_(mk, i - 1, j), C_(mk, i, j), C_(mk, i + 1, j));
    for (int i = Ib; i <= Ie + 1; i++) {
      R v1, v2, a, b;
      vanleer(C_(din, i - 1, j), C_(din, i, j), &v1, &v2, &a, &b);
      X_(outl, i, j) = C_(vin, i - 1, j) + (SPC(1.0) / SPC(6.0)) * (v1 * a + SPC(2.0) * v2 * b);
      vanleer(C_(din, i, j), C_(din, i + 1, j), &v1, &v2, &a, &b);
      X_(outr, i, j) = C_(vin, i, j) - (SPC(1.0) / SPC(6.0)) * (SPC(2.0) * v1 * a + v2 * b);
    }
  }
  free(din);
}
/* CONSTRUCT_HO_Y, fluxes.F:1646-1754 */
static void construct_y(const fwo_params *p, const int *mk, const R *vin, R *outl, R *outr) {
  DIMS;
  R *din = (R *)calloc((size_t)M * N, sizeof(R));
  for (int i = Ib; i <= Ie; i++) {
    for (int j = Jb - 1; j <= Je + 2; j++)
      C_(din, i, j) = din_val(C_(vin, i, j - 2), C_(vin, i, j - 1), C_(vin, i, j), C_(vin, i, j + 1),
                              C_(mk, i, j - 2), C_(mk, i, j - 1), C_(mk, i, j), C_(mk, i, j + 1));
    for (int j = Jb; j <= Je + 1; j++) {
      R v1, v2, a, b;
      vanleer(C_(din, i, j - 1), C_(din, i, j), &v1, &v2, &a, &b);
      Y_(outl, i, j) = C_(vin, i, j - 1) + (SPC(1.0) / SPC(6.0)) * (v1 * a + SPC(2.0) * v2 * b);
      vanleer(C_(din, i, j), C_(din, i, j + 1), &v1, &v2, &a, &b);
      Y_(outr, i, j) = C_(vin, i, j) - (SPC(1.0) / SPC(6.0)) * (SPC(2.0) * v1 * a + v2 * b);
    }
  }
  free(din);
}

/* HLL, fluxes.F:171-194 */
static R hll(R sl, R sr, R fl, R fr, R ul, R ur) {
  if (sl >= ZERO) return fl;
  if (sr <= ZERO) return fr;
  R o = sr * fl - sl * fr + sl * sr * (ur - ul);
  if (RABS(sr - sl) < SMALL) return o / SMALL;
  return o / (sr - sl);
}

/* FLUXES, fluxes.F:52-117 -> CONSTRUCTION_HO :1324-1524, WAVE_SPEED
 * :352-439, FLUX_AT_INTERFACE_HLL :150-161, BOUNDARY_CONDITION bc.F:53-295 */
void fwo_fluxes(const fwo_params *p, fwo_fields *f) {
  DIMS;
  const int *mk = FI(MASK), *m9 = FI(MASK9);
  const R g1 = (R)p->Gamma1, g3 = (R)p->Gamma3;
  /* x faces */
  construct_x(p, mk, F(U), F(UxL), F(UxR));
  construct_x(p, mk, F(V), F(VxL), F(VxR));
  construct_x(p, mk, F(HU), F(HUxL), F(HUxR));
  construct_x(p, mk, F(HV), F(HVxL), F(HVxR));
  construct_x(p, mk, F(Eta), F(EtaRxL), F(EtaRxR));
  for (int j = 1; j <= N; j++)
    for (int i = 1; i <= M1; i++) {
      X_(F(HxL), i, j) = X_(F(EtaRxL), i, j) + X_(F(DepthX), i, j);
      X_(F(HxR), i, j) = X_(F(EtaRxR), i, j) + X_(F(DepthX), i, j);
    }
  if (p->DISPERSION) {
    construct_x(p, mk, F(U4), F(U4xL), F(U4xR));
    construct_x(p, mk, F(V4), F(V4xL), F(V4xR));
  }
  for (int j = 1; j <= N; j++)
    for (int i = 1; i <= M1; i++) {
      R m9u = (R)C_(m9, (i <= M ? i : M), j);
#define XF(n) X_(F(n), i, j)
      XF(PL) = XF(HUxL) + g1 * m9u * XF(HxL) * XF(U4xL);
      XF(PR) = XF(HUxR) + g1 * m9u * XF(HxR) * XF(U4xR);
      XF(FxL) = g3 * XF(PL) * (XF(UxL) + g1 * m9u * XF(U4xL))
                + FLC(0.5) * GRAV * ((XF(EtaRxL)) * (XF(EtaRxL)) * g3 + SPC(2.0) * (XF(EtaRxL)) * (XF(DepthX)));
      XF(FxR) = g3 * XF(PR) * (XF(UxR) + g1 * m9u * XF(U4xR))
                + FLC(0.5) * GRAV * ((XF(EtaRxR)) * (XF(EtaRxR)) * g3 + SPC(2.0) * (XF(EtaRxR)) * (XF(DepthX)));
      XF(GxL) = XF(HxL) * (XF(UxL) + g1 * m9u * XF(U4xL)) * (XF(VxL) + g1 * m9u * XF(V4xL)) * g3;
      XF(GxR) = XF(HxR) * (XF(UxR) + g1 * m9u * XF(U4xR)) * (XF(VxR) + g1 * m9u * XF(V4xR)) * g3;
#undef XF
    }
  /* y faces */
  construct_y(p, mk, F(U), F(UyL), F(UyR));
  construct_y(p, mk, F(V), F(VyL), F(VyR));
  construct_y(p, mk, F(HV), F(HVyL), F(HVyR));
  construct_y(p, mk, F(HU), F(HUyL), F(HUyR));
  construct_y(p, mk, F(Eta), F(EtaRyL), F(EtaRyR));
  for (int j = 1; j <= N1; j++)
    for (int i = 1; i <= M; i++) {
      Y_(F(HyL), i, j) = Y_(F(EtaRyL), i, j) + Y_(F(DepthY), i, j);
      Y_(F(HyR), i, j) = Y_(F(EtaRyR), i, j) + Y_(F(DepthY), i, j);
    }
  if (p->DISPERSION) {
    construct_y(p, mk, F(V4), F(V4yL), F(V4yR));
    construct_y(p, mk, F(U4), F(U4yL), F(U4yR));
  }
  for (int j = 1; j <= N1; j++)
    for (int i = 1; i <= M; i++) {
      R m9v = (R)C_(m9, i, (j <= N ? j : N));
#define YF(n) Y_(F(n), i, j)
      YF(QL) = YF(HVyL) + g1 * m9v * YF(HyL) * YF(V4yL);
      YF(QR) = YF(HVyR) + g1 * m9v * YF(HyR) * YF(V4yR);
      YF(GyL) = g3 * YF(QL) * (YF(VyL) + g1 * m9v * YF(V4yL))
                + FLC(0.5) * GRAV * ((YF(EtaRyL)) * (YF(EtaRyL)) * g3 + SPC(2.0) * (YF(EtaRyL)) * (YF(DepthY)));
      YF(GyR) = g3 * YF(QR) * (YF(VyR) + g1 * m9v * YF(V4yR))
                + FLC(0.5) * GRAV * ((YF(EtaRyR)) * (YF(EtaRyR)) * g3 + SPC(2.0) * (YF(EtaRyR)) * (YF(DepthY)));
      YF(FyL) = YF(HyL) * (YF(UyL) + g1 * m9v * YF(U4yL)) * (YF(VyL) + g1 * m9v * YF(V4yL)) * g3;
      YF(FyR) = YF(HyR) * (YF(UyR) + g1 * m9v * YF(U4yR)) * (YF(VyR) + g1 * m9v * YF(V4yR)) * g3;
#undef YF
    }
  /* WAVE_SPEED, fluxes.F:352-439 */
  {
    R *sxl = F(SxL), *sxr = F(SxR), *syl = F(SyL), *syr = F(SyR);
    for (int j = 4; j <= N - 3; j++)
      for (int i = 4; i <= M1 - 3; i++) {
        R ul = X_(F(UxL), i, j), ur = X_(F(UxR), i, j);
        R spl = RSQRT(GRAV * RABS(X_(F(HxL), i, j)));
        R spr = RSQRT(GRAV * RABS(X_(F(HxR), i, j)));
        R sps = FLC(0.5) * (spl + spr) + FLC(0.25) * (ul - ur);
        R us = FLC(0.5) * (ul + ur) + spl - spr;
        X_(sxl, i, j) = rmin(ul - spl, us - sps);
        X_(sxr, i, j) = rmax(ur + spr, us + sps);
      }
    for (int j = 4; j <= N - 3; j++) {
      for (int i = 1; i <= 3; i++) { X_(sxl, i, j) = X_(sxl, 4, j); X_(sxr, i, j) = X_(sxr, 4, j); }
      for (int i = M1 - 2; i <= M1; i++) { X_(sxl, i, j) = X_(sxl, M1 - 3, j); X_(sxr, i, j) = X_(sxr, M1 - 3, j); }
    }
    for (int i = 1; i <= M1; i++) {
      for (int j = 1; j <= 3; j++) { X_(sxl, i, j) = X_(sxl, i, 4); X_(sxr, i, j) = X_(sxr, i, 4); }
      for (int j = N - 2; j <= N; j++) { X_(sxl, i, j) = X_(sxl, i, N - 3); X_(sxr, i, j) = X_(sxr, i, N - 3); }
    }
    for (int j = 4; j <= N1 - 3; j++)
      for (int i = 4; i <= M - 3; i++) {
        R vl = Y_(F(VyL), i, j), vr = Y_(F(VyR), i, j);
        R spl = RSQRT(GRAV * RABS(Y_(F(HyL), i, j)));
        R spr = RSQRT(GRAV * RABS(Y_(F(HyR), i, j)));
        R sps = FLC(0.5) * (spl + spr) + FLC(0.25) * (vl - vr);
        R us = FLC(0.5) * (vl + vr) + spl - spr;
        Y_(syl, i, j) = rmin(vl - spl, us - sps);
        Y_(syr, i, j) = rmax(vr + spr, us + sps);
      }
    for (int i = 4; i <= M - 3; i++) {
      for (int j = 1; j <= 3; j++) { Y_(syl, i, j) = Y_(syl, i, 4); Y_(syr, i, j) = Y_(syr, i, 4); }
      for (int j = N1 - 2; j <= N1; j++) { Y_(syl, i, j) = Y_(syl, i, N1 - 3); Y_(syr, i, j) = Y_(syr, i, N1 - 3); }
    }
    for (int j = 1; j <= N1; j++) {
      for (int i = 1; i <= 3; i++) { Y_(syl, i, j) = Y_(syl, 4, j); Y_(syr, i, j) = Y_(syr, 4, j); }
      for (int i = M - 2; i <= M; i++) { Y_(syl, i, j) = Y_(syl, M - 3, j); Y_(syr, i, j) = Y_(syr, M - 3, j); }
    }
  }
  /* FLUX_AT_INTERFACE_HLL over all faces */
  for (int j = 1; j <= N; j++)
    for (int i = 1; i <= M1; i++) {
      R sl = X_(F(SxL), i, j), sr = X_(F(SxR), i, j);
      X_(F(P), i, j) = hll(sl, sr, X_(F(PL), i, j), X_(F(PR), i, j), X_(F(EtaRxL), i, j), X_(F(EtaRxR), i, j));
      X_(F(Fx), i, j) = hll(sl, sr, X_(F(FxL), i, j), X_(F(FxR), i, j), X_(F(HUxL), i, j), X_(F(HUxR), i, j));
      X_(F(Gx), i, j) = hll(sl, sr, X_(F(GxL), i, j), X_(F(GxR), i, j), X_(F(HVxL), i, j), X_(F(HVxR), i, j));
    }
  for (int j = 1; j <= N1; j++)
    for (int i = 1; i <= M; i++) {
      R sl = Y_(F(SyL), i, j), sr = Y_(F(SyR), i, j);
      Y_(F(Q), i, j) = hll(sl, sr, Y_(F(QL), i, j), Y_(F(QR), i, j), Y_(F(EtaRyL), i, j), Y_(F(EtaRyR), i, j));
      Y_(F(Fy), i, j) = hll(sl, sr, Y_(F(FyL), i, j), Y_(F(FyR), i, j), Y_(F(HUyL), i, j), Y_(F(HUyR), i, j));
      Y_(F(Gy), i, j) = hll(sl, sr, Y_(F(GyL), i, j), Y_(F(GyR), i, j), Y_(F(HVyL), i, j), Y_(F(HVyR), i, j));
    }
  /* BOUNDARY_CONDITION, bc.F:53-295 (serial, no coupling) */
  {
    R *P = F(P), *Fx = F(Fx), *Gx = F(Gx), *Q = F(Q), *Fy = F(Fy), *Gy = F(Gy);
    R *Dx = F(DepthX), *Dy = F(DepthY);
#define HYD(xi, de) (SPC(0.5) * GRAV * ((xi) * (xi) * g3 + SPC(2.0) * (xi) * (de)))
    for (int j = Jb; j <= Je; j++) {
      R xi = X_(F(EtaRxR), Ib, j), de = X_(Dx, Ib, j);
      X_(P, Ib, j) = ZERO;
      X_(Fx, Ib, j) = HYD(xi, de);
      X_(Gx, Ib, j) = ZERO;
    }
    for (int j = Jb; j <= Je; j++) {
      R xi = X_(F(EtaRxL), Ie + 1, j), de = X_(Dx, Ie + 1, j);
      X_(P, Ie + 1, j) = ZERO;
      X_(Fx, Ie + 1, j) = HYD(xi, de);
      X_(Gx, Ie + 1, j) = ZERO;
    }
    if (!p->PERIODIC) {
      for (int i = Ib; i <= Ie; i++) {
        Y_(Q, i, Jb) = ZERO;
        Y_(Fy, i, Jb) = ZERO;
        R xi = Y_(F(EtaRyR), i, Jb), de = Y_(Dy, i, Jb);
        Y_(Gy, i, Jb) = HYD(xi, de);
      }
      for (int i = Ib; i <= Ie; i++) {
        Y_(Q, i, Je + 1) = ZERO;
        Y_(Fy, i, Je + 1) = ZERO;
        R xi = Y_(F(EtaRyL), i, Je + 1), de = Y_(Dy, i, Je + 1);
        Y_(Gy, i, Je + 1) = HYD(xi, de);
      }
    }
    for (int j = Jb - 1; j <= Je + 1; j++)
      for (int i = Ib - 1; i <= Ie + 1; i++) {
        if (C_(mk, i, j) < 1) {
          X_(P, i, j) = ZERO;
          if (i != Ib) {
            R xi = X_(F(EtaRxL), i, j), de = X_(Dx, i, j);
            X_(Fx, i, j) = HYD(xi, de) * (R)C_(mk, i - 1, j);
          } else {
            X_(Fx, i, j) = ZERO;
          }
          X_(Gx, i, j) = ZERO;
          X_(P, i + 1, j) = ZERO;
          if (i != Ie) {
            R xi = X_(F(EtaRxR), i + 1, j), de = X_(Dx, i + 1, j);
            X_(Fx, i + 1, j) = HYD(xi, de) * (R)C_(mk, i + 1, j);
          } else {
            X_(Fx, i + 1, j) = ZERO;
          }
          X_(Gx, i + 1, j) = ZERO;
          Y_(Q, i, j) = ZERO;
          Y_(Fy, i, j) = ZERO;
          if (j != Jb) {
            R xi = Y_(F(EtaRyL), i, j), de = Y_(Dy, i, j);
            Y_(Gy, i, j) = HYD(xi, de) * (R)C_(mk, i, j - 1);
          } else {
            Y_(Gy, i, j) = ZERO;
          }
          Y_(Q, i, j + 1) = ZERO;
          Y_(Fy, i, j + 1) = ZERO;
          if (j != Je) {
            R xi = Y_(F(EtaRyR), i, j + 1), de = Y_(Dy, i, j + 1);
            Y_(Gy, i, j + 1) = HYD(xi, de) * (R)C_(mk, i, j + 1);
          } else {
            Y_(Gy, i, j + 1) = ZERO;
          }
        }
      }
#undef HYD
  }
}

/* ------------------------------------------------------------------ */
/* SourceTerms, sources.F:56-548 (Cartesian, no MANNING/METEO/VESSEL/MIXING) */
void fwo_sources(const fwo_params *p, fwo_fields *f, double time_) {
  DIMS;
  const R TIME = (R)time_;
  const R dx = (R)p->DX, dy = (R)p->DY, mdf = (R)p->MinDepthFrc;
  const R g1 = (R)p->Gamma1, g2 = (R)p->Gamma2;
  const int *mk = FI(MASK), *m9 = FI(MASK9);
  R *WM = F(WaveMaker_Mass);
  /* wavemaker mass source in the band ilo..ihi x jlo..jhi, :109-184 */
  if (p->wavemaker == FWO_WK_REG) {
    R AA = RTANH(PI / ((R)p->Time_ramp * (R)p->Tperiod) * TIME) * (R)p->D_gen;
    const R *xm = (const R *)f->xmk_wk, *ym = (const R *)f->ymk_wk;
    for (int j = p->jlo; j <= p->jhi; j++)
      for (int i = p->ilo; i <= p->ihi; i++) {
        R d = xm[i - 1] - (R)p->Xc_WK;
        C_(WM, i, j) = AA * REXP(-(R)p->Beta_gen * (d * d))
                       * RSIN((R)p->rlamda * (ym[j - 1] - ZERO) - SPC(2.0) * PI / (R)p->Tperiod * TIME);
      }
  } else if (p->wavemaker == FWO_WK_DATA2D) {
    R AA = RTANH(PI / ((R)p->Time_ramp / (R)p->FreqPeak) * TIME);
    int nf = p->NumFreq;
    const R *om = (const R *)f->OMGN2D, *cm = (const R *)f->Cm, *sm = (const R *)f->Sm;
    R *bb = (R *)malloc(sizeof(R) * (size_t)nf), *cc = (R *)malloc(sizeof(R) * (size_t)nf);
    for (int k = 0; k < nf; k++) {
      bb[k] = RCOS(om[k] * TIME);
      cc[k] = RSIN(om[k] * TIME);
    }
    size_t plane = (size_t)M * N;
    for (int j = p->jlo; j <= p->jhi; j++)
      for (int i = p->ilo; i <= p->ihi; i++) {
        R s = ZERO;
        size_t q = ((size_t)j - 1) * M + (size_t)i - 1;
        for (int k = 0; k < nf; k++) s = s + AA * (cm[q + k * plane] * bb[k] + sm[q + k * plane] * cc[k]);
        C_(WM, i, j) = s;
      }
    free(bb); free(cc);
  }
  /* nu_vis is default REAL (single) even in double builds, :69 */
  size_t ncell = (size_t)M * N;
  float *nuv = (float *)calloc(ncell, sizeof(float));
  if (p->VISCOSITY_BREAKING)
    for (size_t q = 0; q < ncell; q++) nuv[q] = (float)F(nu_break)[q];
  if (p->DIFFUSION_SPONGE)
    for (size_t q = 0; q < ncell; q++) nuv[q] = (float)((R)nuv[q] + F(nu_sponge)[q]);
#define NUV(i, j) nuv[((size_t)(j)-1) * M + (size_t)(i)-1]
  R *U = F(U), *V = F(V), *U4 = F(U4), *V4 = F(V4), *Eta = F(Eta), *H = F(H);
  R *P = F(P), *Q = F(Q), *HU = F(HU), *HV = F(HV);
  for (int j = Jb; j <= Je; j++)
    for (int i = Ib; i <= Ie; i++) {
      R u = C_(U, i, j), v = C_(V, i, j);
      R spd = RSQRT(u * u + v * v);
      R fix = -(C_(F(Cd), i, j) * u * spd);
      C_(F(FrcInsX), i, j) = fix;
      R m9r = (R)C_(m9, i, j);
      R hmax = rmax(C_(H, i, j), mdf);
      R divpq = (X_(P, i + 1, j) - X_(P, i, j)) / dx + (Y_(Q, i, j + 1) - Y_(Q, i, j)) / dy;
      R sx = GRAV * (C_(Eta, i, j)) / dx * (X_(F(DepthX), i + 1, j) - X_(F(DepthX), i, j)) * (R)C_(mk, i, j)
             + fix
             + g1 * m9r * hmax * (u * SPC(0.5) * (C_(U4, i + 1, j) - C_(U4, i - 1, j)) / dx
                                  + v * SPC(0.5) * (C_(U4, i, j + 1) - C_(U4, i, j - 1)) / dy
                                  + C_(U4, i, j) * SPC(0.5) * (C_(U, i + 1, j) - C_(U, i - 1, j)) / dx
                                  + C_(V4, i, j) * SPC(0.5) * (C_(U, i, j + 1) - C_(U, i, j - 1)) / dy
                                  - g2 * m9r * (C_(F(U1pp), i, j) + C_(F(U2), i, j) + C_(F(U3), i, j)))
             + g1 * m9r * divpq * (C_(U4, i, j) - C_(F(U1p), i, j))
             + C_(WM, i, j) * u;
      if (p->FRICTION_SPONGE)
        sx = sx - C_(F(CD_4_SPONGE), i, j) * u * spd * C_(F(Depth), i, j);
      R fiy = -(C_(F(Cd), i, j) * v * spd);
      C_(F(FrcInsY), i, j) = fiy;
      R sy = GRAV * (C_(Eta, i, j)) / dy * (Y_(F(DepthY), i, j + 1) - Y_(F(DepthY), i, j)) * (R)C_(mk, i, j)
             + fiy
             + g1 * m9r * hmax * (u * SPC(0.5) * (C_(V4, i + 1, j) - C_(V4, i - 1, j)) / dx
                                  + v * SPC(0.5) * (C_(V4, i, j + 1) - C_(V4, i, j - 1)) / dy
                                  + C_(U4, i, j) * SPC(0.5) * (C_(V, i + 1, j) - C_(V, i - 1, j)) / dx
                                  + C_(V4, i, j) * SPC(0.5) * (C_(V, i, j + 1) - C_(V, i, j - 1)) / dy
                                  - g2 * m9r * (C_(F(V1pp), i, j) + C_(F(V2), i, j) + C_(F(V3), i, j)))
             + g1 * m9r * divpq * (C_(V4, i, j) - C_(F(V1p), i, j))
             + C_(WM, i, j) * v;
      if (p->FRICTION_SPONGE)
        sy = sy - C_(F(CD_4_SPONGE), i, j) * v * spd * C_(F(Depth), i, j);
      if (p->VISCOSITY_BREAKING || p->DIFFUSION_SPONGE) {
        /* breaking diffusion, :439-458; (nu+nu) is summed in single precision */
        R sxp = (R)(NUV(i + 1, j) + NUV(i, j)), sxm = (R)(NUV(i - 1, j) + NUV(i, j));
        R syp = (R)(NUV(i, j + 1) + NUV(i, j)), sym = (R)(NUV(i, j - 1) + NUV(i, j));
        R bx = SPC(0.5) / dx * (sxp * SPC(1.0) / dx * (C_(HU, i + 1, j) - C_(HU, i, j))
                                - sxm * SPC(1.0) / dx * (C_(HU, i, j) - C_(HU, i - 1, j)))
               + SPC(0.5) / dy * (syp * SPC(1.0) / dy * (C_(HU, i, j + 1) - C_(HU, i, j))
                                  - sym * SPC(1.0) / dy * (C_(HU, i, j) - C_(HU, i, j - 1)));
        C_(F(BreakSourceX), i, j) = bx;
        sx = sx + bx;
        R by = SPC(0.5) / dx * (sxp * SPC(1.0) / dx * (C_(HV, i + 1, j) - C_(HV, i, j))
                                - sxm * SPC(1.0) / dx * (C_(HV, i, j) - C_(HV, i - 1, j)))
               + SPC(0.5) / dy * (syp * SPC(1.0) / dy * (C_(HV, i, j + 1) - C_(HV, i, j))
                                  - sym * SPC(1.0) / dy * (C_(HV, i, j) - C_(HV, i, j - 1)));
        C_(F(BreakSourceY), i, j) = by;
        sy = sy + by;
      }
      C_(F(SourceX), i, j) = sx;
      C_(F(SourceY), i, j) = sy;
    }
#undef NUV
  free(nuv);
}

/* ------------------------------------------------------------------ */
/* TRIDx_ser, tridiagnal.F:375-398 (rows normalised by the diagonal) */
static void trid_x(const fwo_params *p, R *A, R *Cc, R *D, R *Z) {
  DIMS;
  for (int j = Jb; j <= Je; j++) {
    for (int i = Ib + 1; i <= Ie; i++) {
      if (C_(A, i, j) != ZERO) {
        R a = C_(A, i, j);
        R den = SPC(1.0) / a - C_(Cc, i - 1, j);
        C_(Cc, i, j) = C_(Cc, i, j) / a / den;
        C_(D, i, j) = (C_(D, i, j) / a - C_(D, i - 1, j)) / (SPC(1.0) / a - C_(Cc, i - 1, j));
      }
    }
    C_(Z, Ie, j) = C_(D, Ie, j);
    for (int i = Ie - 1; i >= Ib; i--) C_(Z, i, j) = C_(D, i, j) - C_(Cc, i, j) * C_(Z, i + 1, j);
  }
}
/* TRIDy_ser, tridiagnal.F:338-361 */
static void trid_y(const fwo_params *p, R *A, R *Cc, R *D, R *Z) {
  DIMS;
  for (int i = Ib; i <= Ie; i++) {
    for (int j = Jb + 1; j <= Je; j++) {
      if (C_(A, i, j) != ZERO) {
        R a = C_(A, i, j);
        R den = SPC(1.0) / a - C_(Cc, i, j - 1);
        C_(Cc, i, j) = C_(Cc, i, j) / a / den;
        C_(D, i, j) = (C_(D, i, j) / a - C_(D, i, j - 1)) / (SPC(1.0) / a - C_(Cc, i, j - 1));
      }
    }
    C_(Z, i, Je) = C_(D, i, Je);
    for (int j = Je - 1; j >= Jb; j--) C_(Z, i, j) = C_(D, i, j) - C_(Cc, i, j) * C_(Z, i, j + 1);
  }
}
/* TRIDy_periodic (serial path), tridiagnal.F:59-325: Sherman-Morrison with
 * two Thomas solves */
static void trid_y_periodic(const fwo_params *p, R *a, R *c, R *d, R *fo) {
  DIMS;
  size_t ncell = (size_t)M * N;
  R *c1 = (R *)calloc(ncell, sizeof(R)), *y1 = (R *)calloc(ncell, sizeof(R)), *y2 = (R *)calloc(ncell, sizeof(R));
  R *abeg = (R *)calloc((size_t)M + 1, sizeof(R)), *cend = (R *)calloc((size_t)M + 1, sizeof(R));
  for (int i = Ib; i <= Ie; i++) {
    abeg[i] = C_(a, i, Jb);
    cend[i] = C_(c, i, Je);
  }
  for (int i = Ib; i <= Ie; i++) {
    C_(c, i, Jb) = C_(c, i, Jb) / ((R)1 + cend[i]);
    C_(d, i, Jb) = C_(d, i, Jb) / ((R)1 + cend[i]);
    C_(a, i, Je) = C_(a, i, Je) / ((R)1 + abeg[i]);
    C_(d, i, Je) = C_(d, i, Je) / ((R)1 + abeg[i]);
  }
  memcpy(c1, c, ncell * sizeof(R));
  trid_y(p, a, c1, d, y1);
  for (int j = Jb; j <= Je; j++)
    for (int i = Ib; i <= Ie; i++) C_(d, i, j) = ZERO;
  for (int i = Ib; i <= Ie; i++) {
    C_(d, i, Jb) = SPC(1.0) / (SPC(1.0) + cend[i]);
    for (int j = Jb + 1; j <= Je - 1; j++) C_(d, i, j) = ZERO;
    C_(d, i, Je) = -SPC(1.0) / (SPC(1.0) + abeg[i]);
  }
  trid_y(p, a, c, d, y2);
  for (int i = Ib; i <= Ie; i++) {
    R beta = (cend[i] * C_(y1, i, Jb) - abeg[i] * C_(y1, i, Je))
             / (SPC(1.0) - (cend[i] * C_(y2, i, Jb) - abeg[i] * C_(y2, i, Je)));
    for (int j = Jb; j <= Je; j++) C_(fo, i, j) = C_(y1, i, j) + beta * C_(y2, i, j);
  }
  free(c1); free(y1); free(y2); free(abeg); free(cend);
}

/* GET_Eta_U_V_HU_HV, etauv_solver.F:365-604 */
static void get_uv(const fwo_params *p, fwo_fields *f) {
  DIMS;
  const R dx = (R)p->DX, dy = (R)p->DY, mdf = (R)p->MinDepthFrc;
  const R g1 = (R)p->Gamma1, g3 = (R)p->Gamma3, b1 = (R)p->b1, b2 = (R)p->b2;
  const int *mk = FI(MASK), *m9 = FI(MASK9);
  R *H = F(H), *Eta = F(Eta), *Depth = F(Depth), *U = F(U), *V = F(V);
  size_t ncell = (size_t)M * N;
  for (size_t q = 0; q < ncell; q++) H[q] = Eta[q] * g3 + Depth[q];
  if (p->DISPERSION) {
    R *mA = (R *)calloc(ncell, sizeof(R)), *mC = (R *)calloc(ncell, sizeof(R));
    R *mD = (R *)calloc(ncell, sizeof(R)), *mF = (R *)calloc(ncell, sizeof(R));
    for (int j = Jb; j <= Je; j++)
      for (int i = Ib; i <= Ie; i++) {
        R dep = rmax(C_(Depth, i, j), mdf), depl = rmax(C_(Depth, i - 1, j), mdf), depr = rmax(C_(Depth, i + 1, j), mdf);
        R gm = g1 * (R)C_(m9, i, j);
        R t1 = gm * (b1 / SPC(2.0) / dx / dx * dep * dep + b2 / dx / dx * depl * dep);
        R t2 = SPC(1.0) + gm * (-(b1 / dx / dx * dep * dep) - SPC(2.0) * b2 / dx / dx * dep * dep);
        R t3 = gm * (b1 / SPC(2.0) / dx / dx * dep * dep + b2 / dx / dx * dep * depr);
        R t4 = C_(F(Ubar), i, j) * (R)C_(mk, i, j) / rmax(C_(H, i, j), mdf)
               + gm * (-(b1 / SPC(2.0) * dep * dep * C_(F(Vxy), i, j)) - b2 * dep * C_(F(DVxy), i, j));
        if (t2 != ZERO || C_(mk, i, j) > 0) {
          C_(mA, i, j) = t1 / t2; C_(mC, i, j) = t3 / t2; C_(mD, i, j) = t4 / t2;
        } else {
          C_(mA, i, j) = ZERO; C_(mC, i, j) = ZERO; C_(mD, i, j) = ZERO;
        }
      }
    trid_x(p, mA, mC, mD, mF);
    for (int j = Jb; j <= Je; j++)
      for (int i = Ib; i <= Ie; i++) C_(U, i, j) = C_(mF, i, j);
    for (int i = Ib; i <= Ie; i++)
      for (int j = Jb; j <= Je; j++) {
        R dep = rmax(C_(Depth, i, j), mdf), depl = rmax(C_(Depth, i, j - 1), mdf), depr = rmax(C_(Depth, i, j + 1), mdf);
        R gm = g1 * (R)C_(m9, i, j);
        R t1 = gm * (b1 / SPC(2.0) / dy / dy * dep * dep + b2 / dy / dy * depl * dep);
        R t2 = SPC(1.0) + gm * (-(b1 / dy / dy * dep * dep) - SPC(2.0) * b2 / dy / dy * dep * dep);
        R t3 = gm * (b1 / SPC(2.0) / dy / dy * dep * dep + b2 / dy / dy * dep * depr);
        R t4 = C_(F(Vbar), i, j) * (R)C_(mk, i, j) / rmax(C_(H, i, j), mdf)
               + gm * (-(b1 / SPC(2.0) * dep * dep * C_(F(Uxy), i, j)) - b2 * dep * C_(F(DUxy), i, j));
        if (t2 != ZERO || C_(mk, i, j) > 0) {
          C_(mA, i, j) = t1 / t2; C_(mC, i, j) = t3 / t2; C_(mD, i, j) = t4 / t2;
        } else {
          C_(mA, i, j) = ZERO; C_(mC, i, j) = ZERO; C_(mD, i, j) = ZERO;
        }
      }
    if (p->PERIODIC) trid_y_periodic(p, mA, mC, mD, mF);
    else trid_y(p, mA, mC, mD, mF);
    for (int j = Jb; j <= Je; j++)
      for (int i = Ib; i <= Ie; i++) C_(V, i, j) = C_(mF, i, j);
    free(mA); free(mC); free(mD); free(mF);
  } else {
    for (int j = Jb; j <= Je; j++)
      for (int i = Ib; i <= Ie; i++) {
        C_(U, i, j) = C_(F(Ubar), i, j) / rmax(C_(H, i, j), mdf);
        C_(V, i, j) = C_(F(Vbar), i, j) / rmax(C_(H, i, j), mdf);
      }
  }
  const R fcap = (R)p->FroudeCap;
  for (int j = Jb; j <= Je; j++)
    for (int i = Ib; i <= Ie; i++) {
      if (C_(mk, i, j) < 1) {
        C_(F(Ubar), i, j) = ZERO; C_(F(Vbar), i, j) = ZERO;
        C_(U, i, j) = ZERO; C_(V, i, j) = ZERO;
        C_(F(HU), i, j) = ZERO; C_(F(HV), i, j) = ZERO;
      } else {
        R hm = rmax(C_(H, i, j), mdf);
        C_(F(HU), i, j) = hm * C_(U, i, j);
        C_(F(HV), i, j) = hm * C_(V, i, j);
        R ut = RSQRT(C_(U, i, j) * C_(U, i, j) + C_(V, i, j) * C_(V, i, j));
        R fr = RSQRT(GRAV * hm);
        if (ut / fr > fcap) {
          R th = RATAN2(C_(V, i, j), C_(U, i, j));
          C_(U, i, j) = fcap * fr * RCOS(th);
          C_(V, i, j) = fcap * fr * RSIN(th);
          C_(F(HU), i, j) = C_(U, i, j) * hm;
          C_(F(HV), i, j) = C_(V, i, j) * hm;
        }
      }
    }
  if (!p->DISPERSION) {
    memcpy(f->Ubar, f->HU, ncell * sizeof(R));
    memcpy(f->Vbar, f->HV, ncell * sizeof(R));
  }
}

/* ESTIMATE_HUV, etauv_solver.F:53-351 (Cartesian, no ITERATION) */
void fwo_estimate_huv(const fwo_params *p, fwo_fields *f, int istage, double dt_) {
  DIMS;
  const R DT = (R)dt_;
  const R alpha[3] = {SPC(0.0), SPC(3.0) / SPC(4.0), SPC(1.0) / SPC(3.0)};  /* mod_global.F:122 */
  const R beta[3] = {SPC(1.0), SPC(1.0) / SPC(4.0), SPC(2.0) / SPC(3.0)};   /* mod_global.F:123 */
  const R al = alpha[istage - 1], be = beta[istage - 1];
  const R dx = (R)p->DX, dy = (R)p->DY, dxg = (R)p->DXg, dyg = (R)p->DYg, mdf = (R)p->MinDepthFrc;
  R *P = F(P), *Q = F(Q), *H = F(H);
  for (int j = Jb; j <= Je; j++)
    for (int i = Ib; i <= Ie; i++) {
      C_(F(P_center), i, j) = SPC(0.5) * (X_(P, i + 1, j) + X_(P, i, j));
      C_(F(Q_center), i, j) = SPC(0.5) * (Y_(Q, i, j + 1) + Y_(Q, i, j));
      C_(F(U_davg), i, j) = C_(F(P_center), i, j) / rmax(C_(H, i, j), mdf);
      C_(F(V_davg), i, j) = C_(F(Q_center), i, j) / rmax(C_(H, i, j), mdf);
    }
  const R nl = SPC(-1.0), nr = SPC(1.0);
  for (int j = Jb; j <= Je; j++)
    for (int i = Ib; i <= Ie; i++) {
      R fl = X_(P, i, j), fr = X_(P, i + 1, j), fb = Y_(Q, i, j), ft = Y_(Q, i, j + 1);
      R r1 = -(SPC(1.0) / dxg * (fr * nr + fl * nl)) - SPC(1.0) / dyg * (ft * nr + fb * nl);
      if (p->wavemaker != FWO_WK_NONE) {
        R xmk = (R)(i - Ib) * dxg, ymk = (R)(j - Jb) * dyg;
        if (RABS(xmk - (R)p->Xc_WK) < (R)p->Width_WK && RABS(ymk - (R)p->Yc_WK) < (R)p->Ywidth_WK / SPC(2.0))
          r1 = r1 + C_(F(WaveMaker_Mass), i, j);
      }
      C_(F(Wsurf), i, j) = r1;
      R e = al * C_(F(Eta0), i, j) + be * (C_(F(Eta), i, j) + DT * r1);
      if (p->ETA_LIMITER) {
        if (e < (R)p->TroughLimit) e = (R)p->TroughLimit;
        if (e > (R)p->CrestLimit) e = (R)p->CrestLimit;
      }
      C_(F(Eta), i, j) = e;
    }
  for (int j = Jb; j <= Je; j++)
    for (int i = Ib; i <= Ie; i++) {
      R fl = X_(F(Fx), i, j), fr = X_(F(Fx), i + 1, j), fb = Y_(F(Fy), i, j), ft = Y_(F(Fy), i, j + 1);
      R r2 = -(SPC(1.0) / dx * (fr * nr + fl * nl)) - SPC(1.0) / dy * (ft * nr + fb * nl) + C_(F(SourceX), i, j);
      C_(F(Ubar), i, j) = al * C_(F(Ubar0), i, j) + be * (C_(F(Ubar), i, j) + DT * r2);
    }
  for (int j = Jb; j <= Je; j++)
    for (int i = Ib; i <= Ie; i++) {
      R fl = X_(F(Gx), i, j), fr = X_(F(Gx), i + 1, j), fb = Y_(F(Gy), i, j), ft = Y_(F(Gy), i, j + 1);
      R r3 = -(SPC(1.0) / dx * (fr * nr + fl * nl)) - SPC(1.0) / dy * (ft * nr + fb * nl) + C_(F(SourceY), i, j);
      C_(F(Vbar), i, j) = al * C_(F(Vbar0), i, j) + be * (C_(F(Vbar), i, j) + DT * r3);
    }
  get_uv(p, f);
}

/* ------------------------------------------------------------------ */
/* UPDATE_MASK, masks.F:51-250 (no UseEtaScreen, no CHECK_MASS_CONSERVATION) */
void fwo_update_mask(const fwo_params *p, fwo_fields *f) {
  DIMS;
  int *mk = FI(MASK), *m9 = FI(MASK9), *ms = FI(MASK_STRUC);
  R *Eta = F(Eta), *Depth = F(Depth), *Dx = F(DepthX), *Dy = F(DepthY);
  size_t ncell = (size_t)M * N;
  int *mt = (int *)malloc(ncell * sizeof(int));
  memcpy(mt, mk, ncell * sizeof(int));
  for (int j = Jb - 2; j <= Je + 2; j++)
    for (int i = Ib - 2; i <= Ie + 2; i++) {
      if (C_(ms, i, j) == 1) {
        if (C_(mk, i, j) < 1) {
          if (C_(Eta, i, j) > -C_(Depth, i, j)) C_(mt, i, j) = 1;
          if (C_(mk, i - 1, j) == 1 && C_(Eta, i - 1, j) > C_(Eta, i, j)) C_(mt, i, j) = 1;
          if (C_(mk, i + 1, j) == 1 && C_(Eta, i + 1, j) > C_(Eta, i, j)) C_(mt, i, j) = 1;
          if (C_(mk, i, j - 1) == 1 && C_(Eta, i, j - 1) > C_(Eta, i, j)) C_(mt, i, j) = 1;
          if (C_(mk, i, j + 1) == 1 && C_(Eta, i, j + 1) > C_(Eta, i, j)) C_(mt, i, j) = 1;
        } else {
          if (C_(Eta, i, j) < -C_(Depth, i, j)) C_(mt, i, j) = 0;
        }
      }
      if (C_(mk, i, j) < 1) {
        X_(Dx, i, j) = C_(Depth, i - 1, j);
        X_(Dx, i + 1, j) = C_(Depth, i + 1, j);
        Y_(Dy, i, j) = C_(Depth, i, j - 1);
        Y_(Dy, i, j + 1) = C_(Depth, i, j + 1);
      }
    }
  memcpy(mk, mt, ncell * sizeof(int));
  free(mt);
  const R mdf = (R)p->MinDepthFrc, swe = (R)p->SWE_ETA_DEP;
  /* MASK9 = MASK9tmp copies the whole array, masks.F:182, but MASK9tmp is
   * only set on Ibeg-1..Iend+1 x Jbeg-1..Jend+1; the outer two ghost rings
   * come from fresh, uninitialised memory.  They are defined as 0 here (what
   * fresh pages hold; the reference is run with MALLOC_PERTURB_=255 so its
   * fresh allocations read as zero too). */
  memset(m9, 0, ncell * sizeof(int));
  for (int j = Jb - 1; j <= Je + 1; j++)
    for (int i = Ib - 1; i <= Ie + 1; i++) {
      if (p->VISCOSITY_BREAKING) {
        C_(m9, i, j) = 1;
      } else {
        int v = C_(mk, i, j) * C_(mk, i - 1, j) * C_(mk, i + 1, j) * C_(mk, i + 1, j + 1) * C_(mk, i, j + 1)
                * C_(mk, i - 1, j + 1) * C_(mk, i + 1, j - 1) * C_(mk, i, j - 1) * C_(mk, i - 1, j - 1);
        if (RABS(C_(Eta, i, j)) / rmax(C_(Depth, i, j), mdf) > swe) v = 0;
        C_(m9, i, j) = v;
      }
    }
}

/* ------------------------------------------------------------------ */
/* WAVE_BREAKING -> BREAKING, breaker.F:60-353 (default eddy-viscosity
 * branch, with the roller flux and undertow of continuing breaking) */
void fwo_breaking(const fwo_params *p, fwo_fields *f, double dt_) {
  DIMS;
  if (!p->SHOW_BREAKING) return;
  const R DT = (R)dt_, dx = (R)p->DX, dy = (R)p->DY, mdf = (R)p->MinDepthFrc;
  const R c1 = (R)p->Cbrk1, c2 = (R)p->Cbrk2, tbrk = (R)p->T_brk, wkc = (R)p->WAVEMAKER_Cbrk, nub = (R)p->nu_bkg;
  R *AGE = F(AGE_BREAKING), *H = F(H), *ET = F(ETAT), *EX = F(ETAx), *EY = F(ETAy);
  for (int j = Jb - 1; j <= Je + 1; j++)
    for (int i = Ib - 1; i <= Ie + 1; i++) {
      R t3 = RSQRT(GRAV * rmax(mdf, C_(H, i, j)));
      R t1 = c1 * t3, t2 = c2 * t3;
      R ang = RATAN2(-C_(EY, i, j), -C_(EX, i, j)) * SPC(180.0) / PI;
      R ex = C_(EX, i, j), ey = C_(EY, i, j), et = C_(ET, i, j);
      R t4 = rmax(RSQRT(ex * ex + ey * ey), SMALL);
      if (et >= t1 && (C_(AGE, i, j) == ZERO || C_(AGE, i, j) > tbrk)) {
        C_(AGE, i, j) = DT;
      } else if (C_(AGE, i, j) > ZERO) {
        C_(AGE, i, j) = C_(AGE, i, j) + DT;
      } else {
        R c = rmin(RABS(et) / t4, RSQRT(GRAV * RABS(C_(H, i, j))));
        R pxy = RSQRT(dx * dx + dy * dy) / rmax(c, SMALL);
        R px = RSQRT(dx * dx) / rmax(c, SMALL);
        R py = RSQRT(dy * dy) / rmax(c, SMALL);
        if (et >= t2) {
#define TRIG(a1, p1, a2, p2, a3, p3) \
  (((a1) >= DT && (a1) > (p1)) || ((a2) >= DT && (a2) > (p2)) || ((a3) >= DT && (a3) > (p3)))
          if (ang >= ZERO && ang < SPC(90.0)) {
            if (TRIG(C_(AGE, i - 1, j), px, C_(AGE, i - 1, j - 1), pxy, C_(AGE, i, j - 1), py)) C_(AGE, i, j) = DT;
          }
          if (ang >= SPC(90.0) && ang < SPC(180.0)) {
            if (TRIG(C_(AGE, i + 1, j), px, C_(AGE, i + 1, j - 1), pxy, C_(AGE, i, j - 1), py)) C_(AGE, i, j) = DT;
          }
          if (ang >= SPC(-180.0) && ang < SPC(-90.0)) {
            if (TRIG(C_(AGE, i + 1, j), px, C_(AGE, i + 1, j + 1), pxy, C_(AGE, i, j + 1), py)) C_(AGE, i, j) = DT;
          }
          if (ang >= SPC(-90.0) && ang < SPC(0.0)) {
            if (TRIG(C_(AGE, i, j + 1), py, C_(AGE, i - 1, j + 1), pxy, C_(AGE, i - 1, j), px)) C_(AGE, i, j) = DT;
          }
#undef TRIG
        }
      }
      R xmk = (R)(i - Ib) * dx, ymk = (R)(j - Jb) * dy;
      if (RABS(xmk - (R)p->Xc_WK) < (R)p->Width_WK && RABS(ymk - (R)p->Yc_WK) < (R)p->Ywidth_WK / SPC(2.0)) {
        if (et > rmin(t2, wkc * t3)) {
          R cap1 = FLC(1.0) * (rmax(C_(F(Depth), i, j), mdf) + C_(F(Eta), i, j));
          C_(F(nu_break), i, j) = cap1 * wkc * t3 + nub;
        } else {
          C_(F(nu_break), i, j) = ZERO + nub;
        }
      } else {
        if (C_(AGE, i, j) > ZERO && C_(AGE, i, j) < tbrk && et > t2) {
          R cap1 = FLC(1.0) * (rmax(C_(F(Depth), i, j), mdf) + C_(F(Eta), i, j));
          C_(F(nu_break), i, j) = cap1 * t2 + nub;
          /* roller and undertow, breaker.F:326-337 (kept elsewhere) */
          R cr = rmax(RABS(et) / t4, RSQRT(GRAV * RABS(C_(H, i, j))));
          R rr = RABS(FLC(0.45) * et / rmax(cr, SMALL));
          rr = rmin(rr, SPC(0.1638));
          R uu = C_(F(U), i, j), vv = C_(F(V), i, j);
          R rf = RABS(cr - RSQRT(uu * uu + vv * vv)) * rr * (C_(F(Eta), i, j) - C_(F(ETAmean), i, j));
          C_(F(ROLLER_FLUX), i, j) = rf;
          C_(F(UNDERTOW_U), i, j) = -(rf * RCOS(ang * PI / SPC(180.0)));
          C_(F(UNDERTOW_V), i, j) = -(rf * RSIN(ang * PI / SPC(180.0)));
        } else {
          C_(F(nu_break), i, j) = ZERO + nub;
        }
      }
    }
}

/* SPONGE_DAMPING, sponge.F:52-66 */
void fwo_sponge(const fwo_params *p, fwo_fields *f) {
  DIMS;
  size_t ncell = (size_t)M * N;
  const int *mk = FI(MASK);
  R *e = F(Eta), *u = F(U), *v = F(V), *s = F(SPONGE);
  for (size_t q = 0; q < ncell; q++) {
    if ((R)mk[q] > ZERO) e[q] = e[q] / s[q];
    u[q] = u[q] / s[q];
    v[q] = v[q] / s[q];
  }
}

/* CHECK_BLOWUP, misc.F:394-433: max |eta| over the interior */
double fwo_max_abs_eta(const fwo_params *p, const fwo_fields *f) {
  DIMS;
  const R *e = (const R *)f->Eta;
  R m = ZERO;
  for (int j = Jb; j <= Je; j++)
    for (int i = Ib; i <= Ie; i++) {
      R a = RABS(C_(e, i, j));
      if (a > m) m = a;
    }
  return (double)m;
}

/* one or more whole time steps, src/main.F:295-487 + 517 */
int fwo_run(const fwo_params *p, fwo_fields *f, int nsteps, double *dt, double *time) {
  DIMS;
  size_t ncell = (size_t)M * N;
  for (int s = 0; s < nsteps; s++) {
    memcpy(f->Eta0, f->Eta, ncell * sizeof(R));
    memcpy(f->Ubar0, f->Ubar, ncell * sizeof(R));
    memcpy(f->Vbar0, f->Vbar, ncell * sizeof(R));
    fwo_exchange(p, f);
    fwo_estimate_dt(p, f, dt, time);
    memcpy(f->U0, f->U, ncell * sizeof(R));
    memcpy(f->V0, f->V, ncell * sizeof(R));
    for (int st = 1; st <= 3; st++) {
      if (p->DISPERSION) fwo_dispersion(p, f, *dt);
      fwo_fluxes(p, f);
      fwo_sources(p, f, *time);
      fwo_estimate_huv(p, f, st, *dt);
      fwo_update_mask(p, f);
      fwo_breaking(p, f, *dt);
      fwo_exchange(p, f);
      if (p->DIRECT_SPONGE) fwo_sponge(p, f);
    }
    if (fwo_max_abs_eta(p, f) > p->EtaBlowVal) return 1;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* unit-level hooks (tests only) */
void fwo_thomas_line(int n, void *a_, void *c_, void *d_, void *z_) {
  R *a = (R *)a_, *c = (R *)c_, *d = (R *)d_, *z = (R *)z_;
  for (int k = 1; k < n; k++) {
    if (a[k] != ZERO) {
      R den = SPC(1.0) / a[k] - c[k - 1];
      c[k] = c[k] / a[k] / den;
      d[k] = (d[k] / a[k] - d[k - 1]) / (SPC(1.0) / a[k] - c[k - 1]);
    }
  }
  z[n - 1] = d[n - 1];
  for (int k = n - 2; k >= 0; k--) z[k] = d[k] - c[k] * z[k + 1];
}

double fwo_hll(double sl, double sr, double fl, double fr, double ul, double ur) {
  return (double)hll((R)sl, (R)sr, (R)fl, (R)fr, (R)ul, (R)ur);
}

void fwo_muscl_row(int n, const void *vin_, const int *mask, void *outl_, void *outr_) {
  fwo_params p;
  memset(&p, 0, sizeof p);
  p.Mloc = n;
  p.Nloc = 7;
  p.Ibeg = 4;
  p.Iend = n - 3;
  p.Jbeg = 4;
  p.Jend = 4;
  size_t M = (size_t)n, M1 = M + 1;
  R *vin = (R *)calloc(M * 7, sizeof(R));
  int *mk = (int *)calloc(M * 7, sizeof(int));
  R *ol = (R *)calloc(M1 * 7, sizeof(R)), *orr = (R *)calloc(M1 * 7, sizeof(R));
  for (size_t i = 0; i < M; i++) {
    vin[3 * M + i] = ((const R *)vin_)[i];
    mk[3 * M + i] = mask[i];
  }
  construct_x(&p, mk, vin, ol, orr);
  for (size_t i = 0; i < M1; i++) {
    ((R *)outl_)[i] = ol[3 * M1 + i];
    ((R *)outr_)[i] = orr[3 * M1 + i];
  }
  free(vin); free(mk); free(ol); free(orr);
}
