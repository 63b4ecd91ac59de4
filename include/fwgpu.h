/* fwgpu.h -- C ABI of the MI355X (gfx950) FUNWAVE-TVD per-time-step solver.
 *
 * Drop-in boundary for the per-step path of the reference program
 * (src/main.F:284-535).  Every routine the reference main loop calls on
 * that path has one entry point here, with the same meaning.  The
 * reference's routines take no arguments and work on MODULE GLOBAL
 * (src/mod_global.F:52-461); here the state lives in HBM inside an opaque
 * context and crosses the boundary only through fw_upload / fw_download,
 * whose host arrays use the reference's own layout: Fortran column-major,
 * 1-based, three ghost cells, cell arrays (Mloc, Nloc), x-face arrays
 * (Mloc+1, Nloc), y-face arrays (Mloc, Nloc+1), masks INTEGER(4), reals
 * REAL(SP) (4 or 8 bytes, fixed per context).
 *
 * All functions return 0 on success or a negative FW_E* code; fw_strerror()
 * describes the last error of a context.  Calls on one context must come
 * from one host thread at a time.  Work runs on the context's own HIP stream
 * (default priority); calls return after the work is queued, except those
 * that hand a value back to the host (fw_estimate_dt, fw_step,
 * fw_check_blowup, fw_download, fw_reduce_stats), which wait for it.
 */
#ifndef FWGPU_H
#define FWGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FW_OK = 0,
  FW_E_INVALID = -1,    /* bad argument or unsupported configuration */
  FW_E_HIP = -2,        /* HIP runtime error */
  FW_E_NOMEM = -3,      /* device allocation failed */
  FW_E_BLOWUP = -4,     /* max |eta| exceeded EtaBlowVal (misc.F:394-433) */
  FW_E_NOTFOUND = -5,   /* unknown field name */
  FW_E_BREAK_ITER = -6  /* breaking-age propagation did not settle */
};

/* Wavemaker kinds with a per-step source (src/sources.F:109-184).
 * Wavemakers acting only at initialisation (INI_SOL, INI_REC, ...) use
 * FW_WK_NONE. */
enum { FW_WK_NONE = 0, FW_WK_REG = 1, FW_WK_DATA2D = 2 };

/* Scalar configuration: the MODULE GLOBAL scalars the per-step routines
 * read, after READ_INPUT and INITIALIZATION (src/main.F:177-214).
 * Indices are the reference's 1-based ones. */
typedef struct fw_config {
  int real_bytes;            /* 4 (REAL(SP) single) or 8 (-DDOUBLE_PRECISION) */
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
} fw_config;

typedef struct fw_ctx fw_ctx;

/* FP64 budget of the interior (src/statistics.F:52-190 analogue). */
typedef struct fw_stats {
  double mass_volume;   /* sum over wet cells of eta * dx * dy */
  double energy;        /* sum of 0.5*H*(u^2+v^2) + 0.5*g*eta^2, times dx*dy */
  double max_eta, min_eta, max_u, max_v;
  int64_t wet_cells;    /* interior cells with MASK = 1 */
} fw_stats;

/* --- lifetime ---------------------------------------------------------- */
/* Allocates the whole state in HBM (replaces ALLOCATE_VARIABLES,
 * src/init.F:54-269, for the per-step arrays) and zeroes it. */
int fw_create(const fw_config *cfg, fw_ctx **out);
int fw_destroy(fw_ctx *ctx);
const char *fw_strerror(const fw_ctx *ctx);
int fw_synchronize(fw_ctx *ctx);
int fw_device_bytes(const fw_ctx *ctx, int64_t *bytes);

/* --- state transfer ------------------------------------------------------ */
/* Field ids by reference array name ("Eta", "U", "MASK", "DepthX", "Cm",
 * ...); <0 if unknown.  Band tables: "xmk_wk" (Mloc), "ymk_wk" (Nloc),
 * "OMGN2D" (NumFreq), "Cm"/"Sm" (Mloc, Nloc, NumFreq; only the wavemaker
 * band ilo..ihi x jlo..jhi is kept on the device). */
int fw_field_id(const char *name);
int fw_upload(fw_ctx *ctx, int field, const void *host);
int fw_download(fw_ctx *ctx, int field, void *host);
/* model clock (REAL(SP) DT and TIME of MODULE GLOBAL) */
int fw_set_clock(fw_ctx *ctx, double dt, double time);
int fw_get_clock(const fw_ctx *ctx, double *dt, double *time);

/* --- device access --------------------------------------------------------
 * For libraries that work on the HBM state in place, on the context's
 * stream: the device address of a field's plane (element (I,J) at
 * ptr[(J-1)*pitch + (I-1)], pitch in elements; masks are int32, reals
 * REAL(SP)), the stream itself (a hipStream_t), and the configuration the
 * context was created with. */
int fw_field_device(fw_ctx *ctx, int field, void **ptr, int *pitch);
int fw_get_stream(fw_ctx *ctx, void **stream);
int fw_get_config(const fw_ctx *ctx, fw_config *cfg);

/* --- per-step entry points (same order and meaning as src/main.F) ------- */
int fw_exchange(fw_ctx *ctx);                /* EXCHANGE      bc.F:426      */
int fw_estimate_dt(fw_ctx *ctx, double *dt, double *time);
                                             /* ESTIMATE_DT   misc.F:216    */
int fw_save_stage0(fw_ctx *ctx, int which);  /* 1: Eta0,Ubar0,Vbar0 := Eta,Ubar,Vbar
                                                (main.F:295-297); 2: U0,V0 := U,V
                                                (main.F:331-332)            */
int fw_cal_dispersion(fw_ctx *ctx);          /* CAL_DISPERSION dispersion.F:62 */
int fw_fluxes(fw_ctx *ctx);                  /* FLUXES        fluxes.F:52   */
int fw_source_terms(fw_ctx *ctx);            /* SourceTerms   sources.F:56  (incl. wavemaker) */
int fw_estimate_huv(fw_ctx *ctx, int istage);/* ESTIMATE_HUV  etauv_solver.F:53
                                                (incl. GET_Eta_U_V_HU_HV :365
                                                 and TRIDx/TRIDy tridiagnal.F) */
int fw_update_mask(fw_ctx *ctx);             /* UPDATE_MASK   masks.F:51    */
int fw_wave_breaking(fw_ctx *ctx);           /* WAVE_BREAKING breaker.F:60  */
int fw_sponge_damping(fw_ctx *ctx);          /* SPONGE_DAMPING sponge.F:52  */
int fw_check_blowup(fw_ctx *ctx, double *max_abs_eta);
                                             /* CHECK_BLOWUP  misc.F:394    */
/* nsteps whole time steps: main.F:295-487 followed by CHECK_BLOWUP (:517).
 * Returns FW_E_BLOWUP (state kept) if max|eta| > EtaBlowVal. */
int fw_step(fw_ctx *ctx, int nsteps, double *dt, double *time);
/* FP64 mass / energy / extrema budget of the interior. */
int fw_reduce_stats(fw_ctx *ctx, fw_stats *out);
/* The same budget for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): no
 * counterpart in the reference.  It allocates nothing, device or host: the
 * per-row partials go to a static device array of the library, sized for
 * at most FW_NOALLOC_MAX_ROWS interior rows (Nloc - 6); taller domains are
 * refused with FW_E_INVALID.  Every call on one GPU shares that array, so
 * calls that overlap on one GPU (from several threads or contexts) race;
 * each call writes every partial it reads, so nothing carries over from the
 * last call.  Sums are taken in another order than fw_reduce_stats', so the
 * FP64 mass and energy can differ in the last bits; extrema and the wet-cell
 * count are identical. */
#define FW_NOALLOC_MAX_ROWS 8192
int fw_reduce_stats_noalloc(fw_ctx *ctx, fw_stats *out);

/* --- options ------------------------------------------------------------ */
/* FW_OPT_FULL_STATE = 1 keeps every intermediate array of MODULE GLOBAL up to
 * date (face states UxL..GyR, HxL.., SxL.., the wave-speed copies, the
 * first derivatives Ut, Vt, Ux.., Utx.., DUtxx.. of CAL_DISPERSION, its
 * nonlinear terms U1pp, V1pp, U2, V2, U3, V3, and in fw_step the source
 * terms SourceX/Y, FrcInsX/Y, BreakSourceX/Y and WaveMaker_Mass, the
 * mean-flow diagnostics P_center, Q_center, U_davg, V_davg, Wsurf, and the
 * roller flux and undertow ROLLER_FLUX, UNDERTOW_U, UNDERTOW_V), for
 * routine-by-routine comparison with the reference.  Default 0: only the
 * arrays a later routine reads are written (U1pp+U2+U3 and V1pp+V2+V3 only
 * as the sums SourceTerms forms; in fw_step, whose RK update forms the
 * cross derivatives Uxy, Vxy, DUxy, DVxy and U1p, V1p of each cell itself,
 * not those either); results are identical.  fw_source_terms
 * always writes its outputs; the face states, the dispersion first
 * derivatives and nonlinear terms and the mean-flow diagnostics follow this
 * option in the per-routine entry points too (without it, fw_source_terms
 * takes the nonlinear terms from fw_cal_dispersion's sums, not from
 * uploaded U1pp..V3). */
enum { FW_OPT_FULL_STATE = 1 };
/* FW_OPT_HOST_MATH = 1 evaluates the Froude cap's ATAN2, COS and SIN
 * (etauv_solver.F:580-586) with the host's C library, as the reference does,
 * instead of the device math library, which can differ in the last bit:
 * the velocity tail lists the capped cells and the host rewrites them
 * (one host round trip per line solve).  Runs are then bitwise the
 * reference's over any number of steps; default 0. */
enum { FW_OPT_HOST_MATH = 2 };
int fw_set_option(fw_ctx *ctx, int option, int value);

/* --- measurement -------------------------------------------------------- */
/* With profiling on, every kernel launch is bracketed by HIP events on the
 * context stream; fw_profile_get reports launch count and summed device
 * milliseconds for one kernel name. */
int fw_profile_enable(fw_ctx *ctx, int on);
int fw_profile_reset(fw_ctx *ctx);
int fw_profile_get(fw_ctx *ctx, const char *kernel, int64_t *launches, double *ms);
int fw_profile_names(fw_ctx *ctx, char *buf, int buflen);
/* device milliseconds of nsteps whole steps, HIP events on the ctx stream */
int fw_time_steps(fw_ctx *ctx, int nsteps, double *ms);

/* --- ensembles -------------------------------------------------------------
 * No counterpart in the reference, whose main loop advances one domain
 * (src/main.F:284-535).  For callers running many independent domains
 * ("members", each a context created and uploaded as above, of any size):
 * every phase of the step is one launch over all members, whose workgroups
 * take their member and tile from per-member tile queues; each member's DT,
 * TIME, blow-up check and the wavemaker factors of the whole run stay on
 * the device, and one step replays from a captured hipGraph.  Results are
 * bitwise those of fw_step on each member alone.
 *
 * Members need FIXED_DT (the per-step wavemaker factors are evaluated ahead
 * on the host, with the C library, for TIME = TIME + DT_fixed step after
 * step); a member whose time step would halve (misc.F:283-288) stops before
 * that step with status FW_BATCH_DT_CHANGED, one whose max |eta| exceeds
 * EtaBlowVal stops after its step with FW_BATCH_BLOWUP; the others go on.
 * Members share the precision, DISPERSION, PERIODIC and whether DX and DY
 * are powers of two, and run with FW_OPT_FULL_STATE = 0.  The batch keeps
 * pointers to the contexts: they must outlive it, and per-member calls on
 * them may be interleaved with batch calls (each batch call picks up the
 * members' current state, and afterwards fw_get_clock, fw_download, ...
 * see the advanced state).  Calls on one batch come from one host thread. */
typedef struct fw_batch fw_batch;
enum { FW_BATCH_RUNNING = 0, FW_BATCH_BLOWUP = 1, FW_BATCH_DT_CHANGED = 2, FW_BATCH_BREAK_ITER = 3 };
int fw_batch_create(fw_ctx *const *members, int n, fw_batch **out);
int fw_batch_destroy(fw_batch *b);
const char *fw_batch_strerror(const fw_batch *b);
int fw_batch_size(const fw_batch *b, int *n);
/* nsteps whole steps (main.F:295-517) of every running member.  Returns 0,
 * or for the first member that stopped: FW_E_BLOWUP, FW_E_INVALID (time
 * step would change) or FW_E_BREAK_ITER; fw_batch_status tells which. */
int fw_batch_step(fw_batch *b, int nsteps);
int fw_batch_status(const fw_batch *b, int *status /* n entries */);
/* as fw_batch_step, and the device milliseconds of the steps (HIP events
 * around the graph launches on the batch's stream) */
int fw_batch_time_steps(fw_batch *b, int nsteps, double *ms);
/* 1 (default): replay the captured step graphs; 0: launch eagerly */
int fw_batch_set_graph(fw_batch *b, int on);
/* Workgroup shapes of the tunable kernels (the dispersion stencils, the RK
 * update, the velocity tail and mask multiply, UPDATE_MASK and breaking)
 * chosen by timing trial_steps steps per candidate on a snapshot of the
 * members' state, restored afterwards; report: one line per trial. */
int fw_batch_autotune(fw_batch *b, int trial_steps, char *report, int len);
/* the shapes in use: 4 groups x (x, y) */
int fw_batch_get_shapes(const fw_batch *b, int *shapes);

#ifdef __cplusplus
}
#endif
#endif
