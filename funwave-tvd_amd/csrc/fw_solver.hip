// fw_solver.hip -- host side of the MI355X FUNWAVE-TVD per-step solver:
// the HBM-resident context, state transfer in the reference layout, the
// per-routine entry points and the whole-step driver behind include/fwgpu.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "fw_context.h"

using namespace fw;
using namespace fwi;

// ---- the allocation-free budget (fw_reduce_stats_noalloc) ----------------
// Partials of the per-row reduction live in this file-scope device array:
// one row of seven doubles per interior row, up to kNoallocRows rows, and
// the seven totals after them (458 KB, held on every GPU the library runs
// on until the process ends).  Host code never passes the array's name to a
// kernel (in host code it is a stand-in, not a device address): fw_create
// resolves its address on each device once (hipGetSymbolAddress, with that
// device current) and keeps it in g_scratch_ptr[device].
constexpr int kNoallocRows = FW_NOALLOC_MAX_ROWS;
static __device__ __align__(256) double scratch[7 * kNoallocRows + 8];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// the device address of scratch on device dev (current): resolved once
int resolve_scratch(int dev, std::string &err) {
  if (dev < 0 || dev >= kMaxDevices) {
    err = "device index beyond the static scratch table";
    return FW_E_INVALID;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return FW_OK;
  double *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return FW_E_HIP;
  }
  g_scratch_ptr[dev] = ptr;
  return FW_OK;
}
}  // namespace

namespace {

hipEvent_t get_event(fw_ctx *c) {
  if (!c->evpool.empty()) {
    hipEvent_t e = c->evpool.back();
    c->evpool.pop_back();
    return e;
  }
  hipEvent_t e;
  (void)hipEventCreate(&e);
  return e;
}

template <typename K, typename... A>
void launch(fw_ctx *c, const char *name, K k, dim3 grid, dim3 block, A... args) {
  hipEvent_t a = nullptr, b = nullptr;
  if (c->prof) {
    a = get_event(c);
    b = get_event(c);
    (void)hipEventRecord(a, c->cur);
  }
  hipLaunchKernelGGL(k, grid, block, 0, c->cur, args...);
  if (c->prof) {
    (void)hipEventRecord(b, c->cur);
    c->pending.push_back({name, a, b});
  }
}

int collect_profile(fw_ctx *c) {
  if (c->pending.empty()) return FW_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto &p : c->pending) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, p.a, p.b);
    auto &a = c->acc[p.name];
    a.first += 1;
    a.second += ms;
    c->evpool.push_back(p.a);
    c->evpool.push_back(p.b);
  }
  c->pending.clear();
  return FW_OK;
}

inline dim3 grid2(int nx, int ny, dim3 b) { return dim3((nx + b.x - 1) / b.x, (ny + b.y - 1) / b.y); }
const dim3 kB2(64, 4);
// taller tiles for kernels reading j-1/j+1 neighbours: fewer rows fetched
// twice by vertically adjacent workgroups
const dim3 kBT(64, 8);
// the RK update with source terms (~100 VGPRs): four-wave workgroups fill
// the five waves per SIMD its registers allow (two eight-wave workgroups
// would leave one idle)
const dim3 kBH(64, 4);

// both grid spacings powers of two: the stencil kernels multiply by the
// reciprocal spacing instead of dividing (bitwise the same)
template <class R>
bool spacing_pow2(fw_ctx *c) {
  return c->prm<R>().dx_pow2 && c->prm<R>().dy_pow2;
}

template <class R>
void fill_prm(fw_ctx *c) {
  const fw_config &f = c->cfg;
  Prm<R> &p = c->prm<R>();
  p.DX = (R)f.DX; p.DY = (R)f.DY; p.DXg = (R)f.DXg; p.DYg = (R)f.DYg; p.CFL = (R)f.CFL;
  p.MinDepth = (R)f.MinDepth; p.MDF = (R)f.MinDepthFrc;
  p.G1 = (R)f.Gamma1; p.G2 = (R)f.Gamma2; p.G3 = (R)f.Gamma3; p.B1 = (R)f.Beta_1; p.B2 = (R)f.Beta_2;
  p.b1 = (R)f.b1; p.b2 = (R)f.b2; p.SWE = (R)f.SWE_ETA_DEP; p.FCAP = (R)f.FroudeCap;
  p.C1 = (R)f.Cbrk1; p.C2 = (R)f.Cbrk2; p.TBRK = (R)f.T_brk; p.WKC = (R)f.WAVEMAKER_Cbrk; p.NUB = (R)f.nu_bkg;
  p.XC = (R)f.Xc_WK; p.YC = (R)f.Yc_WK; p.WW = (R)f.Width_WK; p.YW = (R)f.Ywidth_WK; p.BLOW = (R)f.EtaBlowVal;
  p.DTF = (R)f.DT_fixed; p.CREST = (R)f.CrestLimit; p.TROUGH = (R)f.TroughLimit;
  p.Dgen = (R)f.D_gen; p.Bgen = (R)f.Beta_gen; p.rlam = (R)f.rlamda; p.Tper = (R)f.Tperiod;
  p.Tramp = (R)f.Time_ramp; p.FreqPeak = (R)f.FreqPeak;
  p.DISPERSION = f.DISPERSION; p.PERIODIC = f.PERIODIC; p.VISC_BRK = f.VISCOSITY_BREAKING;
  p.SHOW_BRK = f.SHOW_BREAKING; p.DIFF_SPONGE = f.DIFFUSION_SPONGE; p.FRIC_SPONGE = f.FRICTION_SPONGE;
  p.DIRECT_SPONGE = f.DIRECT_SPONGE; p.ETA_LIMITER = f.ETA_LIMITER; p.FIXED_DT = f.FIXED_DT;
  p.wavemaker = f.wavemaker; p.NumFreq = f.NumFreq;
  p.ilo = f.ilo; p.ihi = f.ihi; p.jlo = f.jlo; p.jhi = f.jhi;
  // exact reciprocals and constant row factors, in the working precision and
  // the reference's operation order
  auto pow2 = [](R x) {
    int e;
    R m = std::frexp(x, &e);
    return x > (R)0 && m == (R)0.5 && std::isnormal((R)1 / x);
  };
  p.rDX = (R)1 / p.DX;
  p.rDY = (R)1 / p.DY;
  p.dx_pow2 = pow2(p.DX);
  p.dy_pow2 = pow2(p.DY);
  p.rDXg = (R)1 / p.DXg;
  p.rDYg = (R)1 / p.DYg;
  p.h5x = (R)0.5 / p.DX;
  p.h5y = (R)0.5 / p.DY;
  p.b1h = p.b1 / (R)2;
  p.cxa = p.b1 / (R)2 / p.DX / p.DX;
  p.cxb = p.b2 / p.DX / p.DX;
  p.cxc = p.b1 / p.DX / p.DX;
  p.cxd = (R)2 * p.b2 / p.DX / p.DX;
  p.cya = p.b1 / (R)2 / p.DY / p.DY;
  p.cyb = p.b2 / p.DY / p.DY;
  p.cyc = p.b1 / p.DY / p.DY;
  p.cyd = (R)2 * p.b2 / p.DY / p.DY;
}



template <class R>
int alloc_state(fw_ctx *c) {
  Dev<R> &d = c->dev<R>();
  const size_t pe = plane_elems(c->g);
  auto get = [&](size_t bytes, void **out) -> int {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return fail(c, FW_E_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    e = hipMemsetAsync(p, 0, bytes, c->stream);
    if (e != hipSuccess) return fail(c, FW_E_HIP, hipGetErrorString(e));
    c->allocs.push_back(p);
    c->alloc_sizes.push_back(bytes);
    c->bytes += bytes;
    *out = p;
    return FW_OK;
  };
  int rc;
#define FW_ALLOC(n) \
  if ((rc = get(pe * sizeof(R), (void **)&d.n)) != FW_OK) return rc;
  FW_CELL_FIELDS(FW_ALLOC) FW_XFACE_FIELDS(FW_ALLOC) FW_YFACE_FIELDS(FW_ALLOC)
#undef FW_ALLOC
  void **rs[] = {(void **)&d.AGEn, (void **)&d.xa, (void **)&d.xc, (void **)&d.xd, (void **)&d.ya,
                 (void **)&d.yc,   (void **)&d.yd, (void **)&d.c1, (void **)&d.y1, (void **)&d.y2,
                 (void **)&d.sumU, (void **)&d.sumV};
  for (void **q : rs)
    if ((rc = get(pe * sizeof(R), q)) != FW_OK) return rc;
  void **is[] = {(void **)&d.MASK, (void **)&d.MASK9, (void **)&d.MASK_STRUC, (void **)&d.MASKn, (void **)&d.cand};
  for (void **q : is)
    if ((rc = get(pe * sizeof(int), q)) != FW_OK) return rc;
  if ((rc = get(8 * pe * sizeof(R), (void **)&d.cinf)) != FW_OK) return rc;
  if ((rc = get(64 * sizeof(int), (void **)&d.cand_n)) != FW_OK) return rc;
  d.flags = d.cand_n + 16;
  d.fc_n = d.cand_n + 32;  // the Froude-cap list itself is allocated with FW_OPT_HOST_MATH
  if ((rc = get(kRedBlocks * sizeof(R), (void **)&d.red)) != FW_OK) return rc;
  if ((rc = get(kRedBlocks * 7 * sizeof(double), (void **)&d.red64)) != FW_OK) return rc;
  if ((rc = get(64 * sizeof(double), &c->scal)) != FW_OK) return rc;
  const fw_config &f = c->cfg;
  const int M = f.Mloc, N = f.Nloc;
  if ((rc = get((size_t)(M + 8) * sizeof(R), (void **)&d.xmk_wk)) != FW_OK) return rc;
  if ((rc = get((size_t)(N + 8) * sizeof(R), (void **)&d.ymk_wk)) != FW_OK) return rc;
  int nfreq = f.NumFreq > 0 ? f.NumFreq : 1;
  if ((rc = get((size_t)nfreq * sizeof(R), (void **)&d.OMGN2D)) != FW_OK) return rc;
  size_t band = 1;
  if (f.wavemaker != FW_WK_NONE && f.ihi >= f.ilo && f.jhi >= f.jlo)
    band = (size_t)(f.ihi - f.ilo + 1) * (size_t)(f.jhi - f.jlo + 1);
  size_t cmn = band * (f.wavemaker == FW_WK_DATA2D ? (size_t)nfreq : 1);
  if ((rc = get(cmn * sizeof(R), (void **)&d.Cm)) != FW_OK) return rc;
  if ((rc = get(cmn * sizeof(R), (void **)&d.Sm)) != FW_OK) return rc;
  if ((rc = get((size_t)(M + 8) * sizeof(R), (void **)&d.wk_exp)) != FW_OK) return rc;
  size_t stage_elems = (size_t)(N + 8) > (size_t)(2 * nfreq + 8) ? (size_t)(N + 8) : (size_t)(2 * nfreq + 8);
  if ((rc = get(stage_elems * sizeof(R), (void **)&d.wk_stage)) != FW_OK) return rc;
  c->slot_bytes = (stage_elems * sizeof(R) + 255) & ~(size_t)255;
  HIPCHK(c, hipHostMalloc((void **)&c->pinned, c->slot_bytes * kStageSlots + 4096));
  for (int k = 0; k < kStageSlots; k++) HIPCHK(c, hipEventCreateWithFlags(&c->slot_ev[k], hipEventDisableTiming));
  // MASK_STRUC defaults to 1 (no structures); MASK9 / MASK to 1 until uploaded
  return FW_OK;
}

}  // namespace

namespace fwi {

template <class R>
int prepare_wavemaker(fw_ctx *c) {
  if (c->wk_ready || c->cfg.wavemaker != FW_WK_REG) {
    c->wk_ready = true;
    return FW_OK;
  }
  const fw_config &f = c->cfg;
  if ((int)c->xmk.size() < f.Mloc) return fail(c, FW_E_INVALID, "WK_REG needs xmk_wk uploaded");
  std::vector<R> e(f.ihi - f.ilo + 1);
  for (int i = f.ilo; i <= f.ihi; i++) {
    R dxk = (R)c->xmk[i - 1] - (R)f.Xc_WK;
    e[i - f.ilo] = hexp<R>(-(R)f.Beta_gen * (dxk * dxk));
  }
  HIPCHK(c, hipMemcpy(c->dev<R>().wk_exp, e.data(), e.size() * sizeof(R), hipMemcpyHostToDevice));
  c->wk_ready = true;
  return FW_OK;
}

template <class R>
int wavemaker_factors(fw_ctx *c, R T, R &AA, R *h, size_t *n) {
  const fw_config &f = c->cfg;
  AA = (R)0;
  *n = 0;
  if (f.wavemaker == FW_WK_REG) {
    AA = htanh<R>(HPI / ((R)f.Time_ramp * (R)f.Tperiod) * T) * (R)f.D_gen;
    if ((int)c->ymk.size() < f.Nloc) return fail(c, FW_E_INVALID, "WK_REG needs ymk_wk uploaded");
    for (int j = f.jlo; j <= f.jhi; j++)
      h[j - f.jlo] = hsin<R>((R)f.rlamda * ((R)c->ymk[j - 1] - (R)0) - (R)2 * HPI / (R)f.Tperiod * T);
    *n = f.jhi >= f.jlo ? (size_t)(f.jhi - f.jlo + 1) : 0;
  } else if (f.wavemaker == FW_WK_DATA2D) {
    AA = htanh<R>(HPI / ((R)f.Time_ramp / (R)f.FreqPeak) * T);
    if ((int)c->omg.size() < f.NumFreq) return fail(c, FW_E_INVALID, "WK_DATA2D needs OMGN2D uploaded");
    for (int k = 0; k < f.NumFreq; k++) hsincos<R>((R)c->omg[k] * T, h[f.NumFreq + k], h[k]);
    *n = (size_t)(2 * f.NumFreq);
  }
  return FW_OK;
}

template int prepare_wavemaker<float>(fw_ctx *);
template int prepare_wavemaker<double>(fw_ctx *);
template int wavemaker_factors<float>(fw_ctx *, float, float &, float *, size_t *);
template int wavemaker_factors<double>(fw_ctx *, double, double &, double *, size_t *);

}  // namespace fwi

namespace {

// per-stage factors: returns AA; queues the copy of the row/frequency table
template <class R>
int stage_wavemaker(fw_ctx *c, R &AA) {
  AA = (R)0;
  if (c->cfg.wavemaker == FW_WK_NONE) return FW_OK;
  int rc = prepare_wavemaker<R>(c);
  if (rc) return rc;
  const R T = (R)c->time;
  int slot = c->slot_next;
  c->slot_next = (c->slot_next + 1) % kStageSlots;
  HIPCHK(c, hipEventSynchronize(c->slot_ev[slot]));
  R *h = (R *)(c->pinned + (size_t)slot * c->slot_bytes);
  size_t n = 0;
  if ((rc = wavemaker_factors<R>(c, T, AA, h, &n))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->dev<R>().wk_stage, h, n * sizeof(R), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipEventRecord(c->slot_ev[slot], c->stream));
  return FW_OK;
}

// ---- routines ---------------------------------------------------------------
template <class R>
void halo_exchange_list(fw_ctx *c, HaloList &L, HaloList &Lf) {
  const Geo &g = c->g;
  int cnt = 6 * (g.M + 1) + 6 * (g.N + 1) + 64;
  // grid rows = arrays of the list
  dim3 b(256);
  auto grid = [&](const HaloList &l) { return dim3((cnt + 255) / 256, l.n); };
  if (!c->cfg.PERIODIC) {
    for (int k = 0; k < Lf.n; k++) L.a[L.n++] = Lf.a[k];
    if (L.n) launch(c, "halo", k_halo<R>, grid(L), b, L, g, 0);
  } else {
    if (L.n) launch(c, "halo", k_halo<R>, grid(L), b, L, g, 1);
    if (L.n) launch(c, "halo", k_halo<R>, grid(L), b, L, g, 2);
    if (Lf.n) launch(c, "halo", k_halo<R>, grid(Lf), b, Lf, g, 0);
  }
}


// EXCHANGE (bc.F:426-529), optionally followed by SPONGE_DAMPING
template <class R>
int do_exchange(fw_ctx *c, bool sponge) {
  Dev<R> &d = c->dev<R>();
  const fw_config &f = c->cfg;
  HaloList L{}, Lf{};
  hadd(L, d.Eta, 1);
  if (c->prm<R>().full_state) {  // mean-flow diagnostics, read by no routine of the step
    hadd(L, d.Wsurf, 1);
    hadd(L, d.P_center, 1);
    hadd(L, d.Q_center, 1);
    hadd(L, d.U_davg, 1);
    hadd(L, d.V_davg, 1);
  }
  if (f.VISCOSITY_BREAKING) hadd(L, d.AGE_BREAKING, 1);
  if (f.VISCOSITY_BREAKING) hadd(L, d.nu_break, 1);
  if (f.VISCOSITY_BREAKING || f.DIFFUSION_SPONGE) {
    hadd(Lf, d.P, 1, 0, 1);
    hadd(Lf, d.Q, 1, 0, 2);
  }
  hadd(L, d.MASK, 1, 1);
  hadd(L, d.U, 2);
  hadd(L, d.HU, 2);
  hadd(L, d.V, 3);
  hadd(L, d.HV, 3);
  halo_exchange_list<R>(c, L, Lf);
  launch(c, "mask_mul", k_mask_mul<R>, grid2(c->g.M, c->g.N, kB2), kB2, d, c->g, c->prm<R>(), sponge ? 1 : 0, 1);
  return FW_OK;
}

template <class R>
int do_sponge_only(fw_ctx *c) {
  launch(c, "sponge", k_mask_mul<R>, grid2(c->g.M, c->g.N, kB2), kB2, c->dev<R>(), c->g, c->prm<R>(), 1, 0);
  return FW_OK;
}

template <class R>
int do_estimate_dt(fw_ctx *c, bool copy_uv) {
  Dev<R> &d = c->dev<R>();
  int nblk = kRedBlocks;
  long ncell = (long)c->g.M * c->g.N;
  if ((ncell + 255) / 256 < nblk) nblk = (int)((ncell + 255) / 256);
  launch(c, "dt_partial", k_dt_partial<R>, dim3(nblk), dim3(256), d, c->g, c->prm<R>(), copy_uv ? 1 : 0);
  launch(c, "dt_final", k_dt_final<R>, dim3(1), dim3(256), d, c->prm<R>(), nblk, (R *)c->scal);
  R *h = (R *)(c->pinned + (size_t)kStageSlots * c->slot_bytes);
  HIPCHK(c, hipMemcpyAsync(h, c->scal, sizeof(R), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  R dt = h[0];
  c->dt = (double)dt;
  c->time = (double)((R)c->time + dt);
  return FW_OK;
}

template <class R>
int do_max_eta(fw_ctx *c, double *out) {
  Dev<R> &d = c->dev<R>();
  int nblk = 256;
  launch(c, "maxeta", k_maxeta_partial<R>, dim3(nblk), dim3(256), d, c->g);
  launch(c, "max_final", k_max_final<R>, dim3(1), dim3(256), d, nblk, (R *)c->scal + 1);
  R *h = (R *)(c->pinned + (size_t)kStageSlots * c->slot_bytes) + 2;
  HIPCHK(c, hipMemcpyAsync(h, (R *)c->scal + 1, sizeof(R), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *out = (double)h[0];
  return FW_OK;
}

// whether the one-pass dispersion kernel runs (production, non-periodic);
// in fw_step the RK update then forms the cross derivatives itself
template <class R>
bool disp_one_pass(fw_ctx *c) {
  return !c->prm<R>().full_state && !c->cfg.PERIODIC;
}

// in_step: called by fw_step, whose RK update forms the cross derivatives
// (the per-routine ESTIMATE_HUV reads them from the arrays)
template <class R>
int do_dispersion(fw_ctx *c, bool in_step = false) {
  Dev<R> &d = c->dev<R>();
  const Geo &g = c->g;
  if (disp_one_pass<R>(c)) {
    // one pass: the second derivatives stay in LDS (k_disp_fused)
    const dim3 b(kDispTX, kDispTY), gr((g.M + kDispTX - 1) / kDispTX, (g.N + kDispTY - 1) / kDispTY);
    Prm<R> p = c->prm<R>();
    p.xd_in_huv = in_step ? 1 : 0;
    if (spacing_pow2<R>(c))
      launch(c, "disp", k_disp_fused<R, true>, gr, b, d, g, p, (R)c->dt);
    else
      launch(c, "disp", k_disp_fused<R, false>, gr, b, d, g, p, (R)c->dt);
    return FW_OK;
  }
  if (spacing_pow2<R>(c))
    launch(c, "disp_deriv", k_disp_deriv<R, true>, grid2(g.M, g.N, kBT), kBT, d, g, c->prm<R>(), (R)c->dt);
  else
    launch(c, "disp_deriv", k_disp_deriv<R, false>, grid2(g.M, g.N, kBT), kBT, d, g, c->prm<R>(), (R)c->dt);
  HaloList L{}, Lf{};
  hadd(L, d.Uxx, 2); hadd(L, d.DUxx, 2); hadd(L, d.Vyy, 3); hadd(L, d.DVyy, 3);
  hadd(L, d.Uxy, 1); hadd(L, d.DUxy, 1); hadd(L, d.Vxy, 1); hadd(L, d.DVxy, 1);
  if (c->cfg.Gamma2 > 0 && !c->prm<R>().full_state) {
    hadd(L, d.ETAx, 2); hadd(L, d.ETAy, 3);
  } else if (c->cfg.Gamma2 > 0) {
    hadd(L, d.Ut, 2); hadd(L, d.Vt, 3); hadd(L, d.Utx, 1); hadd(L, d.Vty, 1);
    hadd(L, d.Utxx, 2); hadd(L, d.Vtyy, 3); hadd(L, d.Utxy, 1); hadd(L, d.Vtxy, 1);
    hadd(L, d.DUtxx, 2); hadd(L, d.DVtyy, 3); hadd(L, d.DUtxy, 1); hadd(L, d.DVtxy, 1);
    hadd(L, d.Ux, 1); hadd(L, d.DUx, 1); hadd(L, d.Vy, 1); hadd(L, d.DVy, 1);
    hadd(L, d.Uy, 3); hadd(L, d.DUy, 3); hadd(L, d.Vx, 2); hadd(L, d.DVx, 2);
    hadd(L, d.ETAx, 2); hadd(L, d.ETAy, 3);
  }
  halo_exchange_list<R>(c, L, Lf);
  if (spacing_pow2<R>(c))
    launch(c, "disp_terms", k_disp_terms<R, true>, grid2(g.M, g.N, kBT), kBT, d, g, c->prm<R>(), (R)c->dt);
  else
    launch(c, "disp_terms", k_disp_terms<R, false>, grid2(g.M, g.N, kBT), kBT, d, g, c->prm<R>(), (R)c->dt);
  return FW_OK;
}

// production launches (FW_OPT_FULL_STATE 0) take the flux kernels without
// the full-state arrays' code
template <class R, bool PROD>
int do_fluxes_as(fw_ctx *c) {
  Dev<R> &d = c->dev<R>();
  const Geo &g = c->g;
  const dim3 gx((g.M + 1 + kFluxXFaces - 1) / kFluxXFaces, (g.N + kB2.y - 1) / kB2.y);
  const bool disp = c->cfg.DISPERSION;
  if (disp)
    launch(c, "flux_x", k_flux_x<R, true, PROD>, gx, kB2, d, g, c->prm<R>());
  else
    launch(c, "flux_x", k_flux_x<R, false, PROD>, gx, kB2, d, g, c->prm<R>());
  if constexpr (sizeof(R) == 4) {
    // FP32: columns marched in strips of 4 faces (independent waves)
    constexpr int RS = 4;
    const int nt = (g.N + 1 + RS - 1) / RS;
    const dim3 by(64, 4), gy((g.M + 63) / 64, (nt + 3) / 4);
    if (disp)
      launch(c, "flux_y", k_flux_y_march<R, true, RS, PROD>, gy, by, d, g, c->prm<R>());
    else
      launch(c, "flux_y", k_flux_y_march<R, false, RS, PROD>, gy, by, d, g, c->prm<R>());
  } else {
    // FP64: the LDS tile (the marching window would need 216 VGPRs)
    constexpr int KY = FluxYK<R>::K, FY = FluxYTile<KY>::faces;
    const dim3 gy((g.M + 63) / 64, (g.N + 1 + FY - 1) / FY), by(64, kFluxYThreadRows);
    if (disp)
      launch(c, "flux_y", k_flux_y<R, true, KY, PROD>, gy, by, d, g, c->prm<R>());
    else
      launch(c, "flux_y", k_flux_y<R, false, KY, PROD>, gy, by, d, g, c->prm<R>());
  }
  return FW_OK;
}
template <class R>
int do_fluxes(fw_ctx *c) {
  return c->prm<R>().full_state ? do_fluxes_as<R, false>(c) : do_fluxes_as<R, true>(c);
}

template <class R>
int do_sources(fw_ctx *c) {
  R AA;
  int rc = stage_wavemaker<R>(c, AA);
  if (rc) return rc;
  const Geo &g = c->g;
  const dim3 gs = grid2(g.Ie - g.Ib + 1, g.Je - g.Jb + 1, kB2);
  if (spacing_pow2<R>(c))
    launch(c, "sources", k_sources<R, true>, gs, kB2, c->dev<R>(), g, c->prm<R>(), AA);
  else
    launch(c, "sources", k_sources<R, false>, gs, kB2, c->dev<R>(), g, c->prm<R>(), AA);
  return FW_OK;
}

template <class R>
int do_solve(fw_ctx *c);

template <class R>
int do_huv(fw_ctx *c, int istage) {
  if (istage < 1 || istage > 3) return fail(c, FW_E_INVALID, "istage must be 1..3");
  // SSP-RK3 weights, mod_global.F:122-123
  const R alpha[3] = {(R)0, (R)3 / (R)4, (R)1 / (R)3};
  const R beta[3] = {(R)1, (R)1 / (R)4, (R)2 / (R)3};
  Dev<R> &d = c->dev<R>();
  const Geo &g = c->g;
  // (the unfused update divides by no spacing)
  launch(c, "huv", k_huv<R, false, false>, grid2(g.M, g.N, kBH), kBH, d, g, c->prm<R>(), alpha[istage - 1],
         beta[istage - 1], (R)c->dt, (R)0);
  return do_solve<R>(c);
}

// SourceTerms fused into ESTIMATE_HUV (the whole-step path)
template <class R>
int do_mask(fw_ctx *c);
template <class R>
int do_breaking(fw_ctx *c);
template <class R>
int do_roller(fw_ctx *c);

// overlap: UPDATE_MASK and WAVE_BREAKING follow on two more streams,
// concurrently with the line solves
template <class R>
int do_src_huv(fw_ctx *c, int istage, bool overlap) {
  if (istage < 1 || istage > 3) return fail(c, FW_E_INVALID, "istage must be 1..3");
  const R alpha[3] = {(R)0, (R)3 / (R)4, (R)1 / (R)3};
  const R beta[3] = {(R)1, (R)1 / (R)4, (R)2 / (R)3};
  R AA;
  int rc = stage_wavemaker<R>(c, AA);
  if (rc) return rc;
  Dev<R> &d = c->dev<R>();
  const Geo &g = c->g;
  // with the one-pass dispersion of fw_step the update forms the cross
  // derivatives itself
  const bool xd = c->cfg.DISPERSION && disp_one_pass<R>(c);
  if (spacing_pow2<R>(c)) {
    if (xd)
      launch(c, "src_huv", k_huv<R, true, true, true>, grid2(g.M, g.N, kBH), kBH, d, g, c->prm<R>(),
             alpha[istage - 1], beta[istage - 1], (R)c->dt, AA);
    else
      launch(c, "src_huv", k_huv<R, true, true>, grid2(g.M, g.N, kBH), kBH, d, g, c->prm<R>(), alpha[istage - 1],
             beta[istage - 1], (R)c->dt, AA);
  } else {
    if (xd)
      launch(c, "src_huv", k_huv<R, true, false, true>, grid2(g.M, g.N, kBH), kBH, d, g, c->prm<R>(),
             alpha[istage - 1], beta[istage - 1], (R)c->dt, AA);
    else
      launch(c, "src_huv", k_huv<R, true, false>, grid2(g.M, g.N, kBH), kBH, d, g, c->prm<R>(), alpha[istage - 1],
             beta[istage - 1], (R)c->dt, AA);
  }
  if (!overlap) return do_solve<R>(c);
  // UPDATE_MASK reads Eta, Depth, H and the old MASK and writes MASKn,
  // MASK9, DepthX/Y; WAVE_BREAKING reads Eta, Depth, H, the dispersion
  // slopes and AGE and writes AGEn, nu_break; the line solves and the
  // velocity tail read the coefficients, H and the old MASK and write U, V,
  // HU, HV, Ubar, Vbar.  Nothing is shared but the old MASK, which is
  // double-buffered: the tail is launched first and keeps the old buffer;
  // do_mask swaps the host pointers afterwards.  So the three run at once,
  // UPDATE_MASK on the second stream, WAVE_BREAKING on the third.
  HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
  HIPCHK(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
  HIPCHK(c, hipStreamWaitEvent(c->side2, c->ev_fork, 0));
  if ((rc = do_solve<R>(c))) return rc;
  c->cur = c->side;
  rc = do_mask<R>(c);
  if (rc == FW_OK) {
    hipError_t e = hipEventRecord(c->ev_join, c->side);
    if (e != hipSuccess) rc = fail(c, FW_E_HIP, hipGetErrorString(e));
  }
  if (rc == FW_OK) {
    c->cur = c->side2;
    rc = do_breaking<R>(c);
    if (rc == FW_OK) {
      hipError_t e = hipEventRecord(c->ev_brk, c->side2);
      if (e != hipSuccess) rc = fail(c, FW_E_HIP, hipGetErrorString(e));
    }
  }
  c->cur = c->stream;
  if (rc) return rc;
  HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
  HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_brk, 0));
  return do_roller<R>(c);
}

template <class R>
int froude_cap_on_host(fw_ctx *c);

// line solves and the final velocities (GET_Eta_U_V_HU_HV after the
// coefficient assembly)
template <class R>
int do_solve(fw_ctx *c) {
  Dev<R> &d = c->dev<R>();
  const Geo &g = c->g;
  if (c->cfg.DISPERSION) {
    int ny = g.Je - g.Jb + 1, nx = g.Ie - g.Ib + 1;
    if (c->cfg.PERIODIC) {  // short periodic y lines: Sherman-Morrison per thread
      int bs = 64;
      int nxblk = (ny + bs - 1) / bs, nyblk = (nx + bs - 1) / bs;
      launch(c, "trid", k_trid<R>, dim3(nxblk + nyblk), dim3(bs), d, g, c->prm<R>(), nxblk);
    } else {
      // 16 lines per workgroup (2 x 2048 lines: one workgroup per CU);
      // FP64 rows take twice the LDS, so its chunks are 48 rows (125 KB)
      constexpr int L = 16, CH = sizeof(R) == 4 ? 64 : 48;
      int nxblk = (ny + L - 1) / L, nyblk = (nx + L - 1) / L;
      launch(c, "trid", k_trid_lds<R, L, CH>, dim3(nxblk + nyblk), dim3(256), d, g, nxblk);
    }
  }
  launch(c, "uv_final", k_uv_final<R>, grid2(g.M, g.N, kB2), kB2, d, g, c->prm<R>());
  if (c->prm<R>().host_math) {
    int rc = froude_cap_on_host<R>(c);
    if (rc) return rc;
  }
  if (!c->cfg.DISPERSION) {
    size_t n = plane_elems(g);
    launch(c, "ubar_hu", k_ubar_from_hu<R>, dim3(1024), dim3(256), d, n);
  }
  return FW_OK;
}

// The Froude cap (etauv_solver.F:580-586) of the cells k_uv_final listed,
// evaluated with the host's C library as the reference does (ATAN2, COS,
// SIN in the working precision), written back by k_fc_scatter.
template <class R>
int froude_cap_on_host(fw_ctx *c) {
  Dev<R> &d = c->dev<R>();
  int n = 0;
  HIPCHK(c, hipMemcpyAsync(&n, d.fc_n, sizeof(int), hipMemcpyDeviceToHost, c->cur));
  HIPCHK(c, hipStreamSynchronize(c->cur));
  if (n == 0) return FW_OK;
  std::vector<R> v(4 * (size_t)n);
  HIPCHK(c, hipMemcpy(v.data(), d.fc, v.size() * sizeof(R), hipMemcpyDeviceToHost));
  const R fcap = c->prm<R>().FCAP;
  for (int k = 0; k < n; k++) {
    const R u = v[4 * k], vv = v[4 * k + 1], fr = v[4 * k + 2], hm = v[4 * k + 3];
    const R th = hatan2<R>(vv, u);
    // COS(th) and SIN(th) as one sincos call, as the reference's compiler
    // evaluates the pair (the FP64 results of separate calls can differ)
    R sn, cs;
    hsincos<R>(th, sn, cs);
    const R un = fcap * fr * cs;
    const R vn = fcap * fr * sn;
    v[4 * k] = un;
    v[4 * k + 1] = vn;
    v[4 * k + 2] = un * hm;
    v[4 * k + 3] = vn * hm;
  }
  HIPCHK(c, hipMemcpy(d.fc, v.data(), v.size() * sizeof(R), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_fc_scatter<R>, dim3((n + 255) / 256), dim3(256), 0, c->cur, d, (const R *)d.fc, n);
  const int zero = 0;
  HIPCHK(c, hipMemcpyAsync(d.fc_n, &zero, sizeof(int), hipMemcpyHostToDevice, c->cur));
  HIPCHK(c, hipStreamSynchronize(c->cur));
  return FW_OK;
}

template <class R>
int do_mask(fw_ctx *c) {
  Dev<R> &d = c->dev<R>();
  const Geo &g = c->g;
  launch(c, "mask", k_mask<R>, grid2(g.M + 1, g.N + 1, kBT), kBT, d, g, c->prm<R>());
  std::swap(d.MASK, d.MASKn);
  return FW_OK;
}

template <class R>
int do_breaking(fw_ctx *c) {
  if (!c->cfg.SHOW_BREAKING) return FW_OK;
  Dev<R> &d = c->dev<R>();
  const Geo &g = c->g;
  launch(c, "brk1", k_brk1<R>, grid2(g.M, g.N, kBT), kBT, d, g, c->prm<R>(), (R)c->dt);
  launch(c, "brk2", k_brk2<R>, dim3(1), dim3(1024), d, g, c->prm<R>(), (R)c->dt);
  std::swap(d.AGE_BREAKING, d.AGEn);
  return FW_OK;
}

// roller flux and undertow (with every intermediate kept), after
// WAVE_BREAKING and after the line solves
template <class R>
int do_roller(fw_ctx *c) {
  if (!c->cfg.SHOW_BREAKING || !c->prm<R>().full_state) return FW_OK;
  const Geo &g = c->g;
  launch(c, "roller", k_roller<R>, grid2(g.M, g.N, kB2), kB2, c->dev<R>(), g, c->prm<R>());
  return FW_OK;
}

// The breaking kernels raise flags[0] when the in-order age propagation
// does not settle.  Once the host has seen it, it is cleared, so that the
// error is reported for the call that raised it and not for every later one.
int read_break_flag(fw_ctx *c, void *flags) {
  int flag = 0;
  HIPCHK(c, hipMemcpy(&flag, flags, sizeof(int), hipMemcpyDeviceToHost));
  if (!flag) return FW_OK;
  const int zero = 0;
  HIPCHK(c, hipMemcpy(flags, &zero, sizeof(int), hipMemcpyHostToDevice));
  return fail(c, FW_E_BREAK_ITER, "breaking-age propagation did not settle");
}

template <class R>
int do_step(fw_ctx *c, int nsteps) {
  Dev<R> &d = c->dev<R>();
  const fw_config &f = c->cfg;
  const size_t pe = plane_elems(c->g);
  int rc;
  for (int s = 0; s < nsteps; s++) {
    launch(c, "copy3", k_copy3<R>, dim3(2048), dim3(256), d.Eta0, (const R *)d.Eta, d.Ubar0, (const R *)d.Ubar,
           d.Vbar0, (const R *)d.Vbar, pe);
    if ((rc = do_exchange<R>(c, false))) return rc;
    if ((rc = do_estimate_dt<R>(c, true))) return rc;
    for (int st = 1; st <= 3; st++) {
      if (f.DISPERSION && (rc = do_dispersion<R>(c, true))) return rc;
      if ((rc = do_fluxes<R>(c))) return rc;
      if ((rc = do_src_huv<R>(c, st, true))) return rc;
      if ((rc = do_exchange<R>(c, f.DIRECT_SPONGE != 0))) return rc;
    }
    double m;
    if ((rc = do_max_eta<R>(c, &m))) return rc;
    if ((rc = read_break_flag(c, (void *)d.flags))) return rc;
    if (m > f.EtaBlowVal) {
      char buf[128];
      snprintf(buf, sizeof buf, "blow-up: max|eta|=%g > EtaBlowVal=%g at time %g", m, f.EtaBlowVal, c->time);
      return fail(c, FW_E_BLOWUP, buf);
    }
  }
  if (c->prof) return collect_profile(c);
  return FW_OK;
}

template <class R>
int do_stats(fw_ctx *c, fw_stats *out) {
  Dev<R> &d = c->dev<R>();
  int nblk = 256;
  launch(c, "stats", k_stats<R>, dim3(nblk), dim3(256), d, c->g, c->prm<R>());
  std::vector<double> h(nblk * 7);
  HIPCHK(c, hipMemcpyAsync(h.data(), d.red64, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  fw_stats s{};
  s.max_eta = -1e300;
  s.min_eta = 1e300;
  for (int b = 0; b < nblk; b++) {  // fixed order: deterministic
    s.mass_volume += h[b * 7 + 0];
    s.energy += h[b * 7 + 1];
    s.max_eta = std::max(s.max_eta, h[b * 7 + 2]);
    s.min_eta = std::min(s.min_eta, h[b * 7 + 3]);
    s.max_u = std::max(s.max_u, h[b * 7 + 4]);
    s.max_v = std::max(s.max_v, h[b * 7 + 5]);
    s.wet_cells += (int64_t)h[b * 7 + 6];
  }
  *out = s;
  return FW_OK;
}

// fw_reduce_stats without the context's partials buffer: one partial per
// interior row in the static scratch array, summed on the device, seven
// totals copied to the host (no allocation of any kind on this path)
template <class R>
int do_stats_noalloc(fw_ctx *c, fw_stats *out) {
  const Geo &g = c->g;
  const int rows = g.Je - g.Jb + 1;
  if (rows > kNoallocRows) {
    char buf[160];
    snprintf(buf, sizeof buf, "fw_reduce_stats_noalloc: %d interior rows exceed the static scratch of %d rows",
             rows, kNoallocRows);
    return fail(c, FW_E_INVALID, buf);
  }
  double *part = g_scratch_ptr[c->device];
  if (!part) return fail(c, FW_E_INVALID, "static scratch not resolved on this device");
  // plain launches: the profiling wrapper would record events
  hipLaunchKernelGGL(k_stats_rows<R>, dim3(rows), dim3(256), 0, c->stream, c->dev<R>(), g, c->prm<R>(), part);
  hipLaunchKernelGGL(k_stats_sum, dim3(1), dim3(64), 0, c->stream, part, rows);
  HIPCHK(c, hipGetLastError());
  double h[7];
  HIPCHK(c, hipMemcpyAsync(h, part + 7 * rows, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  fw_stats st{};
  st.mass_volume = h[0];
  st.energy = h[1];
  st.max_eta = h[2];
  st.min_eta = h[3];
  st.max_u = h[4];
  st.max_v = h[5];
  st.wet_cells = (int64_t)h[6];
  *out = st;
  return FW_OK;
}

// dispatch on the working precision
#define DISPATCH(c, call) ((c)->rb == 8 ? call<double> : call<float>)

void *field_ptr(fw_ctx *c, int id) {
  if (c->rb == 8) {
    Dev<double> &d = c->dd;
    switch (id) {
#define FW_CASE(n) \
  case F_##n: return d.n;
      FW_CELL_FIELDS(FW_CASE) FW_XFACE_FIELDS(FW_CASE) FW_YFACE_FIELDS(FW_CASE) FW_INT_FIELDS(FW_CASE)
#undef FW_CASE
    }
  } else {
    Dev<float> &d = c->df;
    switch (id) {
#define FW_CASE(n) \
  case F_##n: return d.n;
      FW_CELL_FIELDS(FW_CASE) FW_XFACE_FIELDS(FW_CASE) FW_YFACE_FIELDS(FW_CASE) FW_INT_FIELDS(FW_CASE)
#undef FW_CASE
    }
  }
  return nullptr;
}

}  // namespace

// ============================================================================
// C ABI
extern "C" {

int fw_create(const fw_config *cfg, fw_ctx **out) {
  if (!cfg || !out) return FW_E_INVALID;
  *out = nullptr;
  const fw_config &f = *cfg;
  if (f.real_bytes != 4 && f.real_bytes != 8) return FW_E_INVALID;
  if (f.Mloc < 7 || f.Nloc < 7 || f.Ibeg != 4 || f.Jbeg != 4 || f.Iend != f.Mloc - 3 || f.Jend != f.Nloc - 3)
    return FW_E_INVALID;
  if (f.wavemaker < 0 || f.wavemaker > 2) return FW_E_INVALID;
  if (f.wavemaker == FW_WK_DATA2D && f.NumFreq < 1) return FW_E_INVALID;
  if (f.wavemaker != FW_WK_NONE && (f.ilo < 1 || f.jlo < 1 || f.ihi > f.Mloc || f.jhi > f.Nloc)) {
    // an empty band (ilo > ihi) is allowed
    if (!(f.ilo > f.ihi || f.jlo > f.jhi)) return FW_E_INVALID;
  }
  if (f.PERIODIC && f.Nglob < 3) return FW_E_INVALID;
  fw_ctx *c = new fw_ctx();
  c->cfg = f;
  c->rb = f.real_bytes;
  c->g.M = f.Mloc;
  c->g.N = f.Nloc;
  c->g.P = ((f.Mloc + 1) + 63) / 64 * 64;
  c->g.Ib = f.Ibeg;
  c->g.Ie = f.Iend;
  c->g.Jb = f.Jbeg;
  c->g.Je = f.Jend;
  if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipStreamCreate(&c->side) != hipSuccess || hipStreamCreate(&c->side2) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_brk, hipEventDisableTiming) != hipSuccess) {
    fw_destroy(c);
    return FW_E_HIP;
  }
  c->cur = c->stream;
  int rc = resolve_scratch(c->device, c->err);
  if (rc != FW_OK) {
    fw_destroy(c);
    return rc;
  }
  if (c->rb == 8) {
    fill_prm<double>(c);
    rc = alloc_state<double>(c);
  } else {
    fill_prm<float>(c);
    rc = alloc_state<float>(c);
  }
  if (rc == FW_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = FW_E_HIP;
  if (rc != FW_OK) {
    fw_destroy(c);
    return rc;
  }
  // MASK_STRUC = 1 everywhere until uploaded (no structures)
  std::vector<int> ones(plane_elems(c->g), 1);
  void *ms = field_ptr(c, F_MASK_STRUC);
  if (hipMemcpy(ms, ones.data(), ones.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    fw_destroy(c);
    return FW_E_HIP;
  }
  *out = c;
  return FW_OK;
}

int fw_destroy(fw_ctx *c) {
  if (!c) return FW_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->side) (void)hipStreamSynchronize(c->side);
  if (c->side2) (void)hipStreamSynchronize(c->side2);
  for (auto &p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
  for (int k = 0; k < kStageSlots; k++)
    if (c->slot_ev[k]) (void)hipEventDestroy(c->slot_ev[k]);
  for (void *p : c->allocs) (void)hipFree(p);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->ev_brk) (void)hipEventDestroy(c->ev_brk);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->side2) (void)hipStreamDestroy(c->side2);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return FW_OK;
}

const char *fw_strerror(const fw_ctx *c) { return c ? c->err.c_str() : "null context"; }

int fw_synchronize(fw_ctx *c) {
  if (!c) return FW_E_INVALID;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FW_OK;
}

int fw_device_bytes(const fw_ctx *c, int64_t *bytes) {
  if (!c || !bytes) return FW_E_INVALID;
  *bytes = (int64_t)c->bytes;
  return FW_OK;
}

int fw_field_device(fw_ctx *c, int field, void **ptr, int *pitch) {
  if (!c || !ptr || field < 0 || field >= F_xmk_wk) return fail(c, FW_E_INVALID, "field has no device plane");
  *ptr = field_ptr(c, field);
  if (pitch) *pitch = c->g.P;
  return FW_OK;
}

int fw_get_stream(fw_ctx *c, void **stream) {
  if (!c || !stream) return FW_E_INVALID;
  *stream = (void *)c->stream;
  return FW_OK;
}

int fw_get_config(const fw_ctx *c, fw_config *cfg) {
  if (!c || !cfg) return FW_E_INVALID;
  *cfg = c->cfg;
  return FW_OK;
}

int fw_field_id(const char *name) {
  if (!name) return FW_E_NOTFOUND;
  for (int k = 0; k < F_COUNT; k++)
    if (strcmp(name, kFieldNames[k]) == 0) return k;
  return FW_E_NOTFOUND;
}

static int transfer(fw_ctx *c, int id, void *host, bool up) {
  if (!c || !host || id < 0 || id >= F_COUNT) return fail(c, FW_E_INVALID, "bad field or pointer");
  const fw_config &f = c->cfg;
  const Geo &g = c->g;
  FieldKind k = field_kind(id);
  size_t es = (k == K_INT) ? sizeof(int) : (size_t)c->rb;
  if (k != K_AUX) {
    int w = f.Mloc, h = f.Nloc;
    if (k == K_XFACE) w += 1;
    if (k == K_YFACE) h += 1;
    void *dp = field_ptr(c, id);
    if (up)
      HIPCHK(c, hipMemcpy2DAsync(dp, g.P * es, host, w * es, w * es, h, hipMemcpyHostToDevice, c->stream));
    else
      HIPCHK(c, hipMemcpy2DAsync(host, w * es, dp, g.P * es, w * es, h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FW_OK;
  }
  // band tables
  auto getd = [&](size_t q) -> double {
    return c->rb == 8 ? ((const double *)host)[q] : (double)((const float *)host)[q];
  };
  if (id == F_xmk_wk || id == F_ymk_wk || id == F_OMGN2D) {
    int n = id == F_xmk_wk ? f.Mloc : (id == F_ymk_wk ? f.Nloc : f.NumFreq);
    std::vector<double> &v = id == F_xmk_wk ? c->xmk : (id == F_ymk_wk ? c->ymk : c->omg);
    if (up) {
      v.resize(n);
      for (int q = 0; q < n; q++) v[q] = getd(q);
      c->wk_ready = false;
    } else {
      for (int q = 0; q < n && q < (int)v.size(); q++) {
        if (c->rb == 8) ((double *)host)[q] = v[q];
        else ((float *)host)[q] = (float)v[q];
      }
    }
    return FW_OK;
  }
  // Cm / Sm: (Mloc, Nloc, NumFreq) on the host, band only on the device
  if (f.wavemaker != FW_WK_DATA2D) return fail(c, FW_E_INVALID, "Cm/Sm only with WK_DATA2D");
  const int bw = f.ihi - f.ilo + 1, bh = f.jhi - f.jlo + 1, nf = f.NumFreq;
  if (bw <= 0 || bh <= 0) return FW_OK;
  size_t n = (size_t)bw * bh * nf;
  std::vector<char> tmp(n * c->rb);
  void *dp = (id == F_Cm) ? (c->rb == 8 ? (void *)c->dd.Cm : (void *)c->df.Cm)
                          : (c->rb == 8 ? (void *)c->dd.Sm : (void *)c->df.Sm);
  size_t M = f.Mloc, N = f.Nloc;
  if (up) {
    for (int k = 0; k < nf; k++)
      for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++) {
          size_t src = (size_t)(f.ilo - 1 + i) + (size_t)(f.jlo - 1 + j) * M + (size_t)k * M * N;
          size_t dst = (size_t)i + (size_t)j * bw + (size_t)k * bw * bh;
          memcpy(tmp.data() + dst * c->rb, (const char *)host + src * c->rb, c->rb);
        }
    HIPCHK(c, hipMemcpy(dp, tmp.data(), n * c->rb, hipMemcpyHostToDevice));
  } else {
    HIPCHK(c, hipMemcpy(tmp.data(), dp, n * c->rb, hipMemcpyDeviceToHost));
    memset(host, 0, M * N * nf * c->rb);
    for (int k = 0; k < nf; k++)
      for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++) {
          size_t dst = (size_t)(f.ilo - 1 + i) + (size_t)(f.jlo - 1 + j) * M + (size_t)k * M * N;
          size_t src = (size_t)i + (size_t)j * bw + (size_t)k * bw * bh;
          memcpy((char *)host + dst * c->rb, tmp.data() + src * c->rb, c->rb);
        }
  }
  return FW_OK;
}

int fw_upload(fw_ctx *c, int field, const void *host) { return transfer(c, field, (void *)host, true); }
int fw_download(fw_ctx *c, int field, void *host) { return transfer(c, field, host, false); }

int fw_set_clock(fw_ctx *c, double dt, double time) {
  if (!c) return FW_E_INVALID;
  c->dt = c->rb == 8 ? dt : (double)(float)dt;
  c->time = c->rb == 8 ? time : (double)(float)time;
  return FW_OK;
}
int fw_get_clock(const fw_ctx *c, double *dt, double *time) {
  if (!c) return FW_E_INVALID;
  if (dt) *dt = c->dt;
  if (time) *time = c->time;
  return FW_OK;
}

#define ENTRY(c, expr)                    \
  do {                                    \
    if (!(c)) return FW_E_INVALID;        \
    int rc_ = (expr);                     \
    if (rc_ == FW_OK && (c)->prof) rc_ = collect_profile(c); \
    return rc_;                           \
  } while (0)

int fw_exchange(fw_ctx *c) { ENTRY(c, DISPATCH(c, do_exchange)(c, false)); }
int fw_estimate_dt(fw_ctx *c, double *dt, double *time) {
  if (!c) return FW_E_INVALID;
  int rc = DISPATCH(c, do_estimate_dt)(c, false);
  if (rc) return rc;
  if (dt) *dt = c->dt;
  if (time) *time = c->time;
  return c->prof ? collect_profile(c) : FW_OK;
}
int fw_save_stage0(fw_ctx *c, int which) {
  if (!c) return FW_E_INVALID;
  const size_t pe = plane_elems(c->g);
  size_t bytes = pe * c->rb;
  if (which == 1) {
    void *e0 = field_ptr(c, F_Eta0), *e = field_ptr(c, F_Eta);
    void *u0 = field_ptr(c, F_Ubar0), *u = field_ptr(c, F_Ubar);
    void *v0 = field_ptr(c, F_Vbar0), *v = field_ptr(c, F_Vbar);
    HIPCHK(c, hipMemcpyAsync(e0, e, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(u0, u, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(v0, v, bytes, hipMemcpyDeviceToDevice, c->stream));
  } else if (which == 2) {
    HIPCHK(c, hipMemcpyAsync(field_ptr(c, F_U0), field_ptr(c, F_U), bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(field_ptr(c, F_V0), field_ptr(c, F_V), bytes, hipMemcpyDeviceToDevice, c->stream));
  } else {
    return fail(c, FW_E_INVALID, "which must be 1 or 2");
  }
  return FW_OK;
}
int fw_cal_dispersion(fw_ctx *c) { ENTRY(c, DISPATCH(c, do_dispersion)(c, false)); }
int fw_fluxes(fw_ctx *c) { ENTRY(c, DISPATCH(c, do_fluxes)(c)); }
int fw_source_terms(fw_ctx *c) { ENTRY(c, DISPATCH(c, do_sources)(c)); }
int fw_estimate_huv(fw_ctx *c, int istage) { ENTRY(c, DISPATCH(c, do_huv)(c, istage)); }
int fw_update_mask(fw_ctx *c) { ENTRY(c, DISPATCH(c, do_mask)(c)); }
int fw_wave_breaking(fw_ctx *c) {
  if (!c) return FW_E_INVALID;
  int rc = DISPATCH(c, do_breaking)(c);
  if (rc == FW_OK) rc = DISPATCH(c, do_roller)(c);
  if (rc) return rc;
  if ((rc = read_break_flag(c, c->rb == 8 ? (void *)c->dd.flags : (void *)c->df.flags))) return rc;
  return c->prof ? collect_profile(c) : FW_OK;
}
int fw_sponge_damping(fw_ctx *c) { ENTRY(c, DISPATCH(c, do_sponge_only)(c)); }
int fw_check_blowup(fw_ctx *c, double *max_abs_eta) {
  if (!c) return FW_E_INVALID;
  double m = 0;
  int rc = DISPATCH(c, do_max_eta)(c, &m);
  if (rc) return rc;
  if (max_abs_eta) *max_abs_eta = m;
  if (m > c->cfg.EtaBlowVal) return fail(c, FW_E_BLOWUP, "blow-up");
  return FW_OK;
}
int fw_step(fw_ctx *c, int nsteps, double *dt, double *time) {
  if (!c || nsteps < 0) return FW_E_INVALID;
  int rc = DISPATCH(c, do_step)(c, nsteps);
  if (dt) *dt = c->dt;
  if (time) *time = c->time;
  return rc;
}
int fw_reduce_stats(fw_ctx *c, fw_stats *out) {
  if (!c || !out) return FW_E_INVALID;
  return DISPATCH(c, do_stats)(c, out);
}

int fw_reduce_stats_noalloc(fw_ctx *c, fw_stats *out) {
  if (!c || !out) return FW_E_INVALID;
  return DISPATCH(c, do_stats_noalloc)(c, out);
}

int fw_set_option(fw_ctx *c, int option, int value) {
  if (!c) return FW_E_INVALID;
  if (option == FW_OPT_FULL_STATE) {
    c->pf.full_state = value != 0;
    c->pd.full_state = value != 0;
    return FW_OK;
  }
  if (option == FW_OPT_HOST_MATH) {
    if (value && !c->dd.fc_idx) {
      // list space for every interior cell, allocated on first use
      const size_t cap = (size_t)(c->g.Ie - c->g.Ib + 1) * (size_t)(c->g.Je - c->g.Jb + 1);
      void *idx = nullptr, *val = nullptr;
      hipError_t e = hipMalloc(&idx, cap * sizeof(int));
      if (e == hipSuccess) e = hipMalloc(&val, cap * 4 * (size_t)c->rb);
      if (e != hipSuccess) {
        if (idx) (void)hipFree(idx);
        return fail(c, FW_E_NOMEM, "FW_OPT_HOST_MATH: list allocation failed");
      }
      c->allocs.push_back(idx);
      c->alloc_sizes.push_back(cap * sizeof(int));
      c->allocs.push_back(val);
      c->alloc_sizes.push_back(cap * 4 * (size_t)c->rb);
      c->dd.fc_idx = c->df.fc_idx = (int *)idx;
      c->dd.fc = (double *)val;
      c->df.fc = (float *)val;
    }
    c->pf.host_math = value != 0;
    c->pd.host_math = value != 0;
    return FW_OK;
  }
  return fail(c, FW_E_INVALID, "unknown option");
}

int fw_profile_enable(fw_ctx *c, int on) {
  if (!c) return FW_E_INVALID;
  if (!on) collect_profile(c);
  c->prof = on != 0;
  return FW_OK;
}
int fw_profile_reset(fw_ctx *c) {
  if (!c) return FW_E_INVALID;
  collect_profile(c);
  c->acc.clear();
  return FW_OK;
}
int fw_profile_get(fw_ctx *c, const char *kernel, int64_t *launches, double *ms) {
  if (!c || !kernel) return FW_E_INVALID;
  int rc = collect_profile(c);
  if (rc) return rc;
  auto it = c->acc.find(kernel);
  if (it == c->acc.end()) {
    if (launches) *launches = 0;
    if (ms) *ms = 0;
    return FW_E_NOTFOUND;
  }
  if (launches) *launches = it->second.first;
  if (ms) *ms = it->second.second;
  return FW_OK;
}
int fw_profile_names(fw_ctx *c, char *buf, int buflen) {
  if (!c || !buf || buflen <= 0) return FW_E_INVALID;
  collect_profile(c);
  std::string s;
  for (auto &kv : c->acc) s += kv.first + ",";
  snprintf(buf, buflen, "%s", s.c_str());
  return FW_OK;
}
int fw_time_steps(fw_ctx *c, int nsteps, double *ms) {
  if (!c || !ms) return FW_E_INVALID;
  hipEvent_t a, b;
  HIPCHK(c, hipEventCreate(&a));
  HIPCHK(c, hipEventCreate(&b));
  HIPCHK(c, hipEventRecord(a, c->stream));
  int rc = DISPATCH(c, do_step)(c, nsteps);
  HIPCHK(c, hipEventRecord(b, c->stream));
  HIPCHK(c, hipEventSynchronize(b));
  float t = 0;
  (void)hipEventElapsedTime(&t, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = t;
  return rc;
}

}  // extern "C"
