// fw_batch.hip -- ensembles: many independent domains ("members") of uneven
// size advanced together, for the ensemble entry points of include/fwgpu.h.
//
// The reference advances one domain per program (src/main.F:284-535).  Here
// every phase of the time step (ghost fills, ESTIMATE_DT, CAL_DISPERSION,
// FLUXES, the RK update with SourceTerms, the line solves, UPDATE_MASK,
// WAVE_BREAKING, SPONGE_DAMPING, CHECK_BLOWUP) is one launch over all
// members.  Each launch has a tile queue: entry k names the member and the
// tile of that member's own grid that workgroup k works on, so members of
// any size share one launch with no padding, and the workgroups of one
// member sit in consecutive queue slots ordered for the XCDs (workgroup k
// goes to XCD k mod 8; XCD x gets a contiguous run of tiles, as xcd_tile()
// arranges within one domain).  The kernels are the per-domain kernels' own
// bodies (fw_kernels.h), so every member's arithmetic is exactly that of
// fw_step on the member alone.
//
// Nothing goes back to the host during a run: each member's DT, TIME and
// status live on the device (ESTIMATE_DT, misc.F:216-296, with FIXED_DT;
// CHECK_BLOWUP, misc.F:394-433), and the wavemaker factors of every step
// (sources.F:109-184, functions of TIME only) are evaluated ahead on the host
// with the C library for TIME = TIME + DT step after step, in the working
// precision, as the step forms it.  A member whose time step would change
// (FIXED_DT halving, misc.F:283-288) stops with FW_BATCH_DT_CHANGED before
// the step; one that blows up stops after it.  The UPDATE_MASK and
// WAVE_BREAKING buffers swap each stage; the launches take that parity as an
// argument, so one step, recorded once per starting parity, replays from a
// hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "fw_context.h"

using namespace fw;
using namespace fwi;

namespace fwb {

// Workgroup k of a launch: member m, tile (x, y) of the member's own grid
// for that kernel, aux = a per-kernel count (blocks of x lines, blocks of a
// grid-stride loop).
struct QEnt {
  int m, x, y, aux;
};

enum { ST_RUN = 0, ST_BLOWUP = 1, ST_DT = 2 };

template <class R>
struct Member {
  Dev<R> d;
  Geo g;
  Prm<R> p;
  HaloList hx[2];   // EXCHANGE arrays at buffer parity 0 / 1 (non-periodic: with P, Q)
  HaloList hxf[2];  // periodic: the face arrays P, Q, filled after the cell arrays
  HaloList hd;      // CAL_DISPERSION second derivatives (+ ETAx, ETAy)
  R *clk;           // [0] DT, [1] TIME, [2] max |eta|
  int *st;          // [0] status, [1] steps taken since the wavemaker table
                    // was filled, [2] steps taken in this run
  const R *wk;      // per step: AA, then the band's row / frequency factors
  int wks;          // stride of wk (0: no per-step wavemaker source)
  int show_brk;
  int ndt, nmax;    // workgroups of the ESTIMATE_DT / CHECK_BLOWUP reductions
  size_t pe;        // elements of one plane
};

// The member's arrays with the UPDATE_MASK buffers swapped when pm is set
// and the WAVE_BREAKING age buffers when pa is (and breaking is on).
template <class R>
__device__ __forceinline__ Dev<R> mdev(const Member<R> &M, int pm, int pa) {
  Dev<R> d = M.d;
  if (pm) {
    int *t = d.MASK;
    d.MASK = d.MASKn;
    d.MASKn = t;
  }
  if (pa && M.show_brk) {
    R *t = d.AGE_BREAKING;
    d.AGE_BREAKING = d.AGEn;
    d.AGEn = t;
  }
  return d;
}

#define FWB_HEAD                    \
  const QEnt e = q[blockIdx.x];     \
  const Member<R> &M = ms[e.m];     \
  if (M.st[0] != ST_RUN) return;

#define FWB_ARGS const Member<R> *__restrict__ ms, const QEnt *__restrict__ q

__device__ __forceinline__ Blk tile_blk(const QEnt &e, int gx = 0) { return Blk{e.x, e.y, gx, 1, 0}; }

// Eta0, Ubar0, Vbar0 := Eta, Ubar, Vbar (main.F:295-297)
template <class R>
__global__ void kb_copy3(FWB_ARGS) {
  FWB_HEAD
  const Dev<R> &d = M.d;
  for (size_t k = (size_t)e.x * blockDim.x + threadIdx.x; k < M.pe; k += (size_t)e.aux * blockDim.x) {
    const R a = d.Eta[k], b = d.Ubar[k], c = d.Vbar[k];
    d.Eta0[k] = a;
    d.Ubar0[k] = b;
    d.Vbar0[k] = c;
  }
}

// list 0: EXCHANGE cell arrays (non-periodic: with P, Q), 1: periodic P, Q,
// 2: CAL_DISPERSION's arrays; queue y = array of the list
template <class R>
__global__ void kb_halo(FWB_ARGS, int list, int mode, int pm) {
  FWB_HEAD
  const HaloList &L = list == 0 ? M.hx[pm] : (list == 1 ? M.hxf[pm] : M.hd);
  if (e.y >= L.n) return;
  halo_body<R>(L.a[e.y], M.g, mode, tile_blk(e));
}

// sponge < 0: the member's own DIRECT_SPONGE
template <class R>
__global__ void kb_mask_mul(FWB_ARGS, int pm, int pa, int sponge, int mul) {
  FWB_HEAD
  mask_mul_body<R>(mdev(M, pm, pa), M.g, M.p, sponge < 0 ? M.p.DIRECT_SPONGE : sponge, mul, tile_blk(e));
}

template <class R>
__global__ void kb_dt_partial(FWB_ARGS, int pm, int pa) {
  FWB_HEAD
  dt_partial_body<R>(mdev(M, pm, pa), M.g, M.p, 1, tile_blk(e, M.ndt));
}

// DT of the member (misc.F:216-296); with FIXED_DT it must come out as
// DT_fixed (no halving), else the member stops here with ST_DT; then
// TIME = TIME + DT
template <class R>
__global__ void kb_dt_final(FWB_ARGS) {
  FWB_HEAD
  dt_final_body<R>(M.d, M.p, M.ndt, M.clk, tile_blk(e));
  if (threadIdx.x == 0) {
    const R dt = M.clk[0];
    if (dt != M.p.DTF) {
      M.st[0] = ST_DT;
    } else {
      M.clk[1] = M.clk[1] + dt;
      M.st[1] = M.st[1] + 1;
      M.st[2] = M.st[2] + 1;
    }
  }
}

template <class R, bool P2>
__global__ void __launch_bounds__(WideLB<R>::wg512) kb_disp_deriv(FWB_ARGS, int pm, int pa) {
  FWB_HEAD
  disp_deriv_body<R, P2>(mdev(M, pm, pa), M.g, M.p, M.clk[0], tile_blk(e));
}

template <class R, bool P2>
__global__ void __launch_bounds__(WideLB<R>::wg512) kb_disp_terms(FWB_ARGS, int pm, int pa) {
  FWB_HEAD
  disp_terms_body<R, P2>(mdev(M, pm, pa), M.g, M.p, M.clk[0], tile_blk(e));
}

template <class R, bool P2>
__global__ void __launch_bounds__(kDispTX * kDispTY) kb_disp_fused(FWB_ARGS, int pm, int pa) {
  FWB_HEAD
  disp_fused_body<R, P2>(mdev(M, pm, pa), M.g, M.p, M.clk[0], tile_blk(e));
}

template <class R, bool DISP>
__global__ void __launch_bounds__(256) kb_flux_x(FWB_ARGS, int pm, int pa) {
  FWB_HEAD
  flux_x_body<R, DISP, true>(mdev(M, pm, pa), M.g, M.p, tile_blk(e));
}

template <class R, bool DISP, int RS>
__global__ void __launch_bounds__(256, 4) kb_flux_y_march(FWB_ARGS, int pm, int pa) {
  FWB_HEAD
  flux_y_march_body<R, DISP, RS, true>(mdev(M, pm, pa), M.g, M.p, tile_blk(e));
}

template <class R, bool DISP, int K>
__global__ void __launch_bounds__(64 * kFluxYThreadRows) kb_flux_y(FWB_ARGS, int pm, int pa) {
  FWB_HEAD
  flux_y_body<R, DISP, K, true>(mdev(M, pm, pa), M.g, M.p, tile_blk(e));
}

// SourceTerms + ESTIMATE_HUV of one stage; the wavemaker factors of the
// member's current step come from its table; XD: as in fw_step, the update
// forms the cross derivatives and U1p, V1p itself (the one-pass dispersion
// does not store them)
template <class R, bool P2, bool XD = false>
__global__ void __launch_bounds__(WideLB<R>::wg256) kb_huv(FWB_ARGS, int pm, int pa, R al, R be) {
  FWB_HEAD
  Dev<R> d = mdev(M, pm, pa);
  R AA = (R)0;
  if (M.wks) {
    const R *row = M.wk + (size_t)(M.st[1] - 1) * (size_t)M.wks;
    AA = row[0];
    d.wk_stage = const_cast<R *>(row + 1);
  }
  huv_body<R, true, P2, XD>(d, M.g, M.p, al, be, M.clk[0], AA, tile_blk(e));
}

// lines per workgroup of the line solves, as in fw_step (FP64 in 48-row
// chunks, so that its LDS tile fits)
constexpr int kTridLines = 16;
template <class R, int L, int CH>
__global__ void __launch_bounds__(256) kb_trid_lds(FWB_ARGS, int pm, int pa) {
  FWB_HEAD
  trid_lds_body<R, L, CH>(mdev(M, pm, pa), M.g, e.aux, tile_blk(e));
}

template <class R>
__global__ void kb_trid(FWB_ARGS, int pm, int pa) {
  FWB_HEAD
  trid_body<R>(mdev(M, pm, pa), M.g, M.p, e.aux, tile_blk(e));
}

template <class R>
__global__ void kb_uv_final(FWB_ARGS, int pm, int pa) {
  FWB_HEAD
  uv_final_body<R>(mdev(M, pm, pa), M.g, M.p, tile_blk(e));
}

template <class R>
__global__ void kb_ubar_from_hu(FWB_ARGS) {
  FWB_HEAD
  ubar_from_hu_body<R>(M.d, M.pe, tile_blk(e, e.aux));
}

template <class R>
__global__ void kb_mask(FWB_ARGS, int pm, int pa) {
  FWB_HEAD
  mask_body<R>(mdev(M, pm, pa), M.g, M.p, tile_blk(e));
}

template <class R>
__global__ void kb_brk1(FWB_ARGS, int pm, int pa) {
  FWB_HEAD
  if (!M.show_brk) return;
  brk1_body<R>(mdev(M, pm, pa), M.g, M.p, M.clk[0], tile_blk(e));
}

template <class R>
__global__ void __launch_bounds__(1024) kb_brk2(FWB_ARGS, int pm, int pa) {
  FWB_HEAD
  if (!M.show_brk) return;
  brk2_body<R>(mdev(M, pm, pa), M.g, M.p, M.clk[0], tile_blk(e));
}

template <class R>
__global__ void kb_maxeta_partial(FWB_ARGS) {
  FWB_HEAD
  maxeta_partial_body<R>(M.d, M.g, tile_blk(e, M.nmax));
}

// CHECK_BLOWUP (misc.F:394-433): a member over EtaBlowVal stops here
template <class R>
__global__ void kb_max_final(FWB_ARGS) {
  FWB_HEAD
  max_final_body<R>(M.d, M.nmax, M.clk + 2, tile_blk(e));
  if (threadIdx.x == 0 && M.clk[2] > M.p.BLOW) M.st[0] = ST_BLOWUP;
}

// start of a run: steps counter to zero (a kernel, not a memset, so that
// the call order is the same with and without graphs)
__global__ void kb_reset_steps(int *const *st, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) st[k][1] = 0;
}

// ---------------------------------------------------------------------------
// host side

// tunable workgroup shapes (the kernels take any shape within their
// launch bounds); names as reported by fw_batch_autotune
enum Group { G_DISP = 0, G_HUV, G_CELL, G_BRK, G_COUNT };
const char *kGroupName[G_COUNT] = {"disp (disp_deriv, disp_terms; periodic members)", "src_huv",
                                   "cell (uv_final, mask_mul)",
                                   "mask/brk (mask, brk1)"};

struct Queue {
  std::vector<QEnt> h;
  QEnt *d = nullptr;
  int n = 0;
  dim3 block;
};

}  // namespace fwb

using namespace fwb;

struct fw_batch {
  std::vector<fw_ctx *> mem;
  int n = 0;
  int rb = 4;
  bool periodic = false, disp = false, pow2 = false;
  std::string err;
  hipStream_t stream = nullptr, side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  void *members = nullptr;     // Member<R>[n] on the device
  std::vector<char> members_h;  // host image
  void *clk = nullptr;         // n x 4 R
  int *st = nullptr;           // n x 4 int
  int **st_ptrs = nullptr;     // device array of per-member status pointers
  void *wk = nullptr;          // per member: cap x stride R
  std::vector<int> wk_stride;
  std::vector<size_t> wk_off;
  int cap = 0;                 // steps per wavemaker table fill
  char *pinned = nullptr;      // host staging of the tables
  size_t pinned_bytes = 0;
  // queues
  std::map<std::string, Queue> q;
  dim3 shape[G_COUNT];
  // graphs of one step per starting parity
  hipGraphExec_t gx[2] = {nullptr, nullptr};
  bool use_graph = true;
  int parity = 0;  // stages run so far in this call, mod 2
  std::vector<int> status;
  std::vector<void *> allocs;
};

namespace {

int bfail(fw_batch *b, int code, const std::string &m) {
  if (b) b->err = m;
  return code;
}
#define BCHK(b, x)                                                                   \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) return bfail(b, FW_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

int balloc(fw_batch *b, size_t bytes, void **p) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return bfail(b, FW_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  b->allocs.push_back(*p);
  return FW_OK;
}

// Tiles of a gx x gy grid for every member (member-major, row-major), put in
// XCD order: slot k of the launch runs on XCD k mod 8, and XCD x receives the
// x-th contiguous eighth of the list.
void xcd_order(std::vector<QEnt> &v) {
  const int nb = (int)v.size();
  std::vector<QEnt> o(nb);
  const int per = nb >> 3, rem = nb & 7;
  for (int h = 0; h < nb; h++) {
    const int xcd = h & 7, k = h >> 3;
    o[h] = v[xcd * per + std::min(xcd, rem) + k];
  }
  v.swap(o);
}

int upload_queue(fw_batch *b, Queue &Q) {
  Q.n = (int)Q.h.size();
  if (Q.d) {
    (void)hipFree(Q.d);
    b->allocs.erase(std::remove(b->allocs.begin(), b->allocs.end(), (void *)Q.d), b->allocs.end());
    Q.d = nullptr;
  }
  if (!Q.n) return FW_OK;
  int rc = balloc(b, Q.h.size() * sizeof(QEnt), (void **)&Q.d);
  if (rc) return rc;
  BCHK(b, hipMemcpy(Q.d, Q.h.data(), Q.h.size() * sizeof(QEnt), hipMemcpyHostToDevice));
  return FW_OK;
}

// 2-D tile queue: member m covers ceil(nx(m)/bx) x ceil(ny(m)/by) tiles
template <class FX, class FY>
int make_tiles(fw_batch *b, const std::string &name, dim3 block, int bx, int by, FX nx, FY ny) {
  Queue &Q = b->q[name];
  Q.h.clear();
  Q.block = block;
  for (int m = 0; m < b->n; m++) {
    const Geo &g = b->mem[m]->g;
    const int tx = (nx(g) + bx - 1) / bx, ty = (ny(g) + by - 1) / by;
    for (int y = 0; y < ty; y++)
      for (int x = 0; x < tx; x++) Q.h.push_back(QEnt{m, x, y, 0});
  }
  xcd_order(Q.h);
  return upload_queue(b, Q);
}

// 1-D queue: member m gets cnt(m) workgroups, aux(m) passed along
template <class FC, class FA>
int make_line(fw_batch *b, const std::string &name, dim3 block, FC cnt, FA aux, bool order = true) {
  Queue &Q = b->q[name];
  Q.h.clear();
  Q.block = block;
  for (int m = 0; m < b->n; m++) {
    const int c = cnt(m), a = aux(m);
    for (int x = 0; x < c; x++) Q.h.push_back(QEnt{m, x, 0, a});
  }
  if (order) xcd_order(Q.h);
  return upload_queue(b, Q);
}

template <class R>
Member<R> &mem_h(fw_batch *b, int m) {
  return ((Member<R> *)b->members_h.data())[m];
}

// halo queue of one list: ghost chunks x arrays
template <class R>
int make_halo_queue(fw_batch *b, const std::string &name, int which) {
  Queue &Q = b->q[name];
  Q.h.clear();
  Q.block = dim3(256);
  for (int m = 0; m < b->n; m++) {
    const Member<R> &M = mem_h<R>(b, m);
    const HaloList &L = which == 0 ? M.hx[0] : (which == 1 ? M.hxf[0] : M.hd);
    const Geo &g = M.g;
    const int cnt = 6 * (g.M + 1) + 6 * (g.N + 1) + 64;
    const int nx = (cnt + 255) / 256;
    for (int y = 0; y < L.n; y++)
      for (int x = 0; x < nx; x++) Q.h.push_back(QEnt{m, x, y, 0});
  }
  return upload_queue(b, Q);
}

// EXCHANGE list of a member at buffer parity par (do_exchange's order)
template <class R>
void exchange_lists(fw_ctx *c, const Dev<R> &d, HaloList &L, HaloList &Lf) {
  const fw_config &f = c->cfg;
  L = HaloList{};
  Lf = HaloList{};
  hadd(L, d.Eta, 1);
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
  if (!f.PERIODIC) {
    for (int k = 0; k < Lf.n; k++) L.a[L.n++] = Lf.a[k];
    Lf.n = 0;
  }
}

template <class R>
void swapped(const Dev<R> &in, Dev<R> &out, bool brk) {
  out = in;
  std::swap(out.MASK, out.MASKn);
  if (brk) std::swap(out.AGE_BREAKING, out.AGEn);
}

// the member table from the contexts' current arrays (parity 0 = as the
// contexts hold them now)
template <class R>
int fill_members(fw_batch *b) {
  b->members_h.assign(sizeof(Member<R>) * b->n, 0);
  for (int m = 0; m < b->n; m++) {
    fw_ctx *c = b->mem[m];
    Member<R> &M = mem_h<R>(b, m);
    M.d = c->dev<R>();
    M.g = c->g;
    M.p = c->prm<R>();
    // the one-pass dispersion leaves the cross derivatives and U1p, V1p to
    // the RK update (as in fw_step)
    M.p.xd_in_huv = (c->cfg.DISPERSION && !c->cfg.PERIODIC) ? 1 : 0;
    M.show_brk = c->cfg.SHOW_BREAKING;
    Dev<R> d1;
    swapped(M.d, d1, M.show_brk != 0);
    exchange_lists<R>(c, M.d, M.hx[0], M.hxf[0]);
    exchange_lists<R>(c, d1, M.hx[1], M.hxf[1]);
    M.hd = HaloList{};
    const Dev<R> &d = M.d;
    hadd(M.hd, d.Uxx, 2); hadd(M.hd, d.DUxx, 2); hadd(M.hd, d.Vyy, 3); hadd(M.hd, d.DVyy, 3);
    hadd(M.hd, d.Uxy, 1); hadd(M.hd, d.DUxy, 1); hadd(M.hd, d.Vxy, 1); hadd(M.hd, d.DVxy, 1);
    if (c->cfg.Gamma2 > 0) {
      hadd(M.hd, d.ETAx, 2);
      hadd(M.hd, d.ETAy, 3);
    }
    M.clk = (R *)b->clk + 4 * m;
    M.st = b->st + 4 * m;
    M.wks = b->wk_stride[m];
    M.wk = M.wks ? (const R *)((char *)b->wk + b->wk_off[m]) : nullptr;
    const long ncell = (long)c->g.M * c->g.N;
    M.ndt = (int)std::min<long>(kRedBlocks, (ncell + 255) / 256);
    M.nmax = std::min(256, c->g.Je - c->g.Jb + 1);
    M.pe = plane_elems(c->g);
  }
  BCHK(b, hipMemcpy(b->members, b->members_h.data(), b->members_h.size(), hipMemcpyHostToDevice));
  return FW_OK;
}

inline dim3 shape_or(dim3 s, dim3 dflt) { return s.x ? s : dflt; }

// queues of every launch (after the member table; the tunable ones with
// the current shapes)
template <class R>
int build_queues(fw_batch *b) {
  int rc;
  auto M_ = [](const Geo &g) { return g.M; };
  auto N_ = [](const Geo &g) { return g.N; };
  const dim3 sd = shape_or(b->shape[G_DISP], dim3(64, 8)), sh = shape_or(b->shape[G_HUV], dim3(64, 4));
  const dim3 sc = shape_or(b->shape[G_CELL], dim3(64, 4)), sb = shape_or(b->shape[G_BRK], dim3(64, 8));
  if ((rc = make_tiles(b, "disp", sd, sd.x, sd.y, M_, N_))) return rc;
  if ((rc = make_tiles(b, "dispf", dim3(kDispTX, kDispTY), kDispTX, kDispTY, M_, N_))) return rc;
  if ((rc = make_tiles(b, "huv", sh, sh.x, sh.y, M_, N_))) return rc;
  if ((rc = make_tiles(b, "cell", sc, sc.x, sc.y, M_, N_))) return rc;
  if ((rc = make_tiles(b, "brk1", sb, sb.x, sb.y, M_, N_))) return rc;
  if ((rc = make_tiles(b, "mask", sb, sb.x, sb.y, [](const Geo &g) { return g.M + 1; },
                       [](const Geo &g) { return g.N + 1; })))
    return rc;
  if ((rc = make_tiles(b, "flux_x", dim3(64, 4), kFluxXFaces, 4, [](const Geo &g) { return g.M + 1; }, N_)))
    return rc;
  if (b->rb == 4) {  // columns marched in strips of 4 faces, 64 x 4 threads
    if ((rc = make_tiles(b, "flux_y", dim3(64, 4), 64, 16, M_, [](const Geo &g) { return g.N + 1; }))) return rc;
  } else {  // the LDS tile
    constexpr int FY = FluxYTile<FluxYK<double>::K>::faces;
    if ((rc = make_tiles(b, "flux_y", dim3(64, kFluxYThreadRows), 64, FY, M_, [](const Geo &g) { return g.N + 1; })))
      return rc;
  }
  auto mem = [b](int m) -> const Member<R> & { return mem_h<R>(b, m); };
  if ((rc = make_line(b, "dt", dim3(256), [&](int m) { return mem(m).ndt; }, [](int) { return 0; }))) return rc;
  if ((rc = make_line(b, "max", dim3(256), [&](int m) { return mem(m).nmax; }, [](int) { return 0; }))) return rc;
  if ((rc = make_line(b, "one", dim3(256), [](int) { return 1; }, [](int) { return 0; }))) return rc;
  if ((rc = make_line(b, "one1024", dim3(1024), [](int) { return 1; }, [](int) { return 0; }))) return rc;
  auto chunks = [&](int m) { return (int)std::min<size_t>(2048, (mem(m).pe + 2047) / 2048); };
  if ((rc = make_line(b, "copy", dim3(256), chunks, chunks))) return rc;
  if (b->periodic) {
    auto nx = [&](int m) { const Geo &g = mem(m).g; return (g.Je - g.Jb + 1 + 63) / 64; };
    auto ny = [&](int m) { const Geo &g = mem(m).g; return (g.Ie - g.Ib + 1 + 63) / 64; };
    if ((rc = make_line(b, "trid", dim3(64), [&](int m) { return nx(m) + ny(m); }, nx))) return rc;
  } else {
    constexpr int L = kTridLines;
    auto nx = [&](int m) { const Geo &g = mem(m).g; return (g.Je - g.Jb + 1 + L - 1) / L; };
    auto ny = [&](int m) { const Geo &g = mem(m).g; return (g.Ie - g.Ib + 1 + L - 1) / L; };
    if ((rc = make_line(b, "trid", dim3(256), [&](int m) { return nx(m) + ny(m); }, nx))) return rc;
  }
  if ((rc = make_halo_queue<R>(b, "halo_x", 0))) return rc;
  if ((rc = make_halo_queue<R>(b, "halo_xf", 1))) return rc;
  if ((rc = make_halo_queue<R>(b, "halo_d", 2))) return rc;
  return FW_OK;
}

}  // namespace

// kernels are launched through this helper (the member table type follows
// the precision)
namespace {
template <class R, typename K, typename... A>
void bl(fw_batch *b, hipStream_t s, const char *qname, K k, A... args) {
  Queue &Q = b->q[qname];
  if (!Q.n) return;
  hipLaunchKernelGGL(k, dim3(Q.n), Q.block, 0, s, (const Member<R> *)b->members, (const QEnt *)Q.d, args...);
}

// EXCHANGE (+ SPONGE_DAMPING when sponge) at buffer parity par
template <class R>
void rec_exchange(fw_batch *b, int par, int sponge) {
  hipStream_t s = b->stream;
  if (!b->periodic) {
    bl<R>(b, s, "halo_x", kb_halo<R>, 0, 0, par);
  } else {
    bl<R>(b, s, "halo_x", kb_halo<R>, 0, 1, par);
    bl<R>(b, s, "halo_x", kb_halo<R>, 0, 2, par);
    bl<R>(b, s, "halo_xf", kb_halo<R>, 1, 0, par);
  }
  bl<R>(b, s, "cell", kb_mask_mul<R>, par, par, sponge, 1);
}

// one whole step, starting at buffer parity p0 (do_step's sequence)
template <class R, bool P2>
int rec_step(fw_batch *b, int p0) {
  hipStream_t s = b->stream, sd = b->side;
  const R alpha[3] = {(R)0, (R)3 / (R)4, (R)1 / (R)3};
  const R beta[3] = {(R)1, (R)1 / (R)4, (R)2 / (R)3};
  int par = p0;
  bl<R>(b, s, "copy", kb_copy3<R>);
  rec_exchange<R>(b, par, 0);
  bl<R>(b, s, "dt", kb_dt_partial<R>, par, par);
  bl<R>(b, s, "one", kb_dt_final<R>);
  for (int st = 0; st < 3; st++) {
    if (b->disp && !b->periodic) {
      bl<R>(b, s, "dispf", kb_disp_fused<R, P2>, par, par);
    } else if (b->disp) {
      bl<R>(b, s, "disp", kb_disp_deriv<R, P2>, par, par);
      if (!b->periodic) {
        bl<R>(b, s, "halo_d", kb_halo<R>, 2, 0, par);
      } else {
        bl<R>(b, s, "halo_d", kb_halo<R>, 2, 1, par);
        bl<R>(b, s, "halo_d", kb_halo<R>, 2, 2, par);
      }
      bl<R>(b, s, "disp", kb_disp_terms<R, P2>, par, par);
    }
    if (b->disp) bl<R>(b, s, "flux_x", kb_flux_x<R, true>, par, par);
    else bl<R>(b, s, "flux_x", kb_flux_x<R, false>, par, par);
    if constexpr (sizeof(R) == 4) {
      if (b->disp) bl<R>(b, s, "flux_y", kb_flux_y_march<R, true, 4>, par, par);
      else bl<R>(b, s, "flux_y", kb_flux_y_march<R, false, 4>, par, par);
    } else {
      if (b->disp) bl<R>(b, s, "flux_y", kb_flux_y<R, true, FluxYK<R>::K>, par, par);
      else bl<R>(b, s, "flux_y", kb_flux_y<R, false, FluxYK<R>::K>, par, par);
    }
    if (b->disp && !b->periodic) bl<R>(b, s, "huv", kb_huv<R, P2, true>, par, par, alpha[st], beta[st]);
    else bl<R>(b, s, "huv", kb_huv<R, P2>, par, par, alpha[st], beta[st]);
    // line solves and the velocity tail beside UPDATE_MASK + WAVE_BREAKING
    BCHK(b, hipEventRecord(b->ev_fork, s));
    BCHK(b, hipStreamWaitEvent(sd, b->ev_fork, 0));
    if (b->disp) {
      if (b->periodic) bl<R>(b, s, "trid", kb_trid<R>, par, par);
      else bl<R>(b, s, "trid", kb_trid_lds<R, kTridLines, sizeof(R) == 4 ? 64 : 48>, par, par);
    }
    bl<R>(b, s, "cell", kb_uv_final<R>, par, par);
    if (!b->disp) bl<R>(b, s, "copy", kb_ubar_from_hu<R>);
    bl<R>(b, sd, "mask", kb_mask<R>, par, par);
    bl<R>(b, sd, "brk1", kb_brk1<R>, par ^ 1, par);
    bl<R>(b, sd, "one1024", kb_brk2<R>, par ^ 1, par);
    BCHK(b, hipEventRecord(b->ev_join, sd));
    BCHK(b, hipStreamWaitEvent(s, b->ev_join, 0));
    par ^= 1;
    rec_exchange<R>(b, par, -1);
  }
  bl<R>(b, s, "max", kb_maxeta_partial<R>);
  bl<R>(b, s, "one", kb_max_final<R>);
  return FW_OK;
}

template <class R>
int rec_step_any(fw_batch *b, int p0) {
  return b->pow2 ? rec_step<R, true>(b, p0) : rec_step<R, false>(b, p0);
}

int destroy_graphs(fw_batch *b) {
  for (auto &g : b->gx)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  return FW_OK;
}

template <class R>
int build_graphs(fw_batch *b) {
  destroy_graphs(b);
  for (int p0 = 0; p0 < 2; p0++) {
    hipGraph_t g = nullptr;
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = rec_step_any<R>(b, p0);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return bfail(b, FW_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&b->gx[p0], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return bfail(b, FW_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  return FW_OK;
}

// wavemaker rows of steps k0+1 .. k0+cnt of every member (TIME advanced by
// DT_fixed each step in the working precision, as kb_dt_final does)
template <class R>
int fill_wavemaker(fw_batch *b, const std::vector<R> &t0, int cnt) {
  for (int m = 0; m < b->n; m++) {
    const int ws = b->wk_stride[m];
    if (!ws) continue;
    fw_ctx *c = b->mem[m];
    R *h = (R *)(b->pinned + b->wk_off[m]);
    R T = t0[m];
    const R dt = (R)c->cfg.DT_fixed;
    for (int k = 0; k < cnt; k++) {
      T = T + dt;
      R AA;
      size_t nn = 0;
      int rc = wavemaker_factors<R>(c, T, AA, h + (size_t)k * ws + 1, &nn);
      if (rc) return bfail(b, rc, "member " + std::to_string(m) + ": " + c->err);
      h[(size_t)k * ws] = AA;
    }
  }
  BCHK(b, hipMemcpyAsync(b->wk, b->pinned, b->pinned_bytes, hipMemcpyHostToDevice, b->stream));
  return FW_OK;
}

// the clocks and states from the contexts, before a run
template <class R>
int upload_clocks(fw_batch *b, std::vector<R> &t0) {
  std::vector<R> clk(4 * b->n, (R)0);
  std::vector<int> st(4 * b->n, 0);
  t0.resize(b->n);
  for (int m = 0; m < b->n; m++) {
    clk[4 * m + 0] = (R)b->mem[m]->dt;
    clk[4 * m + 1] = (R)b->mem[m]->time;
    t0[m] = (R)b->mem[m]->time;
    st[4 * m] = b->status[m] == FW_BATCH_BLOWUP ? ST_BLOWUP : (b->status[m] == FW_BATCH_DT_CHANGED ? ST_DT : ST_RUN);
  }
  BCHK(b, hipMemcpyAsync(b->clk, clk.data(), clk.size() * sizeof(R), hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipMemcpyAsync(b->st, st.data(), st.size() * sizeof(int), hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  return FW_OK;
}

// after a run: clocks and statuses back into the contexts, and the buffer
// parity applied to their host-side arrays
template <class R>
int download_clocks(fw_batch *b) {
  std::vector<R> clk(4 * b->n);
  std::vector<int> st(4 * b->n);
  BCHK(b, hipMemcpyAsync(clk.data(), b->clk, clk.size() * sizeof(R), hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipMemcpyAsync(st.data(), b->st, st.size() * sizeof(int), hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  for (int m = 0; m < b->n; m++) {
    fw_ctx *c = b->mem[m];
    // a member stopped by a changing time step keeps DT_fixed, the step of
    // its last completed step
    c->dt = st[4 * m] == ST_DT ? (double)(R)c->cfg.DT_fixed : (double)clk[4 * m + 0];
    c->time = (double)clk[4 * m + 1];
    // the buffers swap once per step taken (three stages); a member that
    // stopped took fewer steps than the others
    if (st[4 * m + 2] & 1) {
      Dev<R> &d = c->dev<R>();
      std::swap(d.MASK, d.MASKn);
      if (c->cfg.SHOW_BREAKING) std::swap(d.AGE_BREAKING, d.AGEn);
    }
    int flag = 0;
    BCHK(b, hipMemcpy(&flag, c->dev<R>().flags, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) {
      const int zero = 0;
      BCHK(b, hipMemcpy(c->dev<R>().flags, &zero, sizeof(int), hipMemcpyHostToDevice));
    }
    const int s = st[4 * m];
    if (s == ST_BLOWUP) b->status[m] = FW_BATCH_BLOWUP;
    else if (s == ST_DT) b->status[m] = FW_BATCH_DT_CHANGED;
    else if (flag) b->status[m] = FW_BATCH_BREAK_ITER;
  }
  b->parity = 0;
  return FW_OK;
}

// nsteps steps of every running member; tables refilled every cap steps
template <class R>
int run_steps(fw_batch *b, int nsteps, double *ms) {
  int rc;
  if ((rc = fill_members<R>(b))) return rc;
  std::vector<R> t0;
  if ((rc = upload_clocks<R>(b, t0))) return rc;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ms) {
    BCHK(b, hipEventCreate(&e0));
    BCHK(b, hipEventCreate(&e1));
  }
  int done = 0;
  b->parity = 0;
  bool timing_started = false;
  while (done < nsteps) {
    const int cnt = std::min(b->cap, nsteps - done);
    if ((rc = fill_wavemaker<R>(b, t0, cnt))) return rc;
    hipLaunchKernelGGL(kb_reset_steps, dim3((b->n + 255) / 256), dim3(256), 0, b->stream, b->st_ptrs, b->n);
    if (ms && !timing_started) {
      BCHK(b, hipEventRecord(e0, b->stream));
      timing_started = true;
    }
    for (int k = 0; k < cnt; k++) {
      const int p0 = b->parity;
      if (b->use_graph) {
        BCHK(b, hipGraphLaunch(b->gx[p0], b->stream));
      } else {
        if ((rc = rec_step_any<R>(b, p0))) return rc;
      }
      b->parity ^= 1;  // three stages: the parity flips once per step
    }
    // TIME at the end of this chunk, as the device forms it
    for (int m = 0; m < b->n; m++) {
      const R dt = (R)b->mem[m]->cfg.DT_fixed;
      for (int k = 0; k < cnt; k++) t0[m] = t0[m] + dt;
    }
    done += cnt;
    if (done < nsteps) BCHK(b, hipStreamSynchronize(b->stream));  // pinned tables reused
  }
  if (ms) {
    BCHK(b, hipEventRecord(e1, b->stream));
    BCHK(b, hipEventSynchronize(e1));
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  return download_clocks<R>(b);
}

int first_error(fw_batch *b) {
  for (int m = 0; m < b->n; m++) {
    const int s = b->status[m];
    if (s == FW_BATCH_BLOWUP) return bfail(b, FW_E_BLOWUP, "member " + std::to_string(m) + " blew up");
    if (s == FW_BATCH_DT_CHANGED)
      return bfail(b, FW_E_INVALID, "member " + std::to_string(m) + ": the fixed time step would halve");
    if (s == FW_BATCH_BREAK_ITER)
      return bfail(b, FW_E_BREAK_ITER, "member " + std::to_string(m) + ": breaking-age propagation did not settle");
  }
  return FW_OK;
}

// --- autotuner ---------------------------------------------------------------
// Snapshot of every device allocation of every member (the state a step
// can change), so that trial steps leave no trace.
struct Snapshot {
  std::vector<void *> buf;
};

int snap_take(fw_batch *b, Snapshot &s) {
  for (fw_ctx *c : b->mem) {
    size_t tot = 0;
    for (size_t z : c->alloc_sizes) tot += z;
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, tot);
    if (e != hipSuccess) return bfail(b, FW_E_NOMEM, "snapshot: hipMalloc failed");
    s.buf.push_back(p);
    size_t off = 0;
    for (size_t k = 0; k < c->allocs.size(); k++) {
      BCHK(b, hipMemcpyAsync((char *)p + off, c->allocs[k], c->alloc_sizes[k], hipMemcpyDeviceToDevice, b->stream));
      off += c->alloc_sizes[k];
    }
  }
  BCHK(b, hipStreamSynchronize(b->stream));
  return FW_OK;
}

int snap_restore(fw_batch *b, Snapshot &s, const std::vector<double> &dt, const std::vector<double> &tm) {
  for (size_t m = 0; m < b->mem.size(); m++) {
    fw_ctx *c = b->mem[m];
    size_t off = 0;
    for (size_t k = 0; k < c->allocs.size(); k++) {
      BCHK(b, hipMemcpyAsync(c->allocs[k], (char *)s.buf[m] + off, c->alloc_sizes[k], hipMemcpyDeviceToDevice,
                             b->stream));
      off += c->alloc_sizes[k];
    }
    c->dt = dt[m];
    c->time = tm[m];
  }
  BCHK(b, hipStreamSynchronize(b->stream));
  return FW_OK;
}

template <class R>
int autotune(fw_batch *b, int trial, std::string &rep) {
  // candidate shapes per group, within each kernel's launch bounds
  const int lim_disp = WideLB<R>::wg512, lim_huv = WideLB<R>::wg256;
  std::vector<dim3> cand[G_COUNT] = {
      {dim3(64, 8), dim3(64, 4), dim3(128, 4), dim3(128, 2), dim3(64, 16)},
      {dim3(64, 4), dim3(64, 2), dim3(128, 2), dim3(64, 8), dim3(256, 1)},
      {dim3(64, 4), dim3(64, 2), dim3(128, 2), dim3(256, 1), dim3(64, 8)},
      {dim3(64, 8), dim3(64, 4), dim3(128, 4), dim3(128, 2)}};
  const int lim[G_COUNT] = {lim_disp, lim_huv, 1024, 1024};
  std::vector<double> dt0, tm0;
  std::vector<int> st0 = b->status;
  std::vector<Dev<R>> dev0;  // the contexts' buffer assignment (trial runs swap it)
  for (fw_ctx *c : b->mem) {
    dt0.push_back(c->dt);
    tm0.push_back(c->time);
    dev0.push_back(c->dev<R>());
  }
  Snapshot snap;
  int rc = snap_take(b, snap);
  auto trial_ms = [&](double &ms) -> int {
    int r;
    if ((r = fill_members<R>(b))) return r;
    if ((r = build_queues<R>(b))) return r;
    if ((r = build_graphs<R>(b))) return r;
    // one untimed step, then the timed ones
    if ((r = run_steps<R>(b, 1, nullptr))) return r;
    if ((r = run_steps<R>(b, trial, &ms))) return r;
    if ((r = snap_restore(b, snap, dt0, tm0))) return r;
    for (int m = 0; m < b->n; m++) b->mem[m]->dev<R>() = dev0[m];
    b->status = st0;
    return FW_OK;
  };
  for (int gi = 0; gi < G_COUNT && rc == FW_OK; gi++) {
    // the two-pass dispersion kernels run only for periodic members (the
    // others take the fused kernel, whose tile is fixed)
    if (gi == G_DISP && !(b->disp && b->periodic)) continue;
    dim3 best = cand[gi][0];
    double best_ms = 1e30;
    for (const dim3 &s : cand[gi]) {
      if ((int)(s.x * s.y) > lim[gi]) continue;
      b->shape[gi] = s;
      double ms = 0;
      if ((rc = trial_ms(ms))) break;
      char line[160];
      snprintf(line, sizeof line, "%s %ux%u: %.3f ms/step\n", kGroupName[gi], s.x, s.y, ms / trial);
      rep += line;
      if (ms < best_ms) {
        best_ms = ms;
        best = s;
      }
    }
    b->shape[gi] = best;
    char line[160];
    snprintf(line, sizeof line, "-> %s: %ux%u\n", kGroupName[gi], best.x, best.y);
    rep += line;
  }
  for (void *p : snap.buf) (void)hipFree(p);
  if (rc) return rc;
  if ((rc = fill_members<R>(b))) return rc;
  if ((rc = build_queues<R>(b))) return rc;
  return build_graphs<R>(b);
}

template <class R>
int batch_setup(fw_batch *b) {
  int rc;
  const int n = b->n;
  if ((rc = balloc(b, sizeof(Member<R>) * n, &b->members))) return rc;
  if ((rc = balloc(b, sizeof(R) * 4 * n, &b->clk))) return rc;
  if ((rc = balloc(b, sizeof(int) * 4 * n, (void **)&b->st))) return rc;
  if ((rc = balloc(b, sizeof(int *) * n, (void **)&b->st_ptrs))) return rc;
  std::vector<int *> sp(n);
  for (int m = 0; m < n; m++) sp[m] = b->st + 4 * m;
  BCHK(b, hipMemcpy(b->st_ptrs, sp.data(), n * sizeof(int *), hipMemcpyHostToDevice));
  // wavemaker tables: cap steps per fill
  b->cap = 256;
  b->wk_stride.assign(n, 0);
  b->wk_off.assign(n, 0);
  size_t off = 0;
  for (int m = 0; m < n; m++) {
    const fw_config &f = b->mem[m]->cfg;
    int w = 0;
    if (f.wavemaker == FW_WK_REG && f.jhi >= f.jlo) w = 1 + (f.jhi - f.jlo + 1);
    if (f.wavemaker == FW_WK_DATA2D) w = 1 + 2 * f.NumFreq;
    if (f.wavemaker == FW_WK_REG && f.jhi < f.jlo) w = 1;
    b->wk_stride[m] = w;
    b->wk_off[m] = off;
    off += ((size_t)w * b->cap * sizeof(R) + 255) & ~(size_t)255;
    if ((rc = prepare_wavemaker<R>(b->mem[m]))) return bfail(b, rc, b->mem[m]->err);
  }
  b->pinned_bytes = off ? off : 256;
  if ((rc = balloc(b, b->pinned_bytes, &b->wk))) return rc;
  BCHK(b, hipHostMalloc((void **)&b->pinned, b->pinned_bytes));
  if ((rc = fill_members<R>(b))) return rc;
  if ((rc = build_queues<R>(b))) return rc;
  return build_graphs<R>(b);
}

}  // namespace

extern "C" {

int fw_batch_create(fw_ctx *const *members, int n, fw_batch **out) {
  if (!members || n < 1 || !out) return FW_E_INVALID;
  *out = nullptr;
  fw_batch *b = new fw_batch();
  b->mem.assign(members, members + n);
  b->n = n;
  b->status.assign(n, FW_BATCH_RUNNING);
  fw_ctx *c0 = members[0];
  auto pow2_of = [](fw_ctx *c) {
    return c->rb == 8 ? (c->pd.dx_pow2 && c->pd.dy_pow2) : (c->pf.dx_pow2 && c->pf.dy_pow2);
  };
  auto full = [](fw_ctx *c) { return c->rb == 8 ? c->pd.full_state : c->pf.full_state; };
  int rc = FW_OK;
  for (int m = 0; m < n && rc == FW_OK; m++) {
    fw_ctx *c = members[m];
    if (!c) { rc = bfail(b, FW_E_INVALID, "null member"); break; }
    for (int k = 0; k < m; k++)
      if (members[k] == c) rc = bfail(b, FW_E_INVALID, "a context appears twice");
    if (c->rb != c0->rb || c->cfg.DISPERSION != c0->cfg.DISPERSION || c->cfg.PERIODIC != c0->cfg.PERIODIC ||
        pow2_of(c) != pow2_of(c0))
      rc = bfail(b, FW_E_INVALID, "members must share precision, DISPERSION, PERIODIC and power-of-two spacing");
    if (!c->cfg.FIXED_DT || !(c->cfg.DT_fixed > 0))
      rc = bfail(b, FW_E_INVALID, "member " + std::to_string(m) + " has no FIXED_DT");
    if (full(c)) rc = bfail(b, FW_E_INVALID, "members run in production mode (FW_OPT_FULL_STATE = 0)");
    if (c->pf.host_math || c->pd.host_math)
      rc = bfail(b, FW_E_INVALID, "FW_OPT_HOST_MATH needs a host round trip per stage: not in a batch");
  }
  if (rc == FW_OK) {
    b->rb = c0->rb;
    b->disp = c0->cfg.DISPERSION;
    b->periodic = c0->cfg.PERIODIC;
    b->pow2 = pow2_of(c0);
    if (hipSetDevice(c0->device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess ||
        hipStreamCreate(&b->side) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess)
      rc = bfail(b, FW_E_HIP, "stream/event creation failed");
  }
  // the members' own queued work (uploads) completes first
  for (int m = 0; m < n && rc == FW_OK; m++)
    if (hipStreamSynchronize(members[m]->stream) != hipSuccess) rc = bfail(b, FW_E_HIP, "member stream");
  if (rc == FW_OK) rc = b->rb == 8 ? batch_setup<double>(b) : batch_setup<float>(b);
  if (rc != FW_OK) {
    // the message stays readable through the first member's fw_strerror
    if (members[0]) members[0]->err = b->err;
    fw_batch_destroy(b);
    return rc;
  }
  *out = b;
  return FW_OK;
}

int fw_batch_destroy(fw_batch *b) {
  if (!b) return FW_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->side) (void)hipStreamSynchronize(b->side);
  destroy_graphs(b);
  for (void *p : b->allocs) (void)hipFree(p);
  if (b->pinned) (void)hipHostFree(b->pinned);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_join) (void)hipEventDestroy(b->ev_join);
  if (b->side) (void)hipStreamDestroy(b->side);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return FW_OK;
}

const char *fw_batch_strerror(const fw_batch *b) { return b ? b->err.c_str() : "null batch"; }

int fw_batch_size(const fw_batch *b, int *n) {
  if (!b || !n) return FW_E_INVALID;
  *n = b->n;
  return FW_OK;
}

int fw_batch_step(fw_batch *b, int nsteps) {
  if (!b || nsteps < 0) return FW_E_INVALID;
  if (nsteps == 0) return FW_OK;
  // members' own queued work first (the batch has its own streams)
  for (fw_ctx *c : b->mem) BCHK(b, hipStreamSynchronize(c->stream));
  int rc = b->rb == 8 ? run_steps<double>(b, nsteps, nullptr) : run_steps<float>(b, nsteps, nullptr);
  if (rc) return rc;
  return first_error(b);
}

int fw_batch_time_steps(fw_batch *b, int nsteps, double *ms) {
  if (!b || nsteps < 1 || !ms) return FW_E_INVALID;
  for (fw_ctx *c : b->mem) BCHK(b, hipStreamSynchronize(c->stream));
  int rc = b->rb == 8 ? run_steps<double>(b, nsteps, ms) : run_steps<float>(b, nsteps, ms);
  if (rc) return rc;
  return first_error(b);
}

int fw_batch_status(const fw_batch *b, int *status) {
  if (!b || !status) return FW_E_INVALID;
  for (int m = 0; m < b->n; m++) status[m] = b->status[m];
  return FW_OK;
}

int fw_batch_set_graph(fw_batch *b, int on) {
  if (!b) return FW_E_INVALID;
  b->use_graph = on != 0;
  return FW_OK;
}

int fw_batch_autotune(fw_batch *b, int trial_steps, char *report, int len) {
  if (!b || trial_steps < 1) return FW_E_INVALID;
  for (fw_ctx *c : b->mem) BCHK(b, hipStreamSynchronize(c->stream));
  std::string rep;
  int rc = b->rb == 8 ? autotune<double>(b, trial_steps, rep) : autotune<float>(b, trial_steps, rep);
  if (report && len > 0) snprintf(report, len, "%s", rep.c_str());
  return rc;
}

int fw_batch_get_shapes(const fw_batch *b, int *shapes) {
  if (!b || !shapes) return FW_E_INVALID;
  const dim3 dflt[G_COUNT] = {dim3(64, 8), dim3(64, 4), dim3(64, 4), dim3(64, 8)};
  for (int k = 0; k < G_COUNT; k++) {
    const dim3 s = b->shape[k].x ? b->shape[k] : dflt[k];
    shapes[2 * k] = (int)s.x;
    shapes[2 * k + 1] = (int)s.y;
  }
  return FW_OK;
}

}  // extern "C"
