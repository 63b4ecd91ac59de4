// fw_context.h -- the solver context (the HBM-resident state of one domain
// and its host-side bookkeeping) and the host helpers shared by the
// per-domain entry points (fw_solver.hip) and the ensemble launches
// (fw_batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "fw_kernels.h"
#include "fwgpu.h"

namespace fwi {
using namespace fw;
inline const char *kFieldNames[F_COUNT] = {
#define FW_NAME(n) #n,
    FW_CELL_FIELDS(FW_NAME) FW_XFACE_FIELDS(FW_NAME) FW_YFACE_FIELDS(FW_NAME) FW_INT_FIELDS(FW_NAME)
        FW_AUX_FIELDS(FW_NAME)
#undef FW_NAME
};

inline FieldKind field_kind(int id) {
  if (id < F_DepthX) return K_CELL;
  if (id < F_DepthY) return K_XFACE;
  if (id < F_MASK) return K_YFACE;
  if (id < F_xmk_wk) return K_INT;
  return K_AUX;
}

struct PendingEvent {
  std::string name;
  hipEvent_t a, b;
};

constexpr int kRedBlocks = 1024;
constexpr int kStageSlots = 8;

}  // namespace fwi

using fw::Dev;
using fw::Geo;
using fw::Prm;
using fwi::PendingEvent;
using fwi::kStageSlots;

struct fw_ctx {
  fw_config cfg{};
  int rb = 4;
  Geo g{};
  hipStream_t stream = nullptr;
  // fw_step runs UPDATE_MASK on a second stream and WAVE_BREAKING on a
  // third while the line solves run (the three share no array; see
  // do_src_huv); cur is the stream launch() uses
  hipStream_t side = nullptr, side2 = nullptr, cur = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_brk = nullptr;
  int device = 0;
  std::vector<void *> allocs;
  std::vector<size_t> alloc_sizes;  // bytes of each allocation (snapshots)
  size_t bytes = 0;
  Dev<float> df{};
  Dev<double> dd{};
  Prm<float> pf{};
  Prm<double> pd{};
  double dt = 0, time = 0;
  // host copies of band tables
  std::vector<double> xmk, ymk, omg;
  bool wk_ready = false;
  // pinned host ring for per-stage wavemaker factors and scalar readback
  char *pinned = nullptr;
  size_t slot_bytes = 0;
  hipEvent_t slot_ev[kStageSlots]{};
  int slot_next = 0;
  void *scal = nullptr;  // device scalars
  std::string err;
  bool prof = false;
  std::vector<PendingEvent> pending;
  std::vector<hipEvent_t> evpool;
  std::map<std::string, std::pair<int64_t, double>> acc;

  template <class R> Dev<R> &dev();
  template <class R> Prm<R> &prm();
};
template <> inline Dev<float> &fw_ctx::dev<float>() { return df; }
template <> inline Dev<double> &fw_ctx::dev<double>() { return dd; }
template <> inline Prm<float> &fw_ctx::prm<float>() { return pf; }
template <> inline Prm<double> &fw_ctx::prm<double>() { return pd; }

namespace fwi {

inline int fail(fw_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}
#define HIPCHK(c, x)                                                                  \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess)                                                             \
      return fail(c, FW_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));      \
  } while (0)

inline size_t plane_elems(const Geo &g) { return (size_t)g.P * (size_t)(g.N + 1); }

inline void hadd(HaloList &L, void *p, int vtype, int isint = 0, int spec = 0) {
  L.a[L.n++] = HaloArr{p, vtype, isint, spec};
}

// ---- wavemaker tables (host, C library, working precision) ----------------
template <class R> R hexp(R x);
template <> inline float hexp(float x) { return expf(x); }
template <> inline double hexp(double x) { return exp(x); }
template <class R> R hsin(R x);
template <> inline float hsin(float x) { return sinf(x); }
template <> inline double hsin(double x) { return sin(x); }
template <class R> R hcos(R x);
template <> inline float hcos(float x) { return cosf(x); }
template <> inline double hcos(double x) { return cos(x); }
// COS and SIN of one argument (sources.F:153-154): the reference's compiler
// evaluates such a pair with one sincos call, whose double result differs
// from separate sin and cos calls of the C library in the last bit for
// about 7 in 10^4 arguments, so the pair is taken from sincos here too
template <class R> void hsincos(R x, R &s, R &c);
template <> inline void hsincos(float x, float &s, float &c) { sincosf(x, &s, &c); }
template <> inline void hsincos(double x, double &s, double &c) { sincos(x, &s, &c); }
template <class R> R hatan2(R y, R x);
template <> inline float hatan2(float y, float x) { return atan2f(y, x); }
template <> inline double hatan2(double y, double x) { return atan2(y, x); }
template <class R> R htanh(R x);
template <> inline float htanh(float x) { return tanhf(x); }
template <> inline double htanh(double x) { return tanh(x); }
#define HPI ((R)3.141592653f)


// WK_REG: the band profile exp(-Beta_gen (x - Xc_WK)^2) uploaded once.
template <class R>
int prepare_wavemaker(fw_ctx *c);
// The per-stage wavemaker factors at model time T (sources.F:109-184):
// returns AA and fills h (row phases sin(...) for WK_REG; cos/sin of
// omega T per frequency for WK_DATA2D), *n = values written.
template <class R>
int wavemaker_factors(fw_ctx *c, R T, R &AA, R *h, size_t *n);

}  // namespace fwi
