// fw_kernels.h -- HIP kernels of the per-step solver (gfx950).
//
// Floating-point contract: every kernel evaluates the reference's
// expressions with the same operand order and no contraction (the library
// is built with -ffp-contract=off), with correctly rounded division and
// square root, so results are bitwise those of the reference's CPU build at
// the same precision.  The exceptions are the transcendental calls that the
// reference evaluates per cell (ATAN2 in the breaking angle and the Froude
// cap, COS/SIN in the Froude cap): those use the device math library and can
// differ from the host C library in the last bit.
// Time-only transcendentals (wavemaker ramps and phases) are evaluated on the
// host with the C library, as the reference does, and passed in.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#include "fw_fields.h"

namespace fw {

#define IX(i, j) ((size_t)((j)-1) * (size_t)g.P + (size_t)((i)-1))

// Tile of this workgroup in a 2-D launch.  Workgroups are dispatched to the
// 8 XCDs in turn (workgroup h of the launch, counted x-fastest, goes to XCD
// h mod 8), and each XCD has its own L2.  Handing XCD x a contiguous run of
// tiles in row-major order keeps a tile's left/right and upper/lower
// neighbours, whose cells its stencils also read, on the same XCD and close
// in time, so those cache lines come from HBM once rather than once per XCD.
struct Tile {
  int x, y;
};

// The workgroup's place in its launch: tile (x, y) of a gx x gy grid.  A
// plain launch takes it from blockIdx/gridDim (grid_blk); a launch over the
// members of an ensemble takes it from its tile queue (the member's own
// grid), already in the XCD order below (remap = 0).
struct Blk {
  int x, y, gx, gy, remap;
};
__device__ __forceinline__ Blk grid_blk() {
  return Blk{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y, 1};
}
__device__ __forceinline__ Tile xcd_tile(const Blk &B) {
  if (!B.remap) return Tile{B.x, B.y};
  const int gx = B.gx;
  const int nb = gx * B.gy;
  const int h = B.x + B.y * gx;
  const int xcd = h & 7, k = h >> 3;
  const int per = nb >> 3, rem = nb & 7;
  const int lin = xcd * per + min(xcd, rem) + k;
  return Tile{lin % gx, lin / gx};
}

template <class R> struct Lit;
template <> struct Lit<float> {
  static __host__ __device__ constexpr float grav() { return 9.81f; }
  static __host__ __device__ constexpr float small() { return 0.000001f; }
  static __host__ __device__ constexpr float large() { return 999999.0f; }
};
template <> struct Lit<double> {
  static __host__ __device__ constexpr double grav() { return 9.81; }
  static __host__ __device__ constexpr double small() { return 0.000001; }
  static __host__ __device__ constexpr double large() { return 999999.0; }
};
// pi is a default-REAL literal in the reference (src/mod_param.F:71)
#define FW_PI ((R)3.141592653f)
#define GRAV (Lit<R>::grav())
#define SMALL (Lit<R>::small())

template <class R> __device__ __forceinline__ R rmax(R a, R b) { return a > b ? a : b; }
template <class R> __device__ __forceinline__ R rmin(R a, R b) { return a < b ? a : b; }
__device__ __forceinline__ float dsqrt(float x) { return __builtin_sqrtf(x); }  // correctly rounded (llvm.sqrt.f32)
__device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float dabs(float x) { return fabsf(x); }
__device__ __forceinline__ double dabs(double x) { return fabs(x); }
// single-precision transcendentals are evaluated in double and rounded once,
// i.e. correctly rounded except in rare hard cases
__device__ __forceinline__ float datan2(float y, float x) { return (float)atan2((double)y, (double)x); }
__device__ __forceinline__ double datan2(double y, double x) { return atan2(y, x); }
__device__ __forceinline__ float dcos(float x) { return (float)cos((double)x); }
__device__ __forceinline__ double dcos(double x) { return cos(x); }
__device__ __forceinline__ float dsin(float x) { return (float)sin((double)x); }
__device__ __forceinline__ double dsin(double x) { return sin(x); }
// SIN and COS of one argument with one range reduction (the same values as
// dsin, dcos)
__device__ __forceinline__ void dsincos(float x, float &s, float &c) {
  double sd, cd;
  sincos((double)x, &sd, &cd);
  s = (float)sd;
  c = (float)cd;
}
__device__ __forceinline__ void dsincos(double x, double &s, double &c) { sincos(x, &s, &c); }
__device__ __forceinline__ float ddiv(float a, float b) { return __fdiv_rn(a, b); }
__device__ __forceinline__ double ddiv(double a, double b) { return a / b; }
// SIGN(1.0, b) with IEEE signed zero
template <class R> __device__ __forceinline__ R sign1(R b) { return signbit(b) ? (R)-1 : (R)1; }
__device__ __forceinline__ float copysign_(float m, float s) { return __builtin_copysignf(m, s); }
__device__ __forceinline__ double copysign_(double m, double s) { return __builtin_copysign(m, s); }
// x / 2 and x * 0.5 are the same correctly rounded value
template <class R> __device__ __forceinline__ R half(R x) { return x * (R)0.5; }
// Division by a grid spacing: a multiplication by the reciprocal when both
// spacings are powers of two (bitwise the same, as both are exact), a
// correctly rounded division otherwise.  The choice is a template argument
// of the kernels (P2), made at launch: a run-time choice puts a branch
// around every derivative, and loads are not moved across branches, so a
// stencil's loads would go out a few at a time.
template <class R>
struct Dv {
  R d, r;
  bool pow2;
};
template <class R> __device__ __forceinline__ R dv(R x, const Dv<R> &s) { return s.pow2 ? x * s.r : ddiv(x, s.d); }
template <bool P2, class R> __device__ __forceinline__ Dv<R> dvx(const Prm<R> &p) { return Dv<R>{p.DX, p.rDX, P2}; }
template <bool P2, class R> __device__ __forceinline__ Dv<R> dvy(const Prm<R> &p) { return Dv<R>{p.DY, p.rDY, P2}; }

// --------------------------------------------------------------------------
// Ghost-cell fills (PHI_COLL bc.F:631-1272, PHI_COLL_VARIABLE_LENGTH
// bc.F:539-617).  The reference fills x ghosts on interior rows, then y
// ghosts over all columns, so in the non-periodic case every ghost value is
// a pure function of one interior value: phi(mi, mj) times the mirror signs.
// With PERIODIC, the y ghost rows are copied first over all columns (the
// x-ghost columns still holding their previous values) and the x ghosts are
// filled afterwards; that takes two ordered launches (modes 1 and 2).
struct HaloArr {
  void *p;
  int vtype;   // 1 even, 2 odd in x, 3 odd in y
  int isint;
  int spec;    // 0 cell, 1 x-face (P), 2 y-face (Q)
};
struct HaloList {
  HaloArr a[32];
  int n;
};

struct HaloSpec {
  int Ma, Na, Ib, Ie, Jb, Je;
};

__device__ __forceinline__ HaloSpec halo_spec(const Geo &g, int spec) {
  HaloSpec s{g.M, g.N, g.Ib, g.Ie, g.Jb, g.Je};
  if (spec == 1) { s.Ma = g.M + 1; s.Ie = g.Ie + 1; }
  if (spec == 2) { s.Na = g.N + 1; s.Je = g.Je + 1; }
  return s;
}

// ghost index t -> (i, j); bottom rows, top rows, then left/right columns of
// interior rows.  Returns false past the end.
__device__ __forceinline__ bool halo_pos(const HaloSpec &s, int t, int mode, int &i, int &j) {
  int rows = 3 * s.Ma;
  int nr = s.Je - s.Jb + 1;
  if (mode != 2) {
    if (t < rows) { j = 1 + t / s.Ma; i = 1 + t % s.Ma; return true; }
    t -= rows;
    if (t < rows) { j = s.Je + 1 + t / s.Ma; i = 1 + t % s.Ma; return true; }
    t -= rows;
    if (mode == 1) return false;
  }
  if (t < 3 * nr) { j = s.Jb + t / 3; i = 1 + t % 3; return true; }
  t -= 3 * nr;
  if (t < 3 * nr) { j = s.Jb + t / 3; i = s.Ie + 1 + t % 3; return true; }
  return false;
}

// Ghost fill of one array A; the workgroup's x tile picks its ghost cells.
template <class R>
__device__ __forceinline__ void halo_body(const HaloArr &A, Geo g, int mode, const Blk &B) {
  int t = B.x * blockDim.x + threadIdx.x;
  HaloSpec s = halo_spec(g, A.spec);
  int i, j;
  if (!halo_pos(s, t, mode, i, j)) return;
  int si = i, sj = j;
  bool neg = false;
  if (mode == 1) {  // periodic y copy (raw, no sign)
    sj = (j <= 3) ? (s.Je - 3 + j) : (3 + (j - s.Je));
  } else {
    if (i < s.Ib) { si = 2 * s.Ib - 1 - i; neg ^= (A.vtype == 2); }
    if (i > s.Ie) { si = 2 * s.Ie + 1 - i; neg ^= (A.vtype == 2); }
    if (mode == 0) {
      if (j < s.Jb) { sj = 2 * s.Jb - 1 - j; neg ^= (A.vtype == 3); }
      if (j > s.Je) { sj = 2 * s.Je + 1 - j; neg ^= (A.vtype == 3); }
    }
  }
  if (A.isint) {
    int *ph = (int *)A.p;
    ph[IX(i, j)] = ph[IX(si, sj)];
  } else {
    R *ph = (R *)A.p;
    R v = ph[IX(si, sj)];
    ph[IX(i, j)] = neg ? -v : v;
  }
}
// One array of the list per grid row (blockIdx.y): within a mode every ghost
// value is read from interior cells only, so the arrays are independent and
// their fills run side by side instead of one load-store round trip after
// another.
template <class R>
__global__ void k_halo(HaloList L, Geo g, int mode) {
  halo_body<R>(L.a[blockIdx.y], g, mode, grid_blk());
}

// --------------------------------------------------------------------------
// ESTIMATE_DT (misc.F:216-296) over all cells including ghosts; the block
// minima are reduced by k_dt_final.  Also U0 := U, V0 := V (main.F:331-332)
// when copy_uv is set.
template <class R>
__device__ __forceinline__ void dt_partial_body(Dev<R> d, Geo g, Prm<R> p, int copy_uv, const Blk &B) {
  __shared__ R sm[256];
  R t3 = Lit<R>::large();
  // workgroups stride over rows, threads over the row (a minimum does not
  // depend on the order it is taken in)
  // four cells per pass: loads first (clamped indices, unconditional), then
  // the copies and the minimum (a store would hold back later loads)
  constexpr int KB = 4;
  for (int j = 1 + B.x; j <= g.N; j += B.gx)
    for (int i0 = 1 + threadIdx.x; i0 <= g.M; i0 += KB * blockDim.x) {
      R u[KB], v[KB], h[KB];
#pragma unroll
      for (int k = 0; k < KB; k++) {
        const int i = i0 + k * blockDim.x;
        const size_t o = IX(i <= g.M ? i : g.M, j);
        u[k] = d.U[o]; v[k] = d.V[o]; h[k] = d.H[o];
      }
#pragma unroll
      for (int k = 0; k < KB; k++) {
        const int i = i0 + k * blockDim.x;
        if (i > g.M) continue;
        if (copy_uv) { d.U0[IX(i, j)] = u[k]; d.V0[IX(i, j)] = v[k]; }
        R c = dsqrt(GRAV * rmax(h[k], p.MDF));
        R t1 = dabs(u[k]) + c;
        R t2 = (t1 < SMALL) ? ddiv(p.DX, SMALL) : ddiv(p.DX, t1);
        if (t2 < t3) t3 = t2;
        t1 = dabs(v[k]) + c;
        t2 = (t1 < SMALL) ? ddiv(p.DY, SMALL) : ddiv(p.DY, t1);
        if (t2 < t3) t3 = t2;
      }
    }
  sm[threadIdx.x] = t3;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = rmin(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) d.red[B.x] = sm[0];
}
template <class R>
__global__ void k_dt_partial(Dev<R> d, Geo g, Prm<R> p, int copy_uv) {
  dt_partial_body<R>(d, g, p, copy_uv, grid_blk());
}

template <class R>
__device__ __forceinline__ void dt_final_body(Dev<R> d, Prm<R> p, int nblk, R *out, const Blk &B) {
  __shared__ R sm[256];
  R m = Lit<R>::large();
  for (int q = threadIdx.x; q < nblk; q += blockDim.x) m = rmin(m, d.red[q]);
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = rmin(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    R dtt = p.CFL * sm[0];
    R dt;
    if (p.FIXED_DT) {
      dt = p.DTF;
      while (dt > dtt) dt = ddiv(dt, (R)2);
    } else {
      dt = dtt;
    }
    out[0] = dt;
  }
}
template <class R>
__global__ void k_dt_final(Dev<R> d, Prm<R> p, int nblk, R *out) {
  dt_final_body<R>(d, p, nblk, out, grid_blk());
}

// max |eta| over the interior (CHECK_BLOWUP, misc.F:394-433)
template <class R>
__device__ __forceinline__ void maxeta_partial_body(Dev<R> d, Geo g, const Blk &B) {
  __shared__ R sm[256];
  R m = (R)0;
  // 8 loads in flight per thread before the comparisons
  for (int j = g.Jb + B.x; j <= g.Je; j += B.gx)
    for (int i0 = g.Ib + threadIdx.x; i0 <= g.Ie; i0 += 8 * blockDim.x) {
      R a[8];
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const int i = i0 + k * blockDim.x;
        const R v = dabs(d.Eta[IX(i <= g.Ie ? i : g.Ie, j)]);  // clamped: unconditional load
        a[k] = i <= g.Ie ? v : (R)0;
      }
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (a[k] > m) m = a[k];
    }
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = rmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) d.red[B.x] = sm[0];
}
template <class R>
__global__ void k_maxeta_partial(Dev<R> d, Geo g) {
  maxeta_partial_body<R>(d, g, grid_blk());
}
template <class R>
__device__ __forceinline__ void max_final_body(Dev<R> d, int nblk, R *out, const Blk &B) {
  __shared__ R sm[256];
  R m = (R)0;
  for (int q = threadIdx.x; q < nblk; q += blockDim.x) m = rmax(m, d.red[q]);
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = rmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sm[0];
}
template <class R>
__global__ void k_max_final(Dev<R> d, int nblk, R *out) {
  max_final_body<R>(d, nblk, out, grid_blk());
}

// plain copies of whole pitched arrays (Eta0 := Eta etc.)
template <class R>
__global__ void k_copy3(R *d0, const R *s0, R *d1, const R *s1, R *d2, const R *s2, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    d0[q] = s0[q];
    d1[q] = s1[q];
    d2[q] = s2[q];
  }
}

// --------------------------------------------------------------------------
// CAL_DISPERSION part 1 (dispersion.F:77-287): masked central derivatives
// (derivatives.F:123-289) on the interior, ETAT on 1..Mloc-1 x 1..Nloc-1,
// Ut/Vt on all cells, and zeroed cross derivatives on the wall lines.
// DU = max(Depth,MinDepthFrc) U and DUt are recomputed at each stencil point.
//
// A stencil's cells are addressed from its centre row and the rows above and
// below as 32-bit byte offsets (C32): with the array's base in scalar
// registers a load then takes the base + 32-bit offset + immediate form, so
// every array read at a stencil point shares that row's offset (the
// columns are immediates, and neighbouring columns merge into one wide
// load) instead of forming a 64-bit address per array and point.  Used
// around cells of the interior and their mirror cells (every offset >= 0),
// for arrays of the working precision on planes below 2^32 bytes.
struct C32 {
  uint32_t bm, b0, bp;  // byte offsets of rows j-1, j, j+1 at column i
  template <class T>
  __device__ __forceinline__ const T &at(const T *a, int di, int dj) const {
    const uint32_t b = dj < 0 ? bm : (dj > 0 ? bp : b0);
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(a) + b + di * (int)sizeof(T));
  }
};
template <class R>
__device__ __forceinline__ C32 c32(const Geo &g, int i, int j) {
  const uint32_t b0 = ((uint32_t)(j - 1) * (uint32_t)g.P + (uint32_t)(i - 1)) * (uint32_t)sizeof(R);
  const uint32_t p = (uint32_t)g.P * (uint32_t)sizeof(R);
  return C32{b0 - p, b0, b0 + p};
}
template <class R>
__device__ __forceinline__ R dfx(const Geo &g, const R *a, int i, int j, const Dv<R> &dx) {
  const C32 c = c32<R>(g, i, j);
  return half(dv(c.at(a, 1, 0) - c.at(a, -1, 0), dx));
}
template <class R>
__device__ __forceinline__ R dfy(const Geo &g, const R *a, int i, int j, const Dv<R> &dy) {
  const C32 c = c32<R>(g, i, j);
  return half(dv(c.at(a, 0, 1) - c.at(a, 0, -1), dy));
}
template <class R>
__device__ __forceinline__ R dfxx(const Geo &g, const R *a, int i, int j, const Dv<R> &dx) {
  const C32 c = c32<R>(g, i, j);
  return dv(dv(c.at(a, 1, 0) - (R)2 * c.at(a, 0, 0) + c.at(a, -1, 0), dx), dx);
}
template <class R>
__device__ __forceinline__ R dfyy(const Geo &g, const R *a, int i, int j, const Dv<R> &dy) {
  const C32 c = c32<R>(g, i, j);
  return dv(dv(c.at(a, 0, 1) - (R)2 * c.at(a, 0, 0) + c.at(a, 0, -1), dy), dy);
}
template <class R>
__device__ __forceinline__ R dfxy(const Geo &g, const R *a, int i, int j, const Dv<R> &dx, const Dv<R> &dy) {
  const C32 c = c32<R>(g, i, j);
  R t1 = dv(half(c.at(a, 1, 1) - c.at(a, 1, -1)), dy);
  R t2 = dv(half(c.at(a, -1, 1) - c.at(a, -1, -1)), dy);
  return dv(half(t1 - t2), dx);
}
// the same with 64-bit addresses (for a kernel with no registers to spare)
template <class R>
__device__ __forceinline__ R dfxx_ix(const Geo &g, const R *a, int i, int j, const Dv<R> &dx) {
  return dv(dv(a[IX(i + 1, j)] - (R)2 * a[IX(i, j)] + a[IX(i - 1, j)], dx), dx);
}
template <class R>
__device__ __forceinline__ R dfyy_ix(const Geo &g, const R *a, int i, int j, const Dv<R> &dy) {
  return dv(dv(a[IX(i, j + 1)] - (R)2 * a[IX(i, j)] + a[IX(i, j - 1)], dy), dy);
}
template <class R>
__device__ __forceinline__ R dfxy_ix(const Geo &g, const R *a, int i, int j, const Dv<R> &dx, const Dv<R> &dy) {
  R t1 = dv(half(a[IX(i + 1, j + 1)] - a[IX(i + 1, j - 1)]), dy);
  R t2 = dv(half(a[IX(i - 1, j + 1)] - a[IX(i - 1, j - 1)]), dy);
  return dv(half(t1 - t2), dx);
}
// stencil of a derived field f(i,j) evaluated on the fly
template <class R, class Fn>
__device__ __forceinline__ R gfx(Fn f, int i, int j, const Dv<R> &dx) { return half(dv(f(i + 1, j) - f(i - 1, j), dx)); }
template <class R, class Fn>
__device__ __forceinline__ R gfy(Fn f, int i, int j, const Dv<R> &dy) { return half(dv(f(i, j + 1) - f(i, j - 1), dy)); }
template <class R, class Fn>
__device__ __forceinline__ R gfxx(Fn f, int i, int j, const Dv<R> &dx) { return dv(dv(f(i + 1, j) - (R)2 * f(i, j) + f(i - 1, j), dx), dx); }
template <class R, class Fn>
__device__ __forceinline__ R gfyy(Fn f, int i, int j, const Dv<R> &dy) { return dv(dv(f(i, j + 1) - (R)2 * f(i, j) + f(i, j - 1), dy), dy); }
template <class R, class Fn>
__device__ __forceinline__ R gfxy(Fn f, int i, int j, const Dv<R> &dx, const Dv<R> &dy) {
  R t1 = dv(half(f(i + 1, j + 1) - f(i + 1, j - 1)), dy);
  R t2 = dv(half(f(i - 1, j + 1) - f(i - 1, j - 1)), dy);
  return dv(half(t1 - t2), dx);
}

// Workgroup-size bounds of the stencil kernels.  Without one the compiler
// assumes 1024-thread workgroups and caps registers at 128; the FP64
// instances need more to keep their loads in flight, so they declare their
// real workgroup size (64x8, 64x4).  The FP32 instances keep the default,
// which is what sizes them for 4-5 waves per SIMD.
template <class R>
struct WideLB {
  static constexpr int wg512 = sizeof(R) == 8 ? 512 : 1024;
  static constexpr int wg256 = sizeof(R) == 8 ? 256 : 1024;
};
template <class R, bool P2>
__device__ __forceinline__ void disp_deriv_body(Dev<R> d, Geo g, Prm<R> p, R DT, const Blk &B) {
  const Tile tb = xcd_tile(B);
  int i = 1 + tb.x * blockDim.x + threadIdx.x;
  int j = 1 + tb.y * blockDim.y + threadIdx.y;
  if (i > g.M || j > g.N) return;
  size_t o = IX(i, j);
  const Dv<R> dx = dvx<P2>(p), dy = dvy<P2>(p);
  const R mdf = p.MDF;
  const bool g2 = p.G2 > (R)0;
  // Every value is formed before the first store: the output arrays could
  // alias the inputs as far as the compiler knows, so a store would hold
  // back the loads after it.
  // ETAT from the previous stage's P, Q (dispersion.F:103-119)
  // (P(i+1), Q(j+1) exist for every cell, so the loads need no branch)
  // (ETAT is formed on each path separately: formed once ahead of the
  // branch, its loads would be waited for there)
  const bool has_etat = i <= g.M - 1 && j <= g.N - 1;
  auto form_etat = [&]() {
    return -dv(d.P[IX(i + 1, j)] - d.P[IX(i, j)], dx) - dv(d.Q[IX(i, j + 1)] - d.Q[IX(i, j)], dy);
  };
  const bool keep_t = g2 && p.full_state;  // :133-142
  R ut = (R)0, vt = (R)0;
  if (keep_t) {
    ut = ddiv(d.U[o] - d.U0[o], DT);
    vt = ddiv(d.V[o] - d.V0[o], DT);
  }
  if (i < g.Ib || i > g.Ie || j < g.Jb || j > g.Je) {
    const R etat = form_etat();
    if (has_etat) d.ETAT[o] = etat;
    if (keep_t) { d.Ut[o] = ut; d.Vt[o] = vt; }
    return;
  }
  const R m9 = (R)d.MASK9[o];
  const R *U = d.U, *V = d.V, *Dp = d.Depth, *U0 = d.U0, *V0 = d.V0;
  const int ci = i, cj = j;
  const C32 cc = c32<R>(g, i, j);
  auto DU = [&](int a, int b) { return rmax(cc.at(Dp, a - ci, b - cj), mdf) * cc.at(U, a - ci, b - cj); };
  auto DV = [&](int a, int b) { return rmax(cc.at(Dp, a - ci, b - cj), mdf) * cc.at(V, a - ci, b - cj); };
  auto UT = [&](int a, int b) { return ddiv(cc.at(U, a - ci, b - cj) - cc.at(U0, a - ci, b - cj), DT); };
  auto VT = [&](int a, int b) { return ddiv(cc.at(V, a - ci, b - cj) - cc.at(V0, a - ci, b - cj), DT); };
  auto DUT = [&](int a, int b) {
    return rmax(cc.at(Dp, a - ci, b - cj), mdf) * ddiv(cc.at(U, a - ci, b - cj) - cc.at(U0, a - ci, b - cj), DT);
  };
  auto DVT = [&](int a, int b) {
    return rmax(cc.at(Dp, a - ci, b - cj), mdf) * ddiv(cc.at(V, a - ci, b - cj) - cc.at(V0, a - ci, b - cj), DT);
  };
  const bool wall = (i == g.Ib || i == g.Ie || j == g.Jb || j == g.Je);
  const R uxx = dfxx(g, U, i, j, dx) * m9;
  // (the cross derivatives are formed on wall cells too and then dropped:
  // their corner loads exist for every interior cell, and unconditional
  // loads can be issued together)
  const R uxy_s = dfxy(g, U, i, j, dx, dy) * m9, vxy_s = dfxy(g, V, i, j, dx, dy) * m9;
  const R uxy = wall ? (R)0 : uxy_s;
  const R vxy = wall ? (R)0 : vxy_s;
  const R vyy = dfyy(g, V, i, j, dy) * m9;
  const R duxx = gfxx<R>(DU, i, j, dx) * m9;
  const R duxy_s = gfxy<R>(DU, i, j, dx, dy) * m9, dvxy_s = gfxy<R>(DV, i, j, dx, dy) * m9;
  const R duxy = wall ? (R)0 : duxy_s;
  const R dvxy = wall ? (R)0 : dvxy_s;
  const R dvyy = gfyy<R>(DV, i, j, dy) * m9;
  const bool slopes = g2 || p.SHOW_BRK;  // read by WAVE_BREAKING
  const R etax = dfx(g, d.Eta, i, j, dx) * m9;
  const R etay = dfy(g, d.Eta, i, j, dy) * m9;
  const R etat = form_etat();
  if (has_etat) d.ETAT[o] = etat;
  if (keep_t) { d.Ut[o] = ut; d.Vt[o] = vt; }
  d.Uxx[o] = uxx; d.Uxy[o] = uxy; d.Vxy[o] = vxy; d.Vyy[o] = vyy;
  d.DUxx[o] = duxx; d.DUxy[o] = duxy; d.DVxy[o] = dvxy; d.DVyy[o] = dvyy;
  if (slopes) { d.ETAx[o] = etax; d.ETAy[o] = etay; }
  // The first derivatives are used only at their own cell by the nonlinear
  // terms, which form them from the stencil themselves; the arrays are
  // filled only when every intermediate is kept.
  if (g2 && p.full_state) {
    d.Ux[o] = dfx(g, U, i, j, dx) * m9;
    d.Vx[o] = dfx(g, V, i, j, dx) * m9;
    d.Uy[o] = dfy(g, U, i, j, dy) * m9;
    d.Vy[o] = dfy(g, V, i, j, dy) * m9;
    d.DUx[o] = gfx<R>(DU, i, j, dx) * m9;
    d.DVx[o] = gfx<R>(DV, i, j, dx) * m9;
    d.DUy[o] = gfy<R>(DU, i, j, dy) * m9;
    d.DVy[o] = gfy<R>(DV, i, j, dy) * m9;
    d.Utx[o] = gfx<R>(UT, i, j, dx) * m9;
    d.Vty[o] = gfy<R>(VT, i, j, dy) * m9;
    d.Utxx[o] = gfxx<R>(UT, i, j, dx) * m9;
    d.Vtyy[o] = gfyy<R>(VT, i, j, dy) * m9;
    d.Utxy[o] = wall ? (R)0 : gfxy<R>(UT, i, j, dx, dy) * m9;
    d.Vtxy[o] = wall ? (R)0 : gfxy<R>(VT, i, j, dx, dy) * m9;
    d.DUtx[o] = gfx<R>(DUT, i, j, dx) * m9;
    d.DVty[o] = gfy<R>(DVT, i, j, dy) * m9;
    d.DUtxx[o] = gfxx<R>(DUT, i, j, dx) * m9;
    d.DVtyy[o] = gfyy<R>(DVT, i, j, dy) * m9;
    d.DUtxy[o] = wall ? (R)0 : gfxy<R>(DUT, i, j, dx, dy) * m9;
    d.DVtxy[o] = wall ? (R)0 : gfxy<R>(DVT, i, j, dx, dy) * m9;
  }
}
template <class R, bool P2>
__global__ void __launch_bounds__(WideLB<R>::wg512) k_disp_deriv(Dev<R> d, Geo g, Prm<R> p, R DT) {
  disp_deriv_body<R, P2>(d, g, p, DT, grid_blk());
}

// CAL_DISPERSION part 2: U4, V4, U1p, V1p on all cells (:293-330) and the
// nonlinear terms U1pp..V3 on the interior (:398-503).
// The terms of cell (i, j).  SD(k, di, dj) gives second derivative k (0
// Uxx, 1 Vxy, 2 Uxy, 3 Vyy, 4 DUxx, 5 DVxy, 6 DUxy, 7 DVyy) at (i+di, j+dj)
// after their ghost fill, and 8, 9, 10 give ETAx, ETAy, ETAT at (i, j): from
// the arrays (k_disp_terms) or from the fused kernel's tile (k_disp_fused).
// TT(di, dj, k) gives max(Depth, MinDepthFrc) (k 0), Ut = (U-U0)/DT (k 1)
// and Vt (k 2) at (i+di, j+dj), or nothing (the terms form them here).
struct NoTT {};
template <class R, bool P2, class SD, class TT = NoTT>
__device__ __forceinline__ void disp_terms_at(Dev<R> &d, const Geo &g, const Prm<R> &p, R DT, int i, int j, SD S,
                                              TT T = TT{}) {
  size_t o = IX(i, j);
  const R B1 = p.B1, B2 = p.B2, b2 = p.b2, g2 = p.G2;
  const Dv<R> dx = dvx<P2>(p), dy = dvy<P2>(p);
  const R c13 = (R)1 / (R)3, c12 = (R)1 / (R)2, c16 = (R)1 / (R)6;
  R r_u4, r_v4, r_u1p, r_v1p;
  {
    R dep = d.Depth[o];
    R uxx = S(0, 0, 0), vxy = S(1, 0, 0), uxy = S(2, 0, 0), vyy = S(3, 0, 0);
    R duxx = S(4, 0, 0), dvxy = S(5, 0, 0), duxy = S(6, 0, 0), dvyy = S(7, 0, 0);
    R k0 = c13 - B1 + (R)0.5 * B1 * B1;
    R u4 = k0 * dep * dep * (uxx + vxy) + (B1 - c12) * dep * (duxx + dvxy);
    R v4 = k0 * dep * dep * (uxy + vyy) + (B1 - c12) * dep * (duxy + dvyy);
    if (g2 > (R)0) {
      R reta = d.Eta[o];
      R ken1 = (c16 - B1 + B1 * B1) * dep * reta * B2 + (c12 * B1 * B1 - c16) * reta * reta * B2 * B2;
      R ken2 = (B1 - c12) * reta * B2;
      R m9 = (R)d.MASK9[o];
      u4 = u4 + g2 * m9 * (ken1 * (uxx + vxy) + ken2 * (duxx + dvxy));
      v4 = v4 + g2 * m9 * (ken1 * (uxy + vyy) + ken2 * (duxy + dvyy));
    }
    r_u4 = u4;
    r_v4 = v4;
    r_u1p = (R)0.5 * ((R)1 - B1) * ((R)1 - B1) * dep * dep * (uxx + vxy) + (B1 - (R)1) * dep * (duxx + dvxy);
    r_v1p = (R)0.5 * ((R)1 - B1) * ((R)1 - B1) * dep * dep * (uxy + vyy) + (B1 - (R)1) * dep * (duxy + dvyy);
  }
  // U4, V4, U1p, V1p are stored after the loads of the nonlinear terms (a
  // store would hold back the loads after it)
  auto store4 = [&]() {
    d.U4[o] = r_u4;
    d.V4[o] = r_v4;
    if (!p.xd_in_huv) { d.U1p[o] = r_u1p; d.V1p[o] = r_v1p; }
  };
  if (!(g2 > (R)0) || i < g.Ib || i > g.Ie || j < g.Jb || j > g.Je) {
    store4();
    return;
  }
  // first derivatives of U, V, Ut, Vt, DU, DV, DUt, DVt at this cell, the
  // same expressions (and masking) as CAL_DISPERSION's arrays
  const R m9 = (R)d.MASK9[o], mdf = p.MDF;
  const bool wall = (i == g.Ib || i == g.Ie || j == g.Jb || j == g.Je);
  const R *Uv = d.U, *Vv = d.V, *Dp = d.Depth, *U0 = d.U0, *V0 = d.V0;
  const int ci = i, cj = j;
  const C32 cc = c32<R>(g, i, j);
  auto DU = [&](int a, int b) { return rmax(cc.at(Dp, a - ci, b - cj), mdf) * cc.at(Uv, a - ci, b - cj); };
  auto DV = [&](int a, int b) { return rmax(cc.at(Dp, a - ci, b - cj), mdf) * cc.at(Vv, a - ci, b - cj); };
  // Ut = (U-U0)/DT and max(Depth,MinDepthFrc) once per stencil point
  // (Ut is not needed at (i,j+-1), Vt not at (i+-1,j)); DUt = hm*Ut
  R hm[3][3], ut[3][3], vt[3][3];
#pragma unroll
  for (int dj = -1; dj <= 1; dj++) {
#pragma unroll
    for (int di = -1; di <= 1; di++) {
      if constexpr (std::is_same<TT, NoTT>::value) {
        hm[dj + 1][di + 1] = rmax(cc.at(Dp, di, dj), mdf);
        ut[dj + 1][di + 1] = (dj == 0 || di != 0) ? ddiv(cc.at(Uv, di, dj) - cc.at(U0, di, dj), DT) : (R)0;
        vt[dj + 1][di + 1] = (di == 0 || dj != 0) ? ddiv(cc.at(Vv, di, dj) - cc.at(V0, di, dj), DT) : (R)0;
      } else {
        hm[dj + 1][di + 1] = T(di, dj, 0);
        ut[dj + 1][di + 1] = T(di, dj, 1);
        vt[dj + 1][di + 1] = T(di, dj, 2);
      }
    }
  }
  auto UT = [&](int a, int b) { return ut[b - j + 1][a - i + 1]; };
  auto VT = [&](int a, int b) { return vt[b - j + 1][a - i + 1]; };
  auto DUT = [&](int a, int b) { return hm[b - j + 1][a - i + 1] * ut[b - j + 1][a - i + 1]; };
  auto DVT = [&](int a, int b) { return hm[b - j + 1][a - i + 1] * vt[b - j + 1][a - i + 1]; };
  const R f_Ux = dfx(g, Uv, i, j, dx) * m9, f_Vx = dfx(g, Vv, i, j, dx) * m9;
  const R f_Uy = dfy(g, Uv, i, j, dy) * m9, f_Vy = dfy(g, Vv, i, j, dy) * m9;
  const R f_DUx = gfx<R>(DU, i, j, dx) * m9, f_DVy = gfy<R>(DV, i, j, dy) * m9;
  const R f_Utx = gfx<R>(UT, i, j, dx) * m9, f_Vty = gfy<R>(VT, i, j, dy) * m9;
  const R f_Utxx = gfxx<R>(UT, i, j, dx) * m9, f_Vtyy = gfyy<R>(VT, i, j, dy) * m9;
  const R f_Utxy = wall ? (R)0 : gfxy<R>(UT, i, j, dx, dy) * m9;
  const R f_Vtxy = wall ? (R)0 : gfxy<R>(VT, i, j, dx, dy) * m9;
  const R f_DUtx = gfx<R>(DUT, i, j, dx) * m9, f_DVty = gfy<R>(DVT, i, j, dy) * m9;
  const R f_DUtxx = gfxx<R>(DUT, i, j, dx) * m9, f_DVtyy = gfyy<R>(DVT, i, j, dy) * m9;
  const R f_DUtxy = wall ? (R)0 : gfxy<R>(DUT, i, j, dx, dy) * m9;
  const R f_DVtxy = wall ? (R)0 : gfxy<R>(DVT, i, j, dx, dy) * m9;
#define A_(n, a, b) cc.at(d.n, (a) - i, (b) - j)
  R UxxVxy = S(0, 0, 0) + S(1, 0, 0);
  R UxyVyy = S(2, 0, 0) + S(3, 0, 0);
  R UxxVxy_x = dv(half(S(0, 1, 0) + S(1, 1, 0) - S(0, -1, 0) - S(1, -1, 0)), dx);
  R UxxVxy_y = dv(half(S(0, 0, 1) + S(1, 0, 1) - S(0, 0, -1) - S(1, 0, -1)), dy);
  R UxyVyy_x = dv(half(S(2, 1, 0) + S(3, 1, 0) - S(2, -1, 0) - S(3, -1, 0)), dx);
  R UxyVyy_y = dv(half(S(2, 0, 1) + S(3, 0, 1) - S(2, 0, -1) - S(3, 0, -1)), dy);
  R HUxxHVxy = S(4, 0, 0) + S(5, 0, 0);
  R HUxyHVyy = S(6, 0, 0) + S(7, 0, 0);
  R HUxxHVxy_x = dv(half(S(4, 1, 0) + S(5, 1, 0) - S(4, -1, 0) - S(5, -1, 0)), dx);
  R HUxxHVxy_y = dv(half(S(4, 0, 1) + S(5, 0, 1) - S(4, 0, -1) - S(5, 0, -1)), dy);
  R HUxyHVyy_x = dv(half(S(6, 1, 0) + S(7, 1, 0) - S(6, -1, 0) - S(7, -1, 0)), dx);
  R HUxyHVyy_y = dv(half(S(6, 0, 1) + S(7, 0, 1) - S(6, 0, -1) - S(7, 0, -1)), dy);
  R rh = A_(Depth, i, j);
  R rhx = dv(half(A_(Depth, i + 1, j) - A_(Depth, i - 1, j)), dx);
  R rhy = dv(half(A_(Depth, i, j + 1) - A_(Depth, i, j - 1)), dy);
  R reta = A_(Eta, i, j);
  R ex = S(8, 0, 0), ey = S(9, 0, 0), et = S(10, 0, 0);
  R Utx_ = f_Utx, Vty_ = f_Vty;
  R u1pp = -reta * B2 * ex * B2 * (Utx_ + Vty_) - (R)0.5 * reta * reta * B2 * B2 * (f_Utxx + f_Vtxy)
           - ex * B2 * (f_DUtx + f_DVty) - reta * B2 * (f_DUtxx + f_DVtxy);
  R v1pp = -reta * B2 * ey * B2 * (Utx_ + Vty_) - (R)0.5 * reta * reta * B2 * B2 * (f_Utxy + f_Vtyy)
           - ey * B2 * (f_DUtx + f_DVty) - reta * B2 * (f_DUtxy + f_DVtyy);
  R ken1 = B1 * ((R)1 - B1) * rh * et * B2 - B1 * B1 * reta * B2 * et * B2;
  R ken2 = B1 * ((R)1 - B1) * rh * reta * B2 - (R)0.5f * B1 * B1 * reta * reta * B2 * B2;
  R ken3 = B1 * et * B2;
  R ken4 = B1 * reta * B2;
  u1pp = u1pp - ken1 * UxxVxy - ken2 * (f_Utxx + f_Vtxy) + ken3 * HUxxHVxy + ken4 * (f_DUtxx + f_DVtxy);
  v1pp = v1pp - ken1 * UxyVyy - ken2 * (f_Utxy + f_Vtyy) + ken3 * HUxyHVyy + ken4 * (f_DUtxy + f_DVtyy);
  R omB = (R)1 - B1;
  ken1 = (B1 - (R)1) * (rhx + ex) * B2;
  ken2 = (B1 - (R)1) * (rh + reta) * B2;
  ken3 = (omB * omB * rh * rhx * B2 * B2 - B1 * omB * (rhx * reta * B2 + rh * ex * B2) + (B1 * B1 - (R)1) * reta * ex * B2 * B2);
  ken4 = ((R)0.5 * omB * omB * rh * rh * B2 * B2 - B1 * omB * rh * reta * B2 + (R)0.5 * (B1 * B1 - (R)1) * reta * reta * B2 * B2);
  R ken5 = (omB * omB * rh * rhy * B2 * B2 - B1 * omB * (rhy * reta * B2 + rh * ey * B2) + (B1 * B1 - (R)1) * reta * ey * B2 * B2);
  R u = A_(U, i, j), v = A_(V, i, j);
  R ux = f_Ux, vx = f_Vx, uy = f_Uy, vy = f_Vy;
  R dux = f_DUx, dvy = f_DVy;
  const R u2 = ken1 * (u * HUxxHVxy + v * HUxyHVyy) + ken2 * (ux * HUxxHVxy + u * HUxxHVxy_x + vx * HUxyHVyy + v * HUxyHVyy_x)
            + ken3 * (u * UxxVxy + v * UxyVyy) + ken4 * (ux * UxxVxy + u * UxxVxy_x + vx * UxyVyy + v * UxyVyy_x)
            + B2 * B2 * (dux + dvy + reta * B2 * (ux + vy)) * (HUxxHVxy + ex * B2 * (ux + vy) + reta * B2 * UxxVxy);
  ken1 = (B1 - (R)1) * (rhy + ey) * B2;
  const R v2 = ken1 * (u * HUxxHVxy + v * HUxyHVyy) + ken2 * (uy * HUxxHVxy + u * HUxxHVxy_y + vy * HUxyHVyy + v * HUxyHVyy_y)
            + ken5 * (u * UxxVxy + v * UxyVyy) + ken4 * (uy * UxxVxy + u * UxxVxy_y + vy * UxyVyy + v * UxyVyy_y)
            + B2 * B2 * (dux + dvy + reta * B2 * (ux + vy)) * (HUxyHVyy + ey * B2 * (ux + vy) + reta * B2 * UxyVyy);
  R omega_0 = vx - uy;
  R omega_1 = (b2 * rhx + B1 * ex) * B2 * (HUxyHVyy + (b2 * rh + B1 * reta) * B2 * UxyVyy)
              - (b2 * rhy + B1 * ey) * B2 * (HUxxHVxy + (b2 * rh + B1 * reta) * B2 * UxxVxy);
  ken1 = ((B1 - c12) * (reta + rh) * B2);
  ken2 = (c13 - B1 + (R)0.5 * B1 * B1) * rh * rh * B2 * B2 + (c16 - B1 + B1 * B1) * rh * reta * B2
         + (c12 * B1 * B1 - c16) * reta * reta * B2 * B2;
  const R u3 = -v * omega_1 - omega_0 * (ken1 * HUxyHVyy + ken2 * UxyVyy);
  const R v3 = u * omega_1 + omega_0 * (ken1 * HUxxHVxy + ken2 * UxxVxy);
  // (the one-pass kernel, TT given, runs in production only)
  if (std::is_same<TT, NoTT>::value && p.full_state) {
    d.U1pp[o] = u1pp; d.V1pp[o] = v1pp;
    d.U2[o] = u2; d.V2[o] = v2;
    d.U3[o] = u3; d.V3[o] = v3;
  }
  // the sums as SourceTerms forms them (sources.F:254, left to right); the
  // arrays themselves are kept only with every intermediate
  d.sumU[o] = u1pp + u2 + u3;
  d.sumV[o] = v1pp + v2 + v3;
  store4();
#undef A_
}
template <class R, bool P2>
__device__ __forceinline__ void disp_terms_body(Dev<R> d, Geo g, Prm<R> p, R DT, const Blk &B) {
  const Tile tb = xcd_tile(B);
  int i = 1 + tb.x * blockDim.x + threadIdx.x;
  int j = 1 + tb.y * blockDim.y + threadIdx.y;
  if (i > g.M || j > g.N) return;
  R *const arr[11] = {d.Uxx, d.Vxy, d.Uxy, d.Vyy, d.DUxx, d.DVxy, d.DUxy, d.DVyy, d.ETAx, d.ETAy, d.ETAT};
  auto S = [&](int k, int di, int dj) { return arr[k][IX(i + di, j + dj)]; };
  disp_terms_at<R, P2>(d, g, p, DT, i, j, S);
}
template <class R, bool P2>
__global__ void __launch_bounds__(WideLB<R>::wg512) k_disp_terms(Dev<R> d, Geo g, Prm<R> p, R DT) {
  disp_terms_body<R, P2>(d, g, p, DT, grid_blk());
}

// --------------------------------------------------------------------------
// CAL_DISPERSION in one pass (dispersion.F:62-506), for the production path
// of a non-periodic domain: the eight second derivatives of a 64 x 8 tile and
// its one-cell ring go to LDS, already as EXCHANGE_DISPERSION leaves them
// (a ghost cell holds the value of its mirror cell, negated for the arrays
// odd in that direction, bc.F:306-529), and the terms of k_disp_terms are
// formed from there.  The derivatives are not written to HBM and read back:
// only what later routines read is stored (ETAT, ETAx, ETAy, U4, V4, the
// nonlinear sums, and the four cross derivatives and U1p, V1p unless the RK
// update forms those itself, p.xd_in_huv).  Same operations on
// the same operands as k_disp_deriv + ghost fill + k_disp_terms.
constexpr int kDispTX = 64, kDispTY = 8;

// the eight second derivatives at (i, j) after the ghost fill (0 Uxx, 1
// Vxy, 2 Uxy, 3 Vyy, 4 DUxx, 5 DVxy, 6 DUxy, 7 DVyy)
template <class R, bool P2>
__device__ __forceinline__ void second_derivs(const Dev<R> &d, const Geo &g, const Prm<R> &p, int i, int j, R *v) {
  const Dv<R> dx = dvx<P2>(p), dy = dvy<P2>(p);
  const R mdf = p.MDF;
  int si = i, sj = j;
  bool nx = false, ny = false;
  if (i < g.Ib) { si = 2 * g.Ib - 1 - i; nx = true; }
  if (i > g.Ie) { si = 2 * g.Ie + 1 - i; nx = true; }
  if (j < g.Jb) { sj = 2 * g.Jb - 1 - j; ny = true; }
  if (j > g.Je) { sj = 2 * g.Je + 1 - j; ny = true; }
  const R m9 = (R)d.MASK9[IX(si, sj)];
  const R *U = d.U, *V = d.V, *Dp = d.Depth;
  const int ci = si, cj = sj;
  const C32 cc = c32<R>(g, si, sj);
  auto DU = [&](int a, int b) { return rmax(cc.at(Dp, a - ci, b - cj), mdf) * cc.at(U, a - ci, b - cj); };
  auto DV = [&](int a, int b) { return rmax(cc.at(Dp, a - ci, b - cj), mdf) * cc.at(V, a - ci, b - cj); };
  const bool wall = (si == g.Ib || si == g.Ie || sj == g.Jb || sj == g.Je);
  const R uxx = dfxx(g, U, si, sj, dx) * m9;
  const R uxy_s = dfxy(g, U, si, sj, dx, dy) * m9, vxy_s = dfxy(g, V, si, sj, dx, dy) * m9;
  const R vyy = dfyy(g, V, si, sj, dy) * m9;
  const R duxx = gfxx<R>(DU, si, sj, dx) * m9;
  const R duxy_s = gfxy<R>(DU, si, sj, dx, dy) * m9, dvxy_s = gfxy<R>(DV, si, sj, dx, dy) * m9;
  const R dvyy = gfyy<R>(DV, si, sj, dy) * m9;
  v[0] = nx ? -uxx : uxx;
  v[1] = wall ? (R)0 : vxy_s;
  v[2] = wall ? (R)0 : uxy_s;
  v[3] = ny ? -vyy : vyy;
  v[4] = nx ? -duxx : duxx;
  v[5] = wall ? (R)0 : dvxy_s;
  v[6] = wall ? (R)0 : duxy_s;
  v[7] = ny ? -dvyy : dvyy;
}

template <class R, bool P2>
__device__ __forceinline__ void disp_fused_body(Dev<R> d, Geo g, Prm<R> p, R DT, const Blk &B) {
  constexpr int EX = kDispTX + 2, EY = kDispTY + 2;
  __shared__ R sS[8][EY][EX];
  // max(Depth, MinDepthFrc), Ut and Vt of the tile and its ring, formed once
  // per cell (the terms read them at the nine points of their stencil)
  __shared__ R sT[3][EY][EX];
  const Tile tb = xcd_tile(B);
  const int x0 = 1 + tb.x * kDispTX, y0 = 1 + tb.y * kDispTY;
  const int tx = threadIdx.x, ty = threadIdx.y;
  // the ring of second derivatives: at most two cells per thread, both
  // formed unconditionally (clamped into the grid, dropped outside it) so
  // that their loads go out together
  {
    const int t = tx + ty * kDispTX;
    constexpr int NT = kDispTX * kDispTY, NE = EX * EY;
    static_assert(NE <= 2 * NT, "two ring cells per thread at most");
    const int qa = t, qb = t + NT < NE ? t + NT : NE - 1;
    const R mdf = p.MDF;
    auto ring = [&](int q, R *v, R *tt, bool &in) {
      const int ex = q % EX, ey = q / EX;
      const int i = x0 - 1 + ex, j = y0 - 1 + ey;
      in = i >= 1 && i <= g.M && j >= 1 && j <= g.N;
      const int ic = i < 1 ? 1 : (i > g.M ? g.M : i), jc = j < 1 ? 1 : (j > g.N ? g.N : j);
      second_derivs<R, P2>(d, g, p, ic, jc, v);
      const C32 c = c32<R>(g, ic, jc);
      tt[0] = rmax(c.at(d.Depth, 0, 0), mdf);
      tt[1] = ddiv(c.at(d.U, 0, 0) - c.at(d.U0, 0, 0), DT);
      tt[2] = ddiv(c.at(d.V, 0, 0) - c.at(d.V0, 0, 0), DT);
    };
    R va[8], vb[8], ta[3], tb_[3];
    bool ina, inb;
    ring(qa, va, ta, ina);
    ring(qb, vb, tb_, inb);
    {
      const int ex = qa % EX, ey = qa / EX;
#pragma unroll
      for (int k = 0; k < 8; k++) sS[k][ey][ex] = ina ? va[k] : (R)0;
#pragma unroll
      for (int k = 0; k < 3; k++) sT[k][ey][ex] = ta[k];
    }
    if (t + NT < NE) {
      const int ex = qb % EX, ey = qb / EX;
#pragma unroll
      for (int k = 0; k < 8; k++) sS[k][ey][ex] = inb ? vb[k] : (R)0;
#pragma unroll
      for (int k = 0; k < 3; k++) sT[k][ey][ex] = tb_[k];
    }
  }
  __syncthreads();
  const int i = x0 + tx, j = y0 + ty;
  if (i > g.M || j > g.N) return;
  const size_t o = IX(i, j);
  const Dv<R> dx = dvx<P2>(p), dy = dvy<P2>(p);
  const bool g2 = p.G2 > (R)0;
  const bool inner = i >= g.Ib && i <= g.Ie && j >= g.Jb && j <= g.Je;
  // ETAT (dispersion.F:103-119) on 1..Mloc-1 x 1..Nloc-1
  const bool has_etat = i <= g.M - 1 && j <= g.N - 1;
  R etat = (R)0;
  if (has_etat) etat = -dv(d.P[IX(i + 1, j)] - d.P[IX(i, j)], dx) - dv(d.Q[IX(i, j + 1)] - d.Q[IX(i, j)], dy);
  // ETAx, ETAy on the interior; their ghost fill with the other arrays
  // when Gamma2 > 0
  int si = i, sj = j;
  bool nx = false, ny = false;
  if (i < g.Ib) { si = 2 * g.Ib - 1 - i; nx = true; }
  if (i > g.Ie) { si = 2 * g.Ie + 1 - i; nx = true; }
  if (j < g.Jb) { sj = 2 * g.Jb - 1 - j; ny = true; }
  if (j > g.Je) { sj = 2 * g.Je + 1 - j; ny = true; }
  const R m9s = (R)d.MASK9[IX(si, sj)];
  R etax = dfx(g, d.Eta, si, sj, dx) * m9s, etay = dfy(g, d.Eta, si, sj, dy) * m9s;
  if (nx) etax = -etax;
  if (ny) etay = -etay;
  const int lx = tx + 1, ly = ty + 1;
  auto S = [&](int k, int di, int dj) -> R {
    if (k == 8) return etax;
    if (k == 9) return etay;
    if (k == 10) return etat;
    return sS[k][ly + dj][lx + di];
  };
  const R uxy = sS[2][ly][lx], vxy = sS[1][ly][lx], duxy = sS[6][ly][lx], dvxy = sS[5][ly][lx];
  auto T = [&](int di, int dj, int k) -> R { return sT[k][ly + dj][lx + di]; };
  disp_terms_at<R, P2>(d, g, p, DT, i, j, S, T);
  if (has_etat) d.ETAT[o] = etat;
  if (!p.xd_in_huv) { d.Uxy[o] = uxy; d.Vxy[o] = vxy; d.DUxy[o] = duxy; d.DVxy[o] = dvxy; }
  if (inner ? (g2 || p.SHOW_BRK) : g2) {
    d.ETAx[o] = etax;
    d.ETAy[o] = etay;
  }
}
template <class R, bool P2>
__global__ void __launch_bounds__(kDispTX * kDispTY) k_disp_fused(Dev<R> d, Geo g, Prm<R> p, R DT) {
  disp_fused_body<R, P2>(d, g, p, DT, grid_blk());
}

// --------------------------------------------------------------------------
// FLUXES (fluxes.F:52-117).  CONSTRUCT_HO_X/Y (:1531-1754): 4th-order
// MUSCL-TVD with compressive minmod on the 4th-order part and van Leer on
// the 3rd-order part.
//
// The limiters are written without branches: both candidate values are
// formed and a select keeps the one the reference's IF takes.  Per lane the
// kept value is the same operation on the same operands; a divergent branch
// would instead cost exec-mask bookkeeping (and hazard waits) on every
// IF, which dominated these VALU-bound kernels.  A discarded candidate may be
// a quotient by zero, which raises nothing on the GPU.
template <class R>
__device__ __forceinline__ R mm3(R a, R b, R c) {
  R tb = (R)2 * b, tc = (R)2 * c;
  const R pos = rmax((R)0, rmin(rmin(a, tb), tc));
  const R neg = rmin((R)0, rmax(rmax(a, tb), tc));
  return a >= (R)0 ? pos : neg;
}
// The same with the hardware minimum and maximum (v_min/v_max, fused into
// v_min3/v_max3): for non-NaN operands they return the same number as the
// selects of rmin/rmax, and differ at most in the sign of a zero result.
__device__ __forceinline__ float hmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double hmin(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ float hmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double hmax(double a, double b) { return __builtin_fmax(a, b); }
template <class R>
__device__ __forceinline__ R mm3_hw(R a, R b, R c) {
  R tb = (R)2 * b, tc = (R)2 * c;
  const R pos = hmax((R)0, hmin(hmin(a, tb), tc));
  const R neg = hmin((R)0, hmax(hmax(a, tb), tc));
  return a >= (R)0 ? pos : neg;
}
__device__ __forceinline__ bool neg_zero(float x) { return __float_as_uint(x) == 0x80000000u; }
__device__ __forceinline__ bool neg_zero(double x) {
  return (unsigned long long)__double_as_longlong(x) == 0x8000000000000000ull;
}
// Din (CONSTRUCT_HO_X fluxes.F:1550-1593) from the slopes t1..t3 with the
// minimum/maximum M given as a functor.
template <class R, class MM>
__device__ __forceinline__ R din_from(R t1, R t2, R t3, bool edge, bool dry, MM m) {
  R d1 = m(t1, t2, t3), d2 = m(t2, t3, t1), d3 = m(t3, t1, t2);
  // next to a dry cell t1 = t3 = t2, so all three limited slopes are mm3(t2, t2, t2)
  const R e = m(t2, t2, t2);
  d1 = edge ? e : d1;
  d2 = edge ? e : d2;
  d3 = edge ? e : d3;
  d1 = dry ? (R)0 : d1;
  d2 = dry ? (R)0 : d2;
  d3 = dry ? (R)0 : d3;
  return t2 - (R)1 / (R)6 * (d3 - (R)2 * d2 + d1);
}
// The limited slopes d1..d3 enter Din = t2 - (1/6)(d3 - 2 d2 + d1) only
// through that sum X.  With the hardware minimum/maximum every operation
// sees the same numbers as with the reference's MIN/MAX, so every nonzero
// intermediate is bitwise the same and only the sign of a zero X can
// differ; t2 - X/6 then differs only when t2 is -0 (-0 - (+0) = -0 but
// -0 - (-0) = +0).  Those lanes (not dry: there d1..d3 are +0 either way)
// take the reference's selects; elsewhere Din is bitwise the same.
// (Non-finite inputs are outside this: a NaN slope can come out different.)
template <class R>
__device__ __forceinline__ R din_val_ref(R vm2, R vm1, R v0, R vp1, int mm2, int mm1, int m0, int mp1) {
  const R t1 = vm1 - vm2, t2 = v0 - vm1, t3 = vp1 - v0;
  return din_from(t1, t2, t3, mm2 == 0 || mp1 == 0, mm1 == 0 || m0 == 0,
                  [](R a, R b, R c) { return mm3(a, b, c); });
}
// FP32 form of the same.  mm3(a, b, c) is the median of min(a,2b,2c),
// max(a,2b,2c) and 0: for a >= 0 the maximum is >= 0, so the median is the
// minimum when that is >= 0 and 0 otherwise (= max(0, min)); for a < 0 the
// minimum is < 0 and the median is min(0, max) the same way.  One v_med3
// after v_min3/v_max3, equal to mm3 up to the sign of a zero like mm3_hw.
// The edge and dry cases need no limited slopes: next to a dry cell
// d1 = d2 = d3 = mm3(t2, t2, t2) = t2 (for t2 != -0), so the sum is
// (t2 - 2 t2) + t2 from the same operations; in a dry cell d1..d3 = +0 and
// the sum is +0.
__device__ __forceinline__ float mm3_med(float a, float tb, float tc) {
  return __builtin_amdgcn_fmed3f(hmin(hmin(a, tb), tc), hmax(hmax(a, tb), tc), 0.0f);
}
__device__ __forceinline__ float din_from_med(float t1, float t2, float t3, bool edge, bool dry) {
  const float tb1 = 2.0f * t1, tb2 = 2.0f * t2, tb3 = 2.0f * t3;
  const float d1 = mm3_med(t1, tb2, tb3), d2 = mm3_med(t2, tb3, tb1), d3 = mm3_med(t3, tb1, tb2);
  float X = d3 - 2.0f * d2 + d1;
  const float Xe = t2 - tb2 + t2;
  X = edge ? Xe : X;
  X = dry ? 0.0f : X;
  return t2 - 1.0f / 6.0f * X;
}
// Din on the hardware minimum/maximum; sets `exact` for a lane that needs
// din_val_ref instead (t2 = -0, not dry)
template <class R>
__device__ __forceinline__ R din_val_hw(R vm2, R vm1, R v0, R vp1, int mm2, int mm1, int m0, int mp1, bool &exact) {
  const R t1 = vm1 - vm2, t2 = v0 - vm1, t3 = vp1 - v0;
  const bool dry = mm1 == 0 || m0 == 0;
  exact = exact || (neg_zero(t2) && !dry);
  if constexpr (sizeof(R) == 4) return din_from_med(t1, t2, t3, mm2 == 0 || mp1 == 0, dry);
  else return din_from(t1, t2, t3, mm2 == 0 || mp1 == 0, dry, [](R a, R b, R c) { return mm3_hw(a, b, c); });
}
template <class R>
__device__ __forceinline__ R din_val(R vm2, R vm1, R v0, R vp1, int mm2, int mm1, int m0, int mp1) {
  bool exact = false;
  R D = din_val_hw(vm2, vm1, v0, vp1, mm2, mm1, m0, mp1, exact);
  if (exact) D = din_val_ref(vm2, vm1, v0, vp1, mm2, mm1, m0, mp1);
  return D;
}
// Quotient a/b by Markstein's method (FP32): y = RN(1/b) from v_rcp_f32 and
// one Newton step, q = RN(a*y), r = a - b*q exactly (fma), q' = RN(q + r*y),
// with q's sign for a zero a.  Given y = RN(1/b), q' is the correctly rounded
// quotient when nothing under- or overflows; checked against the division
// bit for bit on gfx950 (the reciprocal for every significand of every
// exponent in -62..62, quotients on 1.6e9 operand pairs with exponents in
// -62..62) and, on the CPU from y = RN(1/b), exhaustively over the
// significand of a for divisors with an all-ones significand.  Callers keep
// their operands within 2^-62..2^62 and fall back to ddiv otherwise.
__device__ __forceinline__ float mdiv(float a, float b) {
  const float y0 = __builtin_amdgcn_rcpf(b);
  const float y = __builtin_fmaf(__builtin_fmaf(-b, y0, 1.0f), y0, y0);
  const float q = a * y;
  const float q1 = __builtin_fmaf(__builtin_fmaf(-b, q, a), y, q);
  return __builtin_copysignf(q1, q);
}
__device__ __forceinline__ double mdiv(double a, double b) { return a / b; }
// mdiv without the sign fix, for a numerator that is nonzero (the quotient is
// then correctly rounded, sign included) or +0: then q = +0*y has y's sign,
// r = -b*q + 0 = +0 and q' = +0*y + q = q (zeros of one sign).
__device__ __forceinline__ float mdiv_p(float a, float b) {
  const float y0 = __builtin_amdgcn_rcpf(b);
  const float y = __builtin_fmaf(__builtin_fmaf(-b, y0, 1.0f), y0, y0);
  const float q = a * y;
  return __builtin_fmaf(__builtin_fmaf(-b, q, a), y, q);
}
// The van Leer weights with mdiv (FP32): with |a|, |b| in [SMALL, 2^30]
// (SMALL is about 2^-19.9) the ratios lie in 2^-50..2^50 and every
// numerator and denominator in 2^-50..2^51 (or is 0, or a denominator
// 1 + rat is 0: the quotient is then discarded as in the reference), inside
// mdiv's range.  `exact` is set for a lane outside it (or a NaN operand),
// which must use vanleer instead.  The range test on the operands before
// the SMALL clamp is the same test (SMALL < 2^30).  The clamp
// |a| <= SMALL ? SMALL*SIGN(1,a) : a is copysign(med3(|a|, SMALL, 2^30), a)
// for every a of the range (one v_med3, which needs no canonical operand as
// a maximum would).  The numerators are b, a (|.| >= SMALL) and rat + |rat|
// (+0 or positive), so mdiv_p applies.
template <class R>
__device__ __forceinline__ void vanleer_m(R a, R b, R &v1, R &v2, R &ta, R &tb, bool &exact) {
  static_assert(sizeof(R) == 4, "FP32 only");
  exact = exact || !(dabs(a) <= 1073741824.0f && dabs(b) <= 1073741824.0f);
  a = copysign_(__builtin_amdgcn_fmed3f(dabs(a), SMALL, 1073741824.0f), a);
  b = copysign_(__builtin_amdgcn_fmed3f(dabs(b), SMALL, 1073741824.0f), b);
  const float rat = mdiv_p(b, a);
  const float q1 = mdiv_p(rat + dabs(rat), 1.0f + rat);
  v1 = dabs(1.0f + rat) > SMALL ? q1 : 0.0f;
  const float rat2 = mdiv_p(a, b);
  const float q2 = mdiv_p(rat2 + dabs(rat2), 1.0f + rat2);
  v2 = dabs(1.0f + rat2) > SMALL ? q2 : 0.0f;
  ta = a;
  tb = b;
}
template <class R>
__device__ __forceinline__ void vanleer(R a, R b, R &v1, R &v2, R &ta, R &tb) {
  // SMALL * SIGN(1, a) is copysign(SMALL, a) exactly
  a = dabs(a) <= SMALL ? copysign_(SMALL, a) : a;
  b = dabs(b) <= SMALL ? copysign_(SMALL, b) : b;
  const R rat = ddiv(b, a);
  const R q1 = ddiv(rat + dabs(rat), (R)1 + rat);
  v1 = dabs((R)1 + rat) > SMALL ? q1 : (R)0;
  const R rat2 = ddiv(a, b);
  const R q2 = ddiv(rat2 + dabs(rat2), (R)1 + rat2);
  v2 = dabs((R)1 + rat2) > SMALL ? q2 : (R)0;
  ta = a;
  tb = b;
}
// left/right states at x-face (i,j) from cells i-3..i+2 (stride s along the
// line: 1 for x, pitch for y)
template <class R>
__device__ __forceinline__ void muscl(const R *v, const int *mk, size_t o, ptrdiff_t s, R &outL, R &outR) {
  // cells c-3..c+2 relative to face between c-1 and c (o = index of cell c)
  R vv[6];
  int mm[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    vv[k] = v[o + (k - 3) * s];
    mm[k] = mk[o + (k - 3) * s];
  }
  // Din at cells c-1 (uses c-3..c), c (c-2..c+1), c+1 (c-1..c+2)
  R dm1 = din_val(vv[0], vv[1], vv[2], vv[3], mm[0], mm[1], mm[2], mm[3]);
  R d0 = din_val(vv[1], vv[2], vv[3], vv[4], mm[1], mm[2], mm[3], mm[4]);
  R dp1 = din_val(vv[2], vv[3], vv[4], vv[5], mm[2], mm[3], mm[4], mm[5]);
  R w1, w2, a, b;
  vanleer(dm1, d0, w1, w2, a, b);
  outL = vv[2] + ((R)1 / (R)6) * (w1 * a + (R)2 * w2 * b);
  vanleer(d0, dp1, w1, w2, a, b);
  outR = vv[3] - ((R)1 / (R)6) * ((R)2 * w1 * a + w2 * b);
}

template <class R>
__device__ __forceinline__ R hll(R sl, R sr, R fl, R fr, R ul, R ur) {
  const R o = sr * fl - sl * fr + sl * sr * (ur - ul);
  const R mid = ddiv(o, dabs(sr - sl) < SMALL ? SMALL : sr - sl);
  return sl >= (R)0 ? fl : (sr <= (R)0 ? fr : mid);
}
template <class R>
__device__ __forceinline__ R hydro(R xi, R de, R g3) {
  return (R)0.5 * GRAV * (xi * xi * g3 + (R)2 * xi * de);
}

// Wave speeds (WAVE_SPEED fluxes.F:352-439) of one computed face from its own
// per-face reconstruction; ghost faces copy the speeds of the nearest
// computed face and use these.
template <class R>
__device__ __forceinline__ void xface_speed(const Dev<R> &d, const Geo &g, const Prm<R> &p, int i, int j, R &sl, R &sr) {
  size_t o = IX(i, j);
  R ul, ur, el, er;
  muscl(d.U, d.MASK, o, 1, ul, ur);
  muscl(d.Eta, d.MASK, o, 1, el, er);
  R hl = el + d.DepthX[o], hr = er + d.DepthX[o];
  R spl = dsqrt(GRAV * dabs(hl)), spr = dsqrt(GRAV * dabs(hr));
  R sps = (R)0.5 * (spl + spr) + (R)0.25 * (ul - ur);
  R us = (R)0.5 * (ul + ur) + spl - spr;
  sl = rmin(ul - spl, us - sps);
  sr = rmax(ur + spr, us + sps);
}
template <class R>
__device__ __forceinline__ void yface_speed(const Dev<R> &d, const Geo &g, const Prm<R> &p, int i, int j, R &sl, R &sr) {
  size_t o = IX(i, j);
  R vl, vr, el, er;
  muscl(d.V, d.MASK, o, (ptrdiff_t)g.P, vl, vr);
  muscl(d.Eta, d.MASK, o, (ptrdiff_t)g.P, el, er);
  R hl = el + d.DepthY[o], hr = er + d.DepthY[o];
  R spl = dsqrt(GRAV * dabs(hl)), spr = dsqrt(GRAV * dabs(hr));
  R sps = (R)0.5 * (spl + spr) + (R)0.25 * (vl - vr);
  R us = (R)0.5 * (vl + vr) + spl - spr;
  sl = rmin(vl - spl, us - sps);
  sr = rmax(vr + spr, us + sps);
}

// x faces (fluxes.F:52-117 with CONSTRUCT_HO_X :1531-1640, WAVE_SPEED
// :352-439, HLL :150-194) and BOUNDARY_CONDITION (bc.F:53-295) resolved per
// face.
//
// The reconstruction is organised by cell rather than by face.  The limited
// slope D(k) of face k (din_val on cells k-2..k+1) is needed by faces k-1, k
// and k+1, and the van Leer weights of the pair (D(c), D(c+1)) give both the
// right state of face c and the left state of face c+1.  Each is evaluated
// once, from the same operands in the same order as the per-face form, so
// the states are bit-identical to it.
//
// One wavefront = 64 consecutive cells c of one row; lane l holds D(c) and
// the pair of cell c, and takes D(c+1) and the left state of face c from its
// neighbours with a lane shuffle.  Lanes 1..62 own faces; lanes 0 and 63 only
// supply neighbours, so consecutive wavefronts overlap by two cells.
constexpr int kFluxXFaces = 62;

template <class R>
__device__ __forceinline__ void pair_states(R vc, R D, R Dn, R &Lnext, R &Rc) {
  R w1, w2, a, b;
  vanleer(D, Dn, w1, w2, a, b);
  Lnext = vc + ((R)1 / (R)6) * (w1 * a + (R)2 * w2 * b);
  Rc = vc - ((R)1 / (R)6) * ((R)2 * w1 * a + w2 * b);
}
// the same with vanleer_m (sets `exact` where pair_states must be used)
template <class R>
__device__ __forceinline__ void pair_states_m(R vc, R D, R Dn, R &Lnext, R &Rc, bool &exact) {
  R w1, w2, a, b;
  vanleer_m(D, Dn, w1, w2, a, b, exact);
  // where vanleer_m's range holds, w1*a and w2*b are 0 or normal numbers
  // below 2^31 (w in 2^-49..2, |a|, |b| in SMALL..2^30), so (2*w2)*b is
  // 2*(w2*b) exactly, and w1*a + 2*(w2*b) is one fma, zero signs included
  const R p1 = w1 * a, p2 = w2 * b;
  Lnext = vc + ((R)1 / (R)6) * __builtin_fmaf(p2, 2.0f, p1);
  Rc = vc - ((R)1 / (R)6) * __builtin_fmaf(p1, 2.0f, p2);
}

// left/right states of face c along x for field v: lanes with dok hold a
// valid D(c), lanes with pok a valid pair for cell c
// (o is an in-row index with two cells to its left and one to its right for
// every lane, so the loads are unconditional; lanes without dok or pok
// discard what they form)
template <class R, bool M = false>
__device__ __forceinline__ void recon_x(R v0, R Dv, bool dok, bool pok, R &L, R &Rf, bool *exact = nullptr) {
  const R D = dok ? Dv : (R)0, vc = dok ? v0 : (R)0;
  R Dn = __shfl_down(D, 1);
  R Lv, Rv;
  if constexpr (M) pair_states_m(vc, D, Dn, Lv, Rv, *exact);
  else pair_states(vc, D, Dn, Lv, Rv);
  const R Ln = pok ? Lv : (R)0, Rc = pok ? Rv : (R)0;
  L = __shfl_up(Ln, 1);
  Rf = Rc;
}

// Everything after the states: P, F, G, the wave speeds, HLL and the wall
// and dry-cell conditions of one x face.
// PROD: a production-mode launch (FW_OPT_FULL_STATE 0), with no code for
// the full-state arrays
template <class R, bool PROD = false>
__device__ __forceinline__ void xface_finish(const Dev<R> &d, const Geo &g, const Prm<R> &p, int i, int j, bool comp,
                                             R uL, R uR, R vL, R vR, R huL, R huR, R hvL, R hvR, R eL, R eR, R u4L,
                                             R u4R, R v4L, R v4R) {
  const size_t o = IX(i, j);
  const int M1 = g.M + 1;
  const R g1 = p.G1, g3 = p.G3;
  const bool fs = !PROD && p.full_state;
  if (comp && fs) {
    d.UxL[o] = uL; d.UxR[o] = uR; d.VxL[o] = vL; d.VxR[o] = vR;
    d.HUxL[o] = huL; d.HUxR[o] = huR; d.HVxL[o] = hvL; d.HVxR[o] = hvR;
    d.EtaRxL[o] = eL; d.EtaRxR[o] = eR;
    if (p.DISPERSION) { d.U4xL[o] = u4L; d.U4xR[o] = u4R; d.V4xL[o] = v4L; d.V4xR[o] = v4R; }
  }
  const R dpx = d.DepthX[o];
  R hL = eL + dpx, hR = eR + dpx;
  if (fs || !comp) { d.HxL[o] = hL; d.HxR[o] = hR; }
  const R m9u = (R)d.MASK9[IX(i <= g.M ? i : g.M, j)];
  R pl = huL + g1 * m9u * hL * u4L;
  R pr = huR + g1 * m9u * hR * u4R;
  R fxl = g3 * pl * (uL + g1 * m9u * u4L) + (R)0.5 * GRAV * (eL * eL * g3 + (R)2 * eL * dpx);
  R fxr = g3 * pr * (uR + g1 * m9u * u4R) + (R)0.5 * GRAV * (eR * eR * g3 + (R)2 * eR * dpx);
  R gxl = hL * (uL + g1 * m9u * u4L) * (vL + g1 * m9u * v4L) * g3;
  R gxr = hR * (uR + g1 * m9u * u4R) * (vR + g1 * m9u * v4R) * g3;
  if (fs || !comp) { d.PL[o] = pl; d.PR[o] = pr; d.FxL[o] = fxl; d.FxR[o] = fxr; d.GxL[o] = gxl; d.GxR[o] = gxr; }
  // wave speed of the nearest computed face (ghost faces copy it)
  int ci = i < 4 ? 4 : (i > M1 - 3 ? M1 - 3 : i);
  int cj = j < 4 ? 4 : (j > g.N - 3 ? g.N - 3 : j);
  R sl, sr;
  if (ci == i && cj == j) {
    R spl = dsqrt(GRAV * dabs(hL)), spr = dsqrt(GRAV * dabs(hR));
    R sps = (R)0.5 * (spl + spr) + (R)0.25 * (uL - uR);
    R us = (R)0.5 * (uL + uR) + spl - spr;
    sl = rmin(uL - spl, us - sps);
    sr = rmax(uR + spr, us + sps);
  } else {
    xface_speed(d, g, p, ci, cj, sl, sr);
  }
  if (fs) { d.SxL[o] = sl; d.SxR[o] = sr; }
  R P = hll(sl, sr, pl, pr, eL, eR);
  R Fx = hll(sl, sr, fxl, fxr, huL, huR);
  R Gx = hll(sl, sr, gxl, gxr, hvL, hvR);
  // walls (bc.F:86-140)
  if (j >= g.Jb && j <= g.Je) {
    if (i == g.Ib) { P = (R)0; Fx = hydro(eR, dpx, g3); Gx = (R)0; }
    if (i == g.Ie + 1) { P = (R)0; Fx = hydro(eL, dpx, g3); Gx = (R)0; }
  }
  // faces of dry cells (bc.F:221-293); the later cell in loop order wins
  if (j >= g.Jb - 1 && j <= g.Je + 1) {
    int cl = i - 1, cr = i;
    if (cl >= g.Ib - 1 && cl <= g.Ie + 1 && d.MASK[IX(cl, j)] < 1) {
      P = (R)0;
      Fx = (cl != g.Ie) ? hydro(eR, dpx, g3) * (R)d.MASK[IX(cr, j)] : (R)0;
      Gx = (R)0;
    }
    if (cr >= g.Ib - 1 && cr <= g.Ie + 1 && cr <= g.M && d.MASK[IX(cr, j)] < 1) {
      P = (R)0;
      Fx = (cr != g.Ib) ? hydro(eL, dpx, g3) * (R)d.MASK[IX(cl, j)] : (R)0;
      Gx = (R)0;
    }
  }
  d.P[o] = P; d.Fx[o] = Fx; d.Gx[o] = Gx;
}

// blockDim (64, 4): threadIdx.x is the lane, threadIdx.y picks the row.
template <class R, bool DISP, bool PROD = false>
__device__ __forceinline__ void flux_x_body(Dev<R> d, Geo g, Prm<R> p, const Blk &B) {
  const int lane = threadIdx.x;
  const int c = B.x * kFluxXFaces + lane;  // cell / face index of this lane
  const int j = 1 + B.y * blockDim.y + threadIdx.y;
  if (j > g.N) return;  // whole wavefront
  const bool own = lane >= 1 && lane <= kFluxXFaces && c <= g.M + 1;
  const bool row = j >= g.Jb && j <= g.Je;
  const bool comp = row && c >= g.Ib && c <= g.Ie + 1;
  const size_t o = IX(c < 1 ? 1 : c, j);
  R uL, uR, vL, vR, huL, huR, hvL, hvR, eL, eR, u4L, u4R, v4L, v4R;
  if (row) {  // uniform over the wavefront
    // D(c) needs cells c-2..c+1, the pair of cell c needs D(c), D(c+1)
    const bool dok = c >= g.Ib - 1 && c <= g.Ie + 2;
    const bool pok = c >= g.Ib - 1 && c <= g.Ie + 1;
    // (rows Jb..Je have cells on both sides of every lane's index: the
    // loads need no guard)
    constexpr int NV = DISP ? 7 : 5;
    const R *fv[7] = {d.U, d.V, d.HU, d.HV, d.Eta, d.U4, d.V4};
    int mk[4];
#pragma unroll
    for (int k = 0; k < 4; k++) mk[k] = d.MASK[o + k - 2];
    R Lq[NV], Rq[NV];
    if constexpr (sizeof(R) == 8) {
      // field by field (all fields' slopes at once cost FP64 registers)
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const R *v = fv[q];
        const R Dv = din_val(v[o - 2], v[o - 1], v[o], v[o + 1], mk[0], mk[1], mk[2], mk[3]);
        recon_x(v[o], Dv, dok, pok, Lq[q], Rq[q]);
      }
    } else {
      R v0[NV], Dv[NV];
      bool exact = false;
#pragma unroll
      for (int q = 0; q < NV; q++) {
        const R *v = fv[q];
        v0[q] = v[o];
        Dv[q] = din_val_hw(v[o - 2], v[o - 1], v0[q], v[o + 1], mk[0], mk[1], mk[2], mk[3], exact);
      }
      if (exact) {  // rare: a -0 slope t2 in a wet cell (one branch for all fields)
#pragma unroll
        for (int q = 0; q < NV; q++) {
          const R *v = fv[q];
          Dv[q] = din_val_ref(v[o - 2], v[o - 1], v[o], v[o + 1], mk[0], mk[1], mk[2], mk[3]);
        }
      }
      // pairs with mdiv; if any lane of the wave has a slope beyond its range,
      // the whole wave forms them again with the division (the neighbours'
      // states come through shuffles)
      bool mx = false;
#pragma unroll
      for (int q = 0; q < NV; q++) recon_x<R, true>(v0[q], Dv[q], dok, pok, Lq[q], Rq[q], &mx);
      if (__any(mx)) {
#pragma unroll
        for (int q = 0; q < NV; q++) recon_x(v0[q], Dv[q], dok, pok, Lq[q], Rq[q]);
      }
    }
    uL = Lq[0]; uR = Rq[0]; vL = Lq[1]; vR = Rq[1]; huL = Lq[2]; huR = Rq[2];
    hvL = Lq[3]; hvR = Rq[3]; eL = Lq[4]; eR = Rq[4];
    if (DISP) { u4L = Lq[5 % NV]; u4R = Rq[5 % NV]; v4L = Lq[6 % NV]; v4R = Rq[6 % NV]; }
  }
  if (!own) return;
  if (!comp) {
    uL = d.UxL[o]; uR = d.UxR[o]; vL = d.VxL[o]; vR = d.VxR[o];
    huL = d.HUxL[o]; huR = d.HUxR[o]; hvL = d.HVxL[o]; hvR = d.HVxR[o];
    eL = d.EtaRxL[o]; eR = d.EtaRxR[o];
  }
  if (!comp || !DISP) { u4L = d.U4xL[o]; u4R = d.U4xR[o]; v4L = d.V4xL[o]; v4R = d.V4xR[o]; }
  xface_finish<R, PROD>(d, g, p, c, j, comp, uL, uR, vL, vR, huL, huR, hvL, hvR, eL, eR, u4L, u4R, v4L, v4R);
}
template <class R, bool DISP, bool PROD = false>
__global__ void __launch_bounds__(256) k_flux_x(Dev<R> d, Geo g, Prm<R> p) {
  flux_x_body<R, DISP, PROD>(d, g, p, grid_blk());
}

// y faces (CONSTRUCT_HO_Y fluxes.F:1642-1754, bc.F:147-210 and :221-293).
template <class R, bool PROD = false>
__device__ __forceinline__ void yface_finish(const Dev<R> &d, const Geo &g, const Prm<R> &p, int i, int j, bool comp,
                                             R uL, R uR, R vL, R vR, R huL, R huR, R hvL, R hvR, R eL, R eR, R u4L,
                                             R u4R, R v4L, R v4R) {
  const size_t o = IX(i, j);
  const int N1 = g.N + 1;
  const R g1 = p.G1, g3 = p.G3;
  const bool fs = !PROD && p.full_state;
  if (comp && fs) {
    d.UyL[o] = uL; d.UyR[o] = uR; d.VyL[o] = vL; d.VyR[o] = vR;
    d.HUyL[o] = huL; d.HUyR[o] = huR; d.HVyL[o] = hvL; d.HVyR[o] = hvR;
    d.EtaRyL[o] = eL; d.EtaRyR[o] = eR;
    if (p.DISPERSION) { d.U4yL[o] = u4L; d.U4yR[o] = u4R; d.V4yL[o] = v4L; d.V4yR[o] = v4R; }
  }
  const R dpy = d.DepthY[o];
  R hL = eL + dpy, hR = eR + dpy;
  if (fs || !comp) { d.HyL[o] = hL; d.HyR[o] = hR; }
  const R m9v = (R)d.MASK9[IX(i, j <= g.N ? j : g.N)];
  R ql = hvL + g1 * m9v * hL * v4L;
  R qr = hvR + g1 * m9v * hR * v4R;
  R gyl = g3 * ql * (vL + g1 * m9v * v4L) + (R)0.5 * GRAV * (eL * eL * g3 + (R)2 * eL * dpy);
  R gyr = g3 * qr * (vR + g1 * m9v * v4R) + (R)0.5 * GRAV * (eR * eR * g3 + (R)2 * eR * dpy);
  R fyl = hL * (uL + g1 * m9v * u4L) * (vL + g1 * m9v * v4L) * g3;
  R fyr = hR * (uR + g1 * m9v * u4R) * (vR + g1 * m9v * v4R) * g3;
  if (fs || !comp) { d.QL[o] = ql; d.QR[o] = qr; d.GyL[o] = gyl; d.GyR[o] = gyr; d.FyL[o] = fyl; d.FyR[o] = fyr; }
  int ci = i < 4 ? 4 : (i > g.M - 3 ? g.M - 3 : i);
  int cj = j < 4 ? 4 : (j > N1 - 3 ? N1 - 3 : j);
  R sl, sr;
  if (ci == i && cj == j) {
    R spl = dsqrt(GRAV * dabs(hL)), spr = dsqrt(GRAV * dabs(hR));
    R sps = (R)0.5 * (spl + spr) + (R)0.25 * (vL - vR);
    R us = (R)0.5 * (vL + vR) + spl - spr;
    sl = rmin(vL - spl, us - sps);
    sr = rmax(vR + spr, us + sps);
  } else {
    yface_speed(d, g, p, ci, cj, sl, sr);
  }
  if (fs) { d.SyL[o] = sl; d.SyR[o] = sr; }
  R Q = hll(sl, sr, ql, qr, eL, eR);
  R Fy = hll(sl, sr, fyl, fyr, huL, huR);
  R Gy = hll(sl, sr, gyl, gyr, hvL, hvR);
  if (!p.PERIODIC && i >= g.Ib && i <= g.Ie) {  // bc.F:147-210
    if (j == g.Jb) { Q = (R)0; Fy = (R)0; Gy = hydro(eR, dpy, g3); }
    if (j == g.Je + 1) { Q = (R)0; Fy = (R)0; Gy = hydro(eL, dpy, g3); }
  }
  if (i >= g.Ib - 1 && i <= g.Ie + 1) {
    int cb = j - 1, ct = j;
    if (cb >= g.Jb - 1 && cb <= g.Je + 1 && d.MASK[IX(i, cb)] < 1) {
      Q = (R)0;
      Fy = (R)0;
      Gy = (cb != g.Je) ? hydro(eR, dpy, g3) * (R)d.MASK[IX(i, ct)] : (R)0;
    }
    if (ct >= g.Jb - 1 && ct <= g.Je + 1 && ct <= g.N && d.MASK[IX(i, ct)] < 1) {
      Q = (R)0;
      Fy = (R)0;
      Gy = (ct != g.Jb) ? hydro(eL, dpy, g3) * (R)d.MASK[IX(i, cb)] : (R)0;
    }
  }
  d.Q[o] = Q; d.Fy[o] = Fy; d.Gy[o] = Gy;
}

// y faces: the same cell-wise reconstruction as along x, with the
// neighbours exchanged through LDS.  A workgroup of 64 x 16 threads covers 64
// columns and 16*K consecutive faces j of each (thread row t handles tile
// rows t, t+16, ...); tile row r forms D(j) and the pair of cell j, takes
// D(j+1) from row r+1 and the left state of face j from row r-1.  Rows
// 1..16K-2 own their face, so consecutive workgroups overlap by two rows.
constexpr int kFluxYThreadRows = 16;
template <int K>
struct FluxYTile {
  static constexpr int rows = kFluxYThreadRows * K;
  static constexpr int faces = rows - 2;
};
// rows per thread: 2 in FP32, 1 in FP64 (LDS holds 7 fields x rows x 64)
template <class R>
struct FluxYK {
  static constexpr int K = sizeof(R) == 4 ? 2 : 1;
};

template <class R, bool DISP, int K, bool PROD = false>
__device__ __forceinline__ void flux_y_body(Dev<R> d, Geo g, Prm<R> p, const Blk &B) {
  constexpr int NV = DISP ? 7 : 5;
  constexpr int ROWS = FluxYTile<K>::rows, FACES = FluxYTile<K>::faces;
  // per field: D(j) of every tile row, then (after the pairs are formed) the
  // left state of face j+1
  __shared__ R sX[NV][ROWS][64];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i = 1 + B.x * 64 + tx;
  const int N1 = g.N + 1;
  const ptrdiff_t s = (ptrdiff_t)g.P;
  const bool col = i >= g.Ib && i <= g.Ie;
  const R *fv[7] = {d.U, d.V, d.HU, d.HV, d.Eta, d.U4, d.V4};
  R vc[K][NV], D[K][NV];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int r = ty + kFluxYThreadRows * k;
    const int j = B.y * FACES + r;
    // D(j) needs cells j-2..j+1
    const bool dok = col && j >= g.Jb - 1 && j <= g.Je + 2;
    const size_t o = IX(i <= g.M ? i : g.M, j < 1 ? 1 : (j > N1 ? N1 : j));
    int mk[4] = {0, 0, 0, 0};
    R w[NV][4];
    if (dok) {
#pragma unroll
      for (int t = 0; t < 4; t++) mk[t] = d.MASK[o + (t - 2) * s];
#pragma unroll
      for (int q = 0; q < NV; q++)
#pragma unroll
        for (int t = 0; t < 4; t++) w[q][t] = fv[q][o + (t - 2) * s];
    }
#pragma unroll
    for (int q = 0; q < NV; q++) {
      D[k][q] = dok ? din_val(w[q][0], w[q][1], w[q][2], w[q][3], mk[0], mk[1], mk[2], mk[3]) : (R)0;
      vc[k][q] = dok ? w[q][2] : (R)0;
      sX[q][r][tx] = D[k][q];
    }
  }
  __syncthreads();
  R Dn[K][NV];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int r = ty + kFluxYThreadRows * k;
#pragma unroll
    for (int q = 0; q < NV; q++) Dn[k][q] = r + 1 < ROWS ? sX[q][r + 1][tx] : (R)0;
  }
  __syncthreads();
  R Rf[K][NV];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int r = ty + kFluxYThreadRows * k;
    const int j = B.y * FACES + r;
    // the pair of cell j needs D(j), D(j+1)
    const bool pok = col && j >= g.Jb - 1 && j <= g.Je + 1 && r + 1 < ROWS;
#pragma unroll
    for (int q = 0; q < NV; q++) {
      R Lv, Rv;
      pair_states(vc[k][q], D[k][q], Dn[k][q], Lv, Rv);
      Rf[k][q] = pok ? Rv : (R)0;
      sX[q][r][tx] = pok ? Lv : (R)0;
    }
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int r = ty + kFluxYThreadRows * k;
    const int j = B.y * FACES + r;
    const bool own = r >= 1 && r <= FACES && j <= N1 && i <= g.M;
    if (!own) continue;
    const bool comp = col && j >= g.Jb && j <= g.Je + 1;
    const size_t o = IX(i, j);
    R Lf[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) Lf[q] = sX[q][r - 1][tx];
    R uL = Lf[0], uR = Rf[k][0], vL = Lf[1], vR = Rf[k][1], huL = Lf[2], huR = Rf[k][2], hvL = Lf[3],
      hvR = Rf[k][3];
    R eL = Lf[4], eR = Rf[k][4], u4L, u4R, v4L, v4R;
    if (!comp) {
      uL = d.UyL[o]; uR = d.UyR[o]; vL = d.VyL[o]; vR = d.VyR[o];
      huL = d.HUyL[o]; huR = d.HUyR[o]; hvL = d.HVyL[o]; hvR = d.HVyR[o];
      eL = d.EtaRyL[o]; eR = d.EtaRyR[o];
    }
    if (DISP && comp) {
      u4L = Lf[5 % NV]; u4R = Rf[k][5 % NV]; v4L = Lf[6 % NV]; v4R = Rf[k][6 % NV];
    } else {
      u4L = d.U4yL[o]; u4R = d.U4yR[o]; v4L = d.V4yL[o]; v4R = d.V4yR[o];
    }
    yface_finish<R, PROD>(d, g, p, i, j, comp, uL, uR, vL, vR, huL, huR, hvL, hvR, eL, eR, u4L, u4R, v4L, v4R);
  }
}
template <class R, bool DISP, int K, bool PROD = false>
__global__ void __launch_bounds__(64 * kFluxYThreadRows) k_flux_y(Dev<R> d, Geo g, Prm<R> p) {
  flux_y_body<R, DISP, K, PROD>(d, g, p, grid_blk());
}

// y faces by marching: each thread walks one column through a strip of RS
// faces, faces jb+1..jb+RS, loading one cell (every field) per row into a
// four-cell window; the slope D of cell c-1 is formed when cell c arrives,
// the pair of cell c-2 when its second slope is known, and face c-2 is
// finished from the left state the previous pair left and the right state
// of this one.  Consecutive threads take consecutive columns (coalesced
// row loads); no LDS, no barrier; the strip's two end slopes and one pair
// are formed again by the neighbouring strip.  Same operations on the same
// operands as the cell-by-cell reconstruction.
// (four waves per SIMD: with the slopes' -0 fallback taken in one branch
// for all fields, five waves spill and lose 20 %; four are 2 % faster than
// the per-field branches at five)
template <class R, bool DISP, int RS, bool PROD = false>
__device__ __forceinline__ void flux_y_march_body(Dev<R> d, Geo g, Prm<R> p, const Blk &B) {
  constexpr int NV = DISP ? 7 : 5;
  const int i = 1 + B.x * 64 + threadIdx.x;
  const int jb = (B.y * blockDim.y + threadIdx.y) * RS;
  const int N1 = g.N + 1;
  if (i > g.M || jb + 1 > N1) return;
  const bool col = i >= g.Ib && i <= g.Ie;
  const R *fv[7] = {d.U, d.V, d.HU, d.HV, d.Eta, d.U4, d.V4};
  R w[4][NV];   // cells c-3..c
  int mk[4];
  R Dp[NV], Lp[NV];  // slope of cell c-2, left state of face c-2 (from the pair of cell c-3)
#pragma unroll
  for (int q = 0; q < NV; q++) { Dp[q] = (R)0; Lp[q] = (R)0; }
  for (int t = 0; t < 4; t++) {
    mk[t] = 0;
#pragma unroll
    for (int q = 0; q < NV; q++) w[t][q] = (R)0;
  }
  // (kept rolled: unrolled, the window costs registers and occupancy)
#pragma unroll 1
  for (int t = 0; t < RS + 5; t++) {
    const int c = jb - 2 + t;
    // shift the window and load cell c (rows outside 1..N read row 1: no
    // slope that is formed uses them)
#pragma unroll
    for (int t = 0; t < 3; t++) {
      mk[t] = mk[t + 1];
#pragma unroll
      for (int q = 0; q < NV; q++) w[t][q] = w[t + 1][q];
    }
    const size_t oc = IX(i, c >= 1 && c <= g.N ? c : 1);
    mk[3] = d.MASK[oc];
#pragma unroll
    for (int q = 0; q < NV; q++) w[3][q] = fv[q][oc];
    if (t < 3) continue;
    // slope of cell c-1 (cells c-3..c)
    const int cd = c - 1;
    const bool dok = col && cd >= g.Jb - 1 && cd <= g.Je + 2;
    R Dn[NV];
    bool exact = false;
#pragma unroll
    for (int q = 0; q < NV; q++) Dn[q] = din_val_hw(w[0][q], w[1][q], w[2][q], w[3][q], mk[0], mk[1], mk[2], mk[3], exact);
    if (exact) {  // rare: a -0 slope t2 in a wet cell (one branch for all fields)
#pragma unroll
      for (int q = 0; q < NV; q++) Dn[q] = din_val_ref(w[0][q], w[1][q], w[2][q], w[3][q], mk[0], mk[1], mk[2], mk[3]);
    }
#pragma unroll
    for (int q = 0; q < NV; q++) Dn[q] = dok ? Dn[q] : (R)0;
    if (t >= 4) {
      // pair of cell m = c-2: right state of face m, left state of face m+1
      const int m = c - 2;
      const bool pok = col && m >= g.Jb - 1 && m <= g.Je + 1;
      R Rc[NV], Ln[NV];
      bool mx = false;  // a slope beyond mdiv's range: the division instead
#pragma unroll
      for (int q = 0; q < NV; q++) {
        R lv, rv;
        pair_states_m(w[1][q], Dp[q], Dn[q], lv, rv, mx);
        Ln[q] = pok ? lv : (R)0;
        Rc[q] = pok ? rv : (R)0;
      }
      if (mx) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
          R lv, rv;
          pair_states(w[1][q], Dp[q], Dn[q], lv, rv);
          Ln[q] = pok ? lv : (R)0;
          Rc[q] = pok ? rv : (R)0;
        }
      }
      if (t >= 5 && m <= N1) {
        const bool comp = col && m >= g.Jb && m <= g.Je + 1;
        const size_t o = IX(i, m);
        R uL = Lp[0], uR = Rc[0], vL = Lp[1], vR = Rc[1], huL = Lp[2], huR = Rc[2], hvL = Lp[3], hvR = Rc[3];
        R eL = Lp[4], eR = Rc[4], u4L, u4R, v4L, v4R;
        if (!comp) {
          uL = d.UyL[o]; uR = d.UyR[o]; vL = d.VyL[o]; vR = d.VyR[o];
          huL = d.HUyL[o]; huR = d.HUyR[o]; hvL = d.HVyL[o]; hvR = d.HVyR[o];
          eL = d.EtaRyL[o]; eR = d.EtaRyR[o];
        }
        if (DISP && comp) {
          u4L = Lp[5 % NV]; u4R = Rc[5 % NV]; v4L = Lp[6 % NV]; v4R = Rc[6 % NV];
        } else {
          u4L = d.U4yL[o]; u4R = d.U4yR[o]; v4L = d.V4yL[o]; v4R = d.V4yR[o];
        }
        yface_finish<R, PROD>(d, g, p, i, m, comp, uL, uR, vL, vR, huL, huR, hvL, hvR, eL, eR, u4L, u4R, v4L, v4R);
      }
#pragma unroll
      for (int q = 0; q < NV; q++) Lp[q] = Ln[q];
    }
#pragma unroll
    for (int q = 0; q < NV; q++) Dp[q] = Dn[q];
  }
}
template <class R, bool DISP, int RS, bool PROD = false>
__global__ void __launch_bounds__(256, 4) k_flux_y_march(Dev<R> d, Geo g, Prm<R> p) {
  flux_y_march_body<R, DISP, RS, PROD>(d, g, p, grid_blk());
}

// --------------------------------------------------------------------------
// SourceTerms (sources.F:56-548, Cartesian): wavemaker mass source in the
// band, then SourceX/SourceY with slope, friction, dispersion, wavemaker and
// breaking-diffusion terms.  nu_vis is a default REAL (single) array in the
// reference (:69), so the viscosity sums are formed in single precision.
// One interior cell.  With keep, the MODULE GLOBAL outputs (SourceX/Y,
// FrcInsX/Y, BreakSourceX/Y, WaveMaker_Mass in the band) are written; the
// values ESTIMATE_HUV needs are returned either way (wm: the cell's
// WaveMaker_Mass, new in the band, else the stored one).  XD: U1p and V1p
// of the cell are given (u1p_x, v1p_x) instead of read.
template <class R, bool P2, bool XD = false>
__device__ __forceinline__ void source_terms(const Dev<R> &d, const Geo &g, const Prm<R> &p, R AA, int i, int j,
                                             bool keep, bool sums, R &sx_out, R &sy_out, R &wm_out,
                                             R u1p_x = (R)0, R v1p_x = (R)0) {
  const size_t o = IX(i, j), oe = IX(i + 1, j), ow = IX(i - 1, j), on = IX(i, j + 1), os = IX(i, j - 1);
  const Dv<R> dx = dvx<P2>(p), dy = dvy<P2>(p);
  const R mdf = p.MDF, g1 = p.G1, g2 = p.G2;
  const bool fric = p.FRIC_SPONGE, visc = p.VISC_BRK, dsp = p.DIFF_SPONGE, brk = visc || dsp;
  // Every load first, without branches, so that they are in flight
  // together: a term that is switched off reads, in place of its own array,
  // one that this cell reads at the same place anyway (a cache hit), and
  // its value is not used.  (Every plane is allocated.)
  const R wm_old = d.WaveMaker_Mass[o];
  const R u = d.U[o], v = d.V[o], cd = d.Cd[o], hc = d.H[o], eta = d.Eta[o], dep = d.Depth[o];
  const int m9i = d.MASK9[o], mki = d.MASK[o];
  const R pe = d.P[oe], p0 = d.P[o], qn = d.Q[on], q0 = d.Q[o];
  const R dxe = d.DepthX[oe], dx0 = d.DepthX[o], dyn = d.DepthY[on], dy0 = d.DepthY[o];
  const R *U4 = d.U4, *V4 = d.V4, *U = d.U, *V = d.V;
  const R u4o = U4[o], u4e = U4[oe], u4w = U4[ow], u4n = U4[on], u4s = U4[os];
  const R v4o = V4[o], v4e = V4[oe], v4w = V4[ow], v4n = V4[on], v4s = V4[os];
  const R ue = U[oe], uw = U[ow], un = U[on], us = U[os];
  const R ve = V[oe], vw = V[ow], vn = V[on], vs = V[os];
  R u1p, v1p;
  if constexpr (XD) {
    u1p = u1p_x;
    v1p = v1p_x;
  } else {
    u1p = d.U1p[o];
    v1p = d.V1p[o];
  }
  // (U1pp + U2) + U3 (sources.F:254): with every intermediate kept, from the
  // MODULE GLOBAL arrays; otherwise the sums CAL_DISPERSION's kernel formed
  const R *su[3] = {sums ? d.sumU : d.U1pp, sums ? d.sumU : d.U2, sums ? d.sumU : d.U3};
  const R *sv[3] = {sums ? d.sumV : d.V1pp, sums ? d.sumV : d.V2, sums ? d.sumV : d.V3};
  const R su0 = su[0][o], su1 = su[1][o], su2 = su[2][o];
  const R sv0 = sv[0][o], sv1 = sv[1][o], sv2 = sv[2][o];
  const R c4 = (fric ? d.CD_4_SPONGE : d.Cd)[o];
  // viscosity at (i,j), (i+1,j), (i-1,j), (i,j+1), (i,j-1)
  const size_t qs[5] = {o, oe, ow, on, os};
  const R *NB = visc ? d.nu_break : U, *NS = dsp ? d.nu_sponge : U;
  const R *HU = brk ? d.HU : U, *HV = brk ? d.HV : V;
  R nbv[5], nsv[5], hu[5], hv[5];
#pragma unroll
  for (int k = 0; k < 5; k++) {
    nbv[k] = NB[qs[k]]; nsv[k] = NS[qs[k]]; hu[k] = HU[qs[k]]; hv[k] = HV[qs[k]];
  }
  R wm = wm_old;
  bool keep_wm = false;  // the outputs kept for the caller are stored last
  if (p.wavemaker != 0 && i >= p.ilo && i <= p.ihi && j >= p.jlo && j <= p.jhi) {
    int bi = i - p.ilo, bj = j - p.jlo;
    if (p.wavemaker == 1) {
      wm = AA * d.wk_exp[bi] * d.wk_stage[bj];
    } else {
      int bw = p.ihi - p.ilo + 1, bh = p.jhi - p.jlo + 1;
      size_t plane = (size_t)bw * bh, q = (size_t)bj * bw + bi;
      const R *bb = d.wk_stage, *cc = d.wk_stage + p.NumFreq;
      R s = (R)0;
      for (int k = 0; k < p.NumFreq; k++) s = s + AA * (d.Cm[q + k * plane] * bb[k] + d.Sm[q + k * plane] * cc[k]);
      wm = s;
    }
    keep_wm = true;
  }
  R spd = dsqrt(u * u + v * v);
  R fix = -(cd * u * spd);
  R m9r = (R)m9i;
  R hmax = rmax(hc, mdf);
  R divpq = dv(pe - p0, dx) + dv(qn - q0, dy);
  const R dsum_u = sums ? su0 : su0 + su1 + su2;
  const R dsum_v = sums ? sv0 : sv0 + sv1 + sv2;
  R mk = (R)mki;
  R sx = dv(GRAV * (eta), dx) * (dxe - dx0) * mk + fix
         + g1 * m9r * hmax * (dv(u * (R)0.5 * (u4e - u4w), dx)
                              + dv(v * (R)0.5 * (u4n - u4s), dy)
                              + dv(u4o * (R)0.5 * (ue - uw), dx)
                              + dv(v4o * (R)0.5 * (un - us), dy)
                              - g2 * m9r * dsum_u)
         + g1 * m9r * divpq * (u4o - u1p) + wm * u;
  R fiy = -(cd * v * spd);
  R sy = dv(GRAV * (eta), dy) * (dyn - dy0) * mk + fiy
         + g1 * m9r * hmax * (dv(u * (R)0.5 * (v4e - v4w), dx)
                              + dv(v * (R)0.5 * (v4n - v4s), dy)
                              + dv(u4o * (R)0.5 * (ve - vw), dx)
                              + dv(v4o * (R)0.5 * (vn - vs), dy)
                              - g2 * m9r * dsum_v)
         + g1 * m9r * divpq * (v4o - v1p) + wm * v;
  if (fric) {
    sx = sx - c4 * u * spd * dep;
    sy = sy - c4 * v * spd * dep;
  }
  R bx = (R)0, by = (R)0;
  if (brk) {
    // nu_break, plus nu_sponge, in single precision
    float nu[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (visc) {
#pragma unroll
      for (int k = 0; k < 5; k++) nu[k] = (float)nbv[k];
    }
    if (dsp) {
#pragma unroll
      for (int k = 0; k < 5; k++) nu[k] = (float)((R)nu[k] + nsv[k]);
    }
    float n0 = nu[0];
    R sxp = (R)(nu[1] + n0), sxm = (R)(nu[2] + n0);
    R syp = (R)(nu[3] + n0), sym = (R)(nu[4] + n0);
    const R h5x = p.h5x, h5y = p.h5y;
    bx = h5x * (dv(sxp * (R)1, dx) * (hu[1] - hu[0]) - dv(sxm * (R)1, dx) * (hu[0] - hu[2]))
           + h5y * (dv(syp * (R)1, dy) * (hu[3] - hu[0]) - dv(sym * (R)1, dy) * (hu[0] - hu[4]));
    sx = sx + bx;
    by = h5x * (dv(sxp * (R)1, dx) * (hv[1] - hv[0]) - dv(sxm * (R)1, dx) * (hv[0] - hv[2]))
           + h5y * (dv(syp * (R)1, dy) * (hv[3] - hv[0]) - dv(sym * (R)1, dy) * (hv[0] - hv[4]));
    sy = sy + by;
  }
  if (keep) {
    if (keep_wm) d.WaveMaker_Mass[o] = wm;
    d.FrcInsX[o] = fix;
    d.FrcInsY[o] = fiy;
    if (brk) {
      d.BreakSourceX[o] = bx;
      d.BreakSourceY[o] = by;
    }
    d.SourceX[o] = sx;
    d.SourceY[o] = sy;
  }
  sx_out = sx;
  sy_out = sy;
  wm_out = wm;
}
template <class R, bool P2>
__global__ void k_sources(Dev<R> d, Geo g, Prm<R> p, R AA) {
  int i = g.Ib + blockIdx.x * blockDim.x + threadIdx.x;
  int j = g.Jb + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > g.Ie || j > g.Je) return;
  R sx, sy, wm;
  source_terms<R, P2>(d, g, p, AA, i, j, true, p.full_state == 0, sx, sy, wm);
}

// --------------------------------------------------------------------------
// ESTIMATE_HUV (etauv_solver.F:53-351): SSP-RK3 update of Eta, Ubar, Vbar;
// then H = Eta*Gamma3 + Depth on all cells (GET_Eta_U_V_HU_HV :373) and the
// normalised tridiagonal rows of both line systems (:400-539).
// FUSED: SourceTerms of the cell are formed here (source_terms) instead of
// read back, which is the same arithmetic on the same inputs: SourceTerms
// reads nothing ESTIMATE_HUV writes.
// XD: the four cross derivatives Uxy, Vxy, DUxy, DVxy, which only the line
// coefficients of this cell read, and U1p, V1p, which only its source terms
// read, are formed here from U, V, Depth and MASK9 (unchanged since
// CAL_DISPERSION) instead of read from the planes the dispersion kernel
// would otherwise write: the same operations on the same operands as there
// (an interior cell is its own mirror cell; wall cells take 0 for the cross
// derivatives).  XD is used only in production runs (FW_OPT_FULL_STATE 0),
// so that variant carries no code for the full-state arrays.
template <class R, bool FUSED, bool P2, bool XD = false>
__device__ __forceinline__ void huv_body(Dev<R> d, Geo g, Prm<R> p, R al, R be, R DT, R AA, const Blk &B) {
  const Tile tb = xcd_tile(B);
  int i = 1 + tb.x * blockDim.x + threadIdx.x;
  int j = 1 + tb.y * blockDim.y + threadIdx.y;
  if (i > g.M || j > g.N) return;
  size_t o = IX(i, j);
  const bool inner = (i >= g.Ib && i <= g.Ie && j >= g.Jb && j <= g.Je);
  const R mdf = p.MDF;
  R eta = d.Eta[o];
  if (inner) {
    // every input of the update is loaded first, ahead of the source terms
    // and of the first store (the compiler cannot tell the arrays apart, so
    // a store, even one not taken, would hold back the loads after it)
    const R *P = d.P, *Q = d.Q;
    const R hcur = d.H[o];
    const R p0 = P[o], p1 = P[IX(i + 1, j)], q0 = Q[o], q1 = Q[IX(i, j + 1)];
    const R eta0 = d.Eta0[o];
    const R fx0 = d.Fx[o], fx1 = d.Fx[IX(i + 1, j)], fy0 = d.Fy[o], fy1 = d.Fy[IX(i, j + 1)];
    const R gx0 = d.Gx[o], gx1 = d.Gx[IX(i + 1, j)], gy0 = d.Gy[o], gy1 = d.Gy[IX(i, j + 1)];
    const R ubar0 = d.Ubar0[o], ubar = d.Ubar[o], vbar0 = d.Vbar0[o], vbar = d.Vbar[o];
    const R depth = d.Depth[o];
    // the line coefficients' inputs, also without dispersion (then they
    // read this cell's depth and mask again, and are not used)
    const bool disp = p.DISPERSION;
    const int mk = d.MASK[o];
    const R m9 = (R)(disp ? d.MASK9 : d.MASK)[o];
    const R dxl = d.Depth[disp ? IX(i - 1, j) : o], dxr = d.Depth[disp ? IX(i + 1, j) : o];
    const R dyl = d.Depth[disp ? IX(i, j - 1) : o], dyr = d.Depth[disp ? IX(i, j + 1) : o];
    R vxy, dvxy, uxy, duxy, u1p = (R)0, v1p = (R)0;
    if constexpr (XD) {
      // (64-bit addresses here: this kernel is at its register limit, and
      // the 32-bit row offsets of the stencil helpers cost it spills)
      const R *U = d.U, *V = d.V, *Dp = d.Depth;
      auto DU = [&](int a, int b) { size_t q = IX(a, b); return rmax(Dp[q], mdf) * U[q]; };
      auto DV = [&](int a, int b) { size_t q = IX(a, b); return rmax(Dp[q], mdf) * V[q]; };
      const Dv<R> dx = dvx<P2>(p), dy = dvy<P2>(p);
      const bool wall = (i == g.Ib || i == g.Ie || j == g.Jb || j == g.Je);
      const R uxy_s = dfxy_ix(g, U, i, j, dx, dy) * m9, vxy_s = dfxy_ix(g, V, i, j, dx, dy) * m9;
      const R duxy_s = gfxy<R>(DU, i, j, dx, dy) * m9, dvxy_s = gfxy<R>(DV, i, j, dx, dy) * m9;
      uxy = wall ? (R)0 : uxy_s;
      vxy = wall ? (R)0 : vxy_s;
      duxy = wall ? (R)0 : duxy_s;
      dvxy = wall ? (R)0 : dvxy_s;
      // U1p, V1p (dispersion.F:293-330) as k_disp_terms forms them
      const R uxx = dfxx_ix(g, U, i, j, dx) * m9, vyy = dfyy_ix(g, V, i, j, dy) * m9;
      const R duxx = gfxx<R>(DU, i, j, dx) * m9, dvyy = gfyy<R>(DV, i, j, dy) * m9;
      const R B1 = p.B1;
      u1p = (R)0.5 * ((R)1 - B1) * ((R)1 - B1) * depth * depth * (uxx + vxy) + (B1 - (R)1) * depth * (duxx + dvxy);
      v1p = (R)0.5 * ((R)1 - B1) * ((R)1 - B1) * depth * depth * (uxy + vyy) + (B1 - (R)1) * depth * (duxy + dvyy);
    } else {
      vxy = (disp ? d.Vxy : d.Depth)[o];
      dvxy = (disp ? d.DVxy : d.Depth)[o];
      uxy = (disp ? d.Uxy : d.Depth)[o];
      duxy = (disp ? d.DUxy : d.Depth)[o];
    }
    R srcx, srcy, wmm;
    if (FUSED) {
      // (XD is the production form: no full-state arrays, so none of their
      // code or registers)
      constexpr bool prod = XD;
      source_terms<R, P2, XD>(d, g, p, AA, i, j, !prod && p.full_state != 0, prod || p.full_state == 0, srcx, srcy,
                              wmm, u1p, v1p);
    } else {
      srcx = d.SourceX[o];
      srcy = d.SourceY[o];
      wmm = d.WaveMaker_Mass[o];
    }
    R pc = (R)0.5 * (p1 + p0);
    R qc = (R)0.5 * (q1 + q0);
    if (!XD && p.full_state) {  // mean-flow diagnostics (etauv_solver.F:96-99, :247)
      d.P_center[o] = pc;
      d.Q_center[o] = qc;
      d.U_davg[o] = ddiv(pc, rmax(hcur, mdf));
      d.V_davg[o] = ddiv(qc, rmax(hcur, mdf));
    }
    const R nl = (R)-1, nr = (R)1;
    R r1 = -(p.rDXg * (p1 * nr + p0 * nl)) - p.rDYg * (q1 * nr + q0 * nl);
    if (p.wavemaker != 0) {
      R xmk = (R)(i - g.Ib) * p.DXg, ymk = (R)(j - g.Jb) * p.DYg;
      if (dabs(xmk - p.XC) < p.WW && dabs(ymk - p.YC) < half(p.YW)) r1 = r1 + wmm;
    }
    if (!XD && p.full_state) d.Wsurf[o] = r1;
    eta = al * eta0 + be * (eta + DT * r1);
    if (p.ETA_LIMITER) {
      if (eta < p.TROUGH) eta = p.TROUGH;
      if (eta > p.CREST) eta = p.CREST;
    }
    R r2 = -(p.rDX * (fx1 * nr + fx0 * nl)) - p.rDY * (fy1 * nr + fy0 * nl) + srcx;
    R ub = al * ubar0 + be * (ubar + DT * r2);
    R r3 = -(p.rDX * (gx1 * nr + gx0 * nl)) - p.rDY * (gy1 * nr + gy0 * nl) + srcy;
    R vb = al * vbar0 + be * (vbar + DT * r3);
    R h = eta * p.G3 + depth;
    d.Eta[o] = eta;
    d.Ubar[o] = ub;
    d.Vbar[o] = vb;
    d.H[o] = h;
    if (p.DISPERSION) {
      const R b2 = p.b2;
      const R gm = p.G1 * m9;
      R dep = rmax(depth, mdf);
      {
        R depl = rmax(dxl, mdf), depr = rmax(dxr, mdf);
        R t1 = gm * (p.cxa * dep * dep + p.cxb * depl * dep);
        R t2 = (R)1 + gm * (-(p.cxc * dep * dep) - p.cxd * dep * dep);
        R t3 = gm * (p.cxa * dep * dep + p.cxb * dep * depr);
        R t4 = ddiv(ub * (R)mk, rmax(h, mdf)) + gm * (-(p.b1h * dep * dep * vxy) - b2 * dep * dvxy);
        if (t2 != (R)0 || mk > 0) {
          d.xa[o] = ddiv(t1, t2); d.xc[o] = ddiv(t3, t2); d.xd[o] = ddiv(t4, t2);
        } else {
          d.xa[o] = (R)0; d.xc[o] = (R)0; d.xd[o] = (R)0;
        }
      }
      {
        R depl = rmax(dyl, mdf), depr = rmax(dyr, mdf);
        R t1 = gm * (p.cya * dep * dep + p.cyb * depl * dep);
        R t2 = (R)1 + gm * (-(p.cyc * dep * dep) - p.cyd * dep * dep);
        R t3 = gm * (p.cya * dep * dep + p.cyb * dep * depr);
        R t4 = ddiv(vb * (R)mk, rmax(h, mdf)) + gm * (-(p.b1h * dep * dep * uxy) - b2 * dep * duxy);
        if (t2 != (R)0 || mk > 0) {
          d.ya[o] = ddiv(t1, t2); d.yc[o] = ddiv(t3, t2); d.yd[o] = ddiv(t4, t2);
        } else {
          d.ya[o] = (R)0; d.yc[o] = (R)0; d.yd[o] = (R)0;
        }
      }
    } else {
      d.U[o] = ddiv(ub, rmax(h, mdf));
      d.V[o] = ddiv(vb, rmax(h, mdf));
    }
  } else {
    d.H[o] = eta * p.G3 + d.Depth[o];
  }
}
template <class R, bool FUSED, bool P2, bool XD = false>
__global__ void __launch_bounds__(WideLB<R>::wg256) k_huv(Dev<R> d, Geo g, Prm<R> p, R al, R be, R DT, R AA) {
  huv_body<R, FUSED, P2, XD>(d, g, p, al, be, DT, AA, grid_blk());
}

// TRIDx_ser / TRIDy_ser (tridiagnal.F:338-398): Thomas sweeps on rows
// normalised by the diagonal; a row with A == 0 is left as it is.  One
// thread per line; x lines walk along the contiguous direction, y lines are
// coalesced across threads.
template <class R>
__device__ __forceinline__ void thomas(R *A, R *C, R *D, R *Z, size_t o0, ptrdiff_t s, int n) {
  R cprev = C[o0], dprev = D[o0];
  for (int k = 1; k < n; k++) {
    size_t o = o0 + k * s;
    R a = A[o];
    R c = C[o], dd = D[o];
    if (a != (R)0) {
      R den = ddiv((R)1, a) - cprev;
      c = ddiv(ddiv(c, a), den);
      dd = ddiv(ddiv(dd, a) - dprev, ddiv((R)1, a) - cprev);
      C[o] = c;
      D[o] = dd;
    }
    cprev = c;
    dprev = dd;
  }
  size_t ol = o0 + (n - 1) * s;
  R z = D[ol];
  Z[ol] = z;
  for (int k = n - 2; k >= 0; k--) {
    size_t o = o0 + k * s;
    z = D[o] - C[o] * z;
    Z[o] = z;
  }
}

template <class R>
__device__ __forceinline__ void trid_body(Dev<R> d, Geo g, Prm<R> p, int nxblk, const Blk &B) {
  const int nx = g.Ie - g.Ib + 1, ny = g.Je - g.Jb + 1;
  if ((int)B.x < nxblk) {
    int j = g.Jb + B.x * blockDim.x + threadIdx.x;
    if (j > g.Je) return;
    thomas(d.xa, d.xc, d.xd, d.U, IX(g.Ib, j), 1, nx);
    return;
  }
  int i = g.Ib + (B.x - nxblk) * blockDim.x + threadIdx.x;
  if (i > g.Ie) return;
  const ptrdiff_t s = (ptrdiff_t)g.P;
  if (!p.PERIODIC) {
    thomas(d.ya, d.yc, d.yd, d.V, IX(i, g.Jb), s, ny);
    return;
  }
  // TRIDy_periodic (tridiagnal.F:59-325, serial path): Sherman-Morrison
  R *a = d.ya, *c = d.yc, *dd = d.yd;
  size_t ob = IX(i, g.Jb), oe = IX(i, g.Je);
  R abeg = a[ob], cend = c[oe];
  c[ob] = ddiv(c[ob], (R)1 + cend);
  dd[ob] = ddiv(dd[ob], (R)1 + cend);
  a[oe] = ddiv(a[oe], (R)1 + abeg);
  dd[oe] = ddiv(dd[oe], (R)1 + abeg);
  for (int k = 0; k < ny; k++) d.c1[ob + k * s] = c[ob + k * s];
  thomas(a, d.c1, dd, d.y1, ob, s, ny);
  for (int k = 0; k < ny; k++) dd[ob + k * s] = (R)0;
  dd[ob] = ddiv((R)1, (R)1 + cend);
  dd[oe] = -ddiv((R)1, (R)1 + abeg);
  thomas(a, c, dd, d.y2, ob, s, ny);
  R beta = ddiv(cend * d.y1[ob] - abeg * d.y1[oe], (R)1 - (cend * d.y2[ob] - abeg * d.y2[oe]));
  for (int k = 0; k < ny; k++) {
    size_t o = ob + k * s;
    d.V[o] = d.y1[o] + beta * d.y2[o];
  }
}
template <class R>
__global__ void k_trid(Dev<R> d, Geo g, Prm<R> p, int nxblk) {
  trid_body<R>(d, g, p, nxblk, grid_blk());
}

// Line solves staged through LDS.  A workgroup of 256 threads owns L lines
// of one direction, processed in chunks of CH rows.  The elimination step of
// tridiagnal.F:383-388 is
//   c' = (c/a) / (1/a - c'_prev),  d' = (d/a - d'_prev) / (1/a - c'_prev)
// (rows with a == 0, and the first row, keep c, d).  1/a, c/a and d/a do not
// depend on the previous row, so the staging waves form them; what is left
// are two recurrences, c' alone and d' driven by the denominators of c'.
// They run on two waves, lane l on line l:
//   wave 0 eliminates c' of chunk ch and leaves the denominator beside it,
//   wave 1 eliminates d' of chunk ch-1 from those denominators,
//   waves 2-3 put chunk ch+1 (loaded during the previous chunk) into LDS,
//   load chunk ch+2 and write the results of chunk ch-2 back to c, d
// (coalesced: consecutive threads take consecutive addresses, along a row
// for x lines and across columns for y lines), with one barrier per chunk
// and three LDS buffers in rotation.  Back substitution
// (tridiagnal.F:390-394) runs on wave 0 over the chunks in reverse, the
// last three still in LDS, while waves 2-3 restage the others.  Every
// quantity is the same correctly rounded operation on the same operands as
// in the row-by-row sweep, so results are bitwise the same.  The forward
// results of all but the last chunk go back to the c, d arrays.
template <class R>
struct alignas(4 * sizeof(R)) TridRow {
  R r, ca, da, keep;  // 1/a, c/a, d/a (or c, d where keep is set, keep_on)
};
// A kept row (the first, or A == 0) takes its c, d unchanged.  In FP32 the
// flag is an all-ones bit mask and the choice one v_bfi_b32 on the serial
// chain (a compare there would wait for VCC, which the division uses);
// FP64 keeps 1 / 0 and a select.
__device__ __forceinline__ float keep_on(float) { return __uint_as_float(0xffffffffu); }
__device__ __forceinline__ double keep_on(double) { return 1.0; }
__device__ __forceinline__ float keep_sel(float keep, float kept, float x) {
  const uint32_t m = __float_as_uint(keep);
  return __uint_as_float((__float_as_uint(kept) & m) | (__float_as_uint(x) & ~m));
}
__device__ __forceinline__ double keep_sel(double keep, double kept, double x) { return keep != 0.0 ? kept : x; }
template <class R>
struct alignas(2 * sizeof(R)) TridCD {
  R c, d;  // during elimination d first holds the denominator 1/a - c'_prev
};
template <class R, int L, int CH>
struct TridTile {
  TridRow<R> v[3][L][CH + 1];  // forward input
  TridCD<R> o[3][L][CH + 1];   // forward output = backward input
  R z[2][L][CH + 1];           // backward output
};

// The serial lanes issue every instruction on their own (one wave per
// SIMD), so their loops carry no per-row bookkeeping: rows go in groups of
// G, read from LDS one group ahead into two register sets used in turn
// (past the last whole group the reads repeat the first group, so that they
// are unconditional), and the leftover rows of a partial chunk take a plain
// loop.  ld(i) reads row i of the processing order, grp(i, x[G]) processes
// rows i..i+G-1 and op(i, x) one row.
template <int G, class Ld, class Grp, class Op>
__device__ __forceinline__ void trid_rows(int cnt, Ld ld, Grp grp, Op op) {
  using X = decltype(ld(0));
  const int full = cnt - cnt % G;
  X A[G], B[G];
  if (full > 0) {
#pragma unroll
    for (int g = 0; g < G; g++) A[g] = ld(g);
  }
  for (int m = 0; m < full;) {
    int nx = m + G < full ? m + G : 0;
#pragma unroll
    for (int g = 0; g < G; g++) B[g] = ld(nx + g);
    grp(m, A);
    m += G;
    if (m >= full) break;
    nx = m + G < full ? m + G : 0;
#pragma unroll
    for (int g = 0; g < G; g++) A[g] = ld(nx + g);
    grp(m, B);
    m += G;
  }
  for (int i = full; i < cnt; i++) op(i, ld(i));
}

constexpr int TG = 8;

template <class R>
struct TridDIn {
  TridRow<R> v;
  TridCD<R> o;
};

template <class R, int L, int CH>
__device__ __forceinline__ void trid_lds_body(Dev<R> d, Geo g, int nxblk, const Blk &B) {
  __shared__ TridTile<R, L, CH> T;
  const bool xdir = (int)B.x < nxblk;
  const int blk = xdir ? B.x : B.x - nxblk;
  const int nlines_tot = xdir ? (g.Je - g.Jb + 1) : (g.Ie - g.Ib + 1);
  const int n = xdir ? (g.Ie - g.Ib + 1) : (g.Je - g.Jb + 1);
  const int line0 = blk * L;
  const int nl = min(L, nlines_tot - line0);
  const R *A = xdir ? d.xa : d.ya;
  R *CP = xdir ? d.xc : d.yc;
  R *DP = xdir ? d.xd : d.yd;
  R *Z = xdir ? d.U : d.V;
  const int t = threadIdx.x;
  const int w = t >> 6;          // wave
  const int lane = t & 63;
  const int st = t - 128;        // staging thread index (waves 2-3), 0..127
  constexpr int NS = 128;
  constexpr int PER = (L * CH + NS - 1) / NS;
  const int nch = (n + CH - 1) / CH;
  // global index of (line, k)
  auto gix = [&](int l, int k) -> size_t {
    if (xdir) return (size_t)(g.Jb + line0 + l - 1) * g.P + (size_t)(g.Ib + k - 1);
    return (size_t)(g.Jb + k - 1) * g.P + (size_t)(g.Ib + line0 + l - 1);
  };
  // Staging slot q of this thread is (line sl[q], row sk[q]) of every chunk;
  // sl[q] < 0 marks a slot past the block's lines.  A chunk's c, d values
  // are stored and later reloaded by the same thread.
  int sl[PER], sk[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int e = st + q * NS;
    int l, k;
    if (xdir) { l = e / CH; k = e % CH; } else { l = e % L; k = e / L; }
    sl[q] = (e >= 0 && e < L * CH && l < nl) ? l : -1;
    sk[q] = k;
  }
  auto ok = [&](int q, int ch) { return sl[q] >= 0 && ch * CH + sk[q] < n; };
  // Loads are issued for every slot without branching (an unused slot reads
  // the block's first element), so a chunk's loads are all in flight at once.
  const size_t osafe = gix(0, 0);
  auto addr = [&](int q, int ch) -> size_t { return ok(q, ch) ? gix(sl[q], ch * CH + sk[q]) : osafe; };
  R ra[PER], rc[PER], rd[PER];  // one chunk loaded ahead
  auto fwd_load = [&](int ch) {
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const size_t o = addr(q, ch);
      ra[q] = A[o]; rc[q] = CP[o]; rd[q] = DP[o];
    }
  };
  auto fwd_put = [&](int ch, int b) {
#pragma unroll
    for (int q = 0; q < PER; q++) {
      if (ok(q, ch)) {
        const R a = ra[q];
        TridRow<R> v;
        if (ch * CH + sk[q] == 0 || !(a != (R)0)) {
          v.r = (R)0; v.ca = rc[q]; v.da = rd[q]; v.keep = keep_on((R)0);
        } else {
          v.r = ddiv((R)1, a); v.ca = ddiv(rc[q], a); v.da = ddiv(rd[q], a); v.keep = (R)0;
        }
        T.v[b][sl[q]][sk[q]] = v;
      }
    }
  };
  auto fwd_store = [&](int ch, int b) {
#pragma unroll
    for (int q = 0; q < PER; q++) {
      if (ok(q, ch)) {
        const size_t o = gix(sl[q], ch * CH + sk[q]);
        const TridCD<R> cd = T.o[b][sl[q]][sk[q]];
        CP[o] = cd.c; DP[o] = cd.d;
      }
    }
  };
  auto bwd_load = [&](int ch) {
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const size_t o = addr(q, ch);
      rc[q] = CP[o]; rd[q] = DP[o];
    }
  };
  auto bwd_put = [&](int ch, int b) {
#pragma unroll
    for (int q = 0; q < PER; q++)
      if (ok(q, ch)) T.o[b][sl[q]][sk[q]] = TridCD<R>{rc[q], rd[q]};
  };
  auto bwd_store = [&](int ch, int b) {
#pragma unroll
    for (int q = 0; q < PER; q++)
      if (ok(q, ch)) Z[gix(sl[q], ch * CH + sk[q])] = T.z[b][sl[q]][sk[q]];
  };
  auto rows = [&](int ch) { return min(CH, n - ch * CH); };

  // ---- forward elimination, iteration it: wave 0 on chunk it, wave 1 on
  // chunk it-1, staging as above
  if (w >= 2) {
    fwd_load(0);
    fwd_put(0, 0);
    if (nch > 1) fwd_load(1);
  }
  __syncthreads();
  R cprev = (R)0, dprev = (R)0;
  for (int it = 0; it <= nch; it++) {
    if (w == 0) {
      if (it < nch && lane < nl) {
        const TridRow<R> *row = T.v[it % 3][lane];
        TridCD<R> *out = T.o[it % 3][lane];
        auto step = [&](int k, const TridRow<R> &v) {
          const R den = v.r - cprev;
          R c = ddiv(v.ca, den);
          c = keep_sel(v.keep, v.ca, c);
          out[k] = TridCD<R>{c, den};
          cprev = c;
        };
        auto group = [&](int k0, const TridRow<R>(&V)[TG]) {
#pragma unroll
          for (int q = 0; q < TG; q++) step(k0 + q, V[q]);
        };
        trid_rows<TG>(rows(it), [&](int k) { return row[k]; }, group, step);
      }
    } else if (w == 1) {
      if (it >= 1 && lane < nl) {
        const TridRow<R> *row = T.v[(it - 1) % 3][lane];
        TridCD<R> *io = T.o[(it - 1) % 3][lane];
        auto step = [&](int k, const TridDIn<R> &x) {
          R dd = ddiv(x.v.da - dprev, x.o.d);
          dd = keep_sel(x.v.keep, x.v.da, dd);
          io[k].d = dd;
          dprev = dd;
        };
        auto group = [&](int k0, const TridDIn<R>(&X)[TG]) {
#pragma unroll
          for (int q = 0; q < TG; q++) step(k0 + q, X[q]);
        };
        trid_rows<TG>(rows(it - 1), [&](int k) { return TridDIn<R>{row[k], io[k]}; }, group, step);
      }
    } else {
      if (it + 1 < nch) {
        fwd_put(it + 1, (it + 1) % 3);
        if (it + 2 < nch) fwd_load(it + 2);
      }
      if (it >= 2 && it - 2 < nch) fwd_store(it - 2, (it - 2) % 3);
    }
    __syncthreads();
  }
  // ---- back substitution, chunks in reverse; chunks nch-1..nch-3 are in
  // LDS, the staging waves put chunk ch-1 (loaded during the previous chunk)
  // into LDS, load chunk ch-2 and store the solution of chunk ch+1.
  if (w >= 2 && nch >= 4) bwd_load(nch - 4);
  R z = (R)0;
  for (int ch = nch - 1; ch >= 0; ch--) {
    if (w == 0) {
      if (lane < nl) {
        const TridCD<R> *in = T.o[ch % 3][lane];
        R *zo = T.z[ch & 1][lane];
        int top = rows(ch) - 1;
        if (ch == nch - 1) {  // Z(Iend) = D(Iend)
          z = in[top].d;
          zo[top] = z;
          top--;
        }
        auto back = [&](int i, const TridCD<R> &x) {
          z = x.d - x.c * z;
          zo[top - i] = z;
        };
        trid_rows<TG>(
            top + 1, [&](int i) { return in[top - i]; },
            [&](int i0, const TridCD<R>(&X)[TG]) {
#pragma unroll
              for (int q = 0; q < TG; q++) back(i0 + q, X[q]);
            },
            back);
      }
    } else if (w >= 2) {
      if (ch >= 1 && ch <= nch - 3) {
        bwd_put(ch - 1, (ch - 1) % 3);
        if (ch >= 2) bwd_load(ch - 2);
      }
      if (ch + 1 < nch) bwd_store(ch + 1, (ch + 1) & 1);
    }
    __syncthreads();
  }
  if (w >= 2) bwd_store(0, 0);
}
template <class R, int L, int CH>
__global__ void __launch_bounds__(256) k_trid_lds(Dev<R> d, Geo g, int nxblk) {
  trid_lds_body<R, L, CH>(d, g, nxblk, grid_blk());
}

// GET_Eta_U_V_HU_HV tail (etauv_solver.F:568-593): dry cells, HU/HV, Froude
// cap; without dispersion Ubar/Vbar := HU/HV on all cells.
template <class R>
__device__ __forceinline__ void uv_final_body(Dev<R> d, Geo g, Prm<R> p, const Blk &B) {
  int i = 1 + B.x * blockDim.x + threadIdx.x;  // pointwise: launch order
  int j = 1 + B.y * blockDim.y + threadIdx.y;
  if (i > g.M || j > g.N) return;
  size_t o = IX(i, j);
  if (i >= g.Ib && i <= g.Ie && j >= g.Jb && j <= g.Je) {
    if (d.MASK[o] < 1) {
      d.Ubar[o] = (R)0; d.Vbar[o] = (R)0; d.U[o] = (R)0; d.V[o] = (R)0; d.HU[o] = (R)0; d.HV[o] = (R)0;
    } else {
      R hm = rmax(d.H[o], p.MDF);
      R u = d.U[o], v = d.V[o];
      R hu = hm * u, hv = hm * v;
      R ut = dsqrt(u * u + v * v);
      R fr = dsqrt(GRAV * hm);
      if (ddiv(ut, fr) > p.FCAP) {
        if (p.host_math) {
          // the cap's ATAN2/COS/SIN are left to the host's C library (the
          // reference's own math): the cell goes on a list, the host
          // rewrites its U, V, HU, HV before the next kernel reads them
          const int k = atomicAdd(d.fc_n, 1);
          d.fc_idx[k] = (int)o;
          d.fc[4 * k + 0] = u; d.fc[4 * k + 1] = v; d.fc[4 * k + 2] = fr; d.fc[4 * k + 3] = hm;
        } else {
          R th = datan2(v, u);
          u = p.FCAP * fr * dcos(th);
          v = p.FCAP * fr * dsin(th);
          hu = u * hm;
          hv = v * hm;
          d.U[o] = u;
          d.V[o] = v;
        }
      }
      d.HU[o] = hu;
      d.HV[o] = hv;
    }
  }
}
template <class R>
__global__ void k_uv_final(Dev<R> d, Geo g, Prm<R> p) {
  uv_final_body<R>(d, g, p, grid_blk());
}
// the host's Froude-capped U, V, HU, HV back into their cells
template <class R>
__global__ void k_fc_scatter(Dev<R> d, const R *vals, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const size_t o = (size_t)d.fc_idx[k];
  d.U[o] = vals[4 * k + 0];
  d.V[o] = vals[4 * k + 1];
  d.HU[o] = vals[4 * k + 2];
  d.HV[o] = vals[4 * k + 3];
}
template <class R>
__device__ __forceinline__ void ubar_from_hu_body(Dev<R> d, size_t n, const Blk &B) {
  for (size_t q = (size_t)B.x * blockDim.x + threadIdx.x; q < n; q += (size_t)B.gx * blockDim.x) {
    d.Ubar[q] = d.HU[q];
    d.Vbar[q] = d.HV[q];
  }
}
template <class R>
__global__ void k_ubar_from_hu(Dev<R> d, size_t n) {
  ubar_from_hu_body<R>(d, n, grid_blk());
}

// --------------------------------------------------------------------------
// UPDATE_MASK (masks.F:51-250).  Reads the old MASK, writes the new one to
// MASKn (host swaps the buffers), truncates DepthX/DepthY at faces of cells
// that were dry (later cell in loop order wins), and sets MASK9.
template <class R>
__device__ __forceinline__ int new_mask(const Dev<R> &d, const Geo &g, int i, int j) {
  size_t o = IX(i, j);
  int m = d.MASK[o];
  if (i < g.Ib - 2 || i > g.Ie + 2 || j < g.Jb - 2 || j > g.Je + 2) return m;
  if (d.MASK_STRUC[o] != 1) return m;
  R e = d.Eta[o];
  if (m < 1) {
    if (e > -d.Depth[o]) m = 1;
    if (d.MASK[IX(i - 1, j)] == 1 && d.Eta[IX(i - 1, j)] > e) m = 1;
    if (d.MASK[IX(i + 1, j)] == 1 && d.Eta[IX(i + 1, j)] > e) m = 1;
    if (d.MASK[IX(i, j - 1)] == 1 && d.Eta[IX(i, j - 1)] > e) m = 1;
    if (d.MASK[IX(i, j + 1)] == 1 && d.Eta[IX(i, j + 1)] > e) m = 1;
  } else {
    if (e < -d.Depth[o]) m = 0;
  }
  return m;
}

template <class R>
__device__ __forceinline__ void mask_body(Dev<R> d, Geo g, Prm<R> p, const Blk &B) {
  const Tile tb = xcd_tile(B);
  int i = 1 + tb.x * blockDim.x + threadIdx.x;
  int j = 1 + tb.y * blockDim.y + threadIdx.y;
  if (i > g.M + 1 || j > g.N + 1) return;
  const int lo_i = g.Ib - 2, hi_i = g.Ie + 2, lo_j = g.Jb - 2, hi_j = g.Je + 2;
  // Every value the thread may need is loaded first, from indices clamped
  // into the array, and only then used where the reference's conditions
  // take it (a store, or a load behind a branch, would hold back the loads
  // after it): old MASK at the cell and its four neighbours, Eta at the
  // same five, Depth at (i,j), (i-1,j), (i,j-1), MASK_STRUC.
  const int ic = i <= g.M ? i : g.M, jc = j <= g.N ? j : g.N;
  const int im = ic > 1 ? ic - 1 : 1, ip = ic < g.M ? ic + 1 : g.M;
  const int jm = jc > 1 ? jc - 1 : 1, jp = jc < g.N ? jc + 1 : g.N;
  const int im1 = i - 1 >= 1 ? i - 1 : 1, jm1 = j - 1 >= 1 ? j - 1 : 1;
  const size_t o = IX(ic, jc);
  const int m0 = d.MASK[o], mw = d.MASK[IX(im, jc)], me = d.MASK[IX(ip, jc)];
  const int ms = d.MASK[IX(ic, jm)], mn = d.MASK[IX(ic, jp)];
  const int mfw = d.MASK[IX(im1, jc)], mfs = d.MASK[IX(ic, jm1)];
  const R e0 = d.Eta[o], ew = d.Eta[IX(im, jc)], ee = d.Eta[IX(ip, jc)];
  const R es = d.Eta[IX(ic, jm)], en = d.Eta[IX(ic, jp)];
  const R h0 = d.Depth[o], hw = d.Depth[IX(im1, jc)], hs = d.Depth[IX(ic, jm1)];
  const int strc = d.MASK_STRUC[o];
  // DepthX/DepthY truncation (masks.F:152-161): the x face (i, j) is
  // written by cell i-1 (-> Depth(i)) and then by cell i (-> Depth(i-1)),
  // the y face likewise; the later cell wins
  if (j <= g.N && j >= lo_j && j <= hi_j) {
    R v;
    bool w = false;
    if (i - 1 >= lo_i && i - 1 <= hi_i && mfw < 1) { v = h0; w = true; }
    if (i >= lo_i && i <= hi_i && m0 < 1) { v = hw; w = true; }
    if (w) d.DepthX[IX(i, j)] = v;
  }
  if (i <= g.M && i >= lo_i && i <= hi_i) {
    R v;
    bool w = false;
    if (j - 1 >= lo_j && j - 1 <= hi_j && mfs < 1) { v = h0; w = true; }
    if (j >= lo_j && j <= hi_j && m0 < 1) { v = hs; w = true; }
    if (w) d.DepthY[IX(i, j)] = v;
  }
  if (i > g.M || j > g.N) return;
  // the new mask of the cell (masks.F:79-135)
  int m = m0;
  if (i >= lo_i && i <= hi_i && j >= lo_j && j <= hi_j && strc == 1) {
    if (m0 < 1) {
      if (e0 > -h0) m = 1;
      if (mw == 1 && ew > e0) m = 1;
      if (me == 1 && ee > e0) m = 1;
      if (ms == 1 && es > e0) m = 1;
      if (mn == 1 && en > e0) m = 1;
    } else {
      if (e0 < -h0) m = 0;
    }
  }
  d.MASKn[o] = m;
  int m9 = 0;
  if (i >= g.Ib - 1 && i <= g.Ie + 1 && j >= g.Jb - 1 && j <= g.Je + 1) {
    if (p.VISC_BRK) {
      m9 = 1;
    } else {
      m9 = m * new_mask(d, g, i - 1, j) * new_mask(d, g, i + 1, j) * new_mask(d, g, i + 1, j + 1) *
           new_mask(d, g, i, j + 1) * new_mask(d, g, i - 1, j + 1) * new_mask(d, g, i + 1, j - 1) *
           new_mask(d, g, i, j - 1) * new_mask(d, g, i - 1, j - 1);
      if (ddiv(dabs(e0), rmax(h0, p.MDF)) > p.SWE) m9 = 0;
    }
  }
  d.MASK9[o] = m9;
}
template <class R>
__global__ void k_mask(Dev<R> d, Geo g, Prm<R> p) {
  mask_body<R>(d, g, p, grid_blk());
}

// --------------------------------------------------------------------------
// BREAKING (breaker.F:87-353).  The reference updates AGE in place in loop
// order, so a cell reads the new age of its neighbours (I-1,J), (I-1,J-1),
// (I,J-1), (I+1,J-1) and the old age of the others.  Pass 1 evaluates every
// cell with old neighbour ages and lists the cells whose result depends on a
// "new" neighbour, with the inputs of their propagation test and viscosity
// (quadrant, px, py, pxy, ETAt, t2, t3: the square roots, ATAN2 and
// divisions are formed once); pass 2 (one workgroup) iterates those cells to
// the fixed point, which is the in-order result because the dependencies
// form a DAG.
template <class R>
struct BrkIn {
  R t1, t2, t3, ang, et, c, px, py, pxy;
};
template <class R>
__device__ __forceinline__ BrkIn<R> brk_in(const Dev<R> &d, const Geo &g, const Prm<R> &p, size_t o) {
  BrkIn<R> b;
  R h = d.H[o];
  b.t3 = dsqrt(GRAV * rmax(p.MDF, h));
  b.t1 = p.C1 * b.t3;
  b.t2 = p.C2 * b.t3;
  R ex = d.ETAx[o], ey = d.ETAy[o];
  b.ang = ddiv(datan2(-ey, -ex) * (R)180, FW_PI);
  b.et = d.ETAT[o];
  R t4 = rmax(dsqrt(ex * ex + ey * ey), SMALL);
  b.c = rmin(ddiv(dabs(b.et), t4), dsqrt(GRAV * dabs(h)));
  b.pxy = ddiv(dsqrt(p.DX * p.DX + p.DY * p.DY), rmax(b.c, SMALL));
  b.px = ddiv(dsqrt(p.DX * p.DX), rmax(b.c, SMALL));
  b.py = ddiv(dsqrt(p.DY * p.DY), rmax(b.c, SMALL));
  return b;
}
// quadrant of the propagation test: 0 none, 1 [0,90), 2 [90,180),
// 3 [-180,-90) (reads only old ages), 4 [-90,0)
template <class R>
__device__ __forceinline__ int brk_quad(R ang) {
  if (ang >= (R)0 && ang < (R)90) return 1;
  if (ang >= (R)90 && ang < (R)180) return 2;
  if (ang >= (R)-180 && ang < (R)-90) return 3;
  if (ang >= (R)-90 && ang < (R)0) return 4;
  return 0;
}
template <class R>
__device__ __forceinline__ bool brk_trig(R a1, R p1, R a2, R p2, R a3, R p3, R DT) {
  return (a1 >= DT && a1 > p1) || (a2 >= DT && a2 > p2) || (a3 >= DT && a3 > p3);
}
// new age of a cell in the propagation branch; nw/old give neighbour ages
template <class R>
__device__ __forceinline__ R brk_prop(const Dev<R> &d, const Geo &g, const BrkIn<R> &b, int q, int i, int j, const R *nw,
                                      const R *old, R DT, R own) {
  bool t = false;
  if (q == 1) t = brk_trig(nw[IX(i - 1, j)], b.px, nw[IX(i - 1, j - 1)], b.pxy, nw[IX(i, j - 1)], b.py, DT);
  if (q == 2) t = brk_trig(old[IX(i + 1, j)], b.px, nw[IX(i + 1, j - 1)], b.pxy, nw[IX(i, j - 1)], b.py, DT);
  if (q == 3) t = brk_trig(old[IX(i + 1, j)], b.px, old[IX(i + 1, j + 1)], b.pxy, old[IX(i, j + 1)], b.py, DT);
  if (q == 4) t = brk_trig(old[IX(i, j + 1)], b.py, old[IX(i - 1, j + 1)], b.pxy, nw[IX(i - 1, j)], b.px, DT);
  return t ? DT : own;
}
template <class R>
__device__ __forceinline__ R brk_nu(const Dev<R> &d, const Geo &g, const Prm<R> &p, const BrkIn<R> &b, size_t o, int i,
                                    int j, R age) {
  R xmk = (R)(i - g.Ib) * p.DX, ymk = (R)(j - g.Jb) * p.DY;
  if (dabs(xmk - p.XC) < p.WW && dabs(ymk - p.YC) < half(p.YW)) {
    if (b.et > rmin(b.t2, p.WKC * b.t3)) {
      R cap1 = (R)1 * (rmax(d.Depth[o], p.MDF) + d.Eta[o]);
      return cap1 * p.WKC * b.t3 + p.NUB;
    }
    return (R)0 + p.NUB;
  }
  if (age > (R)0 && age < p.TBRK && b.et > b.t2) {
    R cap1 = (R)1 * (rmax(d.Depth[o], p.MDF) + d.Eta[o]);
    return cap1 * b.t2 + p.NUB;
  }
  return (R)0 + p.NUB;
}

// Roller flux and undertow of continuing breaking (breaker.F:326-337): the
// cells outside the wavemaker band whose final age is in (0, T_brk) with
// ETAt above Cbrk2*sqrt(g max(H, MinDepthFrc)) (the branch that sets
// nu_break there); elsewhere the arrays keep their values.  0.45 is a
// default-REAL literal.  They read U and V as WAVE_BREAKING finds them, after
// the line solves, which run beside WAVE_BREAKING on the other stream, so
// they are a kernel of their own, after both; output diagnostics, read by no
// routine of the step, formed only when every intermediate is kept.
template <class R>
__global__ void k_roller(Dev<R> d, Geo g, Prm<R> p) {
  const Tile tb = xcd_tile(grid_blk());
  const int i = 1 + tb.x * blockDim.x + threadIdx.x;
  const int j = 1 + tb.y * blockDim.y + threadIdx.y;
  if (i < g.Ib - 1 || i > g.Ie + 1 || j < g.Jb - 1 || j > g.Je + 1) return;
  const R xmk = (R)(i - g.Ib) * p.DX, ymk = (R)(j - g.Jb) * p.DY;
  if (dabs(xmk - p.XC) < p.WW && dabs(ymk - p.YC) < half(p.YW)) return;
  const size_t o = IX(i, j);
  const R age = d.AGE_BREAKING[o], et = d.ETAT[o], h = d.H[o];
  const R t2 = p.C2 * dsqrt(GRAV * rmax(p.MDF, h));
  if (!(age > (R)0 && age < p.TBRK && et > t2)) return;
  const R u = d.U[o], v = d.V[o], eta = d.Eta[o], em = d.ETAmean[o];
  const R ex = d.ETAx[o], ey = d.ETAy[o];
  const R ang = ddiv(datan2(-ey, -ex) * (R)180, FW_PI);
  const R t4 = rmax(dsqrt(ex * ex + ey * ey), SMALL);
  const R c1 = rmax(ddiv(dabs(et), t4), dsqrt(GRAV * dabs(h)));
  R r = dabs(ddiv((R)0.45f * et, rmax(c1, SMALL)));
  r = rmin(r, (R)(sizeof(R) == 4 ? (R)0.1638f : (R)0.1638));
  const R rf = dabs(c1 - dsqrt(u * u + v * v)) * r * (eta - em);
  R sn, cs;
  dsincos(ang * FW_PI / (R)180, sn, cs);
  d.ROLLER_FLUX[o] = rf;
  d.UNDERTOW_U[o] = -(rf * cs);
  d.UNDERTOW_V[o] = -(rf * sn);
}

template <class R>
__device__ __forceinline__ void brk1_body(Dev<R> d, Geo g, Prm<R> p, R DT, const Blk &B) {
  const Tile tb = xcd_tile(B);
  int i = 1 + tb.x * blockDim.x + threadIdx.x;
  int j = 1 + tb.y * blockDim.y + threadIdx.y;
  if (i > g.M || j > g.N) return;
  size_t o = IX(i, j);
  R age = d.AGE_BREAKING[o];
  if (i < g.Ib - 1 || i > g.Ie + 1 || j < g.Jb - 1 || j > g.Je + 1) {
    d.AGEn[o] = age;
    return;
  }
  BrkIn<R> b = brk_in(d, g, p, o);
  R na;
  if (b.et >= b.t1 && (age == (R)0 || age > p.TBRK)) {
    na = DT;
  } else if (age > (R)0) {
    na = age + DT;
  } else {
    na = age;
    if (b.et >= b.t2) {
      int q = brk_quad(b.ang);
      if (q != 0) {
        na = brk_prop(d, g, b, q, i, j, d.AGE_BREAKING, d.AGE_BREAKING, DT, age);
        if (q != 3) {
          int k = atomicAdd(d.cand_n, 1);
          d.cand[k] = (int)o;
          R *ci = d.cinf + 8 * (size_t)k;
          ci[0] = (R)q; ci[1] = b.px; ci[2] = b.py; ci[3] = b.pxy;
          ci[4] = b.et; ci[5] = b.t2; ci[6] = b.t3;
        }
      }
    }
  }
  d.AGEn[o] = na;
  d.nu_break[o] = brk_nu(d, g, p, b, o, i, j, na);
}
template <class R>
__global__ void k_brk1(Dev<R> d, Geo g, Prm<R> p, R DT) {
  brk1_body<R>(d, g, p, DT, grid_blk());
}

template <class R>
__device__ __forceinline__ void brk2_body(Dev<R> d, Geo g, Prm<R> p, R DT, const Blk &B) {
  __shared__ int changed;
  const int n = *d.cand_n;
  if (n == 0) return;
  // the dependency chains are at most n long, so n+1 sweeps reach the fixed point
  const int maxit = n + 2;
  int it = 0;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) changed = 0;
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      size_t o = (size_t)d.cand[k];
      int i = 1 + (int)(o % g.P), j = 1 + (int)(o / g.P);
      const R *ci = d.cinf + 8 * (size_t)k;
      BrkIn<R> b;
      b.px = ci[1]; b.py = ci[2]; b.pxy = ci[3];
      R na = brk_prop(d, g, b, (int)ci[0], i, j, d.AGEn, d.AGE_BREAKING, DT, d.AGE_BREAKING[o]);
      if (na != d.AGEn[o]) {
        d.AGEn[o] = na;
        changed = 1;
      }
    }
    __syncthreads();
    if (!changed) break;
    if (++it >= maxit) {
      if (threadIdx.x == 0) d.flags[0] = 1;
      break;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    size_t o = (size_t)d.cand[k];
    int i = 1 + (int)(o % g.P), j = 1 + (int)(o / g.P);
    const R *ci = d.cinf + 8 * (size_t)k;
    BrkIn<R> b;
    b.et = ci[4]; b.t2 = ci[5]; b.t3 = ci[6];
    d.nu_break[o] = brk_nu(d, g, p, b, o, i, j, d.AGEn[o]);
  }
  __syncthreads();
  if (threadIdx.x == 0) *d.cand_n = 0;
}
template <class R>
__global__ void __launch_bounds__(1024) k_brk2(Dev<R> d, Geo g, Prm<R> p, R DT) {
  brk2_body<R>(d, g, p, DT, grid_blk());
}

// --------------------------------------------------------------------------
// EXCHANGE tail (bc.F:491-494): U,V,HU,HV *= MASK on all cells, optionally
// followed by SPONGE_DAMPING (sponge.F:52-66).
__device__ __forceinline__ bool bits_equal(float a, float b) { return __float_as_uint(a) == __float_as_uint(b); }
__device__ __forceinline__ bool bits_equal(double a, double b) {
  return __double_as_longlong(a) == __double_as_longlong(b);
}
template <class R>
__device__ __forceinline__ void mask_mul_body(Dev<R> d, Geo g, Prm<R> p, int sponge, int mul, const Blk &B) {
  int i = 1 + B.x * blockDim.x + threadIdx.x;  // pointwise: launch order
  int j = 1 + B.y * blockDim.y + threadIdx.y;
  if (i > g.M || j > g.N) return;
  size_t o = IX(i, j);
  // (MASK and SPONGE loaded together: one round trip)
  const int mk = d.MASK[o];
  const R sp = sponge ? d.SPONGE[o] : (R)1;
  // x*1 and x/1 leave every bit of x unchanged: cells with MASK = 1 and
  // SPONGE = 1 need no work
  if (mk == 1 && sp == (R)1) return;
  // a value whose bits the multiply leaves as they are is not stored again
  // (dry cells already hold the products by 0 of the last multiply)
  const R u0 = d.U[o], v0 = d.V[o];
  R u = u0, v = v0;
  if (mul) {
    R m = (R)mk;
    u = u * m;
    v = v * m;
    const R hu = d.HU[o], hv = d.HV[o];
    const R hum = hu * m, hvm = hv * m;
    if (!bits_equal(hum, hu)) d.HU[o] = hum;
    if (!bits_equal(hvm, hv)) d.HV[o] = hvm;
  }
  if (sponge) {
    if ((R)mk > (R)0) d.Eta[o] = ddiv(d.Eta[o], sp);
    u = ddiv(u, sp);
    v = ddiv(v, sp);
  }
  if (!bits_equal(u, u0)) d.U[o] = u;
  if (!bits_equal(v, v0)) d.V[o] = v;
}
template <class R>
__global__ void k_mask_mul(Dev<R> d, Geo g, Prm<R> p, int sponge, int mul) {
  mask_mul_body<R>(d, g, p, sponge, mul, grid_blk());
}

// FP64 budget (mass, energy, extrema) over the interior.
// FP64 budget terms of one wet interior cell (fw_stats): total depth,
// volume and energy formed in FP64
struct StatAcc {
  double mass = 0, en = 0, mx = -1e300, mn = 1e300, mu = 0, mv = 0, wet = 0;
};
template <class R>
__device__ __forceinline__ void stats_cell(const Dev<R> &d, size_t o, double dxdy, StatAcc &a) {
  if (d.MASK[o] < 1) return;
  a.wet += 1.0;
  double e = (double)d.Eta[o], u = (double)d.U[o], v = (double)d.V[o];
  double h = e + (double)d.Depth[o];  // total depth formed in FP64
  a.mass += e * dxdy;
  a.en += (0.5 * h * (u * u + v * v) + 0.5 * 9.81 * e * e) * dxdy;
  a.mx = e > a.mx ? e : a.mx;
  a.mn = e < a.mn ? e : a.mn;
  a.mu = fabs(u) > a.mu ? fabs(u) : a.mu;
  a.mv = fabs(v) > a.mv ? fabs(v) : a.mv;
}
// workgroup reduction of the seven terms; thread 0 writes them to out[0..6]
__device__ __forceinline__ void stats_block_reduce(const StatAcc &a, double *out) {
  __shared__ double sm[7][256];
  int t = threadIdx.x;
  sm[0][t] = a.mass; sm[1][t] = a.en; sm[2][t] = a.mx; sm[3][t] = a.mn; sm[4][t] = a.mu; sm[5][t] = a.mv;
  sm[6][t] = a.wet;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) {
      sm[0][t] += sm[0][t + s];
      sm[1][t] += sm[1][t + s];
      sm[6][t] += sm[6][t + s];
      sm[2][t] = sm[2][t] > sm[2][t + s] ? sm[2][t] : sm[2][t + s];
      sm[3][t] = sm[3][t] < sm[3][t + s] ? sm[3][t] : sm[3][t + s];
      sm[4][t] = sm[4][t] > sm[4][t + s] ? sm[4][t] : sm[4][t + s];
      sm[5][t] = sm[5][t] > sm[5][t + s] ? sm[5][t] : sm[5][t + s];
    }
    __syncthreads();
  }
  if (t == 0)
    for (int k = 0; k < 7; k++) out[k] = sm[k][0];
}
template <class R>
__global__ void k_stats(Dev<R> d, Geo g, Prm<R> p) {
  StatAcc a;
  int nx = g.Ie - g.Ib + 1, n = nx * (g.Je - g.Jb + 1);
  double dxdy = (double)p.DX * (double)p.DY;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    int i = g.Ib + q % nx, j = g.Jb + q / nx;
    stats_cell(d, IX(i, j), dxdy, a);
  }
  stats_block_reduce(a, d.red64 + blockIdx.x * 7);
}

// The same budget with one partial per interior row (workgroup j - Jbeg
// reduces row j into part[7 * (j - Jbeg)]), then one workgroup sums the
// rows in order into part[7 * rows .. 7 * rows + 6]: the partials live in
// the caller's buffer (fw_reduce_stats_noalloc passes the library's static
// device array).
template <class R>
__global__ void k_stats_rows(Dev<R> d, Geo g, Prm<R> p, double *part) {
  StatAcc a;
  const int j = g.Jb + blockIdx.x;
  const double dxdy = (double)p.DX * (double)p.DY;
  for (int i = g.Ib + threadIdx.x; i <= g.Ie; i += blockDim.x) stats_cell(d, IX(i, j), dxdy, a);
  stats_block_reduce(a, part + 7 * blockIdx.x);
}
static __global__ void k_stats_sum(double *part, int rows) {
  const int k = threadIdx.x;
  if (k >= 7) return;
  double s = k == 2 ? -1e300 : (k == 3 ? 1e300 : 0.0);
  for (int r = 0; r < rows; r++) {
    const double v = part[7 * r + k];
    if (k == 2 || k == 4 || k == 5) s = v > s ? v : s;
    else if (k == 3) s = v < s ? v : s;
    else s += v;
  }
  part[7 * rows + k] = s;
}

#undef IX
}  // namespace fw
