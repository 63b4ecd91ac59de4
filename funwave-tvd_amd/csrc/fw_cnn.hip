// fw_cnn.hip -- learned coarse-grid correction (no counterpart in the
// reference): a four-layer 3x3 convolutional network evaluated on the
// matrix cores of gfx950, its training step, and the in-place correction of
// a coarse solver state.  C ABI in include/fwcnn.h.
//
// Layout.  Activations are bf16, channel-contiguous per cell ("pixel"), on a
// grid with a one-cell zero border: pixel (b, y, x) of an H x W sample sits
// at index b*PS + (y+1)*(W+2) + (x+1), PS = (H+2)*(W+2) + kSlack.  The
// border and the slack are zero and never written, so the convolutions'
// zero padding needs no bounds checks.  Inputs and gradients of the output
// layer have 8 channels (16 B per pixel), hidden layers 32 (64 B).
//
// Convolution = implicit GEMM on v_mfma_f32_16x16x32_bf16: M = 16 pixels of
// a row, N = 16 output channels, K = 32 = (one tap x 32 input channels) for
// 32-channel inputs, or (four taps x 8 channels) for 8-channel inputs.  Lane
// l supplies A[l%16][8(l/16)..+7] (its pixel, shifted by the tap, one 16-B
// load) and B[8(l/16)..+7][l%16] (packed weights, one 16-B load), and holds
// C[4(l/16)..+3][l%16].  A wave computes 64 pixels (4 M-tiles) x all output
// channels with the weights in registers; the 16 x 32 result tile goes
// through wave-private LDS so that stores are 16 B per lane.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "fwcnn.h"

namespace fwl {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16;

constexpr int kSlack = 64;  // pixels past the last row (over-reads of a partial M-tile)
constexpr int kC = 32;      // hidden channels
constexpr int kCin = 5;     // input features
constexpr int kCout = 3;    // outputs

// parameter offsets in the fp32 vector (include/fwcnn.h)
constexpr int oW1 = 0, oB1 = oW1 + 9 * kCin * kC, oW2 = oB1 + kC, oB2 = oW2 + 9 * kC * kC, oW3 = oB2 + kC,
              oB3 = oW3 + 9 * kC * kC, oW4 = oB3 + kC, oB4 = oW4 + 9 * kC * kCout, kNP = oB4 + kCout;
static_assert(kNP == FW_CNN_NPARAMS, "parameter layout");

__host__ __device__ inline bf16 f2bf(float f) {  // round to nearest even
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}
__host__ __device__ inline float bf2f(bf16 h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

struct Grid {
  int H, W, PS;
  __device__ __forceinline__ size_t pix(int b, int y, int x) const {
    return (size_t)b * PS + (size_t)(y + 1) * (W + 2) + (x + 1);
  }
};

enum { MODE_RELU = 0, MODE_LAST = 1, MODE_MASK = 2 };

// Packed weights per K-step s and N-tile nt: 64 lanes x 8 bf16.
//   CIN = 32: s = tap, lane l holds W[tap][ci = 8(l/16)+e][co = 16 nt + l%16]
//   CIN = 8:  s = 0..2, lane l holds W[tap = 4s + l/16][ci = e][co = 16 nt + l%16]
// MODE_RELU: out = bf16(relu(acc + bias)), 32 channels
// MODE_LAST: out = fp32 planar [b][3][H][W] = acc + bias (channels 0..2)
// MODE_MASK: out = bf16(acc) * (act > 0), 32 channels (gradient through ReLU)
//
// A workgroup (4 waves) computes rows y0..y0+3, pixels x0..x0+63.  The 6
// input rows and 66 pixels they read are staged into LDS once (16-B loads;
// outside the padded grid: zeros), and every tap's A fragment is an LDS
// read: each input pixel crosses L1 once instead of nine times.  The pixel
// stride in LDS is 80 B for 32 channels, so the 16 lanes of a fragment read
// hit distinct banks.
template <int CIN, int NT, int MODE>
__global__ void __launch_bounds__(256) k_conv(const bf16 *__restrict__ in, const bf16 *__restrict__ wpk,
                                              const float *__restrict__ bias, void *__restrict__ out,
                                              const bf16 *__restrict__ act, Grid g) {
  constexpr int NS = CIN == 32 ? 9 : 3;
  constexpr int TW = 66, TR = 6;           // staged pixels per row, staged rows
  constexpr int LPX = CIN == 32 ? 40 : 8;  // LDS pixel stride (bf16)
  constexpr int CH16 = CIN / 8;            // 16-B chunks per pixel
  __shared__ __attribute__((aligned(16))) bf16 stg[TR * TW * LPX];
  __shared__ __attribute__((aligned(16))) bf16 tile[4][16][kC + 8];  // per wave: 16 px x 32 ch (+pad)
  const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int row = l & 15, grp = l >> 4;
  const int y0 = blockIdx.y * 4, x0b = blockIdx.x * 64;
  const int y = y0 + wave;
  const int b = blockIdx.z;
  const bool live = y < g.H;
  // padded rows y0..y0+5 (pixel rows y0-1..y0+4), padded columns x0b..x0b+65
  for (int e = threadIdx.x; e < TR * TW * CH16; e += 256) {
    const int c = e % CH16, px = (e / CH16) % TW, r = e / (CH16 * TW);
    const int gy = y0 + r, gx = x0b + px;
    uint4 v{0u, 0u, 0u, 0u};
    if (gy < g.H + 2 && gx < g.W + 2)
      v = *(const uint4 *)(in + ((size_t)b * g.PS + (size_t)gy * (g.W + 2) + gx) * CIN + 8 * c);
    *(uint4 *)&stg[(r * TW + px) * LPX + 8 * c] = v;
  }
  __syncthreads();
  bf16x8 bw[NS][NT];
#pragma unroll
  for (int s = 0; s < NS; s++)
#pragma unroll
    for (int nt = 0; nt < NT; nt++) bw[s][nt] = *(const bf16x8 *)(wpk + ((size_t)(s * NT + nt) * 64 + l) * 8);
  float bs[NT];
#pragma unroll
  for (int nt = 0; nt < NT; nt++) bs[nt] = (MODE == MODE_MASK) ? 0.f : bias[nt * 16 + row];
  // taps outer, the 4 M-tiles inner: 4 x NT independent accumulation chains
  // keep the matrix pipe busy
  f32x4 accs[4][NT];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int nt = 0; nt < NT; nt++) accs[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
  if (live) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const int tap = CIN == 32 ? s : 4 * s + grp;
      const int choff = CIN == 32 ? 8 * grp : 0;
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        bf16x8 a;
        if (CIN == 32 || tap < 9) {
          const int dy = tap / 3 - 1, dx = tap % 3 - 1;
          a = *(const bf16x8 *)&stg[((wave + 1 + dy) * TW + (mt * 16 + row + 1 + dx)) * LPX + choff];
        } else {
#pragma unroll
          for (int e = 0; e < 8; e++) a[e] = (__bf16)0.f;
        }
#pragma unroll
        for (int nt = 0; nt < NT; nt++)
          accs[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[s][nt], accs[mt][nt], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    const int x0 = blockIdx.x * 64 + mt * 16;
    const f32x4 *acc = accs[mt];
    if (MODE == MODE_LAST) {
      if (live && row < kCout) {
        float *o = (float *)out;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int x = x0 + 4 * grp + r;
          if (x < g.W) o[(((size_t)b * kCout + row) * g.H + y) * g.W + x] = acc[0][r] + bs[0];
        }
      }
      continue;
    }
    // 16 x 32 tile through LDS, then 16 B per lane: pixel l/4, channels 8(l%4)..
#pragma unroll
    for (int nt = 0; nt < NT; nt++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        float v = acc[nt][r] + bs[nt];
        if (MODE == MODE_RELU) v = v > 0.f ? v : 0.f;
        tile[wave][4 * grp + r][nt * 16 + row] = f2bf(v);
      }
    // the tile is this wave's own: its LDS operations complete in order, so
    // a wavefront-scope fence (no workgroup barrier) orders write and read
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int q = l >> 2, cg = l & 3;
    const int x = x0 + q;
    if (live && x < g.W) {
      const size_t p = g.pix(b, y, x);
      uint4 v = *(const uint4 *)&tile[wave][q][8 * cg];
      if (MODE == MODE_MASK) {
        const uint4 m = *(const uint4 *)(act + p * kC + 8 * cg);
        uint32_t vv[4] = {v.x, v.y, v.z, v.w}, mm[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
          // bf16 > 0: sign clear and not +0 (activations are never NaN)
          const uint32_t lo = (mm[k] & 0xffffu), hi = mm[k] >> 16;
          const uint32_t keep_lo = (lo != 0u && (lo & 0x8000u) == 0u) ? 0xffffu : 0u;
          const uint32_t keep_hi = (hi != 0u && (hi & 0x8000u) == 0u) ? 0xffff0000u : 0u;
          vv[k] &= (keep_lo | keep_hi);
        }
        v = uint4{vv[0], vv[1], vv[2], vv[3]};
      }
      *(uint4 *)((bf16 *)out + p * kC + 8 * cg) = v;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

// planar fp32 features [B][5][H][W] -> padded bf16 pixels with 8 channels
__global__ void k_pack_input(const float *__restrict__ src, bf16 *__restrict__ dst, Grid g, int B) {
  const size_t n = (size_t)B * g.H * g.W;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(t % g.W), y = (int)((t / g.W) % g.H), b = (int)(t / ((size_t)g.W * g.H));
    uint16_t v[8];
#pragma unroll
    for (int c = 0; c < 8; c++)
      v[c] = c < kCin ? f2bf(src[(((size_t)b * kCin + c) * g.H + y) * g.W + x]) : (uint16_t)0;
    uint4 w{v[0] | ((uint32_t)v[1] << 16), v[2] | ((uint32_t)v[3] << 16), v[4] | ((uint32_t)v[5] << 16),
            v[6] | ((uint32_t)v[7] << 16)};
    *(uint4 *)(dst + g.pix(b, y, x) * 8) = w;
  }
}

// features of a coarse solver state (interior Ib.., Jb.., pitch P)
template <class R>
__global__ void k_features(const R *eta, const R *u, const R *v, const R *dep, const int *mask, int P, int Ib,
                           int Jb, float f0, float f1, float f2, float f3, bf16 *__restrict__ dst, Grid g) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= g.W || y >= g.H) return;
  const size_t o = (size_t)(Jb - 1 + y) * P + (Ib - 1 + x);
  uint16_t w[8] = {f2bf((float)eta[o] * f0), f2bf((float)u[o] * f1), f2bf((float)v[o] * f2),
                   f2bf((float)dep[o] * f3), f2bf((float)mask[o]), 0, 0, 0};
  uint4 q{w[0] | ((uint32_t)w[1] << 16), w[2] | ((uint32_t)w[3] << 16), w[4] | ((uint32_t)w[5] << 16),
          w[6] | ((uint32_t)w[7] << 16)};
  *(uint4 *)(dst + g.pix(0, y, x) * 8) = q;
}

// mean squared error: per-block partial sums of (out - tgt)^2 and the
// output-layer gradient 2 (out - tgt) / count as 8-channel bf16 pixels
__global__ void k_loss(const float *__restrict__ out, const float *__restrict__ tgt, bf16 *__restrict__ g4,
                       float *__restrict__ part, Grid g, int B, float inv_count) {
  __shared__ float red[256];
  const size_t n = (size_t)B * g.H * g.W;
  float s = 0.f;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(t % g.W), y = (int)((t / g.W) % g.H), b = (int)(t / ((size_t)g.W * g.H));
    uint16_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < kCout; c++) {
      const size_t q = (((size_t)b * kCout + c) * g.H + y) * g.W + x;
      const float d = out[q] - tgt[q];
      s += d * d;
      w[c] = f2bf(2.f * d * inv_count);
    }
    uint4 v{w[0] | ((uint32_t)w[1] << 16), w[2] | ((uint32_t)w[3] << 16), 0u, 0u};
    *(uint4 *)(g4 + g.pix(b, y, x) * 8) = v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Weight and bias gradients of one layer over a tile of 4 rows x 64
// pixels: thread (tap, ci) accumulates dW[tap][ci][0..COUT) in fp32;
// partial sums per workgroup, reduced by k_reduce_grad.
//   in: layer input, CIN channels stored (CREAL used); g: gradient of the
//   layer's pre-activation output, GC channels stored (COUT used).
template <int CIN, int CREAL, int GC, int COUT>
__global__ void __launch_bounds__(9 * CIN) k_wgrad(const bf16 *__restrict__ in, const bf16 *__restrict__ gin,
                                                   float *__restrict__ part, Grid g) {
  __shared__ __attribute__((aligned(16))) bf16 sin_[6][66][CIN];
  __shared__ float sg[4][64][COUT];
  const int t = threadIdx.x;
  const int b = blockIdx.z, y0 = blockIdx.y * 4, x0 = blockIdx.x * 64;
  // stage the input tile with a one-pixel halo (rows y0-1..y0+4), and g
  for (int e = t; e < 6 * 66 * CIN / 8; e += blockDim.x) {
    const int c8 = e % (CIN / 8), px = (e / (CIN / 8)) % 66, ry = e / (CIN / 8) / 66;
    const int yy = y0 - 1 + ry, xx = x0 - 1 + px;
    uint4 v{0u, 0u, 0u, 0u};
    if (yy >= -1 && yy <= g.H && xx >= -1 && xx <= g.W)
      v = *(const uint4 *)(in + ((size_t)b * g.PS + (size_t)(yy + 1) * (g.W + 2) + (xx + 1)) * CIN + 8 * c8);
    *(uint4 *)&sin_[ry][px][8 * c8] = v;
  }
  for (int e = t; e < 4 * 64 * COUT; e += blockDim.x) {
    const int co = e % COUT, px = (e / COUT) % 64, ry = e / COUT / 64;
    const int yy = y0 + ry, xx = x0 + px;
    float v = 0.f;
    if (yy < g.H && xx < g.W) v = bf2f(gin[g.pix(b, yy, xx) * GC + co]);
    sg[ry][px][co] = v;
  }
  __syncthreads();
  const int tap = t / CIN, ci = t % CIN;
  const int dy = tap / 3 - 1, dx = tap % 3 - 1;
  float acc[COUT], accb[COUT];
#pragma unroll
  for (int co = 0; co < COUT; co++) acc[co] = accb[co] = 0.f;
  const bool doBias = (t == 0);
  for (int ry = 0; ry < 4; ry++)
    for (int px = 0; px < 64; px++) {
      const float a = bf2f(sin_[ry + 1 + dy][px + 1 + dx][ci]);
#pragma unroll
      for (int co = 0; co < COUT; co++) acc[co] += a * sg[ry][px][co];
      if (doBias) {
#pragma unroll
        for (int co = 0; co < COUT; co++) accb[co] += sg[ry][px][co];
      }
    }
  const size_t nblk = (size_t)blockIdx.x + (size_t)gridDim.x * (blockIdx.y + (size_t)gridDim.y * blockIdx.z);
  float *dst = part + nblk * (9 * CREAL * COUT + COUT);
  if (ci < CREAL) {
#pragma unroll
    for (int co = 0; co < COUT; co++) dst[(tap * CREAL + ci) * COUT + co] = acc[co];
  }
  if (doBias) {
#pragma unroll
    for (int co = 0; co < COUT; co++) dst[9 * CREAL * COUT + co] = accb[co];
  }
}

// grad[off + k] = sum over blocks of part[blk][k]
__global__ void k_reduce_grad(const float *__restrict__ part, int nblk, int len, float *__restrict__ grad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= len) return;
  float s = 0.f;
  for (int q = 0; q < nblk; q++) s += part[(size_t)q * len + k];
  grad[k] = s;
}

__global__ void k_adam(float *p, float *m, float *v, const float *gr, int n, float lr, float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float gk = gr[k];
  const float mk = 0.9f * m[k] + 0.1f * gk;
  const float vk = 0.999f * v[k] + 0.001f * gk * gk;
  m[k] = mk;
  v[k] = vk;
  p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + 1e-8f);
}

// forward and data-gradient weight packs from the fp32 parameters
// (one thread per packed element; see k_conv for the lane layout)
struct Packs {
  bf16 *f1, *f2, *f3, *f4;  // forward
  bf16 *b2, *b3, *b4;       // data gradient of layers 2..4 (flipped taps, ci <-> co)
  float *bias1, *bias2, *bias3, *bias4;
};
__device__ __forceinline__ float wget(const float *p, int off, int tap, int ci, int co, int nci, int nco) {
  return (ci < nci && co < nco && tap < 9) ? p[off + (tap * nci + ci) * nco + co] : 0.f;
}
__global__ void k_pack(const float *__restrict__ p, Packs k) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  // index = ((s * NT + nt) * 64 + lane) * 8 + e
  auto dec = [](int idx, int NT, int &s, int &nt, int &lane, int &e) {
    e = idx % 8; lane = (idx / 8) % 64; nt = (idx / 512) % NT; s = idx / 512 / NT;
  };
  int s, nt, lane, e;
  const int n32 = 9 * 2 * 512, n8 = 3 * 2 * 512, n4 = 9 * 1 * 512;
  if (t < n8) {  // layer 1 forward: CIN 8 (5 real), 32 out
    dec(t, 2, s, nt, lane, e);
    k.f1[t] = f2bf(wget(p, oW1, 4 * s + lane / 16, e, nt * 16 + lane % 16, kCin, kC));
  }
  if (t < n32) {
    dec(t, 2, s, nt, lane, e);
    const int ci = 8 * (lane / 16) + e, co = nt * 16 + lane % 16;
    k.f2[t] = f2bf(wget(p, oW2, s, ci, co, kC, kC));
    k.f3[t] = f2bf(wget(p, oW3, s, ci, co, kC, kC));
    // data gradient: input channel = co of the layer, output = ci; tap flipped
    k.b2[t] = f2bf(wget(p, oW2, 8 - s, co, ci, kC, kC));
    k.b3[t] = f2bf(wget(p, oW3, 8 - s, co, ci, kC, kC));
  }
  if (t < n4) {  // layer 4 forward: 32 in, 16 out (3 real)
    dec(t, 1, s, nt, lane, e);
    k.f4[t] = f2bf(wget(p, oW4, s, 8 * (lane / 16) + e, lane % 16, kC, kCout));
  }
  if (t < n8) {  // layer 4 data gradient: 8-channel input (3 real) -> 32
    dec(t, 2, s, nt, lane, e);
    const int tp = 4 * s + lane / 16;
    k.b4[t] = f2bf(tp < 9 ? wget(p, oW4, 8 - tp, nt * 16 + lane % 16, e, kC, kCout) : 0.f);
  }
  if (t < kC) {
    k.bias1[t] = p[oB1 + t];
    k.bias2[t] = p[oB2 + t];
    k.bias3[t] = p[oB3 + t];
  }
  if (t < 16) k.bias4[t] = t < kCout ? p[oB4 + t] : 0.f;
}

// wet-cell mean of the eta correction (R)(o0*s0), accumulated in FP64 by
// one workgroup in a fixed order (deterministic); 0 without wet cells
template <class R>
__global__ void __launch_bounds__(1024) k_wet_mean(const float *__restrict__ o, const int *mask, int P, int Ib, int Jb,
                                                   int H, int W, float s0, double *mean) {
  __shared__ double ss[1024], sc[1024];
  const int t = threadIdx.x;
  double s = 0, c = 0;
  for (int k = t; k < H * W; k += 1024) {
    const int y = k / W, x = k - y * W;
    if (mask[(size_t)(Jb - 1 + y) * P + (Ib - 1 + x)] >= 1) {
      s += (double)(R)(o[k] * s0);
      c += 1;
    }
  }
  ss[t] = s;
  sc[t] = c;
  __syncthreads();
  for (int w = 512; w > 0; w >>= 1) {
    if (t < w) { ss[t] += ss[t + w]; sc[t] += sc[t + w]; }
    __syncthreads();
  }
  if (t == 0) mean[0] = sc[0] > 0 ? ss[0] / sc[0] : 0.0;
}

// correction of the coarse state from the network output (planar [3][H][W]);
// with mean, the wet-cell mean of the eta correction is taken out of it
template <class R>
__global__ void k_correct(const float *__restrict__ o, R *eta, R *u, R *v, R *h, R *hu, R *hv, R *ubar, R *vbar,
                          const R *dep, const int *mask, int P, int Ib, int Jb, int H, int W, float s0, float s1,
                          float s2, R g3, R mdf, const double *mean) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= W || y >= H) return;
  const size_t q = (size_t)(Jb - 1 + y) * P + (Ib - 1 + x);
  if (mask[q] < 1) return;
  const size_t k = (size_t)y * W + x, hw = (size_t)H * W;
  const R de = (R)(o[k] * s0) - (mean ? (R)mean[0] : (R)0);
  const R du = (R)(o[hw + k] * s1), dv = (R)(o[2 * hw + k] * s2);
  const R e = eta[q] + de;
  const R hh = e * g3 + dep[q];
  const R hm = hh > mdf ? hh : mdf;
  const R uu = u[q] + du, vv = v[q] + dv;
  eta[q] = e;
  h[q] = hh;
  u[q] = uu;
  v[q] = vv;
  hu[q] = hm * uu;
  hv[q] = hm * vv;
  ubar[q] = ubar[q] + hm * du;
  vbar[q] = vbar[q] + hm * dv;
}

}  // namespace fwl

using namespace fwl;

struct fw_cnn {
  int H = 0, W = 0, maxB = 0;
  Grid g{};
  hipStream_t stream = nullptr;
  std::vector<void *> allocs;
  float *params = nullptr, *am = nullptr, *av = nullptr, *grad = nullptr, *part = nullptr, *lpart = nullptr;
  size_t part_len = 0;
  Packs pk{};
  bf16 *a0 = nullptr, *a1 = nullptr, *a2 = nullptr, *a3 = nullptr;
  bf16 *g4 = nullptr, *g3 = nullptr, *g2 = nullptr, *g1 = nullptr;
  float *in = nullptr, *out = nullptr, *tgt = nullptr;
  double *mean = nullptr;  // wet-cell mean of the eta correction
  int conserve = 0;
  int adam_t = 0;
  std::string err;
};

namespace {

int cfail(fw_cnn *n, int code, const std::string &m) {
  if (n) n->err = m;
  return code;
}
#define CHK(n, x)                                                                    \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) return cfail(n, FW_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
int cget(fw_cnn *n, size_t count, T **out) {
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(T));
  if (e != hipSuccess) return cfail(n, FW_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  n->allocs.push_back(p);
  e = hipMemsetAsync(p, 0, count * sizeof(T), n->stream);
  if (e != hipSuccess) return cfail(n, FW_E_HIP, hipGetErrorString(e));
  *out = (T *)p;
  return FW_OK;
}

constexpr int kWgBlockRows = 4;

dim3 conv_grid(const fw_cnn *n, int B) { return dim3((n->W + 63) / 64, (n->H + 3) / 4, B); }
dim3 wgrad_grid(const fw_cnn *n, int B) { return dim3((n->W + 63) / 64, (n->H + kWgBlockRows - 1) / kWgBlockRows, B); }

void forward(fw_cnn *n, int B) {
  const dim3 gr = conv_grid(n, B), bl(256);
  hipLaunchKernelGGL((k_conv<8, 2, MODE_RELU>), gr, bl, 0, n->stream, n->a0, n->pk.f1, n->pk.bias1, n->a1, nullptr, n->g);
  hipLaunchKernelGGL((k_conv<32, 2, MODE_RELU>), gr, bl, 0, n->stream, n->a1, n->pk.f2, n->pk.bias2, n->a2, nullptr, n->g);
  hipLaunchKernelGGL((k_conv<32, 2, MODE_RELU>), gr, bl, 0, n->stream, n->a2, n->pk.f3, n->pk.bias3, n->a3, nullptr, n->g);
  hipLaunchKernelGGL((k_conv<32, 1, MODE_LAST>), gr, bl, 0, n->stream, n->a3, n->pk.f4, n->pk.bias4, n->out, nullptr, n->g);
}

int pack(fw_cnn *n) {
  hipLaunchKernelGGL(k_pack, dim3((9 * 2 * 512 + 255) / 256), dim3(256), 0, n->stream, n->params, n->pk);
  CHK(n, hipGetLastError());
  return FW_OK;
}

int check_batch(fw_cnn *n, int B) {
  if (!n) return FW_E_INVALID;
  if (B < 1 || B > n->maxB) return cfail(n, FW_E_INVALID, "batch size outside 1..max_batch");
  return FW_OK;
}

}  // namespace

extern "C" {

int fw_cnn_create(int H, int W, int max_batch, fw_cnn **out) {
  if (!out) return FW_E_INVALID;
  *out = nullptr;
  if (H < 1 || W < 1 || max_batch < 1 || (size_t)H * W * max_batch > ((size_t)1 << 31)) return FW_E_INVALID;
  fw_cnn *n = new fw_cnn();
  n->H = H; n->W = W; n->maxB = max_batch;
  n->g = Grid{H, W, (H + 2) * (W + 2) + kSlack};
  if (hipStreamCreate(&n->stream) != hipSuccess) { delete n; return FW_E_HIP; }
  const size_t px = (size_t)max_batch * n->g.PS;
  const size_t plane = (size_t)max_batch * H * W;
  const dim3 wg = wgrad_grid(n, max_batch);
  const size_t nblk = (size_t)wg.x * wg.y * wg.z;
  n->part_len = nblk * (9 * kC * kC + kC);
  int rc;
  if ((rc = cget(n, kNP, &n->params)) || (rc = cget(n, kNP, &n->am)) || (rc = cget(n, kNP, &n->av)) ||
      (rc = cget(n, kNP, &n->grad)) || (rc = cget(n, n->part_len, &n->part)) || (rc = cget(n, 4096, &n->lpart)) ||
      (rc = cget(n, 3 * 512 * 2, &n->pk.f1)) || (rc = cget(n, 9 * 512 * 2, &n->pk.f2)) ||
      (rc = cget(n, 9 * 512 * 2, &n->pk.f3)) || (rc = cget(n, 9 * 512 * 1, &n->pk.f4)) ||
      (rc = cget(n, 9 * 512 * 2, &n->pk.b2)) || (rc = cget(n, 9 * 512 * 2, &n->pk.b3)) ||
      (rc = cget(n, 3 * 512 * 2, &n->pk.b4)) || (rc = cget(n, 32, &n->pk.bias1)) ||
      (rc = cget(n, 32, &n->pk.bias2)) || (rc = cget(n, 32, &n->pk.bias3)) || (rc = cget(n, 16, &n->pk.bias4)) ||
      (rc = cget(n, 1, &n->mean)) ||
      (rc = cget(n, px * 8, &n->a0)) || (rc = cget(n, px * kC, &n->a1)) || (rc = cget(n, px * kC, &n->a2)) ||
      (rc = cget(n, px * kC, &n->a3)) || (rc = cget(n, px * 8, &n->g4)) || (rc = cget(n, px * kC, &n->g3)) ||
      (rc = cget(n, px * kC, &n->g2)) || (rc = cget(n, px * kC, &n->g1)) ||
      (rc = cget(n, plane * kCin, &n->in)) || (rc = cget(n, plane * kCout, &n->out)) ||
      (rc = cget(n, plane * kCout, &n->tgt))) {
    fw_cnn_destroy(n);
    return rc;
  }
  if (hipStreamSynchronize(n->stream) != hipSuccess) { fw_cnn_destroy(n); return FW_E_HIP; }
  *out = n;
  return FW_OK;
}

int fw_cnn_destroy(fw_cnn *n) {
  if (!n) return FW_E_INVALID;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  for (void *p : n->allocs) (void)hipFree(p);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
  return FW_OK;
}

const char *fw_cnn_strerror(const fw_cnn *n) { return n ? n->err.c_str() : "null network"; }

int fw_cnn_set_params(fw_cnn *n, const float *p, int count) {
  if (!n || !p) return FW_E_INVALID;
  if (count != kNP) return cfail(n, FW_E_INVALID, "parameter count differs from FW_CNN_NPARAMS");
  CHK(n, hipMemcpyAsync(n->params, p, kNP * sizeof(float), hipMemcpyHostToDevice, n->stream));
  CHK(n, hipMemsetAsync(n->am, 0, kNP * sizeof(float), n->stream));
  CHK(n, hipMemsetAsync(n->av, 0, kNP * sizeof(float), n->stream));
  n->adam_t = 0;
  int rc = pack(n);
  if (rc) return rc;
  CHK(n, hipStreamSynchronize(n->stream));
  return FW_OK;
}

int fw_cnn_get_params(fw_cnn *n, float *p, int count) {
  if (!n || !p) return FW_E_INVALID;
  if (count != kNP) return cfail(n, FW_E_INVALID, "parameter count differs from FW_CNN_NPARAMS");
  CHK(n, hipMemcpyAsync(p, n->params, kNP * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  CHK(n, hipStreamSynchronize(n->stream));
  return FW_OK;
}

int fw_cnn_forward(fw_cnn *n, const float *in, float *out, int B) {
  int rc = check_batch(n, B);
  if (rc) return rc;
  if (!in || !out) return cfail(n, FW_E_INVALID, "null buffer");
  const size_t plane = (size_t)B * n->H * n->W;
  CHK(n, hipMemcpyAsync(n->in, in, plane * kCin * sizeof(float), hipMemcpyHostToDevice, n->stream));
  hipLaunchKernelGGL(k_pack_input, dim3(1024), dim3(256), 0, n->stream, n->in, n->a0, n->g, B);
  forward(n, B);
  CHK(n, hipGetLastError());
  CHK(n, hipMemcpyAsync(out, n->out, plane * kCout * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  CHK(n, hipStreamSynchronize(n->stream));
  return FW_OK;
}

int fw_cnn_train_step(fw_cnn *n, const float *in, const float *target, int B, float lr, float *loss, float *grad) {
  int rc = check_batch(n, B);
  if (rc) return rc;
  if (!in || !target) return cfail(n, FW_E_INVALID, "null buffer");
  const size_t plane = (size_t)B * n->H * n->W;
  CHK(n, hipMemcpyAsync(n->in, in, plane * kCin * sizeof(float), hipMemcpyHostToDevice, n->stream));
  CHK(n, hipMemcpyAsync(n->tgt, target, plane * kCout * sizeof(float), hipMemcpyHostToDevice, n->stream));
  hipLaunchKernelGGL(k_pack_input, dim3(1024), dim3(256), 0, n->stream, n->in, n->a0, n->g, B);
  forward(n, B);
  const int nlb = 1024;
  hipLaunchKernelGGL(k_loss, dim3(nlb), dim3(256), 0, n->stream, n->out, n->tgt, n->g4, n->lpart, n->g, B,
                     1.f / (float)(plane * kCout));
  // data gradients through layers 4, 3, 2 (ReLU masks from the activations)
  const dim3 gr = conv_grid(n, B), bl(256);
  hipLaunchKernelGGL((k_conv<8, 2, MODE_MASK>), gr, bl, 0, n->stream, n->g4, n->pk.b4, nullptr, n->g3, n->a3, n->g);
  hipLaunchKernelGGL((k_conv<32, 2, MODE_MASK>), gr, bl, 0, n->stream, n->g3, n->pk.b3, nullptr, n->g2, n->a2, n->g);
  hipLaunchKernelGGL((k_conv<32, 2, MODE_MASK>), gr, bl, 0, n->stream, n->g2, n->pk.b2, nullptr, n->g1, n->a1, n->g);
  // weight gradients, reduced into the parameter layout
  const dim3 wg = wgrad_grid(n, B);
  const int nblk = (int)(wg.x * wg.y * wg.z);
  auto reduce = [&](int len, int off) {
    hipLaunchKernelGGL(k_reduce_grad, dim3((len + 255) / 256), dim3(256), 0, n->stream, n->part, nblk, len,
                       n->grad + off);
  };
  hipLaunchKernelGGL((k_wgrad<8, kCin, 32, kC>), wg, dim3(9 * 8), 0, n->stream, n->a0, n->g1, n->part, n->g);
  reduce(9 * kCin * kC + kC, oW1);
  hipLaunchKernelGGL((k_wgrad<32, kC, 32, kC>), wg, dim3(9 * 32), 0, n->stream, n->a1, n->g2, n->part, n->g);
  reduce(9 * kC * kC + kC, oW2);
  hipLaunchKernelGGL((k_wgrad<32, kC, 32, kC>), wg, dim3(9 * 32), 0, n->stream, n->a2, n->g3, n->part, n->g);
  reduce(9 * kC * kC + kC, oW3);
  hipLaunchKernelGGL((k_wgrad<32, kC, 8, kCout>), wg, dim3(9 * 32), 0, n->stream, n->a3, n->g4, n->part, n->g);
  reduce(9 * kC * kCout + kCout, oW4);
  CHK(n, hipGetLastError());
  if (grad) CHK(n, hipMemcpyAsync(grad, n->grad, kNP * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  std::vector<float> lp(nlb);
  CHK(n, hipMemcpyAsync(lp.data(), n->lpart, nlb * sizeof(float), hipMemcpyDeviceToHost, n->stream));
  n->adam_t += 1;
  const float c1 = 1.f - powf(0.9f, (float)n->adam_t), c2 = 1.f - powf(0.999f, (float)n->adam_t);
  hipLaunchKernelGGL(k_adam, dim3((kNP + 255) / 256), dim3(256), 0, n->stream, n->params, n->am, n->av, n->grad,
                     kNP, lr, c1, c2);
  if ((rc = pack(n))) return rc;
  CHK(n, hipStreamSynchronize(n->stream));
  if (loss) {
    double s = 0;
    for (float v : lp) s += v;
    *loss = (float)(s / (double)(plane * kCout));
  }
  return FW_OK;
}

int fw_cnn_time_forward(fw_cnn *n, int B, int iters, double *ms) {
  int rc = check_batch(n, B);
  if (rc) return rc;
  if (iters < 1 || !ms) return cfail(n, FW_E_INVALID, "iters must be >= 1");
  hipEvent_t a, b;
  CHK(n, hipEventCreate(&a));
  CHK(n, hipEventCreate(&b));
  forward(n, B);  // warm-up
  CHK(n, hipEventRecord(a, n->stream));
  for (int k = 0; k < iters; k++) forward(n, B);
  CHK(n, hipEventRecord(b, n->stream));
  CHK(n, hipEventSynchronize(b));
  float t = 0.f;
  CHK(n, hipEventElapsedTime(&t, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = (double)t / iters;
  return FW_OK;
}

int fw_cnn_set_conserve(fw_cnn *n, int on) {
  if (!n) return FW_E_INVALID;
  n->conserve = on != 0;
  return FW_OK;
}

int fw_cnn_correct(fw_cnn *n, fw_ctx *solver, const float fs[4], const float os[3]) {
  if (!n || !solver || !fs || !os) return FW_E_INVALID;
  fw_config cfg;
  int rc = fw_get_config(solver, &cfg);
  if (rc) return rc;
  const int Hs = cfg.Jend - cfg.Jbeg + 1, Ws = cfg.Iend - cfg.Ibeg + 1;
  if (Hs != n->H || Ws != n->W) return cfail(n, FW_E_INVALID, "solver interior differs from the network grid");
  void *st = nullptr;
  if ((rc = fw_get_stream(solver, &st))) return rc;
  hipStream_t s = (hipStream_t)st;
  auto fld = [&](const char *name, void **p, int *pitch) { return fw_field_device(solver, fw_field_id(name), p, pitch); };
  void *eta, *u, *v, *dep, *mask, *h, *hu, *hv, *ub, *vb;
  int P = 0;
  if ((rc = fld("Eta", &eta, &P)) || (rc = fld("U", &u, nullptr)) || (rc = fld("V", &v, nullptr)) ||
      (rc = fld("Depth", &dep, nullptr)) || (rc = fld("MASK", &mask, nullptr)) || (rc = fld("H", &h, nullptr)) ||
      (rc = fld("HU", &hu, nullptr)) || (rc = fld("HV", &hv, nullptr)) || (rc = fld("Ubar", &ub, nullptr)) ||
      (rc = fld("Vbar", &vb, nullptr)))
    return cfail(n, rc, "solver field lookup failed");
  // the network runs on the solver's stream so the correction is ordered
  // with the solver's own work
  const dim3 b2(64, 4), g2((Ws + 63) / 64, (Hs + 3) / 4);
  if (cfg.real_bytes == 8) {
    hipLaunchKernelGGL(k_features<double>, g2, b2, 0, s, (const double *)eta, (const double *)u, (const double *)v,
                       (const double *)dep, (const int *)mask, P, cfg.Ibeg, cfg.Jbeg, fs[0], fs[1], fs[2], fs[3],
                       n->a0, n->g);
  } else {
    hipLaunchKernelGGL(k_features<float>, g2, b2, 0, s, (const float *)eta, (const float *)u, (const float *)v,
                       (const float *)dep, (const int *)mask, P, cfg.Ibeg, cfg.Jbeg, fs[0], fs[1], fs[2], fs[3],
                       n->a0, n->g);
  }
  hipStream_t keep = n->stream;
  n->stream = s;
  forward(n, 1);
  n->stream = keep;
  const double *mean = n->conserve ? n->mean : nullptr;
  if (cfg.real_bytes == 8) {
    if (mean)
      hipLaunchKernelGGL(k_wet_mean<double>, dim3(1), dim3(1024), 0, s, n->out, (const int *)mask, P, cfg.Ibeg,
                         cfg.Jbeg, Hs, Ws, os[0], n->mean);
    hipLaunchKernelGGL(k_correct<double>, g2, b2, 0, s, n->out, (double *)eta, (double *)u, (double *)v, (double *)h,
                       (double *)hu, (double *)hv, (double *)ub, (double *)vb, (const double *)dep, (const int *)mask,
                       P, cfg.Ibeg, cfg.Jbeg, Hs, Ws, os[0], os[1], os[2], cfg.Gamma3, cfg.MinDepthFrc, mean);
  } else {
    if (mean)
      hipLaunchKernelGGL(k_wet_mean<float>, dim3(1), dim3(1024), 0, s, n->out, (const int *)mask, P, cfg.Ibeg,
                         cfg.Jbeg, Hs, Ws, os[0], n->mean);
    hipLaunchKernelGGL(k_correct<float>, g2, b2, 0, s, n->out, (float *)eta, (float *)u, (float *)v, (float *)h,
                       (float *)hu, (float *)hv, (float *)ub, (float *)vb, (const float *)dep, (const int *)mask, P,
                       cfg.Ibeg, cfg.Jbeg, Hs, Ws, os[0], os[1], os[2], (float)cfg.Gamma3, (float)cfg.MinDepthFrc,
                       mean);
  }
  CHK(n, hipGetLastError());
  return FW_OK;
}

}  // extern "C"
