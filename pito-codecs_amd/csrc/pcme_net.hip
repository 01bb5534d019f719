// pcme_net.hip -- the learned search pruner on gfx950: a small CNN trained
// and run by hand-written kernels.  No counterpart in the reference.
//
// Every convolution is an implicit GEMM on v_mfma_f32_16x16x32_bf16:
// M = output channels (two 16-row tiles), N = output pixels (16 per tile),
// K = taps x input channels in chunks of 32 (conv1: its 18 = 9 taps x 2 ch
// packed into one zero-padded chunk; conv2-4: one chunk per tap).  Fragment
// layout of the instruction on gfx950: lane l supplies row (l & 15) of A and
// column (l & 15) of B with k = 8*(l >> 4) .. +7, and receives rows
// 4*(l >> 4) .. +3 of column (l & 15) of the fp32 result.  The accumulator
// starts at the bias and the K chunks are issued in the same order in the
// training and the inference kernels, so both produce the same bits.
//
// Activations are channel-last bf16 ([pixel][32]); conv4's output stays
// fp32 and feeds the global average pool; the fc layer is an fp32 fma chain
// over the bf16-rounded fc weights.  Backward: the data gradient is the same
// implicit GEMM with the transposed weights ([ci][tap][co]) and a
// divisibility test for the strided layers; the weight gradient is a GEMM
// over K = output pixels with a column of ones that yields the bias
// gradient, computed in fixed-size pixel chunks and summed in chunk order
// (deterministic).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "pcme_internal.h"
#include "../../include/pcme_net.h"

namespace pcme {
namespace cnn {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));
typedef uint32_t u2 __attribute__((ext_vector_type(2)));

constexpr int CH = 32;
constexpr int NCLS = PCME_NET_CLASSES;   // 293
constexpr int MVC = PCME_NET_MV_CLASSES; // 289
constexpr int NCLS_PAD = 304;
constexpr int IN = PCME_NET_IN;          // 24
constexpr int INPUT = PCME_NET_INPUT;    // 1152 bf16 per MB
constexpr int K1 = 32;                   // packed K of conv1
constexpr int KC = 288;                  // K of conv2-4

// Parameter offsets (floats) of the master copy.
constexpr int OW1 = 0, OB1 = 576, OW2 = 608, OB2 = OW2 + 9216, OW3 = OB2 + 32, OB3 = OW3 + 9216;
constexpr int OW4 = OB3 + 32, OB4 = OW4 + 9216, OWF = OB4 + 32, OBF = OWF + NCLS * CH, NPARAM = OBF + NCLS;
static_assert(NPARAM == PCME_NET_NPARAM, "parameter count");

constexpr int CHUNK = 512;  // pixels per weight-gradient partial

__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf2f(uint32_t b) { return __builtin_bit_cast(float, (b & 0xFFFFu) << 16); }
__device__ __forceinline__ uint32_t pack2(float a, float b) {
  return (uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16);
}
__device__ __forceinline__ f4 mfma(u4 a, u4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0,
                                                  0);
}
__device__ __forceinline__ f4 relu4(f4 a) {
  f4 r;
#pragma unroll
  for (int v = 0; v < 4; v++) r[v] = a[v] > 0.f ? a[v] : 0.f;
  return r;
}

// (v - 128) / 64 as bf16 bits (exact: |v - 128| <= 128 needs 7 mantissa bits).
__device__ __forceinline__ uint32_t byte_bf(uint32_t v) { return f2bf(((float)(int)v - 128.f) * 0.015625f); }

// Network input value pair (channel 0, channel 1) of input pixel (oy, ox) of
// the MB at picture (X, Y): padded planes, pixel (0,0) at PCME_PAD.
typedef const __attribute__((address_space(1))) uint8_t gu8;

__device__ __forceinline__ uint32_t input_px(const uint8_t *src_, const uint8_t *ref_, int stride, int X, int Y,
                                             int oy, int ox) {
  gu8 *src = (gu8 *)src_, *ref = (gu8 *)ref_;
  gu8 *r = ref + (size_t)(Y - 16 + 2 * oy + PCME_PAD) * stride + (X - 16 + 2 * ox + PCME_PAD);
  const uint32_t a = (uint32_t)r[0] + r[1] + r[stride] + r[stride + 1];
  uint32_t lo = byte_bf((a + 2) >> 2), hi = 0;
  if ((oy >= 8) & (oy < 16) & (ox >= 8) & (ox < 16)) {
    gu8 *s = src + (size_t)(Y + 2 * (oy - 8) + PCME_PAD) * stride + (X + 2 * (ox - 8) + PCME_PAD);
    const uint32_t b = (uint32_t)s[0] + s[1] + s[stride] + s[stride + 1];
    hi = byte_bf((b + 2) >> 2);
  }
  return lo | (hi << 16);
}

// Position of 16-byte channel chunk q (channels 8q..8q+7) inside pixel p of
// a [pixel][32] bf16 map.  A 16-byte LDS read of a wave is served 32 lanes
// (512 bytes of banks) at a time; the B-fragment lanes of one half-wave read
// chunks q and q^1 of 16 pixels, so pixels p and p+8 -- 512 bytes apart --
// would share banks.  The inference kernel therefore stores the chunks of
// pixels with bit 3 set in the other half of the pixel (q ^ 2); the training
// kernels keep the plain order.
template <bool SW>
__device__ __forceinline__ int chunk_at(int p, int q) {
  return SW ? (q ^ (((p >> 3) & 1) << 1)) : q;
}

// B fragment of a forward conv: 8 K values of output pixel (oy, ox) for
// K chunk `chunk` and lane group g.  x = one MB's input [HIN*HIN][CIN].
template <int CIN, int HIN, int S, bool SW = false>
__device__ __forceinline__ u4 bfrag_fwd(const uint16_t *x, int oy, int ox, int chunk, int g) {
  u4 v = {0u, 0u, 0u, 0u};
  if constexpr (CIN == CH) {
    const int ty = chunk / 3, tx = chunk - 3 * ty;
    const int iy = oy * S + ty - 1, ix = ox * S + tx - 1;
    const int pi = iy * HIN + ix;
    if ((iy >= 0) & (iy < HIN) & (ix >= 0) & (ix < HIN)) v = *(const u4 *)(x + pi * CH + 8 * chunk_at<SW>(pi, g));
  } else {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int tap = 4 * g + q;
      const int ty = tap / 3, tx = tap - 3 * ty;
      const int iy = oy * S + ty - 1, ix = ox * S + tx - 1;
      uint32_t d = 0u;
      if ((tap < 9) & (iy >= 0) & (iy < HIN) & (ix >= 0) & (ix < HIN)) d = *(const uint32_t *)(x + (iy * HIN + ix) * 2);
      v[q] = d;
    }
  }
  return v;
}

// A fragment of the weights [32][KP] (bf16, rows PITCH apart) for M tile mt.
template <int KP, int PITCH = KP>
__device__ __forceinline__ u4 afrag(const uint16_t *w, int mt, int chunk, int l) {
  return *(const u4 *)(w + (mt * 16 + (l & 15)) * PITCH + chunk * 32 + 8 * (l >> 4));
}

__device__ __forceinline__ f4 bias4(const float *b, int mt, int g) {
  return f4{b[mt * 16 + 4 * g], b[mt * 16 + 4 * g + 1], b[mt * 16 + 4 * g + 2], b[mt * 16 + 4 * g + 3]};
}

// Output pixel tile of a forward conv for one MB: accumulators of both M
// tiles for the lane's pixel (oy, ox).
template <int CIN, int HIN, int S, bool SW = false, int PITCH = (CIN == CH ? KC : K1)>
__device__ __forceinline__ void conv_tile(const uint16_t *x, const uint16_t *w, const float *b, int oy, int ox,
                                          int l, f4 &acc0, f4 &acc1) {
  constexpr int KP = CIN == CH ? KC : K1, NCH = CIN == CH ? 9 : 1;
  const int g = l >> 4;
  acc0 = bias4(b, 0, g);
  acc1 = bias4(b, 1, g);
#pragma unroll
  for (int c = 0; c < NCH; c++) {
    const u4 B = bfrag_fwd<CIN, HIN, S, SW>(x, oy, ox, c, g);
    acc0 = mfma(afrag<KP, PITCH>(w, 0, c, l), B, acc0);
    acc1 = mfma(afrag<KP, PITCH>(w, 1, c, l), B, acc1);
  }
}

// ReLU + bf16 of the lane's 8 outputs (channels 4g..4g+3 and 16+4g..) of
// pixel p; y = the map's base.
template <bool SW = false>
__device__ __forceinline__ void store_bf(uint16_t *y, int p, int g, f4 a0, f4 a1) {
  a0 = relu4(a0);
  a1 = relu4(a1);
  uint16_t *px = y + p * CH;
  *(u2 *)(px + 8 * chunk_at<SW>(p, g >> 1) + 4 * (g & 1)) = u2{pack2(a0[0], a0[1]), pack2(a0[2], a0[3])};
  *(u2 *)(px + 8 * chunk_at<SW>(p, 2 + (g >> 1)) + 4 * (g & 1)) = u2{pack2(a1[0], a1[1]), pack2(a1[2], a1[3])};
}

// ------------------------------------------------------------- training

__global__ void __launch_bounds__(256) inputs_kernel(const uint8_t *__restrict__ src, const uint8_t *__restrict__ ref,
                                                     int mbw, int mbh, const int32_t *__restrict__ mb, int n,
                                                     uint32_t *__restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)n * (IN * IN)) return;
  const int i = (int)(t / (IN * IN)), p = (int)(t - (int64_t)i * (IN * IN));
  const int m = mb[i];
  if (m < 0 || m >= mbw * mbh) return;
  const int my = m / mbw, mx = m - my * mbw;
  const int stride = 16 * mbw + 2 * PCME_PAD;
  x[t] = input_px(src, ref, stride, 16 * mx, 16 * my, p / IN, p % IN);
}

// Gather a batch: xb[i] = x[idx[i]], labels likewise.
__global__ void __launch_bounds__(256) gather_kernel(const u4 *__restrict__ x, const int32_t *__restrict__ mv,
                                                     const int32_t *__restrict__ pt, const int32_t *__restrict__ idx,
                                                     int n, u4 *__restrict__ xb, int32_t *__restrict__ mvb,
                                                     int32_t *__restrict__ ptb) {
  constexpr int V = INPUT * 2 / 16;  // u4 per input (144)
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n * V) return;
  const int i = t / V, k = t - i * V;
  const int s = idx[i];
  xb[t] = x[(size_t)s * V + k];
  if (k == 0) {
    mvb[i] = mv[s];
    ptb[i] = pt[s];
  }
}

// Forward conv over a batch: one wave per 16 output pixels.
template <int CIN, int HIN, int S, int HOUT, bool F32OUT>
__global__ void __launch_bounds__(256) conv_fwd_kernel(const uint16_t *__restrict__ x, const uint16_t *__restrict__ w,
                                                       const float *__restrict__ b, void *__restrict__ y, int npix) {
  const int l = (int)(threadIdx.x & 63u), g = l >> 4;
  const int tile = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (tile * 16 >= npix) return;
  const int gp = tile * 16 + (l & 15);
  const int bi = gp / (HOUT * HOUT), p = gp - bi * (HOUT * HOUT);
  const int oy = p / HOUT, ox = p - oy * HOUT;
  f4 a0, a1;
  conv_tile<CIN, HIN, S>(x + (size_t)bi * HIN * HIN * CIN, w, b, oy, ox, l, a0, a1);
  if constexpr (F32OUT) {
    float *yo = (float *)y + (size_t)gp * CH;
    *(f4 *)(yo + 4 * g) = relu4(a0);
    *(f4 *)(yo + 16 + 4 * g) = relu4(a1);
  } else {
    store_bf((uint16_t *)y + (size_t)gp * CH, 0, g, a0, a1);
  }
}

// Global average pool of conv4 (fp32, pixels in order) -- shared with the
// inference kernel.
__device__ __forceinline__ float pool9(const float *a4, int co) {
  float s = a4[co];
#pragma unroll
  for (int p = 1; p < 9; p++) s += a4[p * CH + co];
  return s * (1.0f / 9.0f);
}

// fc logit of class c: fp32 fma chain over k = 0..31 from the bias; the
// row's 32 bf16 weights are read as 16 dwords (scalar loads where the row is
// wave-uniform).
__device__ __forceinline__ float fc_logit(const float *gv, const uint16_t *wfrow, float bias) {
  const uint32_t *w = (const uint32_t *)wfrow;
  float acc = bias;
#pragma unroll
  for (int k = 0; k < CH; k += 2) {
    const uint32_t d = w[k >> 1];
    acc = __builtin_fmaf(gv[k], __builtin_bit_cast(float, d << 16), acc);
    acc = __builtin_fmaf(gv[k + 1], __builtin_bit_cast(float, d & 0xFFFF0000u), acc);
  }
  return acc;
}

__global__ void __launch_bounds__(256) pool_kernel(const float *__restrict__ a4, float *__restrict__ gv, int n) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n * CH) return;
  const int i = t / CH, co = t - i * CH;
  gv[t] = pool9(a4 + (size_t)i * 9 * CH, co);
}

__global__ void __launch_bounds__(256) fc_kernel(const float *__restrict__ gv, const uint16_t *__restrict__ wfb,
                                                 const float *__restrict__ bf, float *__restrict__ logits, int n) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n * NCLS) return;
  const int i = t / NCLS, c = t - i * NCLS;
  logits[t] = fc_logit(gv + (size_t)i * CH, wfb + (size_t)c * CH, bf[c]);
}

// Softmax cross-entropy of both heads, one thread per MB: loss terms and
// d(mean loss)/d(logits).
__global__ void __launch_bounds__(64) ce_kernel(const float *__restrict__ logits, const int32_t *__restrict__ mv,
                                                const int32_t *__restrict__ pt, int n, float *__restrict__ dlog,
                                                float *__restrict__ lossv) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  const float *z = logits + (size_t)i * NCLS;
  float *d = dlog + (size_t)i * NCLS;
  const float inv = 1.0f / (float)n;
  float m = z[0];
  for (int c = 1; c < MVC; c++) m = fmaxf(m, z[c]);
  float s = 0.f;
  for (int c = 0; c < MVC; c++) s += expf(z[c] - m);
  const int lm = mv[i];
  for (int c = 0; c < MVC; c++) d[c] = (expf(z[c] - m) / s - (c == lm ? 1.f : 0.f)) * inv;
  const float lmv = logf(s) + m - z[lm];
  float m4 = z[MVC];
  for (int c = 1; c < 4; c++) m4 = fmaxf(m4, z[MVC + c]);
  float s4 = 0.f;
  for (int c = 0; c < 4; c++) s4 += expf(z[MVC + c] - m4);
  const int lp = pt[i];
  for (int c = 0; c < 4; c++) d[MVC + c] = (expf(z[MVC + c] - m4) / s4 - (c == lp ? 1.f : 0.f)) * inv;
  const float lpt = logf(s4) + m4 - z[MVC + lp];
  lossv[2 * i] = lmv;
  lossv[2 * i + 1] = lpt;
}

// Mean losses (sequential, deterministic): out = (total, motion, partition).
__global__ void loss_kernel(const float *__restrict__ lossv, int n, float *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float a = 0.f, b = 0.f;
  for (int i = 0; i < n; i++) {
    a += lossv[2 * i];
    b += lossv[2 * i + 1];
  }
  a /= (float)n;
  b /= (float)n;
  out[0] = a + b;
  out[1] = a;
  out[2] = b;
}

// fc backward: dg[i][k] = sum_c dlog[i][c] * wf[c][k]; then the pool and
// conv4 ReLU: dy4[i][p][co] = bf16(a4 > 0 ? dg[i][co] / 9 : 0).
__global__ void __launch_bounds__(256) fc_bwd_kernel(const float *__restrict__ dlog, const uint16_t *__restrict__ wfb,
                                                     const float *__restrict__ a4, uint16_t *__restrict__ dy4, int n) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= n * CH) return;
  const int i = t / CH, k = t - i * CH;
  const float *d = dlog + (size_t)i * NCLS;
  float s = 0.f;
  for (int c = 0; c < NCLS; c++) s = __builtin_fmaf(d[c], bf2f(wfb[c * CH + k]), s);
  const float gk = s * (1.0f / 9.0f);
#pragma unroll
  for (int p = 0; p < 9; p++) {
    const size_t o = ((size_t)i * 9 + p) * CH + k;
    dy4[o] = f2bf(a4[o] > 0.f ? gk : 0.f);
  }
}

// fc weight/bias gradients: grad[OWF + c*32 + k] = sum_i dlog[i][c] g[i][k],
// grad[OBF + c] = sum_i dlog[i][c] (k == 32).
__global__ void __launch_bounds__(256) fc_wgrad_kernel(const float *__restrict__ dlog, const float *__restrict__ gv,
                                                       int n, float *__restrict__ grad) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= NCLS * (CH + 1)) return;
  const int c = t / (CH + 1), k = t - c * (CH + 1);
  float s = 0.f;
  if (k < CH) {
    for (int i = 0; i < n; i++) s = __builtin_fmaf(dlog[(size_t)i * NCLS + c], gv[(size_t)i * CH + k], s);
    grad[OWF + c * CH + k] = s;
  } else {
    for (int i = 0; i < n; i++) s += dlog[(size_t)i * NCLS + c];
    grad[OBF + c] = s;
  }
}

// Data gradient of a 32 -> 32 conv: dx[q][ci] = sum_{tap, co} W[co][tap][ci]
// dy[p(q, tap)][co], masked by the layer input's ReLU (x > 0), as bf16.
// One wave per 16 input pixels.
template <int HIN, int S, int HOUT>
__global__ void __launch_bounds__(256) conv_bwd_data_kernel(const uint16_t *__restrict__ dy,
                                                            const uint16_t *__restrict__ wt,
                                                            const uint16_t *__restrict__ x,
                                                            uint16_t *__restrict__ dx, int npix) {
  const int l = (int)(threadIdx.x & 63u), g = l >> 4;
  const int tile = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (tile * 16 >= npix) return;
  const int gq = tile * 16 + (l & 15);
  const int bi = gq / (HIN * HIN), q = gq - bi * (HIN * HIN);
  const int iy = q / HIN, ix = q - iy * HIN;
  const uint16_t *dyb = dy + (size_t)bi * HOUT * HOUT * CH;
  f4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int tap = 0; tap < 9; tap++) {
    const int ty = tap / 3, tx = tap - 3 * ty;
    const int ny = iy + 1 - ty, nx = ix + 1 - tx;
    u4 B = {0u, 0u, 0u, 0u};
    if ((ny >= 0) & (nx >= 0) & (ny % S == 0) & (nx % S == 0) & (ny / S < HOUT) & (nx / S < HOUT))
      B = *(const u4 *)(dyb + ((ny / S) * HOUT + nx / S) * CH + 8 * g);
    a0 = mfma(*(const u4 *)(wt + ((l & 15) * 9 + tap) * CH + 8 * g), B, a0);
    a1 = mfma(*(const u4 *)(wt + ((16 + (l & 15)) * 9 + tap) * CH + 8 * g), B, a1);
  }
  const uint16_t *xr = x + (size_t)gq * CH;
  uint16_t *o = dx + (size_t)gq * CH;
  float r0[4], r1[4];
#pragma unroll
  for (int v = 0; v < 4; v++) {
    r0[v] = bf2f(xr[4 * g + v]) > 0.f ? a0[v] : 0.f;
    r1[v] = bf2f(xr[16 + 4 * g + v]) > 0.f ? a1[v] : 0.f;
  }
  *(u2 *)(o + 4 * g) = u2{pack2(r0[0], r0[1]), pack2(r0[2], r0[3])};
  *(u2 *)(o + 16 + 4 * g) = u2{pack2(r1[0], r1[1]), pack2(r1[2], r1[3])};
}

// Weight gradient partials of a conv: part[chunk][co][kk] = sum over the
// chunk's output pixels of dy[px][co] * col[px][kk], where col[px][kk] is the
// forward's packed K value (kk = tap*CIN + ci) and col[px][KP] = 1 (bias).
// grid (KPP/16 column tiles, chunks), one wave per block.
template <int CIN, int HIN, int S, int HOUT>
__global__ void __launch_bounds__(64) conv_bwd_w_kernel(const uint16_t *__restrict__ dy,
                                                        const uint16_t *__restrict__ x, int npix,
                                                        float *__restrict__ part) {
  constexpr int KP = CIN == CH ? KC : 18, KPP = CIN == CH ? 304 : 32;
  const int l = (int)threadIdx.x, g = l >> 4;
  const int kk = blockIdx.x * 16 + (l & 15);
  const int chunk = blockIdx.y;
  const int tap = kk < KP ? kk / CIN : 0, ci = kk < KP ? kk - tap * CIN : 0;
  const int ty = tap / 3, tx = tap - 3 * ty;
  const int p0 = chunk * CHUNK, p1 = min(npix, p0 + CHUNK);
  f4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = p0; k0 < p1; k0 += 32) {
    u4 A0, A1, B;
#pragma unroll
    for (int h = 0; h < 4; h++) {
      uint32_t d0[2], d1[2], e[2];
#pragma unroll
      for (int j2 = 0; j2 < 2; j2++) {
        const int px = k0 + 8 * g + 2 * h + j2;
        d0[j2] = dy[(size_t)px * CH + (l & 15)];
        d1[j2] = dy[(size_t)px * CH + 16 + (l & 15)];
        const int bi = px / (HOUT * HOUT), p = px - bi * (HOUT * HOUT);
        const int oy = p / HOUT, ox = p - oy * HOUT;
        const int iy = oy * S + ty - 1, ix = ox * S + tx - 1;
        uint32_t v = 0u;
        if (kk < KP) {
          if ((iy >= 0) & (iy < HIN) & (ix >= 0) & (ix < HIN))
            v = x[((size_t)bi * HIN * HIN + iy * HIN + ix) * CIN + ci];
        } else if (kk == KP) {
          v = 0x3F80u;  // 1.0
        }
        e[j2] = v;
      }
      A0[h] = d0[0] | (d0[1] << 16);
      A1[h] = d1[0] | (d1[1] << 16);
      B[h] = e[0] | (e[1] << 16);
    }
    a0 = mfma(A0, B, a0);
    a1 = mfma(A1, B, a1);
  }
  float *o = part + (size_t)chunk * CH * KPP;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    o[(4 * g + v) * KPP + kk] = a0[v];
    o[(16 + 4 * g + v) * KPP + kk] = a1[v];
  }
}

// Sum the partials in chunk order into the flat gradient (weights [co][kk],
// bias from column KP).
__global__ void __launch_bounds__(256) reduce_w_kernel(const float *__restrict__ part, int nchunks, int KP, int KPP,
                                                       float *__restrict__ gw, float *__restrict__ gb) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= CH * (KP + 1)) return;
  const int co = t / (KP + 1), kk = t - co * (KP + 1);
  float s = 0.f;
  for (int c = 0; c < nchunks; c++) s += part[((size_t)c * CH + co) * KPP + kk];
  if (kk < KP)
    gw[co * KP + kk] = s;
  else
    gb[co] = s;
}

// bf16 copies of master parameter i (conv weights packed for the forward
// A operand and transposed for the data gradient; fc weights).
__device__ __forceinline__ void write_copies(int i, float w, uint16_t *w1b, uint16_t *w2b, uint16_t *w3b,
                                             uint16_t *w4b, uint16_t *w2t, uint16_t *w3t, uint16_t *w4t,
                                             uint16_t *wfb) {
  const uint16_t h = f2bf(w);
  if (i < OB1) {
    w1b[(i / 18) * K1 + i % 18] = h;
  } else if ((i >= OW2 && i < OB2) || (i >= OW3 && i < OB3) || (i >= OW4 && i < OB4)) {
    const int base = i < OB2 ? OW2 : (i < OB3 ? OW3 : OW4);
    uint16_t *wb = i < OB2 ? w2b : (i < OB3 ? w3b : w4b);
    uint16_t *wt = i < OB2 ? w2t : (i < OB3 ? w3t : w4t);
    const int j = i - base, co = j / KC, r = j - co * KC, tap = r / CH, ci = r - tap * CH;
    wb[j] = h;
    wt[(ci * 9 + tap) * CH + co] = h;
  } else if (i >= OWF && i < OBF) {
    wfb[i - OWF] = h;
  }
}

__global__ void __launch_bounds__(256) adam_kernel(float *__restrict__ w, const float *__restrict__ grad,
                                                   float *__restrict__ m, float *__restrict__ v, float lr, float bc1,
                                                   float bc2, uint16_t *w1b, uint16_t *w2b, uint16_t *w3b,
                                                   uint16_t *w4b, uint16_t *w2t, uint16_t *w3t, uint16_t *w4t,
                                                   uint16_t *wfb) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= NPARAM) return;
  const float gr = grad[i];
  const float mi = 0.9f * m[i] + 0.1f * gr;
  const float vi = 0.999f * v[i] + 0.001f * gr * gr;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
  w[i] = wi;
  write_copies(i, wi, w1b, w2b, w3b, w4b, w2t, w3t, w4t, wfb);
}

__global__ void __launch_bounds__(256) refresh_kernel(const float *__restrict__ w, uint16_t *w1b, uint16_t *w2b,
                                                      uint16_t *w3b, uint16_t *w4b, uint16_t *w2t, uint16_t *w3t,
                                                      uint16_t *w4t, uint16_t *wfb) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= NPARAM) return;
  write_copies(i, w[i], w1b, w2b, w3b, w4b, w2t, w3t, w4t, wfb);
}

// ------------------------------------------------------------ inference

struct NetJob {
  const uint8_t *src, *ref;
  pcme_hint *hint;
  int mbw, mbh;
  int64_t mb0;  // first global MB index of the job
};

// Inference runs in two kernels.  conv_kernel: each wave takes one MB at
// a time through conv1-4 and the pool (8 waves per block share the conv
// weights in LDS; each wave owns two activation regions, R0: input / conv2
// out / conv4 out, R1: conv1 out / conv3 out) and writes the 32 pooled
// values.  heads_kernel: one lane per MB runs the fc layer (the fc weights
// are wave-uniform: scalar loads), an online softmax of each head and a
// running top-k of the motion logits, and writes the hint.
constexpr int CW = 8;                           // waves per conv block
constexpr int KCP = KC + 16;                    // LDS pitch of a conv2-4 weight row: 608 bytes put the
                                                // 16 rows of an A fragment on distinct 32-byte bank slots
constexpr int L_W1 = 0;                         // u16 [32][32]
constexpr int L_W2 = L_W1 + CH * K1 * 2;        // u16 [32][KCP]
constexpr int L_W3 = L_W2 + CH * KCP * 2;
constexpr int L_W4 = L_W3 + CH * KCP * 2;
constexpr int L_B = L_W4 + CH * KCP * 2;        // f32 b1..b4 [4][32]
constexpr int L_ZERO = L_B + 4 * CH * 4;         // 1 KB of zeros: out-of-bounds taps (one slot per lane)
constexpr int L_WAVE = L_ZERO + 1024;
constexpr int R0_B = IN * IN * 2 * 2;           // 2304 (>= 36*32*2 and 9*32*4)
constexpr int R1_B = 12 * (12 * 64 + 64);       // 9984: conv1's map with padded rows (map_off<MAP_A>)
constexpr int WAVE_B = R0_B + R1_B;
constexpr int CONV_LDS = L_WAVE + CW * WAVE_B;
static_assert(CONV_LDS <= 160 * 1024, "LDS budget");
// Bank slots (32 bytes of a 256-byte line) are computed from offsets within
// 256-aligned regions.
static_assert(L_ZERO % 256 == 0 && L_WAVE % 256 == 0 && R0_B % 256 == 0 && WAVE_B % 256 == 0, "256-aligned regions");

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ uint16_t q16(float p) {
  const float s = p * 65535.f + 0.5f;
  return (uint16_t)(s >= 65535.f ? 65535u : (s <= 0.f ? 0u : (uint32_t)s));
}

// bf16 bits of (v - 128) / 64 for a byte v held in bits [8k, 8k+8) .. of a
// 16-bit lane: exact (7 significant bits), the low half of the fp32 is zero.
__device__ __forceinline__ uint32_t ds_bf(uint32_t v) {
  return __builtin_bit_cast(uint32_t, (float)((int)v - 128) * 0.015625f) >> 16;
}

// Two input pixels (2c, 2c+1) of input row r from two dwords of each of two
// rows: out = ((a0 + a1 + b0 + b1 + 2) >> 2 per byte pair).
__device__ __forceinline__ uint32_t ds_pair(uint32_t a, uint32_t b) {
  const uint32_t s = (a & 0x00FF00FFu) + ((a >> 8) & 0x00FF00FFu) + (b & 0x00FF00FFu) + ((b >> 8) & 0x00FF00FFu) +
                     0x00020002u;
  return ((s >> 2) & 0x3FFFu) | (((s >> 18) & 0x3FFFu) << 16);
}

// Placement of the fused kernel's activation maps in LDS.  A map pixel is
// 32 bf16 channels in four 16-byte chunks; a B-fragment read takes one chunk
// of 16 pixels per lane group, 8 lanes per LDS cycle, and 8 lanes conflict
// unless they hit 8 distinct 32-byte slots of a 256-byte line.  So each map
// has the placement its reader needs: conv1's 12x12 output (read at stride
// 2) and conv3's 6x6 output (stride 2) store each row's even columns, then
// its odd ones, rows padded by 64 bytes, chunks XOR-swapped on bit 3 of the
// position; conv2's 6x6 output (read at stride 1) is raster with the chunk
// swap on bit 2.  Out-of-bounds taps read zeros in the slots their lane
// group leaves free (tab_group).
template <int HIN, bool SPLIT, int SWB, int RP>
__device__ __forceinline__ int mpos(int iy, int ix, int q) {
  const int x = SPLIT ? ((ix & 1) ? (HIN + 1) / 2 + (ix >> 1) : (ix >> 1)) : ix;
  const int m = iy * HIN + x;
  return iy * (HIN * 64 + RP) + x * 64 + 16 * (q ^ (((m >> SWB) & 1) << 1));
}
enum { MAP_A = 0, MAP_B = 1, MAP_C = 2 };  // conv1 out (R1), conv2 out (R0), conv3 out (R1)
template <int MAP>
__device__ __forceinline__ int map_off(int k, int q) {
  if constexpr (MAP == MAP_A) return mpos<12, true, 3, 64>(k / 12, k % 12, q);
  else if constexpr (MAP == MAP_B) return mpos<6, false, 2, 0>(k / 6, k % 6, q);
  else return mpos<6, true, 3, 64>(k / 6, k % 6, q);
}

// Per-lane read table of the fused convolutions: the LDS byte offset / 4 of
// every B-fragment read of one MB (the same for every MB a wave runs), two
// u16 entries per dword.  Entries: conv1 (tile t, tap group q) at 4t+q,
// conv2 (t, chunk c) at T_C2 + 9t + c, conv3 at T_C3 + 9t + c, conv4 (c) at
// T_C4 + c.
constexpr int T_C2 = 36, T_C3 = 63, T_C4 = 90, T_N = 99, T_W = (T_N + 1) / 2;

__device__ __forceinline__ uint32_t tab(const uint32_t *tb, int i) {
  return ((tb[i >> 1] >> ((i & 1) * 16)) & 0xFFFFu) << 2;
}

// Entry of lane l for the read of tap (ty, tx) by output pixel tile t of a
// HOUT x HOUT output (NPIX pixels) at stride S from the map MAP (HIN x HIN)
// at `base`.  The lane's 8-lane group is evaluated whole: an out-of-bounds
// lane takes the r-th 32-byte slot of the zero line that no in-bounds lane of
// the group uses, r = its rank among the group's out-of-bounds lanes (there
// are at least as many free slots as out-of-bounds lanes).
template <int MAP, int HIN, int HOUT, int S, int NPIX>
__device__ __forceinline__ uint32_t tab_group(int base, int t, int ty, int tx, int l) {
  const int g = l >> 4, l0 = l & ~7;
  uint32_t used = 0u;
  int rank = 0, mine = 0;
  bool oob = false;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int p = t * 16 + ((l0 + j) & 15), oy = p / HOUT, ox = p - HOUT * (p / HOUT);
    const int iy = S * oy + ty - 1, ix = S * ox + tx - 1;
    const bool ok = (p < NPIX) & (iy >= 0) & (iy < HIN) & (ix >= 0) & (ix < HIN);
    const int a = ok ? base + map_off<MAP>(iy * HIN + ix, g) : 0;
    if (ok) used |= 1u << ((a & 255) >> 5);
    if (l0 + j == l) {
      mine = a;
      oob = !ok;
    } else if (!ok && l0 + j < l) {
      rank++;
    }
  }
  if (!oob) return (uint32_t)mine >> 2;
  int slot = 0;
#pragma unroll
  for (int s = 0, n = 0; s < 8; s++) {
    const bool f = ((used >> s) & 1u) == 0u;
    if (f && n == rank) slot = s;
    n += f ? 1 : 0;
  }
  return (uint32_t)(L_ZERO + 32 * slot) >> 2;
}

__device__ __forceinline__ void build_table(int l, int r0, int r1, uint32_t *tb) {
  uint32_t e[T_W * 2];
#pragma unroll
  for (int t = 0; t < 9; t++) {
    const int p = t * 16 + (l & 15), oy = p / 12, ox = p - 12 * (p / 12);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int tap = 4 * (l >> 4) + q, ty = tap / 3, tx = tap - 3 * (tap / 3);
      const int iy = 2 * oy + ty - 1, ix = 2 * ox + tx - 1;
      const bool ok = (tap < 9) & (iy >= 0) & (iy < IN) & (ix >= 0) & (ix < IN);
      e[4 * t + q] = (uint32_t)(ok ? r0 + 4 * (iy * IN + ix) : L_ZERO + 4 * l) >> 2;
    }
  }
#pragma unroll
  for (int t = 0; t < 3; t++) {
#pragma unroll
    for (int c = 0; c < 9; c++) {
      const int ty = c / 3, tx = c - 3 * (c / 3);
      e[T_C2 + 9 * t + c] = tab_group<MAP_A, 12, 6, 2, 36>(r1, t, ty, tx, l);
      e[T_C3 + 9 * t + c] = tab_group<MAP_B, 6, 6, 1, 36>(r0, t, ty, tx, l);
    }
  }
#pragma unroll
  for (int c = 0; c < 9; c++) {
    const int ty = c / 3, tx = c - 3 * (c / 3);
    e[T_C4 + c] = tab_group<MAP_C, 6, 3, 2, 9>(r1, 0, ty, tx, l);
  }
  e[T_N] = 0u;
#pragma unroll
  for (int i = 0; i < T_W; i++) tb[i] = e[2 * i] | (e[2 * i + 1] << 16);
}

// ReLU + bf16 of the lane's 8 outputs (channels 4g..4g+3 and 16+4g..) of a
// pixel stored at index k of a fused-kernel map (swizzled chunks).
template <int MAP>
__device__ __forceinline__ void store_map(uint16_t *y, int k, int g, f4 a0, f4 a1) {
  a0 = relu4(a0);
  a1 = relu4(a1);
  uint8_t *b = (uint8_t *)y;
  *(u2 *)(b + map_off<MAP>(k, g >> 1) + 8 * (g & 1)) = u2{pack2(a0[0], a0[1]), pack2(a0[2], a0[3])};
  *(u2 *)(b + map_off<MAP>(k, 2 + (g >> 1)) + 8 * (g & 1)) = u2{pack2(a1[0], a1[1]), pack2(a1[2], a1[3])};
}

// conv2 / conv3 of one MB: 36 output pixels in three tiles, reads through
// the table from entry T0, writes Y ([pixel][32] bf16, swizzled).
template <int T0, int MAP>
__device__ __forceinline__ void conv_mid(const uint8_t *lds, const uint32_t *tb, uint16_t *Y, const uint16_t *W,
                                         const float *b, int l) {
  const int g = l >> 4;
  f4 acc[3][2];
#pragma unroll
  for (int t = 0; t < 3; t++) {
    acc[t][0] = bias4(b, 0, g);
    acc[t][1] = bias4(b, 1, g);
  }
#pragma unroll
  for (int c = 0; c < 9; c++) {
    const u4 A0 = afrag<KC, KCP>(W, 0, c, l), A1 = afrag<KC, KCP>(W, 1, c, l);
#pragma unroll
    for (int t = 0; t < 3; t++) {
      const u4 B = *(const u4 *)(lds + tab(tb, T0 + 9 * t + c));
      acc[t][0] = mfma(A0, B, acc[t][0]);
      acc[t][1] = mfma(A1, B, acc[t][1]);
    }
  }
  wave_sync();  // every read of the input precedes the writes of Y
#pragma unroll
  for (int t = 0; t < 3; t++) {
    const int p = t * 16 + (l & 15);
    if (p < 36) store_map<MAP>(Y, p, g, acc[t][0], acc[t][1]);
  }
  wave_sync();
}

typedef const __attribute__((address_space(1))) uint32_t gu32c;

// Raw bytes of one MB's network input: item (r, c) of the lane = input row r
// (0..23), dword c (0..11) of the 48x48 reference window (input pixels 2c,
// 2c+1), rows 2r and 2r+1; the source MB likewise where the item lies in
// rows/cols 8..15.
struct InputRegs {
  uint32_t ra[5], rb[5], sa[5], sb[5];
};

__device__ __forceinline__ void load_input(const NetJob &J, int li, int l, InputRegs &in) {
  const int my = li / J.mbw, mx = li - my * J.mbw;
  const int sdw = (16 * J.mbw + 2 * PCME_PAD) >> 2;  // stride in dwords
  gu32c *ref = (gu32c *)J.ref;
  gu32c *src = (gu32c *)J.src;
  const int64_t rb = (int64_t)(16 * my - 16 + PCME_PAD) * sdw + ((16 * mx - 16 + PCME_PAD) >> 2);
  const int64_t sb = (int64_t)(16 * my + PCME_PAD) * sdw + ((16 * mx + PCME_PAD) >> 2);
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const int it = min(l + 64 * q, 287);
    const int r = it / 12, c = it - 12 * (it / 12);
    in.ra[q] = ref[rb + (int64_t)(2 * r) * sdw + c];
    in.rb[q] = ref[rb + (int64_t)(2 * r + 1) * sdw + c];
    const int rs = min(max(r - 8, 0), 7), cs = min(max(c - 4, 0), 3);
    in.sa[q] = src[sb + (int64_t)(2 * rs) * sdw + cs];
    in.sb[q] = src[sb + (int64_t)(2 * rs + 1) * sdw + cs];
  }
}

// The 24x24x2 input map (one dword per pixel: channel 0 | channel 1 << 16).
__device__ __forceinline__ void store_input(const InputRegs &in, int l, uint16_t *R0) {
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const int it = l + 64 * q;
    if (it < 288) {
      const int r = it / 12, c = it - 12 * (it / 12);
      const uint32_t d = ds_pair(in.ra[q], in.rb[q]);
      uint32_t e0 = 0u, e1 = 0u;
      if ((r >= 8) & (r < 16) & (c >= 4) & (c < 8)) {
        const uint32_t e = ds_pair(in.sa[q], in.sb[q]);
        e0 = ds_bf(e & 0xFFFFu) << 16;
        e1 = ds_bf(e >> 16) << 16;
      }
      *(u2 *)(R0 + (r * IN + 2 * c) * 2) = u2{ds_bf(d & 0xFFFFu) | e0, ds_bf(d >> 16) | e1};
    }
  }
}

__global__ void __launch_bounds__(64 * CW) conv_kernel(const NetJob *__restrict__ jobs, int njobs, int64_t total,
                                                       const uint16_t *__restrict__ w1b,
                                                       const uint16_t *__restrict__ w2b,
                                                       const uint16_t *__restrict__ w3b,
                                                       const uint16_t *__restrict__ w4b,
                                                       const float *__restrict__ prm, float *__restrict__ gout) {
  extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
  const int tid = (int)threadIdx.x;
  {
    const u4 *s;
    u4 *d;
    s = (const u4 *)w1b; d = (u4 *)(lds + L_W1);
    for (int i = tid; i < CH * K1 * 2 / 16; i += 64 * CW) d[i] = s[i];
    // conv2-4 rows: 36 u4 of weights each, placed KCP apart
    constexpr int RU = KC * 2 / 16, PU = KCP * 2 / 16;
    for (int i = tid; i < 3 * CH * RU; i += 64 * CW) {
      const int k = i / (CH * RU), r = (i / RU) % CH, j = i % RU;
      const u4 *w = (const u4 *)(k == 0 ? w2b : (k == 1 ? w3b : w4b));
      ((u4 *)(lds + L_W2))[k * CH * PU + r * PU + j] = w[r * RU + j];
    }
    float *bb = (float *)(lds + L_B);
    for (int i = tid; i < 4 * CH; i += 64 * CW) {
      const int k = i / CH;
      bb[i] = prm[(k == 0 ? OB1 : k == 1 ? OB2 : k == 2 ? OB3 : OB4) + (i - k * CH)];
    }
  }
  __syncthreads();
  const uint16_t *W1 = (const uint16_t *)(lds + L_W1);
  const uint16_t *W2 = (const uint16_t *)(lds + L_W2);
  const uint16_t *W3 = (const uint16_t *)(lds + L_W3);
  const uint16_t *W4 = (const uint16_t *)(lds + L_W4);
  const float *BB = (const float *)(lds + L_B);
  const int wv = tid >> 6, l = tid & 63, g = l >> 4;
  uint8_t *wbase = lds + L_WAVE + wv * WAVE_B;
  uint16_t *R0 = (uint16_t *)wbase;
  uint16_t *R1 = (uint16_t *)(wbase + R0_B);
  if (tid < 256) ((uint32_t *)(lds + L_ZERO))[tid] = 0u;
  uint32_t tb[T_W];
  build_table(l, L_WAVE + wv * WAVE_B, L_WAVE + wv * WAVE_B + R0_B, tb);
  __syncthreads();  // the zero block
  const int64_t nw = (int64_t)gridDim.x * CW;
  // The input loads of the next MB are issued before the current MB's
  // convolutions (software pipelining over the wave's MB sequence).
  InputRegs in;
  int jb = 0;
  int64_t m = (int64_t)blockIdx.x * CW + wv;
  if (m < total) {
    while (jb + 1 < njobs && m >= jobs[jb + 1].mb0) jb++;
    load_input(jobs[jb], (int)(m - jobs[jb].mb0), l, in);
  }
  for (; m < total; m += nw) {
    store_input(in, l, R0);
    const int64_t mn = m + nw;
    if (mn < total) {
      while (jb + 1 < njobs && mn >= jobs[jb + 1].mb0) jb++;
      load_input(jobs[jb], (int)(mn - jobs[jb].mb0), l, in);
    }
    wave_sync();
    // conv1: R0 -> R1 [144][32]
    {
      const u4 A0 = afrag<K1>(W1, 0, 0, l), A1 = afrag<K1>(W1, 1, 0, l);
#pragma unroll
      for (int t = 0; t < 9; t++) {
        const int p = t * 16 + (l & 15);
        u4 B;
#pragma unroll
        for (int q = 0; q < 4; q++) B[q] = *(const uint32_t *)(lds + tab(tb, 4 * t + q));
        const f4 a0 = mfma(A0, B, bias4(BB, 0, g));
        const f4 a1 = mfma(A1, B, bias4(BB, 1, g));
        store_map<MAP_A>(R1, p, g, a0, a1);
      }
    }
    wave_sync();
    conv_mid<T_C2, MAP_B>(lds, tb, R0, W2, BB + CH, l);
    conv_mid<T_C3, MAP_C>(lds, tb, R1, W3, BB + 2 * CH, l);
    // conv4: R1 -> fp32 [9][32] in R0, then the pool
    {
      const int p = l & 15;
      f4 a0 = bias4(BB + 3 * CH, 0, g), a1 = bias4(BB + 3 * CH, 1, g);
#pragma unroll
      for (int c = 0; c < 9; c++) {
        const u4 B = *(const u4 *)(lds + tab(tb, T_C4 + c));
        a0 = mfma(afrag<KC, KCP>(W4, 0, c, l), B, a0);
        a1 = mfma(afrag<KC, KCP>(W4, 1, c, l), B, a1);
      }
      float *A4 = (float *)R0;
      if (p < 9) {
        *(f4 *)(A4 + p * CH + 4 * g) = relu4(a0);
        *(f4 *)(A4 + p * CH + 16 + 4 * g) = relu4(a1);
      }
      wave_sync();
      if (l < CH) gout[m * CH + l] = pool9(A4, l);
      wave_sync();
    }
  }
}

// The fc layer, both heads and the hint of one MB per lane.  z is computed
// class by class with the same fma chain as fc_kernel; the motion head keeps
// a running max / sum of exp and the k best (logit desc, class asc: classes
// arrive in ascending order and only a strictly larger logit moves ahead).
__global__ void __launch_bounds__(256) heads_kernel(const NetJob *__restrict__ jobs, int njobs, int64_t total,
                                                    const float *__restrict__ gin, const uint16_t *__restrict__ wfb,
                                                    const float *__restrict__ prm, int ncand, float thr_part,
                                                    float thr_mv) {
  const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (m >= total) return;
  int jb = 0;
  while (jb + 1 < njobs && m >= jobs[jb + 1].mb0) jb++;
  float gv[CH];
#pragma unroll
  for (int k = 0; k < CH; k += 4) {
    const f4 v = *(const f4 *)(gin + m * CH + k);
    gv[k] = v[0]; gv[k + 1] = v[1]; gv[k + 2] = v[2]; gv[k + 3] = v[3];
  }
  const float *bf = prm + OBF;
  float zmax = -INFINITY, ssum = 0.f;
  float tz[8];
  int tc[8];
#pragma unroll
  for (int i = 0; i < 8; i++) { tz[i] = -INFINITY; tc[i] = 0; }
  for (int c = 0; c < MVC; c++) {
    const float z = fc_logit(gv, wfb + c * CH, bf[c]);
    // online softmax: ssum = sum exp(z_i - zmax)
    if (z > zmax) {
      ssum = ssum * expf(zmax - z) + 1.f;
      zmax = z;
    } else {
      ssum += expf(z - zmax);
    }
#pragma unroll
    for (int i = 7; i >= 1; i--) {
      const bool above = z > tz[i - 1];
      const bool here = !above && z > tz[i];
      tz[i] = above ? tz[i - 1] : (here ? z : tz[i]);
      tc[i] = above ? tc[i - 1] : (here ? c : tc[i]);
    }
    const bool h0 = z > tz[0];
    tz[0] = h0 ? z : tz[0];
    tc[0] = h0 ? c : tc[0];
  }
  float q[4];
#pragma unroll
  for (int i = 0; i < 4; i++) q[i] = fc_logit(gv, wfb + (MVC + i) * CH, bf[MVC + i]);
  int pa = 0;
  float qm = q[0];
  if (q[1] > qm) { qm = q[1]; pa = 1; }
  if (q[2] > qm) { qm = q[2]; pa = 2; }
  if (q[3] > qm) { qm = q[3]; pa = 3; }
  const float conf_part = 1.f / (expf(q[0] - qm) + expf(q[1] - qm) + expf(q[2] - qm) + expf(q[3] - qm));
  float mass = 0.f;
  uint32_t cw[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    const bool on = i < ncand;
    mass += on ? expf(tz[i] - zmax) : 0.f;
    const int cy = tc[i] / PCME_NET_MV_SIDE, cx = tc[i] - cy * PCME_NET_MV_SIDE;
    cw[i] = on ? (((uint32_t)(8 * cx - 64) & 0xFFFFu) | ((uint32_t)(8 * cy - 64) << 16)) : 0u;
  }
  const float conf_mv = mass / ssum;
  const bool prune = (conf_part >= thr_part) & (conf_mv >= thr_mv);
  const NetJob &J = jobs[jb];
  __attribute__((address_space(1))) uint32_t *hw =
      (__attribute__((address_space(1))) uint32_t *)(J.hint + (m - J.mb0));
#pragma unroll
  for (int i = 0; i < 8; i++) hw[i] = cw[i];
  hw[8] = (uint32_t)q16(conf_part) | ((uint32_t)q16(conf_mv) << 16);
  hw[9] = (uint32_t)pa | ((uint32_t)ncand << 8) | ((prune ? 1u : 0u) << 16);
}

}  // namespace cnn
}  // namespace pcme

// ------------------------------------------------------------ host side

using namespace pcme::cnn;

namespace {

thread_local char t_err[256];

int nfail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_err, sizeof t_err, fmt, ap);
  va_end(ap);
  pcme::set_error(t_err);
  return code;
}

#define NCHK(x)                                                                  \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) return nfail(PCME_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double gauss(uint64_t &s) {
  const double u1 = ((double)(splitmix(s) >> 11) + 1.0) * (1.0 / 9007199254740993.0);
  const double u2 = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

}  // namespace

struct pcme_net {
  int dev = 0;
  int maxb = 0;
  int64_t step = 0;
  float *master = nullptr, *grad = nullptr, *am = nullptr, *av = nullptr;
  uint16_t *w1b = nullptr, *w2b = nullptr, *w3b = nullptr, *w4b = nullptr;
  uint16_t *w2t = nullptr, *w3t = nullptr, *w4t = nullptr, *wfb = nullptr;
  uint16_t *xb = nullptr, *a1 = nullptr, *a2 = nullptr, *a3 = nullptr;
  uint16_t *dy1 = nullptr, *dy2 = nullptr, *dy3 = nullptr, *dy4 = nullptr;
  float *a4 = nullptr, *gv = nullptr, *logits = nullptr, *dlog = nullptr, *lossv = nullptr;
  int32_t *mvb = nullptr, *ptb = nullptr;
  float *part1 = nullptr, *part2 = nullptr, *part3 = nullptr, *part4 = nullptr;
  NetJob *djobs = nullptr;
  std::vector<NetJob> hjobs;
  int conv_blocks = 0;
  float *gbuf = nullptr;  // pooled vectors of the last inference (fp32 [MBs][32])
  int64_t gcap = 0;
  std::vector<void *> allocs;
};

namespace {

template <typename T>
int nalloc(pcme_net *n, T **p, size_t count) {
  void *q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess)
    return nfail(PCME_ENOMEM, "hipMalloc of %zu bytes failed", count * sizeof(T));
  n->allocs.push_back(q);
  *p = (T *)q;
  return PCME_OK;
}

void net_free(pcme_net *n) {
  if (!n) return;
  for (void *p : n->allocs) (void)hipFree(p);
  if (n->djobs) (void)hipFree(n->djobs);
  if (n->gbuf) (void)hipFree(n->gbuf);
  delete n;
}

int check_dev(pcme_net *n) {
  if (!n) return nfail(PCME_EINVAL, "null net");
  int d = -1;
  NCHK(hipGetDevice(&d));
  if (d != n->dev) return nfail(PCME_EINVAL, "net lives on device %d, current device is %d", n->dev, d);
  return PCME_OK;
}

inline unsigned blocks(int64_t threads, int per) { return (unsigned)((threads + per - 1) / per); }

int refresh(pcme_net *n, hipStream_t s) {
  hipLaunchKernelGGL(refresh_kernel, dim3(blocks(NPARAM, 256)), dim3(256), 0, s, n->master, n->w1b, n->w2b, n->w3b,
                     n->w4b, n->w2t, n->w3t, n->w4t, n->wfb);
  NCHK(hipGetLastError());
  return PCME_OK;
}

int forward(pcme_net *n, const uint16_t *x, int nb, hipStream_t s) {
  const float *P = n->master;
  hipLaunchKernelGGL((conv_fwd_kernel<2, 24, 2, 12, false>), dim3(blocks((int64_t)nb * 144 / 16, 4)), dim3(256), 0,
                     s, x, n->w1b, P + OB1, (void *)n->a1, nb * 144);
  hipLaunchKernelGGL((conv_fwd_kernel<CH, 12, 2, 6, false>), dim3(blocks((int64_t)nb * 36 / 16, 4)), dim3(256), 0,
                     s, n->a1, n->w2b, P + OB2, (void *)n->a2, nb * 36);
  hipLaunchKernelGGL((conv_fwd_kernel<CH, 6, 1, 6, false>), dim3(blocks((int64_t)nb * 36 / 16, 4)), dim3(256), 0,
                     s, n->a2, n->w3b, P + OB3, (void *)n->a3, nb * 36);
  hipLaunchKernelGGL((conv_fwd_kernel<CH, 6, 2, 3, true>), dim3(blocks((int64_t)nb * 9 / 16, 4)), dim3(256), 0, s,
                     n->a3, n->w4b, P + OB4, (void *)n->a4, nb * 9);
  hipLaunchKernelGGL(pool_kernel, dim3(blocks((int64_t)nb * CH, 256)), dim3(256), 0, s, n->a4, n->gv, nb);
  hipLaunchKernelGGL(fc_kernel, dim3(blocks((int64_t)nb * NCLS, 256)), dim3(256), 0, s, n->gv, n->wfb, P + OBF,
                     n->logits, nb);
  NCHK(hipGetLastError());
  return PCME_OK;
}

template <int CIN, int HIN, int S, int HOUT>
int wgrad(const uint16_t *dy, const uint16_t *x, int nb, float *part, float *gw, float *gb, hipStream_t s) {
  constexpr int KP = CIN == CH ? KC : 18, KPP = CIN == CH ? 304 : 32;
  const int npix = nb * HOUT * HOUT;
  const int nch = (npix + CHUNK - 1) / CHUNK;
  hipLaunchKernelGGL((conv_bwd_w_kernel<CIN, HIN, S, HOUT>), dim3(KPP / 16, nch), dim3(64), 0, s, dy, x, npix, part);
  hipLaunchKernelGGL(reduce_w_kernel, dim3(blocks((int64_t)CH * (KP + 1), 256)), dim3(256), 0, s, part, nch, KP,
                     KPP, gw, gb);
  NCHK(hipGetLastError());
  return PCME_OK;
}

int backward(pcme_net *n, const uint16_t *x, int nb, hipStream_t s) {
  float *G = n->grad;
  hipLaunchKernelGGL(fc_bwd_kernel, dim3(blocks((int64_t)nb * CH, 256)), dim3(256), 0, s, n->dlog, n->wfb, n->a4,
                     n->dy4, nb);
  hipLaunchKernelGGL(fc_wgrad_kernel, dim3(blocks((int64_t)NCLS * (CH + 1), 256)), dim3(256), 0, s, n->dlog, n->gv,
                     nb, G);
  NCHK(hipGetLastError());
  int rc;
  if ((rc = wgrad<CH, 6, 2, 3>(n->dy4, n->a3, nb, n->part4, G + OW4, G + OB4, s))) return rc;
  hipLaunchKernelGGL((conv_bwd_data_kernel<6, 2, 3>), dim3(blocks((int64_t)nb * 36 / 16, 4)), dim3(256), 0, s,
                     n->dy4, n->w4t, n->a3, n->dy3, nb * 36);
  if ((rc = wgrad<CH, 6, 1, 6>(n->dy3, n->a2, nb, n->part3, G + OW3, G + OB3, s))) return rc;
  hipLaunchKernelGGL((conv_bwd_data_kernel<6, 1, 6>), dim3(blocks((int64_t)nb * 36 / 16, 4)), dim3(256), 0, s,
                     n->dy3, n->w3t, n->a2, n->dy2, nb * 36);
  if ((rc = wgrad<CH, 12, 2, 6>(n->dy2, n->a1, nb, n->part2, G + OW2, G + OB2, s))) return rc;
  hipLaunchKernelGGL((conv_bwd_data_kernel<12, 2, 6>), dim3(blocks((int64_t)nb * 144 / 16, 4)), dim3(256), 0, s,
                     n->dy2, n->w2t, n->a1, n->dy1, nb * 144);
  if ((rc = wgrad<2, 24, 2, 12>(n->dy1, x, nb, n->part1, G + OW1, G + OB1, s))) return rc;
  NCHK(hipGetLastError());
  return PCME_OK;
}

int grad_impl(pcme_net *n, const uint16_t *x, const int32_t *mv, const int32_t *pt, const int32_t *idx, int nb,
              float *loss_out, hipStream_t s) {
  if (nb < 32 || nb > n->maxb || nb % 32) return nfail(PCME_EINVAL, "batch %d: need 32 <= n <= %d, n %% 32 == 0", nb, n->maxb);
  if (!x || !mv || !pt || !idx || !loss_out) return nfail(PCME_EINVAL, "null argument");
  hipLaunchKernelGGL(gather_kernel, dim3(blocks((int64_t)nb * (INPUT * 2 / 16), 256)), dim3(256), 0, s,
                     (const u4 *)x, mv, pt, idx, nb, (u4 *)n->xb, n->mvb, n->ptb);
  NCHK(hipGetLastError());
  int rc;
  if ((rc = forward(n, n->xb, nb, s))) return rc;
  hipLaunchKernelGGL(ce_kernel, dim3(blocks(nb, 64)), dim3(64), 0, s, n->logits, n->mvb, n->ptb, nb, n->dlog,
                     n->lossv);
  hipLaunchKernelGGL(loss_kernel, dim3(1), dim3(64), 0, s, n->lossv, nb, loss_out);
  NCHK(hipGetLastError());
  return backward(n, n->xb, nb, s);
}

}  // namespace

extern "C" {

int pcme_net_create(uint64_t seed, int32_t max_batch, pcme_net **out) {
  if (!out) return nfail(PCME_EINVAL, "null out");
  *out = nullptr;
  if (max_batch < 32 || max_batch % 32 || max_batch > (1 << 20))
    return nfail(PCME_EINVAL, "max_batch %d: need a multiple of 32 in [32, 2^20]", max_batch);
  if (pcme_device_count() < 1) return nfail(PCME_ENODEV, "no gfx950 device");
  int dev = -1;
  NCHK(hipGetDevice(&dev));
  pcme_net *n = new pcme_net();
  n->dev = dev;
  n->maxb = max_batch;
  const size_t B = (size_t)max_batch;
  int rc = 0;
  float *all = nullptr;
  if ((rc = nalloc(n, &all, 4 * (size_t)NPARAM)) ||
      (rc = nalloc(n, &n->w1b, (size_t)CH * K1)) || (rc = nalloc(n, &n->w2b, (size_t)CH * KC)) ||
      (rc = nalloc(n, &n->w3b, (size_t)CH * KC)) || (rc = nalloc(n, &n->w4b, (size_t)CH * KC)) ||
      (rc = nalloc(n, &n->w2t, (size_t)CH * KC)) || (rc = nalloc(n, &n->w3t, (size_t)CH * KC)) ||
      (rc = nalloc(n, &n->w4t, (size_t)CH * KC)) || (rc = nalloc(n, &n->wfb, (size_t)NCLS_PAD * CH)) ||
      (rc = nalloc(n, &n->xb, B * INPUT)) || (rc = nalloc(n, &n->a1, B * 144 * CH)) ||
      (rc = nalloc(n, &n->a2, B * 36 * CH)) || (rc = nalloc(n, &n->a3, B * 36 * CH)) ||
      (rc = nalloc(n, &n->dy1, B * 144 * CH)) || (rc = nalloc(n, &n->dy2, B * 36 * CH)) ||
      (rc = nalloc(n, &n->dy3, B * 36 * CH)) || (rc = nalloc(n, &n->dy4, B * 9 * CH)) ||
      (rc = nalloc(n, &n->a4, B * 9 * CH)) || (rc = nalloc(n, &n->gv, B * CH)) ||
      (rc = nalloc(n, &n->logits, B * NCLS)) || (rc = nalloc(n, &n->dlog, B * NCLS)) ||
      (rc = nalloc(n, &n->lossv, B * 2)) || (rc = nalloc(n, &n->mvb, B)) || (rc = nalloc(n, &n->ptb, B)) ||
      (rc = nalloc(n, &n->part1, (B * 144 / CHUNK + 1) * CH * 32)) ||
      (rc = nalloc(n, &n->part2, (B * 36 / CHUNK + 1) * CH * 304)) ||
      (rc = nalloc(n, &n->part3, (B * 36 / CHUNK + 1) * CH * 304)) ||
      (rc = nalloc(n, &n->part4, (B * 9 / CHUNK + 1) * CH * 304))) {
    net_free(n);
    return rc;
  }
  n->master = all;
  n->grad = all + NPARAM;
  n->am = all + 2 * NPARAM;
  n->av = all + 3 * NPARAM;
  if (hipMemset(n->w1b, 0, (size_t)CH * K1 * 2) != hipSuccess ||
      hipMemset(n->wfb, 0, (size_t)NCLS_PAD * CH * 2) != hipSuccess) {
    net_free(n);
    return nfail(PCME_EHIP, "hipMemset failed");
  }
  // He-normal init of the conv weights, N(0, 1/32) fc weights, zero biases.
  std::vector<float> h((size_t)NPARAM, 0.f);
  uint64_t st = seed;
  auto fill = [&](int off, int cnt, double sd) {
    for (int i = 0; i < cnt; i++) h[(size_t)(off + i)] = (float)(sd * gauss(st));
  };
  fill(OW1, 576, sqrt(2.0 / 18.0));
  fill(OW2, 9216, sqrt(2.0 / 288.0));
  fill(OW3, 9216, sqrt(2.0 / 288.0));
  fill(OW4, 9216, sqrt(2.0 / 288.0));
  fill(OWF, NCLS * CH, sqrt(1.0 / 32.0));
  if ((rc = pcme_net_set_params(n, h.data()))) {
    net_free(n);
    return rc;
  }
  int per_cu = 0;
  if (hipFuncSetAttribute((const void *)conv_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, CONV_LDS) !=
          hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, conv_kernel, 64 * CW, CONV_LDS) != hipSuccess) {
    net_free(n);
    return nfail(PCME_EHIP, "conv kernel set-up failed");
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, n->dev) != hipSuccess) {
    net_free(n);
    return nfail(PCME_EHIP, "hipGetDeviceProperties failed");
  }
  n->conv_blocks = std::max(1, per_cu) * prop.multiProcessorCount;
  *out = n;
  return PCME_OK;
}

void pcme_net_destroy(pcme_net *net) {
  if (!net) return;
  (void)hipDeviceSynchronize();
  net_free(net);
}

int pcme_net_get_params(pcme_net *net, float *host) {
  int rc;
  if ((rc = check_dev(net))) return rc;
  if (!host) return nfail(PCME_EINVAL, "null host buffer");
  NCHK(hipDeviceSynchronize());
  NCHK(hipMemcpy(host, net->master, (size_t)NPARAM * 4, hipMemcpyDeviceToHost));
  return PCME_OK;
}

int pcme_net_set_params(pcme_net *net, const float *host) {
  int rc;
  if ((rc = check_dev(net))) return rc;
  if (!host) return nfail(PCME_EINVAL, "null host buffer");
  NCHK(hipDeviceSynchronize());
  NCHK(hipMemcpy(net->master, host, (size_t)NPARAM * 4, hipMemcpyHostToDevice));
  NCHK(hipMemset(net->am, 0, (size_t)NPARAM * 4));
  NCHK(hipMemset(net->av, 0, (size_t)NPARAM * 4));
  net->step = 0;
  if ((rc = refresh(net, nullptr))) return rc;
  NCHK(hipDeviceSynchronize());
  return PCME_OK;
}

int pcme_net_inputs(const uint8_t *src, const uint8_t *ref, int32_t mbw, int32_t mbh, const int32_t *mb, int32_t n,
                    uint16_t *x, void *stream) {
  if (n == 0) return PCME_OK;
  if (!src || !ref || !mb || !x || n < 0 || mbw < 1 || mbh < 1) return nfail(PCME_EINVAL, "bad argument");
  if (pcme_device_count() < 1) return nfail(PCME_ENODEV, "no gfx950 device");
  hipLaunchKernelGGL(inputs_kernel, dim3(blocks((int64_t)n * IN * IN, 256)), dim3(256), 0, (hipStream_t)stream, src,
                     ref, mbw, mbh, mb, n, (uint32_t *)x);
  NCHK(hipGetLastError());
  return PCME_OK;
}

int pcme_net_forward(pcme_net *net, const uint16_t *x, int32_t n, float *logits, void *stream) {
  int rc;
  if ((rc = check_dev(net))) return rc;
  if (n < 32 || n > net->maxb || n % 32) return nfail(PCME_EINVAL, "batch %d: need 32 <= n <= %d, n %% 32 == 0", n, net->maxb);
  if (!x || !logits) return nfail(PCME_EINVAL, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if ((rc = forward(net, x, n, s))) return rc;
  NCHK(hipMemcpyAsync(logits, net->logits, (size_t)n * NCLS * 4, hipMemcpyDeviceToDevice, s));
  return PCME_OK;
}

int pcme_net_grad(pcme_net *net, const uint16_t *x, const int32_t *mv_cls, const int32_t *part, const int32_t *idx,
                  int32_t n, float *loss_out, float *grad, void *stream) {
  int rc;
  if ((rc = check_dev(net))) return rc;
  hipStream_t s = (hipStream_t)stream;
  if ((rc = grad_impl(net, x, mv_cls, part, idx, n, loss_out, s))) return rc;
  if (grad) NCHK(hipMemcpyAsync(grad, net->grad, (size_t)NPARAM * 4, hipMemcpyDeviceToDevice, s));
  return PCME_OK;
}

int pcme_net_train_step(pcme_net *net, const uint16_t *x, const int32_t *mv_cls, const int32_t *part,
                        const int32_t *idx, int32_t n, float lr, float *loss_out, void *stream) {
  int rc;
  if ((rc = check_dev(net))) return rc;
  if (!(lr > 0.f) || lr > 1.f) return nfail(PCME_EINVAL, "learning rate %g outside (0, 1]", (double)lr);
  hipStream_t s = (hipStream_t)stream;
  if ((rc = grad_impl(net, x, mv_cls, part, idx, n, loss_out, s))) return rc;
  net->step++;
  const float bc1 = 1.f - powf(0.9f, (float)net->step), bc2 = 1.f - powf(0.999f, (float)net->step);
  hipLaunchKernelGGL(adam_kernel, dim3(blocks(NPARAM, 256)), dim3(256), 0, s, net->master, net->grad, net->am,
                     net->av, lr, bc1, bc2, net->w1b, net->w2b, net->w3b, net->w4b, net->w2t, net->w3t, net->w4t,
                     net->wfb);
  NCHK(hipGetLastError());
  return PCME_OK;
}

int pcme_net_hints(pcme_net *net, const pcme_frame_job *jobs, int32_t n, int32_t ncand, float thr_part,
                   float thr_mv, void *stream) {
  int rc;
  if ((rc = check_dev(net))) return rc;
  if (n < 1 || !jobs) return nfail(PCME_EINVAL, "bad job array");
  if (ncand < 0 || ncand > PCME_NET_MAXCAND) return nfail(PCME_EINVAL, "ncand %d outside [0, 8]", ncand);
  std::vector<NetJob> hj((size_t)n);
  int64_t total = 0;
  for (int i = 0; i < n; i++) {
    const pcme_frame_job &J = jobs[i];
    if (!J.src || !J.ref || !J.hint || J.prm.mbw < 1 || J.prm.mbh < 1)
      return nfail(PCME_EINVAL, "job %d: needs src, ref, hint and a positive size", i);
    if (((uintptr_t)J.hint & 3u) != 0) return nfail(PCME_EINVAL, "job %d: hints must be 4-byte aligned", i);
    hj[(size_t)i] = NetJob{J.src, J.ref, const_cast<pcme_hint *>(J.hint), J.prm.mbw, J.prm.mbh, total};
    total += (int64_t)J.prm.mbw * J.prm.mbh;
  }
  hipStream_t s = (hipStream_t)stream;
  const bool same = net->hjobs.size() == hj.size() &&
                    memcmp(net->hjobs.data(), hj.data(), hj.size() * sizeof(NetJob)) == 0;
  if (!same) {
    NCHK(hipStreamSynchronize(s));
    if (net->djobs) NCHK(hipFree(net->djobs));
    net->djobs = nullptr;
    net->hjobs.clear();
    if (hipMalloc((void **)&net->djobs, hj.size() * sizeof(NetJob)) != hipSuccess)
      return nfail(PCME_ENOMEM, "job table allocation failed");
    NCHK(hipMemcpy(net->djobs, hj.data(), hj.size() * sizeof(NetJob), hipMemcpyHostToDevice));
    net->hjobs = hj;
  }
  if (total > net->gcap) {
    NCHK(hipStreamSynchronize(s));
    if (net->gbuf) NCHK(hipFree(net->gbuf));
    net->gbuf = nullptr;
    net->gcap = 0;
    if (hipMalloc((void **)&net->gbuf, (size_t)total * CH * sizeof(float)) != hipSuccess)
      return nfail(PCME_ENOMEM, "pooled-vector buffer of %lld MBs failed", (long long)total);
    net->gcap = total;
  }
  const int64_t need = (total + CW - 1) / CW;
  const unsigned grid = (unsigned)std::min<int64_t>(need, net->conv_blocks);
  hipLaunchKernelGGL(conv_kernel, dim3(grid), dim3(64 * CW), CONV_LDS, s, net->djobs, n, total, net->w1b, net->w2b,
                     net->w3b, net->w4b, net->master, net->gbuf);
  hipLaunchKernelGGL(heads_kernel, dim3(blocks(total, 256)), dim3(256), 0, s, net->djobs, n, total, net->gbuf,
                     net->wfb, net->master, ncand, thr_part, thr_mv);
  NCHK(hipGetLastError());
  return PCME_OK;
}

}  // extern "C"
