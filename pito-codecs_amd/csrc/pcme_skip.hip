// pcme_skip.hip -- gfx950 kernel of the early P-skip gate around the motion
// search (analyse_p, slice.rs:1884-1920): predict the MB at its pskip_mv
// (pred_block = mc_luma + mc_chroma), compare sad16 with the QP's bound,
// and when below it find code_inter's coded block pattern (slice.rs:943-1001)
// -- the 4x4 and/or 8x8 luma transform with the inter dead-zone quantiser
// and decimation (code_luma4 / code_luma8, slice.rs:813-941; residual.rs),
// chroma DC and AC (code_chroma, slice.rs:679-811) -- so that a cbp of 0
// marks the MB an early PSkip.
//
// One wave per MB.  The prediction and the SAD use all 64 lanes (4 luma
// pixels, 2 chroma pixels each); the transforms give each lane a block:
// lanes 0-15 the 4x4 luma blocks, lanes 0-3 the 8x8 ones, lanes 0-7 the
// chroma 4x4 blocks, exchanging scores and flags through LDS.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "pcme_internal.h"
#include "pcme_dev.h"

namespace pcme {

// ZIGZAG4 / ZIGZAG8 (syntax tables.rs:1-20) and the 4x4 block order
// BLK4_XY (dec/tables.rs:95-112).
__constant__ uint8_t kZz4[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};
__constant__ uint8_t kZz8[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                 12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};
__constant__ uint8_t kB4x[16] = {0, 1, 0, 1, 2, 3, 2, 3, 0, 1, 0, 1, 2, 3, 2, 3};
__constant__ uint8_t kB4y[16] = {0, 0, 1, 1, 0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 3, 3};
// QPC_TABLE (dec/tables.rs:3-14) beyond index 29
__constant__ uint8_t kQpcTail[22] = {29, 30, 31, 32, 32, 33, 34, 34, 35, 35, 36,
                                     36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};

#define SKIP_DZ_INTER 43  // Deadzone::INTER.f256
#define SKIP_MAX_LEVEL 2047

// decimate_score (slice.rs:149-165) of n levels.
__device__ __forceinline__ uint32_t decimate_score(const int16_t *lv, int n) {
  uint32_t score = 0;
  int run = 0;
  for (int i = 0; i < n; i++) {
    const int v = lv[i];
    if (v == 0) {
      run++;
      continue;
    }
    if (v > 1 || v < -1) return 64;
    // T = 3, 2, 2, 1, 1, 1, 0...
    score += run == 0 ? 3u : (run <= 2 ? 2u : (run <= 5 ? 1u : 0u));
    run = 0;
  }
  return score;
}

__device__ __forceinline__ uint8_t clip8(int v) { return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

// transform.rs fdct8_1d / dec/transform.rs idct8_1d
__device__ __forceinline__ void fdct8_1d(const int32_t *s, int32_t *o) {
  const int32_t a0 = s[0] + s[7], a1 = s[1] + s[6], a2 = s[2] + s[5], a3 = s[3] + s[4];
  const int32_t b0 = s[0] - s[7], b1 = s[1] - s[6], b2 = s[2] - s[5], b3 = s[3] - s[4];
  const int32_t e0 = a0 + a3, e1 = a1 + a2, e2 = a0 - a3, e3 = a1 - a2;
  o[0] = 8 * (e0 + e1);
  o[1] = 12 * b0 + 10 * b1 + 6 * b2 + 3 * b3;
  o[2] = 8 * e2 + 4 * e3;
  o[3] = 10 * b0 - 3 * b1 - 12 * b2 - 6 * b3;
  o[4] = 8 * (e0 - e1);
  o[5] = 6 * b0 - 12 * b1 + 3 * b2 + 10 * b3;
  o[6] = 4 * e2 - 8 * e3;
  o[7] = 3 * b0 - 6 * b1 + 10 * b2 - 12 * b3;
}
__device__ __forceinline__ void idct8_1d(const int32_t *d, int32_t *o) {
  const int32_t a0 = d[0] + d[4], a4 = d[0] - d[4];
  const int32_t a2 = (d[2] >> 1) - d[6], a6 = d[2] + (d[6] >> 1);
  const int32_t b0 = a0 + a6, b2 = a4 + a2, b4 = a4 - a2, b6 = a0 - a6;
  const int32_t a1 = -d[3] + d[5] - d[7] - (d[7] >> 1);
  const int32_t a3 = d[1] + d[7] - d[3] - (d[3] >> 1);
  const int32_t a5 = -d[1] + d[7] + d[5] + (d[5] >> 1);
  const int32_t a7 = d[3] + d[5] + d[1] + (d[1] >> 1);
  const int32_t b1 = a1 + (a7 >> 2), b7 = a7 - (a1 >> 2);
  const int32_t b3 = a3 + (a5 >> 2), b5 = (a3 >> 2) - a5;
  o[0] = b0 + b7; o[1] = b2 + b5; o[2] = b4 + b3; o[3] = b6 + b1;
  o[4] = b6 - b1; o[5] = b4 - b3; o[6] = b2 - b5; o[7] = b0 - b7;
}

// Sum over the wave (every lane gets it).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  v = red8(v);
#pragma unroll
  for (int m = 8; m < 64; m <<= 1) v += (uint32_t)__shfl_xor((int)v, m);
  return v;
}

__global__ void __launch_bounds__(64) skip_gate_kernel(const pcme_skip_job *__restrict__ jobs, int njobs,
                                                       const SkipTables *__restrict__ T, pcme_skip_out *out) {
  __shared__ uint8_t s_sy[256], s_py[256], s_r4[256], s_r8[256];
  __shared__ uint8_t s_sc[128], s_pc[128];  // chroma: U rows 0-7, then V
  __shared__ uint32_t s_score[16];
  __shared__ int s_nz[16];
  __shared__ double s_b4[16], s_b8[4];  // per-block rates (adaptive choice)
  __shared__ int32_t s_dc[8];
  const int j = blockIdx.x;
  if (j >= njobs) return;
  const pcme_skip_job &J = jobs[j];
  const int l = (int)lane_id();
  const int qp = J.qp, px = 16 * J.mx, py = 16 * J.my;
  const int mvx = J.skip_mv.x, mvy = J.skip_mv.y;
  pcme_skip_out o;
  o.sad = 0;
  o.thresh = ((uint32_t)(T->v4[qp % 6][0] << (qp / 6)) * 16u) * 2u;
  o.mv_ok = o.coded = o.skip = o.t8x8 = 0;
  o.cbp = 0xFF;
  o.reserved[0] = o.reserved[1] = o.reserved[2] = 0;
  // mv_ok (slice.rs:496-507)
  {
    const int x0 = px + (mvx >> 2), y0 = py + (mvy >> 2);
    o.mv_ok = (mvx < 2000 && mvx > -2000 && mvy < 2000 && mvy > -2000 && x0 >= -44 && y0 >= -44 &&
               x0 + 16 <= 16 * J.mbw + 44 && (int64_t)(y0 + 16) < J.vlim)
                  ? 1
                  : 0;
  }
  if (o.mv_ok) {
    // mc_luma (picture.rs:415-449) of the 16x16 block, 4 pixels per lane,
    // and sad16 against the source (dsp.rs:15-33)
    const int stride = 16 * J.mbw + 2 * PCME_PAD;
    const uint32_t e = qpel_entry((uint32_t)((mvy & 3) * 4 + (mvx & 3)));
    const int r = l >> 2, c = (l & 3) * 4;
    const uint8_t *pa = J.ref[e & 3u] +
                        (size_t)(py + r + (mvy >> 2) + (int)((e >> 3) & 1u) + PCME_PAD) * stride + px + c +
                        (mvx >> 2) + (int)((e >> 2) & 1u) + PCME_PAD;
    const uint8_t *pb = J.ref[(e >> 4) & 3u] +
                        (size_t)(py + r + (mvy >> 2) + (int)((e >> 7) & 1u) + PCME_PAD) * stride + px + c +
                        (mvx >> 2) + (int)((e >> 6) & 1u) + PCME_PAD;
    const uint8_t *sy = J.src_y + (size_t)r * J.ystride + c;
    uint32_t sad = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int p = ((int)pa[i] + (int)pb[i] + 1) >> 1;
      const int s = sy[i];
      s_py[r * 16 + c + i] = (uint8_t)p;
      s_sy[r * 16 + c + i] = (uint8_t)s;
      sad += (uint32_t)(s > p ? s - p : p - s);
    }
    o.sad = wave_sum(sad);
  }
  o.coded = (o.mv_ok && o.sad < o.thresh) ? 1 : 0;
  if (o.coded) {
    // mc_chroma (picture.rs:451-487) of the two 8x8 blocks, 2 pixels per lane
    {
      const int cstride = 8 * J.mbw + 2 * PCME_CPAD;
      const int comp = l >> 5, r = (l >> 2) & 7, c = (l & 3) * 2;
      const int xi = px / 2 + (mvx >> 3), yi = py / 2 + (mvy >> 3);
      const uint32_t fx = (uint32_t)(mvx & 7), fy = (uint32_t)(mvy & 7);
      const uint32_t w00 = (8 - fx) * (8 - fy), w01 = fx * (8 - fy), w10 = (8 - fx) * fy, w11 = fx * fy;
      const uint8_t *a = (comp ? J.ref_v : J.ref_u) + (size_t)(yi + r + PCME_CPAD) * cstride + xi + c + PCME_CPAD;
      const uint8_t *b = a + cstride;
      const uint8_t *sc = (comp ? J.src_v : J.src_u) + (size_t)r * J.cstride + c;
#pragma unroll
      for (int i = 0; i < 2; i++) {
        s_pc[comp * 64 + r * 8 + c + i] = (uint8_t)((w00 * a[i] + w01 * a[i + 1] + w10 * b[i] + w11 * b[i + 1] + 32) >> 6);
        s_sc[comp * 64 + r * 8 + c + i] = sc[i];
      }
    }
    __syncthreads();
    const bool cabac = (J.flags & PCME_SKIP_CABAC) != 0;
    const bool t8 = (J.flags & PCME_SKIP_T8X8) != 0;
    const bool adaptive = t8 && (J.flags & PCME_SKIP_ADAPTIVE_T8) != 0;
    // ---- code_luma4 (slice.rs:813-878, residual.rs luma4_fwd/rec_body)
    int cbp4 = 0;
    if (!t8 || adaptive) {
      int16_t lv[16];
      int16_t cf[16];  // raster positions
      if (l < 16) {
        const int bx = kB4x[l], by = kB4y[l];
        const int qbits = 15 + qp / 6;
        int32_t lim = cabac ? 32767 : SKIP_MAX_LEVEL;
        const int32_t f = (int32_t)((((int64_t)1 << qbits) * SKIP_DZ_INTER) >> 8);
        const int32_t *mf = T->mf4[qp % 6];
        int32_t d[16], t[16];
#pragma unroll
        for (int k = 0; k < 16; k++) {
          const int i = (by * 4 + k / 4) * 16 + bx * 4 + k % 4;
          d[k] = (int32_t)s_sy[i] - (int32_t)s_py[i];
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int32_t a = d[i * 4], b = d[i * 4 + 1], c = d[i * 4 + 2], e = d[i * 4 + 3];
          t[i * 4] = a + e + b + c;
          t[i * 4 + 1] = 2 * (a - e) + (b - c);
          t[i * 4 + 2] = a + e - b - c;
          t[i * 4 + 3] = (a - e) - 2 * (b - c);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int32_t a = t[i], b = t[4 + i], c = t[8 + i], e = t[12 + i];
          const int32_t ov[4] = {a + e + b + c, 2 * (a - e) + (b - c), a + e - b - c, (a - e) - 2 * (b - c)};
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const int pos = k * 4 + i;
            const int32_t v = ov[k];
            int32_t q = ((v < 0 ? -v : v) * mf[pos] + f) >> qbits;
            q = q > lim ? lim : q;
            cf[pos] = (int16_t)(v < 0 ? -q : q);
          }
        }
#pragma unroll
        for (int k = 0; k < 16; k++) lv[k] = cf[kZz4[k]];
        s_score[l] = decimate_score(lv, 16);
      }
      __syncthreads();
      uint32_t sc8[4], total = 0;
#pragma unroll
      for (int b8 = 0; b8 < 4; b8++) {
        sc8[b8] = s_score[4 * b8] + s_score[4 * b8 + 1] + s_score[4 * b8 + 2] + s_score[4 * b8 + 3];
        total += sc8[b8];
      }
      __syncthreads();
      if (l < 16) {
        if (sc8[l >> 2] < 4 || total < 6) {
#pragma unroll
          for (int k = 0; k < 16; k++) lv[k] = cf[k] = 0;
        }
        int nz = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) nz |= lv[k] != 0;
        s_nz[l] = nz;
        if (adaptive) {
          // the block's rate, summed in scan order (slice.rs:984)
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < 16; k++) s = __dadd_rn(s, T->lb[lv[k] < 0 ? -(int)lv[k] : (int)lv[k]]);
          s_b4[l] = s;
          // luma4_rec_body: dequantise, inverse transform, add the prediction
          const int32_t *v = T->v4[qp % 6];
          const int sh = qp / 6;
          int32_t dq[16], fr[16];
#pragma unroll
          for (int k = 0; k < 16; k++) dq[k] = ((int32_t)cf[k] * v[k]) << sh;
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const int32_t d0 = dq[i * 4], d1 = dq[i * 4 + 1], d2 = dq[i * 4 + 2], d3 = dq[i * 4 + 3];
            const int32_t e = d0 + d2, ff = d0 - d2, g = (d1 >> 1) - d3, h = d1 + (d3 >> 1);
            fr[i * 4] = e + h;
            fr[i * 4 + 1] = ff + g;
            fr[i * 4 + 2] = ff - g;
            fr[i * 4 + 3] = e - h;
          }
          const int bx = kB4x[l], by = kB4y[l];
#pragma unroll
          for (int jj = 0; jj < 4; jj++) {
            const int32_t f0 = fr[jj], f1 = fr[4 + jj], f2 = fr[8 + jj], f3 = fr[12 + jj];
            const int32_t e = f0 + f2, ff = f0 - f2, g = (f1 >> 1) - f3, h = f1 + (f3 >> 1);
            const int32_t rr[4] = {e + h, ff + g, ff - g, e - h};
#pragma unroll
            for (int y = 0; y < 4; y++) {
              const int i = (by * 4 + y) * 16 + bx * 4 + jj;
              s_r4[i] = clip8((int)s_py[i] + ((rr[y] + 32) >> 6));
            }
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int b = 0; b < 16; b++) cbp4 |= s_nz[b] ? (1 << (b >> 2)) : 0;
    }
    // ---- code_luma8 (slice.rs:880-941, residual.rs luma8_fwd/rec_body)
    int cbp8 = 0;
    double bits8 = 0.0;
    if (t8) {
      int16_t lv[64];
      int32_t cf[64];
      __syncthreads();
      if (l < 4) {
        const int ox = (l & 1) * 8, oy = (l >> 1) * 8;
        const int qbits = 22 + qp / 6;
        const int64_t f = (((int64_t)1 << qbits) * SKIP_DZ_INTER) >> 8;
        const int32_t *mf = T->mf8[qp % 6];
        int32_t tt[8][8];
        for (int y = 0; y < 8; y++) {
          int32_t s[8];
#pragma unroll
          for (int x = 0; x < 8; x++) {
            const int i = (oy + y) * 16 + ox + x;
            s[x] = (int32_t)s_sy[i] - (int32_t)s_py[i];
          }
          fdct8_1d(s, tt[y]);
        }
        for (int x = 0; x < 8; x++) {
          int32_t col[8], ov[8];
#pragma unroll
          for (int y = 0; y < 8; y++) col[y] = tt[y][x];
          fdct8_1d(col, ov);
#pragma unroll
          for (int k = 0; k < 8; k++) {
            const int32_t v = ov[k];
            const uint64_t av = (uint64_t)(v < 0 ? -(int64_t)v : (int64_t)v);
            const int32_t a = (int32_t)((av * (uint64_t)mf[k * 8 + x] + (uint64_t)f) >> qbits);
            cf[k * 8 + x] = v < 0 ? -a : a;
          }
        }
        for (int k = 0; k < 64; k++) lv[k] = (int16_t)cf[kZz8[k]];
        s_score[l] = decimate_score(lv, 64);
      }
      __syncthreads();
      const uint32_t total = s_score[0] + s_score[1] + s_score[2] + s_score[3];
      if (l < 4) {
        const int ox = (l & 1) * 8, oy = (l >> 1) * 8;
        if (s_score[l] < 4 || total < 6) {
          for (int k = 0; k < 64; k++) lv[k] = 0;
          for (int k = 0; k < 64; k++) cf[k] = 0;
        }
        int nz = 0;
        for (int k = 0; k < 64; k++) nz |= lv[k] != 0;
        if (adaptive) {
          double s = 0.0;
          for (int k = 0; k < 64; k++) s = __dadd_rn(s, T->lb[lv[k] < 0 ? -(int)lv[k] : (int)lv[k]]);
          s_b8[l] = s;
          if (!nz) {
            for (int y = 0; y < 8; y++)
#pragma unroll
              for (int x = 0; x < 8; x++) s_r8[(oy + y) * 16 + ox + x] = s_py[(oy + y) * 16 + ox + x];
          } else {
            const int32_t *v = T->v8[qp % 6];
            const int q6 = qp / 6;
            int32_t g[8][8];
            for (int i = 0; i < 8; i++) {
              int32_t row[8];
#pragma unroll
              for (int k = 0; k < 8; k++) {
                const int32_t lev = cf[i * 8 + k], ls = 16 * v[i * 8 + k];
                row[k] = q6 >= 6 ? (lev * ls) << (q6 - 6) : (lev * ls + (1 << (5 - q6))) >> (6 - q6);
              }
              idct8_1d(row, g[i]);
            }
            for (int x = 0; x < 8; x++) {
              int32_t col[8], ov[8];
#pragma unroll
              for (int i = 0; i < 8; i++) col[i] = g[i][x];
              idct8_1d(col, ov);
#pragma unroll
              for (int y = 0; y < 8; y++) {
                const int i = (oy + y) * 16 + ox + x;
                s_r8[i] = clip8((int)s_py[i] + ((ov[y] + 32) >> 6));
              }
            }
          }
        }
        s_nz[l] = nz;
      }
      __syncthreads();
#pragma unroll
      for (int b = 0; b < 4; b++) cbp8 |= s_nz[b] ? (1 << b) : 0;
      if (adaptive) {
#pragma unroll
        for (int b = 0; b < 4; b++)
          if ((cbp8 >> b) & 1) bits8 = __dadd_rn(bits8, s_b8[b]);
      }
    }
    // ---- code_inter's luma choice (slice.rs:952-998)
    int cbpl, use8;
    if (!t8) {
      cbpl = cbp4;
      use8 = 0;
    } else if (!adaptive) {
      cbpl = cbp8;
      use8 = 1;
    } else {
      // j8 <= j4 with j = ssd(src, reconstruction) + lam * rate, the rates
      // summed in the reference's order (slice.rs:975-988), every float
      // operation rounded on its own
      uint32_t e8 = 0, e4 = 0;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int p = l * 4 + i;
        const int d8 = (int)s_sy[p] - (int)s_r8[p], d4 = (int)s_sy[p] - (int)s_r4[p];
        e8 += (uint32_t)(d8 * d8);
        e4 += (uint32_t)(d4 * d4);
      }
      const uint64_t ssd8 = wave_sum(e8), ssd4 = wave_sum(e4);
      double bits4 = 0.0;
#pragma unroll
      for (int b = 0; b < 16; b++)
        if ((cbp4 >> (b / 4)) & 1) bits4 = __dadd_rn(bits4, __dadd_rn(1.0, s_b4[b]));
      const double lam = T->lam_ssd[qp];
      const double j8 = __dadd_rn((double)ssd8, __dmul_rn(lam, bits8));
      const double j4 = __dadd_rn((double)ssd4, __dmul_rn(lam, bits4));
      use8 = j8 <= j4 ? 1 : 0;
      cbpl = use8 ? cbp8 : cbp4;
    }
    // ---- code_chroma (slice.rs:679-811): DC and AC of U and V
    __syncthreads();
    const int qpc = (J.qp + J.chroma_qp_offset) < 30 ? max(J.qp + J.chroma_qp_offset, 0)
                                                     : kQpcTail[min(J.qp + J.chroma_qp_offset, 51) - 30];
    if (l < 8) {
      const int comp = l >> 2, b = l & 3, bx = (b & 1) * 4, by = (b >> 1) * 4;
      int32_t d[16], t[16];
#pragma unroll
      for (int y = 0; y < 4; y++)
#pragma unroll
        for (int x = 0; x < 4; x++) {
          const int i = comp * 64 + (by + y) * 8 + bx + x;
          d[y * 4 + x] = (int32_t)s_sc[i] - (int32_t)s_pc[i];
        }
      // fdct4 (transform.rs)
      int32_t u[16];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int32_t a = d[i * 4], b1 = d[i * 4 + 1], c = d[i * 4 + 2], e = d[i * 4 + 3];
        const int32_t s03 = a + e, d03 = a - e, s12 = b1 + c, d12 = b1 - c;
        u[i * 4] = s03 + s12;
        u[i * 4 + 1] = 2 * d03 + d12;
        u[i * 4 + 2] = s03 - s12;
        u[i * 4 + 3] = d03 - 2 * d12;
      }
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int32_t a = u[i], b1 = u[4 + i], c = u[8 + i], e = u[12 + i];
        const int32_t s03 = a + e, d03 = a - e, s12 = b1 + c, d12 = b1 - c;
        t[i] = s03 + s12;
        t[4 + i] = 2 * d03 + d12;
        t[8 + i] = s03 - s12;
        t[12 + i] = d03 - 2 * d12;
      }
      s_dc[l] = t[0];
      // quant4 with skip_dc (the AC), then the CAVLC clamp
      const int qbits = 15 + qpc / 6;
      const int64_t f = (((int64_t)1 << qbits) * SKIP_DZ_INTER) >> 8;
      const int32_t *mf = T->mf4[qpc % 6];
      int16_t sc[15];
      int acnz = 0;
#pragma unroll
      for (int k = 1; k < 16; k++) {
        const int32_t v = t[kZz4[k]];
        int32_t a = (int32_t)(((int64_t)(v < 0 ? -(int64_t)v : v) * mf[kZz4[k]] + f) >> qbits);
        a = v < 0 ? -a : a;
        if (!cabac) a = a < -SKIP_MAX_LEVEL ? -SKIP_MAX_LEVEL : (a > SKIP_MAX_LEVEL ? SKIP_MAX_LEVEL : a);
        sc[k - 1] = (int16_t)a;
        acnz |= a != 0;
      }
      s_score[l] = decimate_score(sc, 15);
      s_nz[l] = acnz;
    }
    __syncthreads();
    bool ac_any = false, dc_any = false;
#pragma unroll
    for (int comp = 0; comp < 2; comp++) {
      const uint32_t ac_score = s_score[comp * 4] + s_score[comp * 4 + 1] + s_score[comp * 4 + 2] + s_score[comp * 4 + 3];
      if (ac_score >= 7) ac_any |= (s_nz[comp * 4] | s_nz[comp * 4 + 1] | s_nz[comp * 4 + 2] | s_nz[comp * 4 + 3]) != 0;
      // quant_chroma_dc (transform.rs)
      const int32_t *dc = s_dc + comp * 4;
      const int32_t h[4] = {dc[0] + dc[1] + dc[2] + dc[3], dc[0] - dc[1] + dc[2] - dc[3],
                            dc[0] + dc[1] - dc[2] - dc[3], dc[0] - dc[1] - dc[2] + dc[3]};
      const int qbits = 16 + qpc / 6;
      const int64_t f = (((int64_t)1 << qbits) * SKIP_DZ_INTER) >> 8;
      const int64_t mf0 = T->mf4[qpc % 6][0];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int32_t v = h[i];
        const int32_t a = (int32_t)(((int64_t)(v < 0 ? -(int64_t)v : v) * mf0 + f) >> qbits);
        dc_any |= a != 0;
      }
    }
    const int cbpc = ac_any ? 2 : (dc_any ? 1 : 0);
    const int cbp = cbpl | (cbpc << 4);
    o.cbp = (uint8_t)cbp;
    o.t8x8 = (uint8_t)(use8 && cbpl != 0);
    o.skip = cbp == 0 ? 1 : 0;
  }
  if (l == 0) out[j] = o;
}

// ---------------------------------------------------------------- tables

// QUANT4 (enc/tables.rs) by position class: (even, even), (odd, odd), mixed
static const int32_t kQuant4[6][3] = {{13107, 5243, 8066}, {11916, 4660, 7490}, {10082, 4194, 6554},
                                      {9362, 3647, 5825},  {8192, 3355, 5243},  {7282, 2893, 4559}};
// NORM_ADJUST4 / NORM_ADJUST8 (dec/tables.rs) by the same classes
static const int32_t kNorm4[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16},
                                     {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};
static const int32_t kNorm8[6][6] = {{20, 18, 32, 19, 25, 24}, {22, 19, 35, 21, 28, 26},
                                     {26, 23, 42, 24, 33, 31}, {28, 25, 45, 26, 35, 33},
                                     {32, 28, 51, 30, 40, 38}, {36, 32, 58, 34, 46, 43}};

static int class4(int pos) {
  const int i = pos / 4, j = pos % 4;
  return (i % 2 == 0 && j % 2 == 0) ? 0 : ((i % 2 == 1 && j % 2 == 1) ? 1 : 2);
}

static int class8(int pos) {
  const int i = pos / 8, j = pos % 8;
  if (i % 4 == 0 && j % 4 == 0) return 0;
  if (i % 2 == 1 && j % 2 == 1) return 1;
  if (i % 4 == 2 && j % 4 == 2) return 2;
  if ((i % 4 == 0 && j % 2 == 1) || (i % 2 == 1 && j % 4 == 0)) return 3;
  if ((i % 4 == 0 && j % 4 == 2) || (i % 4 == 2 && j % 4 == 0)) return 4;
  return 5;
}

// QuantTables::new (transform.rs): MF4 from QUANT4, the 8x8 MF as
// round(2^36 / (|row x|^2 |row y|^2 V)) over the 8x8 basis's squared row
// norms; Lambdas::new().ssd = 0.85 * 2^((qp - 12) / 3); the level rate
// 3 + 2 log2|v|.  Each sum and product is rounded on its own (no fused
// multiply-add), as the reference's f64 arithmetic is.
void build_skip_tables(SkipTables *t) {
#pragma clang fp contract(off)
  static const double norm8[8] = {512.0, 578.0, 320.0, 578.0, 512.0, 578.0, 320.0, 578.0};
  for (int m = 0; m < 6; m++) {
    for (int i = 0; i < 16; i++) {
      t->mf4[m][i] = kQuant4[m][class4(i)];
      t->v4[m][i] = kNorm4[m][class4(i)];
    }
    for (int i = 0; i < 64; i++) {
      const int v = kNorm8[m][class8(i)];
      t->v8[m][i] = v;
      t->mf8[m][i] = (int32_t)round((double)(1ULL << 36) / (norm8[i & 7] * norm8[i >> 3] * (double)v));
    }
  }
  for (int qp = 0; qp < 52; qp++) t->lam_ssd[qp] = 0.85 * pow(2.0, ((double)qp - 12.0) / 3.0);
  t->lb[0] = 0.4;
  for (int k = 1; k <= 32768; k++) t->lb[k] = 3.0 + 2.0 * log2((double)k);
}

}  // namespace pcme
