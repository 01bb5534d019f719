/*
 * skip_oracle.c -- CPU oracle of the early P-skip gate around the motion
 * search (TEST INFRASTRUCTURE ONLY: the checker of libpcme's
 * pcme_skip_gate_batch, never part of the product path).
 *
 * Restates, line by line:
 *   the gate of analyse_p             src/h264/enc/slice.rs:1884-1920
 *   code_inter (cbp only)             slice.rs:943-1001
 *   code_luma4 / code_luma8           slice.rs:813-941
 *   code_chroma (cbp only)            slice.rs:679-811
 *   decimate_score                    slice.rs:149-165
 *   Lambdas::new (ssd)                slice.rs:52-68
 *   fdct4, idct4, fdct8_1d, quant4, dequant8_one, quant_chroma_dc,
 *   QuantTables::new, Deadzone        src/h264/enc/transform.rs:1-296
 *   luma4_fwd/rec, luma8_fwd/rec      src/h264/enc/residual.rs:1-133
 *   idct8_1d                          src/h264/dec/transform.rs:133-160
 *   QUANT4, pos_class4, chroma_qp     src/h264/enc/tables.rs:117-140
 *   QPC_TABLE, NORM_ADJUST4/8, norm_adjust4/8, BLK4_XY
 *                                     src/h264/dec/tables.rs:1-110
 *   ZIGZAG4, ZIGZAG8                  src/syntax/h264/tables.rs:1-20
 *   ssd                               src/h264/enc/dsp.rs:90-102
 * Built with -ffp-contract=off: the adaptive transform choice compares
 * float sums that must round operation by operation as in the reference.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pcme_oracle.h"

static const int32_t QUANT4[6][3] = {
    {13107, 5243, 8066}, {11916, 4660, 7490}, {10082, 4194, 6554},
    {9362, 3647, 5825},  {8192, 3355, 5243},  {7282, 2893, 4559},
};
static const uint16_t NORM_ADJUST4[6][3] = {
    {10, 16, 13}, {11, 18, 14}, {13, 20, 16}, {14, 23, 18}, {16, 25, 20}, {18, 29, 23},
};
static const uint16_t NORM_ADJUST8[6][6] = {
    {20, 18, 32, 19, 25, 24}, {22, 19, 35, 21, 28, 26}, {26, 23, 42, 24, 33, 31},
    {28, 25, 45, 26, 35, 33}, {32, 28, 51, 30, 40, 38}, {36, 32, 58, 34, 46, 43},
};
static const uint8_t ZIGZAG4[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};
static const uint8_t BLK4_X[16] = {0, 1, 0, 1, 2, 3, 2, 3, 0, 1, 0, 1, 2, 3, 2, 3};
static const uint8_t BLK4_Y[16] = {0, 0, 1, 1, 0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 3, 3};
#define DZ_INTER 43 /* Deadzone::INTER.f256 (transform.rs) */
#define MAX_LEVEL 2047 /* cavlc::MAX_LEVEL (cavlc.rs:11) */

/* tables.rs/dec: the 8x8 zigzag built by diagonals (syntax tables.rs:5-20) */
static void zigzag8(uint8_t out[64]) {
  int k = 0;
  for (int d = 0; d < 15; d++)
    for (int i = 0; i < 8; i++) {
      int row = d % 2 == 0 ? d - i : i, col = d % 2 == 0 ? i : d - i;
      if (row >= 0 && row < 8 && col >= 0 && col < 8) out[k++] = (uint8_t)(row * 8 + col);
    }
}

int o_chroma_qp(int qpi) {
  static const uint8_t tail[22] = {29, 30, 31, 32, 32, 33, 34, 34, 35, 35, 36,
                                   36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};
  int q = qpi < 0 ? 0 : (qpi > 51 ? 51 : qpi);
  return q < 30 ? q : tail[q - 30];
}

double o_lambda_ssd(int qp) { return 0.85 * pow(2.0, ((double)qp - 12.0) / 3.0); }

static uint16_t norm_adjust4(int m, int pos) {
  int i = pos / 4, j = pos % 4;
  if (i % 2 == 0 && j % 2 == 0) return NORM_ADJUST4[m][0];
  if (i % 2 == 1 && j % 2 == 1) return NORM_ADJUST4[m][1];
  return NORM_ADJUST4[m][2];
}

static uint16_t norm_adjust8(int m, int pos) {
  int i = pos / 8, j = pos % 8, k;
  if (i % 4 == 0 && j % 4 == 0) k = 0;
  else if (i % 2 == 1 && j % 2 == 1) k = 1;
  else if (i % 4 == 2 && j % 4 == 2) k = 2;
  else if ((i % 4 == 0 && j % 2 == 1) || (i % 2 == 1 && j % 4 == 0)) k = 3;
  else if ((i % 4 == 0 && j % 4 == 2) || (i % 4 == 2 && j % 4 == 0)) k = 4;
  else k = 5;
  return NORM_ADJUST8[m][k];
}

static int pos_class4(int i) {
  int x = i & 3, y = i >> 2;
  if (x % 2 == 0 && y % 2 == 0) return 0;
  if (x % 2 == 1 && y % 2 == 1) return 1;
  return 2;
}

/* transform.rs QuantTables::new */
void o_quant_tables(oquant *q) {
  static const double norm8[8] = {512.0, 578.0, 320.0, 578.0, 512.0, 578.0, 320.0, 578.0};
  for (int m = 0; m < 6; m++) {
    for (int i = 0; i < 16; i++) {
      q->mf4[m][i] = QUANT4[m][pos_class4(i)];
      q->v4[m][i] = norm_adjust4(m, i);
    }
    for (int i = 0; i < 64; i++) {
      int v = norm_adjust8(m, i);
      q->v8[m][i] = v;
      int x = i & 7, y = i >> 3;
      q->mf8[m][i] = (int32_t)round((double)(1ULL << 36) / (norm8[x] * norm8[y] * (double)v));
    }
  }
}

static const oquant *tables(void) {
  static oquant q;
  static int done = 0;
  if (!done) {
    o_quant_tables(&q);
    done = 1;
  }
  return &q;
}

/* transform.rs fdct4 */
void o_fdct4(const int32_t d[16], int32_t o[16]) {
  int32_t t[16];
  for (int i = 0; i < 4; i++) {
    int32_t a = d[i * 4], b = d[i * 4 + 1], c = d[i * 4 + 2], e = d[i * 4 + 3];
    int32_t s03 = a + e, d03 = a - e, s12 = b + c, d12 = b - c;
    t[i * 4] = s03 + s12;
    t[i * 4 + 1] = 2 * d03 + d12;
    t[i * 4 + 2] = s03 - s12;
    t[i * 4 + 3] = d03 - 2 * d12;
  }
  for (int i = 0; i < 4; i++) {
    int32_t a = t[i], b = t[4 + i], c = t[8 + i], e = t[12 + i];
    int32_t s03 = a + e, d03 = a - e, s12 = b + c, d12 = b - c;
    o[i] = s03 + s12;
    o[4 + i] = 2 * d03 + d12;
    o[8 + i] = s03 - s12;
    o[12 + i] = d03 - 2 * d12;
  }
}

/* transform.rs fdct8_1d */
static void fdct8_1d(const int32_t s[8], int32_t o[8]) {
  int32_t a0 = s[0] + s[7], a1 = s[1] + s[6], a2 = s[2] + s[5], a3 = s[3] + s[4];
  int32_t b0 = s[0] - s[7], b1 = s[1] - s[6], b2 = s[2] - s[5], b3 = s[3] - s[4];
  int32_t e0 = a0 + a3, e1 = a1 + a2, e2 = a0 - a3, e3 = a1 - a2;
  o[0] = 8 * (e0 + e1);
  o[1] = 12 * b0 + 10 * b1 + 6 * b2 + 3 * b3;
  o[2] = 8 * e2 + 4 * e3;
  o[3] = 10 * b0 - 3 * b1 - 12 * b2 - 6 * b3;
  o[4] = 8 * (e0 - e1);
  o[5] = 6 * b0 - 12 * b1 + 3 * b2 + 10 * b3;
  o[6] = 4 * e2 - 8 * e3;
  o[7] = 3 * b0 - 6 * b1 + 10 * b2 - 12 * b3;
}

/* dec/transform.rs idct8_1d */
static void idct8_1d(const int32_t d[8], int32_t o[8]) {
  int32_t a0 = d[0] + d[4], a4 = d[0] - d[4];
  int32_t a2 = (d[2] >> 1) - d[6], a6 = d[2] + (d[6] >> 1);
  int32_t b0 = a0 + a6, b2 = a4 + a2, b4 = a4 - a2, b6 = a0 - a6;
  int32_t a1 = -d[3] + d[5] - d[7] - (d[7] >> 1);
  int32_t a3 = d[1] + d[7] - d[3] - (d[3] >> 1);
  int32_t a5 = -d[1] + d[7] + d[5] + (d[5] >> 1);
  int32_t a7 = d[3] + d[5] + d[1] + (d[1] >> 1);
  int32_t b1 = a1 + (a7 >> 2), b7 = a7 - (a1 >> 2);
  int32_t b3 = a3 + (a5 >> 2), b5 = (a3 >> 2) - a5;
  o[0] = b0 + b7;
  o[1] = b2 + b5;
  o[2] = b4 + b3;
  o[3] = b6 + b1;
  o[4] = b6 - b1;
  o[5] = b4 - b3;
  o[6] = b2 - b5;
  o[7] = b0 - b7;
}

/* slice.rs:149-165 */
uint32_t o_decimate_score(const int16_t *levels, int n) {
  static const uint32_t T[16] = {3, 2, 2, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t score = 0;
  int run = 0;
  for (int i = 0; i < n; i++) {
    int16_t l = levels[i];
    if (l == 0) {
      run++;
      continue;
    }
    if ((l < 0 ? -l : l) > 1) return 64;
    score += T[run < 15 ? run : 15];
    run = 0;
  }
  return score;
}

static inline uint8_t clip8(int32_t v) { return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }

/* slice.rs:813-878 with residual.rs luma4_fwd_body / luma4_rec_body:
 * levels[blk][k] in zigzag order per 4x4 block (BLK4 order), the
 * reconstruction into rec.  Returns cbp (bits per 8x8). */
int o_code_luma4(const uint8_t src[256], const uint8_t pred[256], int qp, int decimate, int cabac,
                 int16_t levels[16][16], uint8_t rec[256]) {
  const oquant *qt = tables();
  const int qbits = 15 + qp / 6;
  int32_t lim = cabac ? INT32_MAX : MAX_LEVEL;
  if (lim > 32767) lim = 32767;
  const int32_t f = (int32_t)((((int64_t)1 << qbits) * DZ_INTER) >> 8);
  const int32_t *mf = qt->mf4[qp % 6];
  int16_t coefs[16][16]; /* [pos][raster block] */
  for (int rb = 0; rb < 16; rb++) {
    int by = rb / 4, bx = rb % 4;
    int32_t d[16], t[16];
    for (int k = 0; k < 16; k++) {
      int i = (by * 4 + k / 4) * 16 + bx * 4 + k % 4;
      d[k] = (int32_t)src[i] - (int32_t)pred[i];
    }
    for (int i = 0; i < 4; i++) {
      int32_t a = d[i * 4], b = d[i * 4 + 1], c = d[i * 4 + 2], e = d[i * 4 + 3];
      t[i * 4] = a + e + b + c;
      t[i * 4 + 1] = 2 * (a - e) + (b - c);
      t[i * 4 + 2] = a + e - b - c;
      t[i * 4 + 3] = (a - e) - 2 * (b - c);
    }
    for (int i = 0; i < 4; i++) {
      int32_t a = t[i], b = t[4 + i], c = t[8 + i], e = t[12 + i];
      int32_t o[4] = {a + e + b + c, 2 * (a - e) + (b - c), a + e - b - c, (a - e) - 2 * (b - c)};
      for (int k = 0; k < 4; k++) {
        int pos = k * 4 + i;
        int32_t v = o[k];
        int32_t q = ((v < 0 ? -v : v) * mf[pos] + f) >> qbits;
        if (q > lim) q = lim;
        coefs[pos][rb] = (int16_t)(v < 0 ? -q : q);
      }
    }
  }
  uint32_t score8[4] = {0, 0, 0, 0};
  for (int blk = 0; blk < 16; blk++) {
    int rb = BLK4_Y[blk] * 4 + BLK4_X[blk];
    for (int k = 0; k < 16; k++) levels[blk][k] = coefs[ZIGZAG4[k]][rb];
    score8[blk / 4] += o_decimate_score(levels[blk], 16);
  }
  uint32_t total = score8[0] + score8[1] + score8[2] + score8[3];
  int cbp = 0;
  for (int b8 = 0; b8 < 4; b8++) {
    int kill = decimate && (score8[b8] < 4 || total < 6);
    for (int i = 0; i < 4; i++) {
      int blk = b8 * 4 + i;
      if (kill) {
        memset(levels[blk], 0, sizeof levels[blk]);
        int rb = BLK4_Y[blk] * 4 + BLK4_X[blk];
        for (int pos = 0; pos < 16; pos++) coefs[pos][rb] = 0;
      }
      for (int k = 0; k < 16; k++)
        if (levels[blk][k] != 0) cbp |= 1 << b8;
    }
  }
  /* residual.rs luma4_rec_body */
  const int32_t *v = qt->v4[qp % 6];
  const int sh = qp / 6;
  for (int rb = 0; rb < 16; rb++) {
    int32_t d[16], fr[16];
    for (int k = 0; k < 16; k++) d[k] = ((int32_t)coefs[k][rb] * v[k]) << sh;
    for (int i = 0; i < 4; i++) {
      int32_t d0 = d[i * 4], d1 = d[i * 4 + 1], d2 = d[i * 4 + 2], d3 = d[i * 4 + 3];
      int32_t e = d0 + d2, ff = d0 - d2, g = (d1 >> 1) - d3, h = d1 + (d3 >> 1);
      fr[i * 4] = e + h;
      fr[i * 4 + 1] = ff + g;
      fr[i * 4 + 2] = ff - g;
      fr[i * 4 + 3] = e - h;
    }
    int by = rb / 4, bx = rb % 4;
    for (int j = 0; j < 4; j++) {
      int32_t f0 = fr[j], f1 = fr[4 + j], f2 = fr[8 + j], f3 = fr[12 + j];
      int32_t e = f0 + f2, ff = f0 - f2, g = (f1 >> 1) - f3, h = f1 + (f3 >> 1);
      int32_t r[4] = {e + h, ff + g, ff - g, e - h};
      for (int y = 0; y < 4; y++) {
        int i = (by * 4 + y) * 16 + bx * 4 + j;
        rec[i] = clip8((int32_t)pred[i] + ((r[y] + 32) >> 6));
      }
    }
  }
  return cbp;
}

/* transform.rs dequant8_one */
static int32_t dequant8_one(int32_t level, int32_t v, int q6) {
  int32_t ls = 16 * v;
  if (q6 >= 6) return (level * ls) << (q6 - 6);
  return (level * ls + (1 << (5 - q6))) >> (6 - q6);
}

/* slice.rs:880-941 with residual.rs luma8_fwd_body / luma8_rec_body:
 * levels[b8][k] in 8x8 zigzag order.  Returns cbp (bits per 8x8). */
int o_code_luma8(const uint8_t src[256], const uint8_t pred[256], int qp, int decimate, int16_t levels[4][64],
                 uint8_t rec[256]) {
  const oquant *qt = tables();
  uint8_t zz8[64];
  zigzag8(zz8);
  int32_t coefs[4][64];
  uint32_t score[4];
  for (int b8 = 0; b8 < 4; b8++) {
    const int qbits = 22 + qp / 6;
    const int64_t f = (((int64_t)1 << qbits) * DZ_INTER) >> 8;
    const int32_t *mf = qt->mf8[qp % 6];
    int ox = (b8 & 1) * 8, oy = (b8 >> 1) * 8;
    int32_t t[8][8];
    for (int y = 0; y < 8; y++) {
      int32_t s[8];
      for (int x = 0; x < 8; x++) {
        int i = (oy + y) * 16 + ox + x;
        s[x] = (int32_t)src[i] - (int32_t)pred[i];
      }
      fdct8_1d(s, t[y]);
    }
    for (int x = 0; x < 8; x++) {
      int32_t col[8], o[8];
      for (int y = 0; y < 8; y++) col[y] = t[y][x];
      fdct8_1d(col, o);
      for (int k = 0; k < 8; k++) {
        int32_t v = o[k];
        uint64_t av = (uint64_t)(v < 0 ? -(int64_t)v : (int64_t)v);
        int32_t a = (int32_t)((av * (uint64_t)mf[k * 8 + x] + (uint64_t)f) >> qbits);
        coefs[b8][k * 8 + x] = v < 0 ? -a : a;
      }
    }
    for (int k = 0; k < 64; k++) levels[b8][k] = (int16_t)coefs[b8][zz8[k]];
    score[b8] = o_decimate_score(levels[b8], 64);
  }
  uint32_t total = score[0] + score[1] + score[2] + score[3];
  int cbp = 0;
  const int32_t *v = qt->v8[qp % 6];
  const int q6 = qp / 6;
  for (int b8 = 0; b8 < 4; b8++) {
    int ox = (b8 & 1) * 8, oy = (b8 >> 1) * 8;
    if (decimate && (score[b8] < 4 || total < 6)) {
      memset(levels[b8], 0, sizeof levels[b8]);
      memset(coefs[b8], 0, sizeof coefs[b8]);
    }
    int nz = 0;
    for (int k = 0; k < 64; k++) nz += levels[b8][k] != 0;
    if (nz == 0) {
      for (int y = 0; y < 8; y++) memcpy(rec + (oy + y) * 16 + ox, pred + (oy + y) * 16 + ox, 8);
      continue;
    }
    cbp |= 1 << b8;
    int32_t g[8][8];
    for (int i = 0; i < 8; i++) {
      int32_t row[8];
      for (int k = 0; k < 8; k++) row[k] = dequant8_one(coefs[b8][i * 8 + k], v[i * 8 + k], q6);
      idct8_1d(row, g[i]);
    }
    for (int x = 0; x < 8; x++) {
      int32_t col[8], o[8];
      for (int i = 0; i < 8; i++) col[i] = g[i][x];
      idct8_1d(col, o);
      for (int y = 0; y < 8; y++) {
        int i = (oy + y) * 16 + ox + x;
        rec[i] = clip8((int32_t)pred[i] + ((o[y] + 32) >> 6));
      }
    }
  }
  return cbp;
}

/* slice.rs:679-811, the coded block pattern of the chroma residual of an
 * inter MB (the reconstruction does not change it). */
int o_code_chroma(const uint8_t su[64], const uint8_t sv[64], const uint8_t pu[64], const uint8_t pv[64], int qp,
                  int chroma_qp_offset, int intra, int cabac) {
  const oquant *qt = tables();
  const int qpc = o_chroma_qp(qp + chroma_qp_offset);
  const int64_t fdz = intra ? 85 : DZ_INTER;
  int dc_any = 0, ac_any = 0;
  for (int comp = 0; comp < 2; comp++) {
    const uint8_t *s = comp == 0 ? su : sv, *p = comp == 0 ? pu : pv;
    int32_t dc[4];
    uint32_t ac_score = 0;
    int16_t cac[4][16];
    for (int b = 0; b < 4; b++) {
      int bx = (b & 1) * 4, by = (b >> 1) * 4;
      int32_t d[16], t[16];
      for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) d[y * 4 + x] = (int32_t)s[(by + y) * 8 + bx + x] - (int32_t)p[(by + y) * 8 + bx + x];
      o_fdct4(d, t);
      dc[b] = t[0];
      /* transform.rs quant4 with skip_dc */
      const int qbits = 15 + qpc / 6;
      const int64_t f = (((int64_t)1 << qbits) * fdz) >> 8;
      const int32_t *mf = qt->mf4[qpc % 6];
      for (int i = 1; i < 16; i++) {
        int32_t v = t[i];
        int32_t a = (int32_t)(((int64_t)(v < 0 ? -(int64_t)v : v) * mf[i] + f) >> qbits);
        t[i] = v < 0 ? -a : a;
      }
      if (!cabac)
        for (int i = 1; i < 16; i++) t[i] = t[i] < -MAX_LEVEL ? -MAX_LEVEL : (t[i] > MAX_LEVEL ? MAX_LEVEL : t[i]);
      cac[b][0] = 0;
      for (int k = 1; k < 16; k++) cac[b][k] = (int16_t)t[ZIGZAG4[k]];
      ac_score += o_decimate_score(cac[b] + 1, 15);
    }
    if (!intra && ac_score < 7) memset(cac, 0, sizeof cac);
    /* transform.rs quant_chroma_dc */
    int32_t h[4] = {dc[0] + dc[1] + dc[2] + dc[3], dc[0] - dc[1] + dc[2] - dc[3], dc[0] + dc[1] - dc[2] - dc[3],
                    dc[0] - dc[1] - dc[2] + dc[3]};
    const int qbits = 16 + qpc / 6;
    const int64_t f = (((int64_t)1 << qbits) * fdz) >> 8;
    const int64_t mf0 = qt->mf4[qpc % 6][0];
    for (int i = 0; i < 4; i++) {
      int32_t v = h[i];
      int32_t a = (int32_t)(((int64_t)(v < 0 ? -(int64_t)v : v) * mf0 + f) >> qbits);
      int32_t lv = v < 0 ? -a : a;
      if (!cabac) lv = lv < -MAX_LEVEL ? -MAX_LEVEL : (lv > MAX_LEVEL ? MAX_LEVEL : lv);
      dc_any |= lv != 0;
    }
    for (int b = 0; b < 4; b++)
      for (int k = 0; k < 16; k++) ac_any |= cac[b][k] != 0;
  }
  return ac_any ? 2 : (dc_any ? 1 : 0);
}

/* slice.rs:968-974 */
static double level_bits(int16_t v) { return v == 0 ? 0.4 : 3.0 + 2.0 * log2((double)(v < 0 ? -(int)v : (int)v)); }

/* dsp.rs:90-102 */
static uint64_t ssd16(const uint8_t *a, const uint8_t *b) {
  uint64_t s = 0;
  for (int r = 0; r < 16; r++) {
    uint32_t row = 0;
    for (int i = 0; i < 16; i++) {
      int32_t d = (int32_t)a[r * 16 + i] - (int32_t)b[r * 16 + i];
      row += (uint32_t)(d * d);
    }
    s += row;
  }
  return s;
}

/* slice.rs:943-1001, the coded block pattern (luma bits 0-3, chroma << 4)
 * and the 8x8-transform flag of code_inter. */
int o_code_inter(const uint8_t sy[256], const uint8_t su[64], const uint8_t sv[64], const uint8_t py[256],
                 const uint8_t pu[64], const uint8_t pv[64], int qp, int flags, int chroma_qp_offset,
                 int *t8x8_out) {
  const int cabac = (flags & OSKIP_CABAC) != 0;
  int use8, cbpl;
  int16_t l4[16][16], l8[4][64];
  uint8_t r4[256], r8[256];
  if (!(flags & OSKIP_T8X8)) {
    use8 = 0;
    cbpl = o_code_luma4(sy, py, qp, 1, cabac, l4, r4);
  } else if (!(flags & OSKIP_ADAPTIVE_T8)) {
    use8 = 1;
    cbpl = o_code_luma8(sy, py, qp, 1, l8, r8);
  } else {
    int cbp8 = o_code_luma8(sy, py, qp, 1, l8, r8);
    int cbp4 = o_code_luma4(sy, py, qp, 1, cabac, l4, r4);
    double lam = o_lambda_ssd(qp);
    double bits8 = 0.0;
    for (int b = 0; b < 4; b++)
      if ((cbp8 >> b) & 1) {
        double s = 0.0;
        for (int k = 0; k < 64; k++) s += level_bits(l8[b][k]);
        bits8 += s;
      }
    double bits4 = 0.0;
    for (int b = 0; b < 16; b++)
      if ((cbp4 >> (b / 4)) & 1) {
        double s = 0.0;
        for (int k = 0; k < 16; k++) s += level_bits(l4[b][k]);
        bits4 += 1.0 + s;
      }
    double j8 = (double)ssd16(sy, r8) + lam * bits8;
    double j4 = (double)ssd16(sy, r4) + lam * bits4;
    if (j8 <= j4) {
      use8 = 1;
      cbpl = cbp8;
    } else {
      use8 = 0;
      cbpl = cbp4;
    }
  }
  if (t8x8_out) *t8x8_out = use8 && cbpl != 0;
  int cbpc = o_code_chroma(su, sv, pu, pv, qp, chroma_qp_offset, 0, cabac);
  return cbpl | (cbpc << 4);
}

/* slice.rs:1884-1920: the early P-skip gate of the MB at (mx, my) for its
 * pskip_mv.  src_y 16x16 (stride 16), src_u / src_v 8x8 (stride 8); r the
 * reference luma, ru / rv its padded chroma planes (stride cstride, pad
 * O_CHROMA_PAD); p->vlim the vlim of the MB row. */
void o_skip_gate(const oparams *p, const oref *r, const uint8_t *ru, const uint8_t *rv, int cstride, int mx, int my,
                 omv skip_mv, const uint8_t *src_y, const uint8_t *src_u, const uint8_t *src_v, int qp, int flags,
                 int chroma_qp_offset, oskip_out *out) {
  const oquant *qt = tables();
  memset(out, 0, sizeof *out);
  out->cbp = 0xFF;
  out->thresh = ((uint32_t)(qt->v4[qp % 6][0] << (qp / 6)) * 16u) * 2u;
  const int px = mx * 16, py = my * 16;
  out->mv_ok = (uint8_t)o_mv_ok(p, px, py, 16, 16, skip_mv);
  if (!out->mv_ok) return;
  uint8_t y[256], u[64], v[64];
  o_mc_luma(r, px, py, skip_mv, 16, 16, y, 16);
  o_mc_chroma(ru, rv, cstride, O_CHROMA_PAD, px / 2, py / 2, skip_mv, 8, 8, u, v, 8);
  out->sad = o_sad16(src_y, 16, y, 16, 16);
  if (out->sad >= out->thresh) return;
  out->coded = 1;
  int t8 = 0;
  int cbp = o_code_inter(src_y, src_u, src_v, y, u, v, qp, flags, chroma_qp_offset, &t8);
  out->cbp = (uint8_t)cbp;
  out->t8x8 = (uint8_t)t8;
  out->skip = cbp == 0;
}
