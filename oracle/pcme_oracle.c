/*
 * pcme_oracle.c -- CPU oracle for the H.264 P-frame motion search.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU
 * baseline of the benchmark.  Nothing in the product library calls it.
 *
 * A plain-C restatement of pito-codecs 5.3.0 (Rust, not buildable here: no
 * cargo/rustc in the image), file by file:
 *   src/h264/enc/dsp.rs:1-88          sad, sad16, satd_strip, satd
 *   src/h264/enc/slice.rs:52-68       Lambdas::new
 *   src/h264/enc/slice.rs:130-147     se_bits, mv_bits, median
 *   src/h264/enc/slice.rs:392-507     nb_part, mvp, pskip_mv, mv_ok
 *   src/h264/enc/slice.rs:1590-1723   search
 *   src/h264/enc/slice.rs:1751-1868   cand_inter_info, mask_of, shape_of,
 *                                     search_part
 *   src/h264/enc/slice.rs:1870-1945   analyse_p (skip_mv, 16x16 search,
 *                                     partition choice)
 *   src/h264/enc/picture.rs:7-51,285-487  planes, extend, hpel_rows,
 *                                     QPEL_SRC, mc_luma, mc_chroma
 *   src/h264/enc/frame.rs:13,58-97,152-156  VMAX, extend, the ref view row
 *                                     limit; picture.rs:211-268 (lim)
 *   src/h264/enc/mod.rs:842-873       import (alignment replication)
 *   src/h264/enc/lookahead.rs:10-33   LowRes::new (o_lowres)
 *   src/h264/enc/lookahead.rs:52-108  complexity (o_complexity)
 *   src/h264/enc/mbtree.rs:21-49      intra_cost (o_lowres_intra)
 *   src/h264/enc/mbtree.rs:51-159     search, fetch, analyse (o_tree_search,
 *                                     o_tree_analyse)
 *   src/h264/enc/slice.rs:2030-2084   direct_spatial (o_direct_spatial)
 *   src/h264/enc/slice.rs:2086-2377   b_partition, analyse_b's search
 *                                     (o_b_analyse; without BSkip, RD, intra)
 *   src/h264/enc/dsp.rs:104-114       block_var (o_block_var)
 *   src/h264/enc/lookahead.rs:116-139 aq_offsets (o_aq_offsets)
 *   src/h264/enc/mbtree.rs:162-231    Node, propagate, offsets
 *                                     (o_mbtree_propagate, o_mbtree_offsets)
 *
 * How each part is pinned (tests/):
 *   o_sad/o_sad16            closed form (numpy |a-b| sum), random + extremes
 *   o_satd                   textbook 4x4 Hadamard sum(|H D H^T|)/2 in numpy,
 *                            constant-residual KAT (8a per 4x4 block)
 *   o_lambda                 closed form and the published table of values
 *   o_se_bits/o_mv_bits      Exp-Golomb se(v) length from its definition
 *   o_hpel/o_mc_luma         H.264 8.4.2.2.1 equations (b, h, j and the
 *                            quarter-sample averages) in numpy with
 *                            clamp-to-plane; impulse KAT
 *   o_mc_chroma              H.264 8.4.2.2.2 bilinear equation in numpy
 *   o_nb_part/o_mvp/o_pskip  H.264 8.4.1.3 / 8.4.1.1 restated in Python in
 *                            the decoder's structure (dec/slicedec.rs:581-681)
 *                            over random neighbour configurations
 *   o_search/o_search_part/  a second, independently written Python
 *   o_analyse_mb/            restatement on tiny frames (golden vectors
 *   o_search_frame           under tests/golden/), plus invariants: returned
 *                            cost == SATD(mc(best)) + lambda*bits, mv_ok(best),
 *                            brute-force recovery of a pure integer shift
 *   o_import_plane/o_extend  numpy edge replication
 *   o_lowres                 constant and 2x2-constant frames, the
 *                            rounding of (s + 2) >> 2 on planted blocks
 *   o_lowres_intra           predictions that a block matches exactly
 *                            (flat 128, left column, top row: cost 16),
 *                            a constant residual over a mode's prediction
 *                            (32|d| + 16), and the minimum of textbook
 *                            Hadamard SATDs over the three predictions
 *   o_complexity/            identical frames (every inter cost 8), no
 *   o_tree_analyse           previous frame (inter == intra), pure shifts
 *                            recovered exactly, an exact bi-average win
 *                            (cost 12), and a second, independently written
 *                            Python restatement on small frames
 *   o_direct_spatial         spatial direct restated in the reference
 *                            decoder's structure (dec/slicedec.rs:718-800)
 *   o_b_analyse              every cost recomputed from its MVs with the
 *                            H.264 quarter-sample equations and textbook
 *                            Hadamard; first-minimum directions
 *   o_block_var/o_aq_offsets closed forms (flat, one textured MB, clamp)
 *                            and a float32 numpy restatement (1e-5)
 *   o_mbtree_propagate       closed forms: a zero-MV edge moves intra -
 *                            min(inter, intra), whole-cell shifts, the
 *                            four-cell area split of asymmetric and negative
 *                            fractional MVs (floor, not truncation), the
 *                            half split of bi blocks, a three-node chain,
 *                            skipped references; mass conservation on random
 *                            nodes; a float64 scatter (np.add.at) on random
 *                            multi-node inputs
 *   o_mbtree_offsets         closed forms (prop 0, prop = intra, intra 0)
 */
#include "pcme_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

int o_version(void) { return 1; }

/* ------------------------------------------------------------------ dsp */

/* dsp.rs:1-14 */
uint32_t o_sad(const uint8_t *a, int sa, const uint8_t *b, int sb, int w,
               int h) {
  uint32_t s = 0;
  for (int r = 0; r < h; r++)
    for (int i = 0; i < w; i++) {
      int d = (int)a[r * sa + i] - (int)b[r * sb + i];
      s += (uint32_t)(d < 0 ? -d : d);
    }
  return s;
}

/* dsp.rs:16-33: per-row u16 accumulator, 16 wide */
uint32_t o_sad16(const uint8_t *a, int sa, const uint8_t *b, int sb, int h) {
  uint32_t s = 0;
  for (int r = 0; r < h; r++) {
    uint16_t acc = 0;
    for (int i = 0; i < 16; i++) {
      int d = (int)a[r * sa + i] - (int)b[r * sb + i];
      acc = (uint16_t)(acc + (d < 0 ? -d : d));
    }
    s += acc;
  }
  return s;
}

static inline uint32_t uabs16(int16_t v) {
  return (uint32_t)(v < 0 ? -(int32_t)v : (int32_t)v);
}

/* dsp.rs:35-70: one 4-row strip of width n (4, 8 or 16), i16 internals */
static uint32_t satd_strip(const uint8_t *a, int sa, const uint8_t *b, int sb,
                           int n) {
  int16_t d[4][16], v[4][16];
  for (int r = 0; r < 4; r++)
    for (int i = 0; i < n; i++)
      d[r][i] = (int16_t)((int16_t)a[r * sa + i] - (int16_t)b[r * sb + i]);
  for (int i = 0; i < n; i++) {
    int16_t s01 = (int16_t)(d[0][i] + d[1][i]);
    int16_t d01 = (int16_t)(d[0][i] - d[1][i]);
    int16_t s23 = (int16_t)(d[2][i] + d[3][i]);
    int16_t d23 = (int16_t)(d[2][i] - d[3][i]);
    v[0][i] = (int16_t)(s01 + s23);
    v[1][i] = (int16_t)(s01 - s23);
    v[2][i] = (int16_t)(d01 + d23);
    v[3][i] = (int16_t)(d01 - d23);
  }
  uint32_t total = 0;
  int hh = n / 2;
  for (int r = 0; r < 4; r++) {
    int16_t hs[8], hd[8];
    for (int k = 0; k < hh; k++) {
      hs[k] = (int16_t)(v[r][2 * k] + v[r][2 * k + 1]);
      hd[k] = (int16_t)(v[r][2 * k] - v[r][2 * k + 1]);
    }
    for (int k = 0; k < hh / 2; k++) {
      uint32_t x0 = uabs16(hs[2 * k]), x1 = uabs16(hs[2 * k + 1]);
      uint32_t y0 = uabs16(hd[2 * k]), y1 = uabs16(hd[2 * k + 1]);
      total += x0 > x1 ? x0 : x1;
      total += y0 > y1 ? y0 : y1;
    }
  }
  return total;
}

/* dsp.rs:72-88 */
uint32_t o_satd(const uint8_t *a, int sa, const uint8_t *b, int sb, int w,
                int h) {
  uint32_t total = 0;
  for (int by = 0; by < h; by += 4) {
    const uint8_t *ao = a + by * sa, *bo = b + by * sb;
    if (w == 16 || w == 8 || w == 4)
      total += satd_strip(ao, sa, bo, sb, w);
    else
      for (int bx = 0; bx < w; bx += 4)
        total += satd_strip(ao + bx, sa, bo + bx, sb, 4);
  }
  return total;
}

/* ------------------------------------------------------------ constants */

/* slice.rs:52-68: sad[qp] = max(16, round(256*sqrt(0.85*2^((qp-12)/3)))) */
uint32_t o_lambda(int qp) {
  double l = 0.85 * pow(2.0, ((double)qp - 12.0) / 3.0);
  double v = round(sqrt(l) * 256.0);
  if (v < 16.0) v = 16.0;
  return (uint32_t)v;
}

/* slice.rs:130-139 */
uint32_t o_se_bits(int32_t v) {
  uint32_t k = v > 0 ? 2u * (uint32_t)v - 1u : (uint32_t)(-2 * v);
  uint32_t kp = k + 1;
  int lz = __builtin_clz(kp);
  return 2u * (uint32_t)(31 - lz) + 1u;
}

/* slice.rs:141-143 */
uint32_t o_mv_bits(omv mv, omv mvp) {
  return o_se_bits((int32_t)mv.x - (int32_t)mvp.x) +
         o_se_bits((int32_t)mv.y - (int32_t)mvp.y);
}

/* slice.rs:145-147 */
int16_t o_median(int16_t a, int16_t b, int16_t c) {
  int16_t mx = a > b ? a : b, mn = a < b ? a : b;
  int16_t t = mn > c ? mn : c;
  return mx < t ? mx : t;
}

/* --------------------------------------------------------------- planes */

/* mod.rs:842-873: rows 0..ah copy source row min(y, h-1); columns w..aw
 * repeat the last copied pixel.  dst = first byte of the padded plane. */
void o_import_plane(const uint8_t *src, int sstride, int w, int h,
                    uint8_t *dst, int dstride, int pad, int aw, int ah) {
  for (int y = 0; y < ah; y++) {
    int sy = y < h - 1 ? y : h - 1;
    uint8_t *o = dst + (size_t)(y + pad) * dstride + pad;
    memcpy(o, src + (size_t)sy * sstride, (size_t)w);
    uint8_t last = o[w - 1];
    for (int x = w; x < aw; x++) o[x] = last;
  }
}

/* picture.rs:285-316 + frame.rs:58-97: horizontal pad fill of every picture
 * row, then the top pad rows copy the first row and the bottom pad rows copy
 * the last one (both including their already-filled side pads). */
void o_extend_plane(uint8_t *plane, int stride, int pad, int w, int h) {
  for (int r = pad; r < pad + h; r++) {
    uint8_t *row = plane + (size_t)r * stride;
    uint8_t l = row[pad], rr = row[pad + w - 1];
    memset(row, l, (size_t)pad);
    memset(row + pad + w, rr, (size_t)(stride - pad - w));
  }
  for (int r = 0; r < pad; r++)
    memcpy(plane + (size_t)r * stride, plane + (size_t)pad * stride,
           (size_t)stride);
  int last = pad + h - 1, rows = h + 2 * pad;
  for (int r = last + 1; r < rows; r++)
    memcpy(plane + (size_t)r * stride, plane + (size_t)last * stride,
           (size_t)stride);
}

static inline uint8_t clip8(int32_t v) {
  return (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
}

/* picture.rs:318-380 over rows 0..rows: H = horizontal 6-tap, V = vertical
 * 6-tap, C = vertical 6-tap over the i16 horizontal intermediates.  Row and
 * column indices clamp to the plane. */
void o_hpel(const uint8_t *src, int stride, int rows, uint8_t *hout,
            uint8_t *vout, uint8_t *cout) {
  int16_t *hs = (int16_t *)malloc(sizeof(int16_t) * (size_t)stride * rows);
#define AT(row, xx) \
  ((int16_t)(row)[(xx) < 0 ? 0 : ((xx) > stride - 1 ? stride - 1 : (xx))])
  for (int r = 0; r < rows; r++) {
    const uint8_t *row = src + (size_t)r * stride;
    int16_t *o = hs + (size_t)r * stride;
    for (int x = 0; x < stride; x++) {
      o[x] = (int16_t)(AT(row, x - 2) + AT(row, x + 3) -
                       5 * (AT(row, x - 1) + AT(row, x + 2)) +
                       20 * (AT(row, x) + AT(row, x + 1)));
    }
  }
#undef AT
  for (int r = 0; r < rows; r++) {
    int rr[6];
    for (int i = 0; i < 6; i++) {
      int q = r - 2 + i;
      rr[i] = q < 0 ? 0 : (q > rows - 1 ? rows - 1 : q);
    }
    for (int x = 0; x < stride; x++) {
      size_t o = (size_t)r * stride + x;
      hout[o] = clip8(((int32_t)hs[o] + 16) >> 5);
      int32_t s = (int32_t)src[(size_t)rr[0] * stride + x] +
                  (int32_t)src[(size_t)rr[5] * stride + x] -
                  5 * ((int32_t)src[(size_t)rr[1] * stride + x] +
                       (int32_t)src[(size_t)rr[4] * stride + x]) +
                  20 * ((int32_t)src[(size_t)rr[2] * stride + x] +
                        (int32_t)src[(size_t)rr[3] * stride + x]);
      vout[o] = clip8((s + 16) >> 5);
      int32_t c = (int32_t)hs[(size_t)rr[0] * stride + x] +
                  (int32_t)hs[(size_t)rr[5] * stride + x] -
                  5 * ((int32_t)hs[(size_t)rr[1] * stride + x] +
                       (int32_t)hs[(size_t)rr[4] * stride + x]) +
                  20 * ((int32_t)hs[(size_t)rr[2] * stride + x] +
                        (int32_t)hs[(size_t)rr[3] * stride + x]);
      cout[o] = clip8((c + 512) >> 10);
    }
  }
  free(hs);
}

/* picture.rs:387-404: per q = (mv.y&3)*4 + (mv.x&3), two (plane, dx, dy)
 * sources; plane 0 = full, 1 = H, 2 = V, 3 = C. */
static const uint8_t QPEL_SRC[16][2][3] = {
    {{0, 0, 0}, {0, 0, 0}}, {{0, 0, 0}, {1, 0, 0}}, {{1, 0, 0}, {1, 0, 0}},
    {{1, 0, 0}, {0, 1, 0}}, {{0, 0, 0}, {2, 0, 0}}, {{1, 0, 0}, {2, 0, 0}},
    {{1, 0, 0}, {3, 0, 0}}, {{1, 0, 0}, {2, 1, 0}}, {{2, 0, 0}, {2, 0, 0}},
    {{2, 0, 0}, {3, 0, 0}}, {{3, 0, 0}, {3, 0, 0}}, {{3, 0, 0}, {2, 1, 0}},
    {{2, 0, 0}, {0, 0, 1}}, {{2, 0, 0}, {1, 0, 1}}, {{3, 0, 0}, {1, 0, 1}},
    {{2, 1, 0}, {1, 0, 1}},
};

static inline size_t vidx(const oref *r, long x, long y) {
  return (size_t)((y + r->pad) * (long)r->stride + (x + r->pad));
}

/* picture.rs:415-449 */
void o_mc_luma(const oref *r, int x, int y, omv mv, int w, int h,
               uint8_t *out, int ostride) {
  long xi = (long)x + (mv.x >> 2);
  long yi = (long)y + (mv.y >> 2);
  int q = (mv.y & 3) * 4 + (mv.x & 3);
  const uint8_t *sa = QPEL_SRC[q][0], *sb = QPEL_SRC[q][1];
  const uint8_t *a = r->pl[sa[0]] + vidx(r, xi + sa[1], yi + sa[2]);
  if (sa[0] == sb[0] && sa[1] == sb[1] && sa[2] == sb[2]) {
    for (int rr = 0; rr < h; rr++)
      memcpy(out + rr * ostride, a + (size_t)rr * r->stride, (size_t)w);
  } else {
    const uint8_t *b = r->pl[sb[0]] + vidx(r, xi + sb[1], yi + sb[2]);
    for (int rr = 0; rr < h; rr++)
      for (int i = 0; i < w; i++)
        out[rr * ostride + i] =
            (uint8_t)(((uint16_t)a[(size_t)rr * r->stride + i] +
                       (uint16_t)b[(size_t)rr * r->stride + i] + 1) >> 1);
  }
}

/* picture.rs:451-487 */
void o_mc_chroma(const uint8_t *u, const uint8_t *v, int cstride, int cpad,
                 int x, int y, omv mv, int w, int h, uint8_t *out_u,
                 uint8_t *out_v, int ostride) {
  long xi = (long)x + (mv.x >> 3);
  long yi = (long)y + (mv.y >> 3);
  uint32_t fx = (uint32_t)(mv.x & 7), fy = (uint32_t)(mv.y & 7);
  uint32_t w00 = (8 - fx) * (8 - fy), w01 = fx * (8 - fy);
  uint32_t w10 = (8 - fx) * fy, w11 = fx * fy;
  const uint8_t *pls[2] = {u, v};
  uint8_t *outs[2] = {out_u, out_v};
  for (int k = 0; k < 2; k++) {
    const uint8_t *p = pls[k] + (size_t)((yi + cpad) * cstride + xi + cpad);
    for (int r = 0; r < h; r++) {
      const uint8_t *a = p + (size_t)r * cstride;
      const uint8_t *b = a + cstride;
      for (int i = 0; i < w; i++)
        outs[k][r * ostride + i] =
            (uint8_t)((w00 * a[i] + w01 * a[i + 1] + w10 * b[i] +
                       w11 * b[i + 1] + 32) >> 6);
    }
  }
}

/* --------------------------------------------------------------- search */

/* frame.rs:152 (need = 16*(my+1) + VMAX) with picture.rs:211-268: a view
 * of a picture of coded height 16*mbh gives lim = need while need < h, else
 * h + LUMA_PAD - 20. */
long o_vlim(int my, int mbh) {
  long need = 16L * (my + 1) + O_VMAX;
  long h = 16L * mbh;
  return need >= h ? h + O_LUMA_PAD - 20 : need;
}

/* slice.rs:496-507 */
int o_mv_ok(const oparams *p, int x, int y, int w, int h, omv mv) {
  long pw = (long)p->mbw * 16;
  long m = O_LUMA_PAD - 20;
  long x0 = (long)x + (mv.x >> 2);
  long y0 = (long)y + (mv.y >> 2);
  int ax = mv.x < 0 ? -(int)mv.x : mv.x;
  int ay = mv.y < 0 ? -(int)mv.y : mv.y;
  return ax < 2000 && ay < 2000 && x0 >= -m && y0 >= -m && x0 + w <= pw + m &&
         y0 + h < p->vlim;
}

static inline int is_intra(int8_t k) {
  return k == OK_I4x4 || k == OK_I8x8 || k == OK_I16x16;
}

/* mbinfo.rs:109-115 */
static inline int8_t ref_at(const ombinfo *m, int r) {
  int x = r & 3, y = r >> 2;
  return m->ref_idx[(y >> 1) * 2 + (x >> 1)];
}

/* slice.rs:392-432 */
int o_nb_part(const oparams *p, const ombinfo *mbs, const ombinfo *cur,
              uint16_t mask, int mx, int my, int bx, int by, int8_t *ref,
              omv *mv) {
  static const omv ZERO = {0, 0};
  if (bx >= 0 && bx < 4 && by >= 0 && by < 4) {
    int r = by * 4 + bx;
    if (mask & (1u << r)) {
      *ref = ref_at(cur, r);
      *mv = cur->mv[r];
      return 1;
    }
    return 0;
  }
  int dx = bx < 0 ? -1 : (bx >= 4 ? 1 : 0);
  int dy = by < 0 ? -1 : 0;
  if (by >= 4 || (dy == 0 && dx == 1)) return 0;
  int nx = mx + dx, ny = my + dy;
  if (nx < 0 || nx >= p->mbw || ny < 0) return 0;
  const ombinfo *m = &mbs[ny * p->mbw + nx];
  int r = ((by + 4) % 4) * 4 + (bx + 4) % 4;
  if (is_intra(m->kind)) {
    *ref = -1;
    *mv = ZERO;
    return 1;
  }
  int8_t rf = ref_at(m, r);
  *ref = rf;
  *mv = rf >= 0 ? m->mv[r] : ZERO;
  return 1;
}

/* slice.rs:434-482 */
omv o_mvp(const oparams *p, const ombinfo *mbs, const ombinfo *cur,
          uint16_t mask, int mx, int my, int x, int y, int w, int8_t rf,
          int shape) {
  int8_t ar = -1, br = -1, cr = -1;
  omv am = {0, 0}, bm = {0, 0}, cm = {0, 0};
  int ha = o_nb_part(p, mbs, cur, mask, mx, my, x - 1, y, &ar, &am);
  int hb = o_nb_part(p, mbs, cur, mask, mx, my, x, y - 1, &br, &bm);
  int hc = o_nb_part(p, mbs, cur, mask, mx, my, x + w, y - 1, &cr, &cm);
  if (!hc) hc = o_nb_part(p, mbs, cur, mask, mx, my, x - 1, y - 1, &cr, &cm);
  if (!hb && !hc && ha) {
    hb = hc = 1;
    br = cr = ar;
    bm = cm = am;
  }
  if (!ha) { ar = -1; am.x = am.y = 0; }
  if (!hb) { br = -1; bm.x = bm.y = 0; }
  if (!hc) { cr = -1; cm.x = cm.y = 0; }
  switch (shape) {
  case 1: if (br == rf) return bm; break;
  case 2: if (ar == rf) return am; break;
  case 3: if (ar == rf) return am; break;
  case 4: if (cr == rf) return cm; break;
  default: break;
  }
  int n = (ar == rf) + (br == rf) + (cr == rf);
  if (n == 1) {
    if (ar == rf) return am;
    if (br == rf) return bm;
    return cm;
  }
  omv o;
  o.x = o_median(am.x, bm.x, cm.x);
  o.y = o_median(am.y, bm.y, cm.y);
  return o;
}

/* slice.rs:484-494 */
omv o_pskip_mv(const oparams *p, const ombinfo *mbs, int mx, int my) {
  ombinfo cur;
  memset(&cur, 0, sizeof cur);
  for (int i = 0; i < 4; i++) cur.ref_idx[i] = -1;
  omv z = {0, 0};
  int8_t ar, br;
  omv am, bm;
  int ha = o_nb_part(p, mbs, &cur, 0, mx, my, -1, 0, &ar, &am);
  int hb = o_nb_part(p, mbs, &cur, 0, mx, my, 0, -1, &br, &bm);
  if (!ha || !hb) return z;
  if (ar == 0 && am.x == 0 && am.y == 0) return z;
  if (br == 0 && bm.x == 0 && bm.y == 0) return z;
  return o_mvp(p, mbs, &cur, 0, mx, my, 0, 0, 4, 0, 0);
}

static inline omv mvnew(int32_t x, int32_t y) {
  omv m;
  m.x = (int16_t)x;
  m.y = (int16_t)y;
  return m;
}
static inline int mveq(omv a, omv b) { return a.x == b.x && a.y == b.y; }

typedef struct {
  const oparams *p;
  const oref *r;
  int bx, by, w, h;
  const uint8_t *src;
  int sstride;
  omv mvp;
  uint32_t lam;
  ocounters *ctr;
} sctx;

/* slice.rs:1605-1616 (fcost) */
static uint32_t fcost(const sctx *s, omv mv) {
  const uint8_t *pl =
      s->r->pl[0] + vidx(s->r, (long)s->bx + (mv.x >> 2), (long)s->by + (mv.y >> 2));
  uint32_t d = s->w == 16 ? o_sad16(s->src, s->sstride, pl, s->r->stride, s->h)
                          : o_sad(s->src, s->sstride, pl, s->r->stride, s->w, s->h);
  if (s->ctr) {
    s->ctr->sad_px += (uint64_t)s->w * s->h;
    s->ctr->sad_evals++;
  }
  return d + ((s->lam * o_mv_bits(mv, s->mvp)) >> 8);
}

/* slice.rs:1684-1687 (scost) */
static uint32_t scost(const sctx *s, omv mv) {
  uint8_t buf[256];
  o_mc_luma(s->r, s->bx, s->by, mv, s->w, s->h, buf, 16);
  if (s->ctr) {
    s->ctr->satd_px += (uint64_t)s->w * s->h;
    s->ctr->satd_evals++;
  }
  return o_satd(s->src, s->sstride, buf, 16, s->w, s->h) +
         ((s->lam * o_mv_bits(mv, s->mvp)) >> 8);
}

static const int SQ8[8][2] = {{-1, 0}, {1, 0},  {0, -1}, {0, 1},
                              {-1, -1}, {1, -1}, {-1, 1}, {1, 1}};

/* slice.rs:1590-1723 */
omv o_search(const oparams *p, const oref *r, int bx, int by, int w, int h,
             const uint8_t *src, int sstride, omv mvp, const omv *cands,
             int ncands, uint32_t lam, uint32_t *cost_out, ocounters *ctr) {
  return o_search_ex(p, r, bx, by, w, h, src, sstride, mvp, cands, ncands,
                     NULL, -1, lam, cost_out, ctr);
}

/* slice.rs:1590-1723; with nh >= 0 the hint scan replaces lines 1632-1656 */
omv o_search_ex(const oparams *p, const oref *r, int bx, int by, int w, int h,
                const uint8_t *src, int sstride, omv mvp, const omv *cands,
                int ncands, const omv *hc, int nh, uint32_t lam,
                uint32_t *cost_out, ocounters *ctr) {
  sctx s = {p, r, bx, by, w, h, src, sstride, mvp, lam, ctr};
  if (ctr) ctr->searches++;
  omv best = mvnew(((int32_t)mvp.x + 2) & ~3, ((int32_t)mvp.y + 2) & ~3);
  if (!o_mv_ok(p, bx, by, w, h, best)) best = mvnew(0, 0);
  uint32_t bcost = fcost(&s, best);
  for (int i = 0; i < ncands; i++) {
    omv c = mvnew(((int32_t)cands[i].x + 2) & ~3, ((int32_t)cands[i].y + 2) & ~3);
    if (!mveq(c, best) && o_mv_ok(p, bx, by, w, h, c)) {
      uint32_t k = fcost(&s, c);
      if (k < bcost) {
        bcost = k;
        best = c;
      }
    }
  }
  for (int i = 0; i < nh; i++) {
    omv c = mvnew(((int32_t)hc[i].x + 2) & ~3, ((int32_t)hc[i].y + 2) & ~3);
    if (!mveq(c, best) && o_mv_ok(p, bx, by, w, h, c)) {
      uint32_t k = fcost(&s, c);
      if (k < bcost) {
        bcost = k;
        best = c;
      }
    }
  }
  int range = nh < 0 ? p->me_range : 0;
  omv centre0 = best;
  static const int HEX[6][2] = {{-2, 0}, {2, 0}, {-1, -2}, {1, -2}, {-1, 2}, {1, 2}};
  for (int it = 0; it < range; it++) {
    int moved = 0;
    omv c = best;
    for (int j = 0; j < 6; j++) {
      omv m = mvnew((int32_t)c.x + HEX[j][0] * 4, (int32_t)c.y + HEX[j][1] * 4);
      int ddx = (int32_t)m.x - (int32_t)centre0.x;
      int ddy = (int32_t)m.y - (int32_t)centre0.y;
      if (((ddx < 0 ? -ddx : ddx) >> 2) > range ||
          ((ddy < 0 ? -ddy : ddy) >> 2) > range ||
          !o_mv_ok(p, bx, by, w, h, m))
        continue;
      uint32_t k = fcost(&s, m);
      if (k < bcost) {
        bcost = k;
        best = m;
        moved = 1;
      }
    }
    if (!moved) break;
  }
  for (int it = 0; it < 2; it++) {
    omv c = best;
    for (int j = 0; j < 8; j++) {
      omv m = mvnew((int32_t)c.x + SQ8[j][0] * 4, (int32_t)c.y + SQ8[j][1] * 4);
      if (!o_mv_ok(p, bx, by, w, h, m)) continue;
      uint32_t k = fcost(&s, m);
      if (k < bcost) {
        bcost = k;
        best = m;
      }
    }
    if (mveq(best, c)) break;
  }
  uint32_t bs = scost(&s, best);
  for (int step = 2; step >= 1; step--) {
    int rounds = step == 1 ? 1 + p->subpel_rounds : 1;
    for (int it = 0; it < rounds; it++) {
      omv c = best;
      for (int j = 0; j < 8; j++) {
        omv m = mvnew((int32_t)c.x + SQ8[j][0] * step, (int32_t)c.y + SQ8[j][1] * step);
        if (!o_mv_ok(p, bx, by, w, h, m)) continue;
        uint32_t k = scost(&s, m);
        if (k < bs) {
          bs = k;
          best = m;
        }
      }
      if (mveq(best, c)) break;
    }
  }
  if (cost_out) *cost_out = bs;
  return best;
}

/* mbinfo.rs:64-71 */
static void part_rect(int part, int i, int *x, int *y, int *w, int *h) {
  switch (part) {
  case OP_16x16: *x = 0; *y = 0; *w = 16; *h = 16; break;
  case OP_16x8: *x = 0; *y = 8 * i; *w = 16; *h = 8; break;
  case OP_8x16: *x = 8 * i; *y = 0; *w = 8; *h = 16; break;
  default: *x = 8 * (i & 1); *y = 8 * (i >> 1); *w = 8; *h = 8; break;
  }
}
static int part_count(int part) {
  return part == OP_16x16 ? 1 : (part == OP_8x8 ? 4 : 2);
}

/* slice.rs:1787-1798 */
static uint16_t mask_of(int part, int upto) {
  uint16_t m = 0;
  for (int q = 0; q < upto; q++) {
    int x, y, w, h;
    part_rect(part, q, &x, &y, &w, &h);
    for (int by = y / 4; by < (y + h) / 4; by++)
      for (int bx = x / 4; bx < (x + w) / 4; bx++) m |= (uint16_t)(1u << (by * 4 + bx));
  }
  return m;
}

/* slice.rs:1800-1808 */
static int shape_of(int part, int q) {
  if (part == OP_16x8) return q == 0 ? 1 : 2;
  if (part == OP_8x16) return q == 0 ? 3 : 4;
  return 0;
}

static uint32_t search_part_ex(const oparams *p, const ombinfo *mbs, int mx,
                               int my, const oref *r, int rf, int part,
                               const uint8_t *src16, uint32_t lam, omv seed,
                               const omv *hc, int nh, omv mvs[4],
                               omv mvps[4], ocounters *ctr);

/* slice.rs:1810-1868 (list 0) */
uint32_t o_search_part(const oparams *p, const ombinfo *mbs, int mx, int my,
                       const oref *r, int part, const uint8_t *src16,
                       uint32_t lam, omv seed, omv mvs[4], omv mvps[4],
                       ocounters *ctr) {
  return search_part_ex(p, mbs, mx, my, r, 0, part, src16, lam, seed, NULL, -1,
                        mvs, mvps, ctr);
}

/* With reference index rf (list 0): the predictor takes the neighbours of
 * that index (o_mvp's rf) and the searched partitions are committed with
 * it; rf 0 is the reference's search_part. */
static uint32_t search_part_ex(const oparams *p, const ombinfo *mbs, int mx,
                               int my, const oref *r, int rf, int part,
                               const uint8_t *src16, uint32_t lam, omv seed,
                               const omv *hc, int nh, omv mvs[4],
                               omv mvps[4], ocounters *ctr) {
  int px = mx * 16, py = my * 16;
  ombinfo info;
  memset(&info, 0, sizeof info);
  info.kind = OK_NONE;
  for (int i = 0; i < 4; i++) info.ref_idx[i] = -1;
  for (int i = 0; i < 4; i++) mvs[i] = mvps[i] = mvnew(0, 0);
  uint32_t total = 0;
  for (int q = 0; q < part_count(part); q++) {
    int x, y, w, h;
    part_rect(part, q, &x, &y, &w, &h);
    uint16_t mask = mask_of(part, q);
    omv mvp = o_mvp(p, mbs, &info, mask, mx, my, x / 4, y / 4, w / 4, (int8_t)rf,
                    shape_of(part, q));
    omv cands[2] = {seed, mvs[0]};
    uint32_t cost;
    omv mv = o_search_ex(p, r, px + x, py + y, w, h, src16 + y * 16 + x, 16,
                         mvp, cands, 2, hc, nh, lam, &cost, ctr);
    mvs[q] = mv;
    mvps[q] = mvp;
    total += cost;
    for (int by = y / 4; by < (y + h) / 4; by++)
      for (int bx = x / 4; bx < (x + w) / 4; bx++) info.mv[by * 4 + bx] = mv;
    for (int by8 = y / 8; by8 < (y + h + 7) / 8; by8++)
      for (int bx8 = x / 8; bx8 < (x + w + 7) / 8; bx8++)
        info.ref_idx[by8 * 2 + bx8] = (int8_t)rf;
  }
  return total;
}

/* slice.rs:1884-1945 without early skip / intra / RD, and the commit of
 * cand_inter_info (slice.rs:1751-1785) with refs [[0;4],[-1;4]]. */
void o_analyse_mb(const oparams *p, ombinfo *mbs, int mx, int my,
                  const oref *r, const uint8_t *src_plane, int sstride,
                  int spad, uint32_t lam, omb_out *out, ocounters *ctr) {
  o_analyse_mb_ex(p, mbs, mx, my, r, src_plane, sstride, spad, lam, NULL, out,
                  ctr);
}

/* Length of ref_idx's te(v) code with num_ref_idx_active = nref (H.264
 * 7.4.5.1, 9.1): none with one reference, one bit with two, else ue(v). */
uint32_t o_ref_bits(int r, int nref) {
  if (nref <= 1) return 0;
  if (nref == 2) return 1;
  uint32_t k = (uint32_t)r + 1, n = 0;
  while (k > 1) {
    k >>= 1;
    n++;
  }
  return 2 * n + 1;
}

/* o_analyse_mb; a pruned MB (h->prune) passes the hint candidates to every
 * search and searches only the hinted modes. */
void o_analyse_mb_ex(const oparams *p, ombinfo *mbs, int mx, int my,
                     const oref *r, const uint8_t *src_plane, int sstride,
                     int spad, uint32_t lam, const ohint *h, omb_out *out,
                     ocounters *ctr) {
  o_analyse_mb_mr(p, mbs, mx, my, r, 1, src_plane, sstride, spad, lam, h, out, ctr);
}

/* o_analyse_mb_ex over nref references of list 0 (refs[0] = ref_idx 0, the
 * newest).  No counterpart in the reference, whose P slices use one
 * reference: the P16x16 search runs on every reference, each with the
 * predictor of its index and with (lam * te(v) bits of the index) >> 8
 * added to its cost; the least total (the lower index on ties) gives the
 * MB's reference, and the partition searches run on it, each partition
 * paying the index's rate.  A pruned MB's hint candidates (motion over one
 * picture) are scaled by the reference's distance, index + 1.  A reference
 * other than the newest is read without the row limit (it is complete).
 * With nref = 1 this is o_analyse_mb_ex. */
void o_analyse_mb_mr(const oparams *p, ombinfo *mbs, int mx, int my,
                     const oref *refs, int nref, const uint8_t *src_plane,
                     int sstride, int spad, uint32_t lam, const ohint *h,
                     omb_out *out, ocounters *ctr) {
  const int pr = h != NULL && h->prune != 0;
  const omv *hc = pr ? h->cand : NULL;
  const int nh = pr ? (h->ncand > 8 ? 8 : (int)h->ncand) : -1;
  uint8_t src16[256];
  int px = mx * 16, py = my * 16;
  for (int rr = 0; rr < 16; rr++)
    memcpy(src16 + rr * 16,
           src_plane + (size_t)(py + rr + spad) * sstride + px + spad, 16);
  omv skip = o_pskip_mv(p, mbs, mx, my);
  oparams pall = *p;
  pall.vlim = o_vlim(p->mbh - 1, p->mbh);
  omv mv16[4], mvp16[4];
  uint32_t cost16 = 0;
  int rs = 0;
  omv hs[4][8];  /* the hint candidates scaled to each reference's distance */
  for (int rf = 0; rf < nref && pr; rf++)
    for (int i = 0; i < 8; i++) hs[rf][i] = mvnew((int32_t)hc[i].x * (rf + 1), (int32_t)hc[i].y * (rf + 1));
  for (int rf = 0; rf < nref; rf++) {
    omv m[4], pp[4];
    const uint32_t c =
        search_part_ex(rf == 0 ? p : &pall, mbs, mx, my, &refs[rf], rf, OP_16x16, src16, lam, skip,
                       pr ? hs[rf] : NULL, nh, m, pp, ctr) +
        ((lam * o_ref_bits(rf, nref)) >> 8);
    if (rf == 0 || c < cost16) {
      cost16 = c;
      rs = rf;
      memcpy(mv16, m, sizeof m);
      memcpy(mvp16, pp, sizeof pp);
    }
  }
  const oparams *ps = rs == 0 ? p : &pall;
  const uint32_t rcost = (lam * o_ref_bits(rs, nref)) >> 8;
  int best_part = OP_16x16;
  omv best_mvs[4], best_mvps[4];
  memcpy(best_mvs, mv16, sizeof mv16);
  memcpy(best_mvps, mvp16, sizeof mvp16);
  uint64_t best_cost = (uint64_t)cost16 + (uint64_t)(lam >> 8);
  out->cost[0] = cost16;
  out->cost[1] = out->cost[2] = out->cost[3] = 0xFFFFFFFFu;
  if (p->partitions && (uint64_t)cost16 > 256u * 4u) {
    static const int parts[3] = {OP_16x8, OP_8x16, OP_8x8};
    for (int k = 0; k < 3; k++) {
      if (pr && (h->part == OP_8x8 ? parts[k] != OP_8x8
                                   : (h->part == OP_16x16 || parts[k] == OP_8x8)))
        continue;
      omv m[4], pp[4];
      uint32_t c = search_part_ex(ps, mbs, mx, my, &refs[rs], rs, parts[k], src16, lam,
                                  mv16[0], pr ? hs[rs] : NULL, nh, m, pp, ctr) +
                   (uint32_t)part_count(parts[k]) * rcost;
      out->cost[parts[k]] = c;
      uint32_t extra = parts[k] == OP_8x8 ? 10 : 3;
      uint64_t cc = (uint64_t)c + (uint64_t)((lam * extra) >> 8);
      if (cc < best_cost) {
        best_cost = cc;
        best_part = parts[k];
        memcpy(best_mvs, m, sizeof m);
        memcpy(best_mvps, pp, sizeof pp);
      }
    }
  }
  ombinfo *mi = &mbs[my * p->mbw + mx];
  memset(mi, 0, sizeof *mi);
  mi->kind = OK_PINTER;
  for (int q = 0; q < part_count(best_part); q++) {
    int x, y, w, h;
    part_rect(best_part, q, &x, &y, &w, &h);
    for (int by = y / 4; by < (y + h) / 4; by++)
      for (int bx = x / 4; bx < (x + w) / 4; bx++) mi->mv[by * 4 + bx] = best_mvs[q];
    for (int by8 = y / 8; by8 < (y + h + 7) / 8; by8++)
      for (int bx8 = x / 8; bx8 < (x + w + 7) / 8; bx8++)
        mi->ref_idx[by8 * 2 + bx8] = (int8_t)rs;
  }
  memcpy(out->mv, mi->mv, sizeof mi->mv);
  memcpy(out->mvs, best_mvs, sizeof best_mvs);
  memcpy(out->mvps, best_mvps, sizeof best_mvps);
  out->best_cost = (uint32_t)best_cost;
  out->skip_mv = skip;
  out->mv16 = mv16[0];
  out->part = (uint8_t)best_part;
  out->ref = (uint8_t)rs;
  out->pad_[0] = out->pad_[1] = 0;
  if (ctr) ctr->mbs++;
}

void o_search_frame(int mbw, int mbh, int me_range, int subpel_rounds,
                    int partitions, uint32_t lam, const uint8_t *src,
                    int sstride, int spad, const oref *r, omb_out *out,
                    ocounters *ctr) {
  o_search_frame_ex(mbw, mbh, me_range, subpel_rounds, partitions, lam, src,
                    sstride, spad, r, NULL, out, ctr);
}

void o_search_frame_ex(int mbw, int mbh, int me_range, int subpel_rounds,
                       int partitions, uint32_t lam, const uint8_t *src,
                       int sstride, int spad, const oref *r,
                       const ohint *hints, omb_out *out, ocounters *ctr) {
  o_search_frame_mr(mbw, mbh, me_range, subpel_rounds, partitions, lam, src, sstride, spad, r, 1, hints, out,
                    ctr);
}

void o_search_frame_mr(int mbw, int mbh, int me_range, int subpel_rounds,
                       int partitions, uint32_t lam, const uint8_t *src,
                       int sstride, int spad, const oref *refs, int nref,
                       const ohint *hints, omb_out *out, ocounters *ctr) {
  ombinfo *mbs = (ombinfo *)calloc((size_t)mbw * mbh, sizeof(ombinfo));
  oparams p = {mbw, mbh, me_range, subpel_rounds, partitions, 0};
  for (int my = 0; my < mbh; my++) {
    p.vlim = o_vlim(my, mbh);
    for (int mx = 0; mx < mbw; mx++)
      o_analyse_mb_mr(&p, mbs, mx, my, refs, nref, src, sstride, spad, lam,
                      hints ? &hints[my * mbw + mx] : NULL,
                      &out[my * mbw + mx], ctr);
  }
  free(mbs);
}

/* ------------------------------------------------- frame-thread driver */

typedef struct {
  const ojob *jobs;
  int njobs;
  int next;
  int count; /* 0: no counters (a timed run counts nothing) */
  pthread_mutex_t mu;
  ocounters total;
} pool_t;

static void *worker(void *arg) {
  pool_t *pl = (pool_t *)arg;
  ocounters c;
  memset(&c, 0, sizeof c);
  for (;;) {
    pthread_mutex_lock(&pl->mu);
    int j = pl->next++;
    pthread_mutex_unlock(&pl->mu);
    if (j >= pl->njobs) break;
    const ojob *jb = &pl->jobs[j];
    int stride = 16 * jb->mbw + 2 * O_LUMA_PAD;
    int rows = 16 * jb->mbh + 2 * O_LUMA_PAD;
    size_t n = (size_t)stride * rows;
    const int nref = jb->nref < 1 ? 1 : (jb->nref > 4 ? 4 : jb->nref);
    uint8_t *hp = (uint8_t *)malloc(3 * n * (size_t)nref);
    oref r[4];
    for (int k = 0; k < nref; k++) {
      const uint8_t *luma = k == 0 ? jb->ref : jb->older[k - 1];
      uint8_t *h = hp + 3 * n * (size_t)k;
      o_hpel(luma, stride, rows, h, h + n, h + 2 * n);
      oref rk = {{luma, h, h + n, h + 2 * n}, stride, O_LUMA_PAD};
      r[k] = rk;
    }
    o_search_frame_mr(jb->mbw, jb->mbh, jb->me_range, jb->subpel_rounds,
                      jb->partitions, o_lambda(jb->qp), jb->src, stride,
                      O_LUMA_PAD, r, nref, NULL, jb->out, pl->count ? &c : NULL);
    free(hp);
  }
  pthread_mutex_lock(&pl->mu);
  pl->total.sad_px += c.sad_px;
  pl->total.satd_px += c.satd_px;
  pl->total.sad_evals += c.sad_evals;
  pl->total.satd_evals += c.satd_evals;
  pl->total.searches += c.searches;
  pl->total.mbs += c.mbs;
  pthread_mutex_unlock(&pl->mu);
  return NULL;
}

double o_search_jobs_mt(const ojob *jobs, int njobs, int threads,
                        ocounters *ctr) {
  if (threads < 1) threads = 1;
  pool_t pl;
  memset(&pl, 0, sizeof pl);
  pl.jobs = jobs;
  pl.njobs = njobs;
  pl.count = ctr != NULL;
  pthread_mutex_init(&pl.mu, NULL);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)threads);
  for (int i = 0; i < threads; i++) pthread_create(&th[i], NULL, worker, &pl);
  for (int i = 0; i < threads; i++) pthread_join(th[i], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  free(th);
  pthread_mutex_destroy(&pl.mu);
  if (ctr) *ctr = pl.total;
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ------------------------------------------- lookahead complexity (C6) */

/* LowRes::new (lookahead.rs:10-33): the coded luma plane (pw x ph inside a
 * plane of the given stride and pad) 2x2-averaged with rounding. */
void o_lowres(const uint8_t *plane, int stride, int pad, int pw, int ph, uint8_t *out) {
  const int w = pw / 2, h = ph / 2;
  for (int y = 0; y < h; y++) {
    const uint8_t *r0 = plane + (size_t)(2 * y + pad) * stride + pad;
    const uint8_t *r1 = r0 + stride;
    for (int x = 0; x < w; x++) {
      const uint32_t s = (uint32_t)r0[2 * x] + r0[2 * x + 1] + r1[2 * x] + r1[2 * x + 1];
      out[(size_t)y * w + x] = (uint8_t)((s + 2) >> 2);
    }
  }
}

/* mbtree.rs:21-49: least SATD of the 8x8 block at (x0, y0) against the
 * average of its top row and left column (128 at the picture's top or left
 * edge), its left column and its top row (the last two only where they
 * exist), plus 16. */
uint32_t o_lowres_intra(const uint8_t *cur, int w, int x0, int y0) {
  uint8_t blk[64];
  uint32_t best = 0xFFFFFFFFu;
  const uint8_t *base = cur + (size_t)y0 * w + x0;
  for (int mode = 0; mode < 3; mode++) {
    if ((mode == 1 && x0 == 0) || (mode == 2 && y0 == 0)) continue;
    for (int y = 0; y < 8; y++)
      for (int x = 0; x < 8; x++) {
        uint8_t v;
        if (mode == 0)
          v = (x0 > 0 && y0 > 0)
                  ? (uint8_t)(((uint32_t)cur[(size_t)(y0 - 1) * w + x0 + x] + cur[(size_t)(y0 + y) * w + x0 - 1]) / 2)
                  : 128;
        else if (mode == 1)
          v = cur[(size_t)(y0 + y) * w + x0 - 1];
        else
          v = cur[(size_t)(y0 - 1) * w + x0 + x];
        blk[y * 8 + x] = v;
      }
    const uint32_t s = o_satd(base, w, blk, 8, 8, 8);
    if (s < best) best = s;
  }
  return best + 16;
}

/* lookahead.rs:52-108 (complexity): for every 8x8 block of the w x h lowres
 * frame cur, the intra cost; with a previous lowres frame, a diamond SAD
 * search from (0, 0) -- up to 8 rounds of (-1,0),(1,0),(0,-1),(0,1),(-2,0),
 * (2,0),(0,-2),(0,2) around the round's centre, strict-< in that order,
 * candidates outside the frame never taken, stop when the best stays -- then
 * SATD at the best + 8, capped by the intra cost.  Sums into *intra and
 * *inter; per-block intra, inter and MV (x, y) when the arrays are given. */
void o_complexity(const uint8_t *cur, const uint8_t *prev, int w, int h, uint64_t *intra, uint64_t *inter,
                  uint32_t *blk_intra, uint32_t *blk_inter, int16_t *blk_mv) {
  static const int DX[8] = {-1, 1, 0, 0, -2, 2, 0, 0};
  static const int DY[8] = {0, 0, -1, 1, 0, 0, -2, 2};
  const int bw = w / 8, bh = h / 8;
  uint64_t ci = 0, cp = 0;
  for (int by = 0; by < bh; by++)
    for (int bx = 0; bx < bw; bx++) {
      const int x0 = 8 * bx, y0 = 8 * by, b = by * bw + bx;
      const uint8_t *base = cur + (size_t)y0 * w + x0;
      const uint32_t ic = o_lowres_intra(cur, w, x0, y0);
      ci += ic;
      uint32_t pc = ic;
      int mx = 0, my = 0;
      if (prev) {
        uint32_t bc = o_sad(base, w, prev + (size_t)y0 * w + x0, w, 8, 8);
        for (int it = 0; it < 8; it++) {
          const int cx = mx, cy = my;
          for (int k = 0; k < 8; k++) {
            const int x = x0 + cx + DX[k], y = y0 + cy + DY[k];
            if (x < 0 || y < 0 || x + 8 > w || y + 8 > h) continue;
            const uint32_t c = o_sad(base, w, prev + (size_t)y * w + x, w, 8, 8);
            if (c < bc) {
              bc = c;
              mx = cx + DX[k];
              my = cy + DY[k];
            }
          }
          if (mx == cx && my == cy) break;
        }
        const uint32_t sc = o_satd(base, w, prev + (size_t)(y0 + my) * w + x0 + mx, w, 8, 8) + 8;
        pc = sc < ic ? sc : ic;
      }
      cp += pc;
      if (blk_intra) blk_intra[b] = ic;
      if (blk_inter) blk_inter[b] = pc;
      if (blk_mv) {
        blk_mv[2 * b] = (int16_t)mx;
        blk_mv[2 * b + 1] = (int16_t)my;
      }
    }
  *intra = ci;
  *inter = cp;
}

/* mbtree.rs:51-102 (search): SAD + |mx| + |my| from (0, 0), then the seed
 * when strictly cheaper, then up to 16 rounds of (-1,0),(1,0),(0,-1),(0,1),
 * (-2,-2),(2,-2),(-2,2),(2,2) around the round's centre, strict-< in that
 * order, MVs whose block leaves the frame never taken. */
static uint32_t o_tree_cost(const uint8_t *cur, const uint8_t *r, int w, int h, int x0, int y0, int mx, int my) {
  const int x = x0 + mx, y = y0 + my;
  if (x < 0 || y < 0 || x > w - 8 || y > h - 8) return 0xFFFFFFFFu;
  const uint32_t mvc = (uint32_t)(abs(mx) + abs(my));
  return o_sad(cur + (size_t)y0 * w + x0, w, r + (size_t)y * w + x, w, 8, 8) + mvc;
}

omv o_tree_search(const uint8_t *cur, const uint8_t *r, int w, int h, int x0, int y0, omv seed) {
  static const int DX[8] = {-1, 1, 0, 0, -2, 2, -2, 2};
  static const int DY[8] = {0, 0, -1, 1, -2, -2, 2, 2};
  int bx = 0, by = 0;
  uint32_t bc = o_tree_cost(cur, r, w, h, x0, y0, 0, 0);
  const uint32_t sc = o_tree_cost(cur, r, w, h, x0, y0, seed.x, seed.y);
  if (sc < bc) {
    bc = sc;
    bx = seed.x;
    by = seed.y;
  }
  for (int it = 0; it < 16; it++) {
    const int cx = bx, cy = by;
    for (int k = 0; k < 8; k++) {
      const uint32_t c = o_tree_cost(cur, r, w, h, x0, y0, cx + DX[k], cy + DY[k]);
      if (c < bc) {
        bc = c;
        bx = cx + DX[k];
        by = cy + DY[k];
      }
    }
    if (bx == cx && by == cy) break;
  }
  omv m = {(int16_t)bx, (int16_t)by};
  return m;
}

/* mbtree.rs:104-111 (fetch) */
static void o_tree_fetch(const uint8_t *r, int w, int x0, int y0, omv m, uint8_t out[64]) {
  for (int row = 0; row < 8; row++) memcpy(out + 8 * row, r + (size_t)(y0 + m.y + row) * w + x0 + m.x, 8);
}

/* mbtree.rs:113-159 (analyse): per 8x8 block of the w x h lowres frame cur,
 * the intra cost and the best of list 0, list 1 (each the search above,
 * seeded with the left block's MV of that list, then SATD + 8) and their
 * rounded average (SATD + 12).  A list whose reference is NULL is skipped.
 * out[b] gets {intra, inter, list (0 intra, 1 L0, 2 L1, 3 bi), mv[2]}, with
 * mv as the reference records it: at a list-0 win mv[1] is still (0, 0). */
void o_tree_analyse(const uint8_t *cur, const uint8_t *l0, const uint8_t *l1, int w, int h, otree_block *out) {
  const int bw = w / 8, bh = h / 8;
  const uint8_t *refs[2] = {l0, l1};
  for (int by = 0; by < bh; by++) {
    omv left[2] = {{0, 0}, {0, 0}};
    for (int bx = 0; bx < bw; bx++) {
      const int x0 = 8 * bx, y0 = 8 * by;
      const uint8_t *base = cur + (size_t)y0 * w + x0;
      const uint32_t intra = o_lowres_intra(cur, w, x0, y0);
      uint32_t best = intra;
      uint8_t list = 0;
      omv bmv[2] = {{0, 0}, {0, 0}}, mvs[2] = {{0, 0}, {0, 0}};
      uint8_t p[2][64], bi[64];
      for (int l = 0; l < 2; l++) {
        if (!refs[l]) continue;
        const omv m = o_tree_search(cur, refs[l], w, h, x0, y0, left[l]);
        mvs[l] = m;
        o_tree_fetch(refs[l], w, x0, y0, m, p[l]);
        const uint32_t c = o_satd(base, w, p[l], 8, 8, 8) + 8;
        if (c < best) {
          best = c;
          list = (uint8_t)(1 + l);
          bmv[0] = mvs[0];
          bmv[1] = mvs[1];
        }
      }
      if (l0 && l1) {
        for (int i = 0; i < 64; i++) bi[i] = (uint8_t)(((uint32_t)p[0][i] + p[1][i] + 1) >> 1);
        const uint32_t c = o_satd(base, w, bi, 8, 8, 8) + 12;
        if (c < best) {
          best = c;
          list = 3;
          bmv[0] = mvs[0];
          bmv[1] = mvs[1];
        }
      }
      left[0] = mvs[0];
      left[1] = mvs[1];
      otree_block *o = &out[by * bw + bx];
      o->intra = intra;
      o->inter = best;
      o->mv[0] = bmv[0];
      o->mv[1] = bmv[1];
      o->list = list;
      o->reserved[0] = o->reserved[1] = o->reserved[2] = 0;
    }
  }
}

/* The CPU baseline of the lookahead: complexity (against the previous
 * frame) and analyse (list 0 = the previous frame) of frames 1..n-1 of n
 * w x h lowres frames, frames split over `threads` threads.  Returns
 * seconds of wall time. */
typedef struct {
  const uint8_t *lr;
  int w, h, n, threads, t;
  uint64_t *sums;
  otree_block *blocks;
} o_la_arg;

static void *o_la_worker(void *p) {
  o_la_arg *a = (o_la_arg *)p;
  const size_t fs = (size_t)a->w * a->h, nb = (size_t)(a->w / 8) * (a->h / 8);
  for (int f = 1 + a->t; f < a->n; f += a->threads) {
    o_complexity(a->lr + f * fs, a->lr + (f - 1) * fs, a->w, a->h, &a->sums[2 * f], &a->sums[2 * f + 1], NULL, NULL,
                 NULL);
    o_tree_analyse(a->lr + f * fs, a->lr + (f - 1) * fs, NULL, a->w, a->h, a->blocks + (size_t)f * nb);
  }
  return NULL;
}

double o_lookahead_mt(const uint8_t *lr, int w, int h, int n, int threads, uint64_t *sums, otree_block *blocks) {
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  if (threads < 1) threads = 1;
  pthread_t th[64];
  o_la_arg args[64];
  if (threads > 64) threads = 64;
  for (int t = 0; t < threads; t++) {
    args[t] = (o_la_arg){lr, w, h, n, threads, t, sums, blocks};
    pthread_create(&th[t], NULL, o_la_worker, &args[t]);
  }
  for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* ------------------------------------------------------ B macroblocks (C7) */

/* The list-l view of a two-list MbInfo array: nb_part (slice.rs:392-432)
 * reads only kind, ref_idx[l] and mv[l], so the list-0 functions above run
 * any list on its projection. */
static void project_list(const ombinfo2 *mbs, int n, int l, ombinfo *out) {
  for (int i = 0; i < n; i++) {
    out[i].kind = mbs[i].kind;
    memcpy(out[i].ref_idx, mbs[i].ref_idx[l], 4);
    memcpy(out[i].mv, mbs[i].mv[l], sizeof out[i].mv);
  }
}

static int8_t minp(int8_t x, int8_t y) { return (x >= 0 && y >= 0) ? (x < y ? x : y) : (x > y ? x : y); }

/* slice.rs:2030-2084 direct_spatial; col = the list-1 picture's MbInfo at
 * (mx, my). */
static void direct_spatial_proj(const oparams *p, const ombinfo *const proj2[2], const ombinfo2 *col, int mx,
                                int my, int8_t out_ref[2][4], omv out_mv[2][4]) {
  ombinfo cur;
  memset(&cur, 0, sizeof cur);
  for (int i = 0; i < 4; i++) cur.ref_idx[i] = -1;
  int8_t refs[2] = {-1, -1};
  omv mvp[2] = {{0, 0}, {0, 0}};
  for (int l = 0; l < 2; l++) {
    const ombinfo *proj = proj2[l];
    int8_t a = -1, b = -1, c = -1, r;
    omv m;
    if (o_nb_part(p, proj, &cur, 0, mx, my, -1, 0, &r, &m)) a = r;
    if (o_nb_part(p, proj, &cur, 0, mx, my, 0, -1, &r, &m)) b = r;
    if (o_nb_part(p, proj, &cur, 0, mx, my, 4, -1, &r, &m)) c = r;
    else if (o_nb_part(p, proj, &cur, 0, mx, my, -1, -1, &r, &m)) c = r;
    refs[l] = minp(a, minp(b, c));
  }
  const int zero = refs[0] < 0 && refs[1] < 0;
  if (zero) {
    refs[0] = refs[1] = 0;
  } else {
    for (int l = 0; l < 2; l++)
      if (refs[l] >= 0) mvp[l] = o_mvp(p, proj2[l], &cur, 0, mx, my, 0, 0, 4, refs[l], 0);
  }
  static const int corner[4] = {0, 3, 12, 15};
  for (int q = 0; q < 4; q++) {
    int col_zero = 0;
    if (!is_intra(col->kind)) {
      int8_t rc;
      omv mvc;
      if (col->ref_idx[0][q] >= 0) {
        rc = col->ref_idx[0][q];
        mvc = col->mv[0][corner[q]];
      } else {
        rc = col->ref_idx[1][q];
        mvc = col->mv[1][corner[q]];
      }
      col_zero = rc == 0 && abs(mvc.x) <= 1 && abs(mvc.y) <= 1;
    }
    for (int l = 0; l < 2; l++) {
      out_ref[l][q] = refs[l];
      out_mv[l][q] = mvnew(0, 0);
      if (refs[l] >= 0) out_mv[l][q] = (zero || (refs[l] == 0 && col_zero)) ? mvnew(0, 0) : mvp[l];
    }
  }
}

void o_direct_spatial(const oparams *p, const ombinfo2 *mbs, const ombinfo2 *col, int mx, int my,
                      int8_t out_ref[2][4], omv out_mv[2][4]) {
  const int n = p->mbw * p->mbh;
  ombinfo *proj = (ombinfo *)malloc(sizeof(ombinfo) * 2 * (size_t)n);
  project_list(mbs, n, 0, proj);
  project_list(mbs, n, 1, proj + n);
  const ombinfo *pl[2] = {proj, proj + n};
  direct_spatial_proj(p, pl, col, mx, my, out_ref, out_mv);
  free(proj);
}

/* picture.rs:491-495 avg_into */
static void avg_into(uint8_t *d, const uint8_t *s, int n) {
  for (int i = 0; i < n; i++) d[i] = (uint8_t)(((uint32_t)d[i] + s[i] + 1) >> 1);
}

/* The motion search of analyse_b (slice.rs:2189-2377) and b_partition
 * (slice.rs:2086-2187) for the B MB at (mx, my), without the parts that
 * need the residual coder or RD (BSkip, rd_cost, intra): direct_spatial
 * and the direct prediction's cost, the list-0/list-1 16x16 search_part
 * seeded with the direct MVs, the three direction costs, and the 16x8 and
 * 8x16 b_partition searches (seeded with the 16x16 MVs) with their
 * per-partition directions and costs.  The caller's decision picks among
 * them (direct, 16x16 dir, partitions when best.cost > 1024). */
static void b_analyse_proj(const oparams *p, const ombinfo *const proj2[2], const ombinfo2 *col, int mx, int my,
                           const oref *r0, const oref *r1, const uint8_t *src16, uint32_t lam, ob_out *o) {
  memset(o, 0, sizeof *o);
  const int px = 16 * mx, py = 16 * my;
  const oref *rp[2] = {r0, r1};
  direct_spatial_proj(p, proj2, col, mx, my, o->direct_ref, o->direct_mv);
  int ok = 1;
  for (int q = 0; q < 4; q++)
    for (int l = 0; l < 2; l++)
      if (o->direct_ref[l][q] >= 0 && !o_mv_ok(p, px + 8 * (q & 1), py + 8 * (q >> 1), 8, 8, o->direct_mv[l][q]))
        ok = 0;
  o->direct_ok = (uint32_t)ok;
  o->direct_cost = 0xFFFFFFFFu;
  if (ok) {
    uint8_t y[256], t[256];
    for (int q = 0; q < 4; q++) {
      const int x = 8 * (q & 1), yy = 8 * (q >> 1);
      const int u0 = o->direct_ref[0][q] >= 0, u1 = o->direct_ref[1][q] >= 0;
      if (u0) o_mc_luma(r0, px + x, py + yy, o->direct_mv[0][q], 8, 8, y + yy * 16 + x, 16);
      if (u1) {
        if (u0) {
          o_mc_luma(r1, px + x, py + yy, o->direct_mv[1][q], 8, 8, t + yy * 16 + x, 16);
          for (int rr = 0; rr < 8; rr++) avg_into(y + (yy + rr) * 16 + x, t + (yy + rr) * 16 + x, 8);
        } else {
          o_mc_luma(r1, px + x, py + yy, o->direct_mv[1][q], 8, 8, y + yy * 16 + x, 16);
        }
      }
    }
    o->direct_cost = o_satd(src16, 16, y, 16, 16, 16) + ((lam * 2) >> 8);
  }
  omv mvs[4], mvps[4];
  for (int l = 0; l < 2; l++) {
    o->cost16[l] = o_search_part(p, proj2[l], mx, my, rp[l], OP_16x16, src16, lam, o->direct_mv[l][0], mvs, mvps,
                                 NULL);
    o->mv16[l] = mvs[0];
    o->mvp16[l] = mvps[0];
  }
  uint8_t pr[2][256], bi[256];
  for (int l = 0; l < 2; l++) o_mc_luma(rp[l], px, py, o->mv16[l], 16, 16, pr[l], 16);
  memcpy(bi, pr[0], 256);
  avg_into(bi, pr[1], 256);
  const uint32_t bits = o_mv_bits(o->mv16[0], o->mvp16[0]) + o_mv_bits(o->mv16[1], o->mvp16[1]);
  o->cost[0] = o->cost16[0] + ((lam * 3) >> 8);
  o->cost[1] = o->cost16[1] + ((lam * 3) >> 8);
  o->cost[2] = o_satd(src16, 16, bi, 16, 16, 16) + ((lam * (bits + 5)) >> 8);
  o->dir = 0;
  for (int d = 1; d < 3; d++)
    if (o->cost[d] < o->cost[o->dir]) o->dir = (uint32_t)d;
  for (int k = 0; k < 2; k++) {
    const int part = k == 0 ? OP_16x8 : OP_8x16;
    for (int l = 0; l < 2; l++) {
      o_search_part(p, proj2[l], mx, my, rp[l], part, src16, lam, o->mv16[l], mvs, mvps, NULL);
      o->part_mv[k][l][0] = mvs[0];
      o->part_mv[k][l][1] = mvs[1];
    }
    for (int l = 0; l < 2; l++)
      for (int pi = 0; pi < 2; pi++) {
        int x, y, w, h;
        part_rect(part, pi, &x, &y, &w, &h);
        o_mc_luma(rp[l], px + x, py + y, o->part_mv[k][l][pi], w, h, pr[l] + y * 16 + x, 16);
      }
    memcpy(bi, pr[0], 256);
    avg_into(bi, pr[1], 256);
    const uint8_t *srcs[3] = {pr[0], pr[1], bi};
    uint32_t cost = (lam * 8) >> 8;
    for (int pi = 0; pi < 2; pi++) {
      int x, y, w, h;
      part_rect(part, pi, &x, &y, &w, &h);
      uint32_t bc = 0xFFFFFFFFu;
      int bd = 0;
      for (int d = 0; d < 3; d++) {
        const uint32_t b = d == 2 ? 14 : 7;
        const uint32_t c = o_satd(src16 + y * 16 + x, 16, srcs[d] + y * 16 + x, 16, w, h) + ((lam * b) >> 8);
        if (c < bc) {
          bc = c;
          bd = d;
        }
      }
      cost += bc;
      o->part_dir[k][pi] = (uint8_t)bd;
    }
    o->part_cost[k] = cost;
  }
}

void o_b_analyse(const oparams *p, const ombinfo2 *mbs, const ombinfo2 *col, int mx, int my, const oref *r0,
                 const oref *r1, const uint8_t *src16, uint32_t lam, ob_out *o) {
  const int n = p->mbw * p->mbh;
  ombinfo *proj = (ombinfo *)malloc(sizeof(ombinfo) * 2 * (size_t)n);
  project_list(mbs, n, 0, proj);
  project_list(mbs, n, 1, proj + n);
  const ombinfo *pl[2] = {proj, proj + n};
  b_analyse_proj(p, pl, col, mx, my, r0, r1, src16, lam, o);
  free(proj);
}

/* The CPU baseline of the B search: o_b_analyse of every MB of a frame
 * (source plane padded like the references; cols[mb] the co-located MBs),
 * MB rows split over `threads` threads.  Returns seconds of wall time. */
typedef struct {
  const oparams *p;
  const ombinfo *proj[2];
  const ombinfo2 *cols;
  const oref *r0, *r1;
  const uint8_t *src;
  int sstride, pad, threads, t;
  uint32_t lam;
  ob_out *out;
} o_b_arg;

static void *o_b_worker(void *a_) {
  o_b_arg *a = (o_b_arg *)a_;
  for (int my = a->t; my < a->p->mbh; my += a->threads) {
    oparams q = *a->p;
    q.vlim = o_vlim(my, q.mbh);
    for (int mx = 0; mx < q.mbw; mx++) {
      uint8_t s16[256];
      for (int r = 0; r < 16; r++)
        memcpy(s16 + 16 * r, a->src + (size_t)(a->pad + 16 * my + r) * a->sstride + a->pad + 16 * mx, 16);
      b_analyse_proj(&q, a->proj, &a->cols[my * q.mbw + mx], mx, my, a->r0, a->r1, s16, a->lam,
                     &a->out[my * q.mbw + mx]);
    }
  }
  return NULL;
}

double o_b_frame_mt(const oparams *p, const ombinfo2 *mbs, const ombinfo2 *cols, const oref *r0, const oref *r1,
                    const uint8_t *src, int sstride, int pad, uint32_t lam, int threads, ob_out *out) {
  struct timespec t0, t1;
  const int n = p->mbw * p->mbh;
  ombinfo *proj = (ombinfo *)malloc(sizeof(ombinfo) * 2 * (size_t)n);
  project_list(mbs, n, 0, proj);
  project_list(mbs, n, 1, proj + n);
  clock_gettime(CLOCK_MONOTONIC, &t0);
  if (threads < 1) threads = 1;
  if (threads > 64) threads = 64;
  pthread_t th[64];
  o_b_arg args[64];
  for (int t = 0; t < threads; t++) {
    args[t] = (o_b_arg){p, {proj, proj + n}, cols, r0, r1, src, sstride, pad, threads, t, lam, out};
    pthread_create(&th[t], NULL, o_b_worker, &args[t]);
  }
  for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  free(proj);
  return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}

/* slice.rs:2189-2377 with rd off and without BSkip and intra: the decision
 * among the searches of b_analyse_proj, then cand_inter_info's commit
 * (slice.rs:1751-1785): ref_idx per 8x8 and MV per 4x4 and list, zero MVs
 * where a list is unused. */
static void b_decide(int partitions, const ob_out *o, ob_mb *m) {
  m->search = *o;
  uint64_t best = 0;
  int mode = -1;
  if (o->direct_ok) {
    best = o->direct_cost;
    mode = 0;
  }
  if (mode < 0 || (uint64_t)o->cost[o->dir] < best) {
    best = o->cost[o->dir];
    mode = 1;
  }
  if (partitions && best > 256u * 4u)
    for (int k = 0; k < 2; k++)
      if ((uint64_t)o->part_cost[k] < best) {
        best = o->part_cost[k];
        mode = 2 + k;
      }
  m->mode = (uint8_t)mode;
  m->dir = (uint8_t)o->dir;
  m->best_cost = (uint32_t)best;
  m->reserved[0] = m->reserved[1] = 0;
  for (int l = 0; l < 2; l++)
    for (int r = 0; r < 16; r++) {
      const int q = ((r >> 3) << 1) | ((r & 3) >> 1);  /* the 8x8 of 4x4 block r */
      int8_t rf;
      omv mv;
      if (mode == 0) {
        rf = o->direct_ref[l][q];
        mv = o->direct_mv[l][q];
      } else if (mode == 1) {
        rf = (o->dir == 2 || o->dir == (uint32_t)l) ? 0 : -1;
        mv = o->mv16[l];
      } else {
        const int k = mode - 2;
        const int pi = k == 0 ? (r >> 3) : ((r & 3) >> 1);  /* 16x8: top/bottom; 8x16: left/right */
        const int d = o->part_dir[k][pi];
        rf = (d == 2 || d == l) ? 0 : -1;
        mv = o->part_mv[k][l][pi];
      }
      if (rf < 0) mv = mvnew(0, 0);
      m->mv[l][r] = mv;
      m->ref[l][q] = rf;
    }
}

void o_b_field(const oparams *p0, const ombinfo2 *cols, const oref *r0, const oref *r1, const uint8_t *src,
               int sstride, int pad, uint32_t lam, ob_mb *out) {
  const int n = p0->mbw * p0->mbh;
  ombinfo *proj = (ombinfo *)calloc(2 * (size_t)n, sizeof(ombinfo));
  const ombinfo *pl[2] = {proj, proj + n};
  oparams p = *p0;
  for (int my = 0; my < p.mbh; my++) {
    p.vlim = o_vlim(my, p.mbh);
    for (int mx = 0; mx < p.mbw; mx++) {
      const int i = my * p.mbw + mx;
      uint8_t s16[256];
      for (int r = 0; r < 16; r++)
        memcpy(s16 + 16 * r, src + (size_t)(pad + 16 * my + r) * sstride + pad + 16 * mx, 16);
      ob_out o;
      b_analyse_proj(&p, pl, &cols[i], mx, my, r0, r1, s16, lam, &o);
      b_decide(p.partitions, &o, &out[i]);
      for (int l = 0; l < 2; l++) {
        ombinfo *c = proj + (size_t)l * n + i;
        c->kind = OK_PINTER;
        memcpy(c->ref_idx, out[i].ref[l], 4);
        memcpy(c->mv, out[i].mv[l], sizeof c->mv);
      }
    }
  }
  free(proj);
}

/* ------------------------------------------------------------- AQ (C8) */

/* dsp.rs:104-114 block_var: sum of squares minus square of sum over w*h,
 * the quotient in 64 bits. */
uint32_t o_block_var(const uint8_t *a, int sa, int w, int h) {
  uint32_t s = 0, ss = 0;
  for (int r = 0; r < h; r++)
    for (int x = 0; x < w; x++) {
      const uint32_t v = a[(size_t)r * sa + x];
      s += v;
      ss += v * v;
    }
  return ss - (uint32_t)(((uint64_t)s * (uint64_t)s) / (uint64_t)(w * h));
}

/* lookahead.rs:116-139 aq_offsets: per MB log2(var_y/256 + 0.5*(var_u/64 +
 * var_v/64) + 1) in f32, minus the mean over the picture (summed in MB
 * order), times strength, clamped to +-8.  y, u, v point at picture (0, 0)
 * of planes with strides ys, cs.  Built with -ffp-contract=off: every f32
 * operation rounds on its own, as in the reference. */
void o_aq_offsets(const uint8_t *y, int ys, const uint8_t *u, const uint8_t *v, int cs, int mbw, int mbh,
                  float strength, float *out) {
  const int n = mbw * mbh;
  for (int my = 0; my < mbh; my++)
    for (int mx = 0; mx < mbw; mx++) {
      const float vy = (float)o_block_var(y + (size_t)(16 * my) * ys + 16 * mx, ys, 16, 16) / 256.0f;
      const float vu = (float)o_block_var(u + (size_t)(8 * my) * cs + 8 * mx, cs, 8, 8) / 64.0f;
      const float vv = (float)o_block_var(v + (size_t)(8 * my) * cs + 8 * mx, cs, 8, 8) / 64.0f;
      out[my * mbw + mx] = log2f(vy + 0.5f * (vu + vv) + 1.0f);
    }
  float sum = 0.0f;
  for (int i = 0; i < n; i++) sum += out[i];
  const float mean = sum / (float)n;
  for (int i = 0; i < n; i++) {
    float o = strength * (out[i] - mean);
    out[i] = o < -8.0f ? -8.0f : (o > 8.0f ? 8.0f : o);
  }
}

/* ------------------------------------------------- macroblock tree (C9) */

static int div_floor(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

/* mbtree.rs:162-223 propagate: nodes n-1..0; node i's blocks move
 * (intra + prop_i) * (intra - min(inter, intra)) / intra of their cost into
 * the 8x8 cells their MV lands on in the earlier nodes they reference,
 * split by area (list 3: half to each list).  prop receives n * bw * bh
 * floats (zeroed first).  refs[2*i + l] = node index or -1. */
void o_mbtree_propagate(const otree_block *const *blocks, const int *refs, int n, int bw, int bh, float *prop) {
  const int nb = bw * bh;
  memset(prop, 0, sizeof(float) * (size_t)n * (size_t)nb);
  for (int i = n - 1; i >= 0; i--) {
    const float *mine = prop + (size_t)i * nb;
    for (int by = 0; by < bh; by++)
      for (int bx = 0; bx < bw; bx++) {
        const int k = by * bw + bx;
        const otree_block *b = &blocks[i][k];
        if (b->list == 0 || b->intra == 0) continue;
        const float intra = (float)b->intra;
        const float inter0 = (float)b->inter;
        const float inter = inter0 < intra ? inter0 : intra;
        const float amount = (intra + mine[k]) * (intra - inter) / intra;
        if (amount <= 0.0f) continue;
        const int nl = b->list == 3 ? 2 : 1;
        for (int j = 0; j < nl; j++) {
          const int l = b->list == 3 ? j : b->list - 1;
          const float w = b->list == 3 ? 0.5f : 1.0f;
          const int r = refs[2 * i + l];
          if (r < 0 || r >= i) continue;
          float *dst = prop + (size_t)r * nb;
          const int x = bx * 8 + b->mv[l].x, y = by * 8 + b->mv[l].y;
          const int cx = div_floor(x, 8), cy = div_floor(y, 8);
          const int fx = x - 8 * cx, fy = y - 8 * cy;
          const float a = amount * w / 64.0f;
          const int area[4] = {(8 - fx) * (8 - fy), fx * (8 - fy), (8 - fx) * fy, fx * fy};
          for (int c = 0; c < 4; c++) {
            const int tx = cx + (c & 1), ty = cy + (c >> 1);
            if (area[c] == 0 || tx < 0 || ty < 0 || tx >= bw || ty >= bh) continue;
            dst[ty * bw + tx] += a * (float)area[c];
          }
        }
      }
  }
}

/* mbtree.rs:225-231 offsets: -2 * log2((max(intra, 1) + prop) / max(intra, 1)). */
void o_mbtree_offsets(const otree_block *blocks, const float *prop, int nb, float *out) {
  for (int k = 0; k < nb; k++) {
    const float intra = (float)(blocks[k].intra > 1 ? blocks[k].intra : 1);
    out[k] = -2.0f * log2f((intra + prop[k]) / intra);
  }
}
