/* pcme_dec.h -- a minimal H.264 decoder (host C++, libpcme_dec.so): the
 * Baseline-profile subset that checks encoded bitstreams decode bit-exact.
 *
 * Scope: progressive 4:2:0 8-bit pictures, CAVLC, I and P slices (I_NxN,
 * I_16x16, I_PCM, P_L0 16x16/16x8/8x16/8x8 with every sub-partition,
 * P_Skip), several slices per picture, several reference frames (sliding
 * window and memory-management operations, reference list
 * modification), the deblocking filter, frame cropping.  Anything else
 * (CABAC, B slices, interlace, FMO/ASO, 8x8 transform, scaling matrices,
 * weighted prediction) is rejected with PCME_EINVAL and a message.
 *
 * Replaces, for that subset, the reference's decoder
 * (src/h264/dec/decoder.rs Decoder::decode / flush, slicedec.rs,
 * inter.rs, intra.rs, cavlc.rs, deblock.rs, transform.rs) as its test
 * harness uses it (src/h264/tests.rs:121-189).  Status codes are those of
 * pcme.h; messages via pcme_h264dec_last_error(). */
#ifndef PCME_DEC_H
#define PCME_DEC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pcme_h264dec pcme_h264dec;

/* A decoded picture's geometry: the cropped luma size (chroma is half in
 * both directions) and its picture order count. */
typedef struct pcme_h264_frame_info {
  int32_t width, height;
  int32_t poc;
  int32_t frame_num;
} pcme_h264_frame_info;

/* A decoded macroblock, as the next queued picture holds it (raster MB
 * order).  144 bytes. */
typedef struct pcme_h264_mb {
  int8_t kind;        /* 0 I_4x4, 1 I_16x16, 2 I_PCM, 3 P_L0 (any partition), 4 P_Skip */
  int8_t part;        /* P_L0: 0 16x16, 1 16x8, 2 8x16, 3 8x8 (P_8x8 and P_8x8ref0); else -1 */
  int8_t sub[4];      /* P_8x8: sub_mb_type per 8x8 (0 8x8, 1 8x4, 2 4x8, 3 4x4); else -1 */
  int8_t ref[4];      /* refIdxL0 per 8x8 (raster); -1 for intra */
  uint8_t qp;         /* QP_Y */
  uint8_t reserved;
  int16_t mv[16][2];  /* mvL0 per 4x4 block (raster), quarter samples */
  int16_t mvp[16][2]; /* the predictor (8.4.1.3, or the P_Skip MV of 8.4.1.1) of the
                         partition covering each 4x4 block; 0 for intra */
  int32_t slice;      /* slice index within the picture */
} pcme_h264_mb;

/* Decoder::new (decoder.rs). */
int pcme_h264dec_create(pcme_h264dec **out);
void pcme_h264dec_destroy(pcme_h264dec *dec);

/* Decoder::decode (decoder.rs): feed Annex-B bytes holding whole NAL units
 * (any number; an access unit per call like the reference's packets, or a
 * whole stream).  Pictures completed by this call are queued for output in
 * output order. */
int pcme_h264dec_decode(pcme_h264dec *dec, const uint8_t *data, size_t size);

/* Decoder::flush (decoder.rs): complete the picture in progress and queue
 * every remaining picture. */
int pcme_h264dec_flush(pcme_h264dec *dec);

/* The next queued picture: returns 1 and fills *info (a peek: planes may be
 * null) or copies the picture and dequeues it when y, u and v are all
 * given (rows ystride / cstride bytes apart); returns 0 when none is
 * queued, a negative status on bad arguments. */
int pcme_h264dec_frame(pcme_h264dec *dec, pcme_h264_frame_info *info, uint8_t *y, int32_t ystride, uint8_t *u,
                       uint8_t *v, int32_t cstride);

/* The macroblocks of the next queued picture (the one pcme_h264dec_frame
 * would return): copies min(n, MBs) records to out (may be null) and
 * returns the picture's MB count, 0 when none is queued. */
int pcme_h264dec_frame_mbs(pcme_h264dec *dec, pcme_h264_mb *out, int32_t n);

/* 8.4.2.2.1 (inter.rs:73-186): the luma prediction of a bw x bh block at
 * (x, y) of a w x h picture (rows `stride` bytes apart) displaced by the
 * quarter-sample MV (mvx, mvy), reference samples clamped to the picture;
 * into out (rows ostride apart). */
int pcme_h264_mc_luma(const uint8_t *plane, int32_t stride, int32_t w, int32_t h, int32_t x, int32_t y,
                      int32_t mvx, int32_t mvy, int32_t bw, int32_t bh, uint8_t *out, int32_t ostride);

/* The message of the last failure on this thread. */
const char *pcme_h264dec_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
