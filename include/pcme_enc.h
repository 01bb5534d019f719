/* pcme_enc.h -- the field coder (host C++, libpcme_h264.so): a Baseline
 * CAVLC H.264 encoder that codes a given motion field -- the ME-only field
 * of pcme_search_frames, plain or pruned -- into a real bitstream, so that a
 * field's rate-distortion cost is measured on bits that the decoder of
 * pcme_dec.h reproduces exactly.
 *
 * No counterpart in the reference: its encoder (src/h264/enc/) decides with
 * RD, CABAC and intra analysis.  This coder makes no decision of its own
 * beyond the residual: one slice per picture; the first picture an IDR of
 * I_PCM macroblocks; every later picture P with one reference (the previous
 * reconstruction, deblocked), each MB coded with the field's partition and
 * MVs (P_L0 16x16 / 16x8 / 8x16, P_8x8 with 8x8 sub-partitions), its
 * residual by the 4x4 transform with a dead-zone quantiser (rounding offset
 * 1/6 of a step, luma and chroma), and P_Skip where the field's 16x16 MV is
 * the skip MV and no level survives.  Constant QP per picture, deblocking
 * on with zero offsets.  Two fields coded at the same QPs differ only in
 * their motion, which is what a BD-rate between them measures. */
#ifndef PCME_ENC_H
#define PCME_ENC_H

#include <stddef.h>
#include <stdint.h>

#include "pcme.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pcme_h264enc pcme_h264enc;

/* A coder for width x height 4:2:0 pictures (even sizes; coded size the
 * next multiple of 16, cropped in the SPS). */
int pcme_h264enc_create(int32_t width, int32_t height, pcme_h264enc **out);
void pcme_h264enc_destroy(pcme_h264enc *enc);

/* Codes one picture: y (rows ystride apart), u and v (half size, rows
 * cstride apart), host memory.  mbs == null: an IDR picture (preceded by
 * the SPS and PPS the first time); else a P picture coded with the field
 * (mbw*mbh pcme_mb_out records in raster order, host memory: part, mv[16]).
 * Appends the picture's NAL units (Annex B) to out, at most cap bytes;
 * *len = bytes written.  PCME_EINVAL on bad arguments, a field MV outside
 * +-2048 quarter samples or an output buffer too small. */
int pcme_h264enc_picture(pcme_h264enc *enc, const uint8_t *y, int32_t ystride, const uint8_t *u,
                         const uint8_t *v, int32_t cstride, const pcme_mb_out *mbs, int32_t qp,
                         uint8_t *out, size_t cap, size_t *len);

/* The reconstruction of the last coded picture (width x height, cropped),
 * as a decoder of the stream outputs it. */
int pcme_h264enc_recon(pcme_h264enc *enc, uint8_t *y, int32_t ystride, uint8_t *u, uint8_t *v,
                       int32_t cstride);

/* Reference indices of the following P pictures (1..PCME_MAX_REFS, default
 * 1): list 0 is the last nref reconstructions, newest first (fewer after an
 * IDR picture), num_ref_idx_l0_active the count in use, and each MB's
 * pcme_mb_out.ref names its reference (coded as te(v); P_Skip only with
 * index 0).  A record naming an index beyond those in use is refused. */
int pcme_h264enc_set_refs(pcme_h264enc *enc, int32_t nref);

/* Macroblock counts of the last coded picture: P_Skip, other P, I_PCM. */
int pcme_h264enc_stats(pcme_h264enc *enc, int32_t *skipped, int32_t *inter, int32_t *pcm);

#ifdef __cplusplus
}
#endif

#endif
