/*
 * pcme_synth.h -- seeded synthetic I420 video for the motion-search
 * workloads.  No counterpart in the reference (its encoder tests use hash
 * noise and a moving checkerboard, src/h264/tests_enc.rs:16-62); this
 * generator makes the benchmark content: blurred-noise background under a
 * global pan (and optional zoom), flat rectangles moving at fractional
 * velocities, additive Gaussian noise; chroma is blurred noise around 128.
 * Host code; frames are written to host memory.
 */
#ifndef PCME_SYNTH_H
#define PCME_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pcme_synth_cfg {
  int32_t w, h;          /* luma size, even */
  int32_t frames;
  int32_t nrect;         /* number of rectangles */
  int32_t rect_lo, rect_hi;  /* side lengths, uniform integers */
  float vmax;            /* rectangle velocity bound, px/frame */
  float pan_max;         /* global pan bound, px/frame */
  float zoom;            /* zoom per frame about the centre (0.002 = 0.2%) */
  float noise_sigma;     /* additive luma noise */
  uint64_t seed;
} pcme_synth_cfg;

/* Writes `frames` I420 frames: frame t's luma at y + t*w*h (stride w), its
 * chroma at u/v + t*(w/2)*(h/2) (stride w/2).  threads <= 0: all cores.
 * Returns 0 or a negative error (PCME_EINVAL). */
int pcme_synth_clip(const pcme_synth_cfg *cfg, uint8_t *y, uint8_t *u,
                    uint8_t *v, int32_t threads);

#ifdef __cplusplus
}
#endif
#endif
