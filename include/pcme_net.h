/*
 * pcme_net.h -- C ABI of the learned search pruner (MI355X / gfx950).
 *
 * No counterpart in the reference: pito-codecs has no learned part.  The
 * pruner is a small CNN that looks at a macroblock and a downsampled
 * reference window and proposes candidate motion vectors and a partition
 * mode; pcme_search_frames (include/pcme.h) then runs the exact search only
 * around them for the macroblocks where the network is confident, and the
 * full search elsewhere.  Training and inference run on hand-written HIP
 * kernels: every convolution is a bf16 MFMA GEMM with fp32 accumulation.
 *
 * Network (per macroblock):
 *   input  24x24x2: channel 0 = the 48x48 reference window centred on the
 *          macroblock (picture rect x-16..x+31, y-16..y+31 of the padded
 *          reference luma), 2x2-averaged with (a+b+c+d+2)>>2; channel 1 =
 *          the 16x16 source macroblock averaged the same way (8x8) at
 *          rows/cols 8..15, zero elsewhere.  A byte v becomes (v-128)/64
 *          (exact in bf16).
 *   conv1  3x3 stride 2 pad 1, 2 -> 32 ch, ReLU   (12x12)
 *   conv2  3x3 stride 2 pad 1, 32 -> 32, ReLU     (6x6)
 *   conv3  3x3 stride 1 pad 1, 32 -> 32, ReLU     (6x6)
 *   conv4  3x3 stride 2 pad 1, 32 -> 32, ReLU     (3x3)
 *   pool   global average (fp32)
 *   fc     32 -> 293 logits (fp32): 289 motion classes (a 17x17 grid of
 *          integer-pel offsets (2i-16, 2j-16), class 17*(j+8)+(i+8) for
 *          offset (2i, 2j)) then 4 partition classes (PCME_P*).
 *   Conv weights and activations are bf16 (round to nearest even), biases
 *   and accumulation fp32; the fc weights are used as bf16 values.
 *
 * Parameters (fp32 master copy, PCME_NET_NPARAM floats, in this order):
 *   w1[32][9][2] b1[32] w2[32][9][32] b2[32] w3[32][9][32] b3[32]
 *   w4[32][9][32] b4[32] wf[293][32] bf[293]
 *   (conv weight [co][tap][ci], tap = 3*ky + kx.)
 *
 * Errors and devices as in pcme.h.  A net lives on the device that was
 * current at pcme_net_create; calls must be made with that device current.
 * Calls on one net must not overlap (they share its workspace).
 */
#ifndef PCME_NET_H
#define PCME_NET_H

#include <stdint.h>

#include "pcme.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PCME_NET_IN 24          /* input side */
#define PCME_NET_INPUT 1152     /* bf16 values per macroblock input (24*24*2) */
#define PCME_NET_MV_SIDE 17
#define PCME_NET_MV_CLASSES 289
#define PCME_NET_CLASSES 293    /* 289 motion + 4 partition */
#define PCME_NET_NPARAM 38021
#define PCME_NET_MAXCAND 8

typedef struct pcme_net pcme_net;

/* Create a net for training batches of up to max_batch macroblocks
 * (a multiple of 32) with He-normal weights drawn from `seed` (splitmix64 +
 * Box-Muller; biases zero).  Allocates the parameters, Adam moments and the
 * training workspace on the current device. */
int pcme_net_create(uint64_t seed, int32_t max_batch, pcme_net **out);
void pcme_net_destroy(pcme_net *net);
/* Copy the fp32 parameters to / from host memory (PCME_NET_NPARAM floats).
 * set_params also resets the Adam state.  Both synchronise. */
int pcme_net_get_params(pcme_net *net, float *host);
int pcme_net_set_params(pcme_net *net, const float *host);

/* Network inputs of macroblocks: for i < n, macroblock mb[i] (raster index,
 * device int32) of the frame with padded source luma `src` and padded
 * reference luma `ref` (same geometry, mbw x mbh) is written to
 * x + i*PCME_NET_INPUT (device, bf16 bit patterns, [24][24][2]). */
int pcme_net_inputs(const uint8_t *src, const uint8_t *ref, int32_t mbw,
                    int32_t mbh, const int32_t *mb, int32_t n, uint16_t *x,
                    void *stream);

/* Forward pass of n inputs (device, n <= max_batch, n % 32 == 0) with the
 * training kernels: writes n*PCME_NET_CLASSES fp32 logits (device). */
int pcme_net_forward(pcme_net *net, const uint16_t *x, int32_t n,
                     float *logits, void *stream);

/* Loss and gradients of a batch: inputs x[idx[i]] (idx device int32, n of
 * them, n % 32 == 0, n <= max_batch), motion class mv_cls[idx[i]] and
 * partition part[idx[i]] (device int32).  loss = mean cross-entropy of the
 * motion head + mean cross-entropy of the partition head.  Writes
 * loss_out[0..2) = (total, motion, partition) and, if grad is not NULL,
 * the fp32 gradient of the loss w.r.t. the parameters (PCME_NET_NPARAM,
 * parameter order; device).  No update. */
int pcme_net_grad(pcme_net *net, const uint16_t *x, const int32_t *mv_cls,
                  const int32_t *part, const int32_t *idx, int32_t n,
                  float *loss_out, float *grad, void *stream);

/* One training step: pcme_net_grad, then Adam (beta1 0.9, beta2 0.999,
 * eps 1e-8, bias-corrected) with learning rate lr on the fp32 master
 * parameters, then the bf16 copies.  loss_out (device, 3 floats) as above. */
int pcme_net_train_step(pcme_net *net, const uint16_t *x,
                        const int32_t *mv_cls, const int32_t *part,
                        const int32_t *idx, int32_t n, float lr,
                        float *loss_out, void *stream);

/* Inference for every macroblock of n frame jobs (host array; src, ref,
 * prm.mbw, prm.mbh and hint are used): one fused kernel computes the
 * network per macroblock and writes jobs[i].hint[mb] (device; the field is
 * written through despite the const).  cand = the top-k motion classes
 * (k = ncand <= 8, by logit, ties to the lower class) as quarter-pel MVs;
 * part = argmax of the partition logits (ties lower); conf_part = its
 * softmax probability, conf_mv = the softmax mass of the k classes;
 * prune = conf_part >= thr_part && conf_mv >= thr_mv.  Logits are bit-equal
 * to pcme_net_forward's. */
int pcme_net_hints(pcme_net *net, const pcme_frame_job *jobs, int32_t n,
                   int32_t ncand, float thr_part, float thr_mv, void *stream);

#ifdef __cplusplus
}
#endif
#endif
