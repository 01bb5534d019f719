/*
 * pcme.h -- C ABI of the MI355X (gfx950) H.264 P-frame motion search.
 *
 * The reference (pito-codecs 5.3.0) has no FFI on this path: its search is
 * private Rust inside the CPU H.264 encoder.  Each entry point below names
 * the Rust item it replaces (file:line under src/h264/enc/) and keeps that
 * item's argument meaning.  Plain pointers and sizes only.
 *
 * Device pointers: every pointer marked (device) must be device memory
 * reachable by the current HIP device.  Launches go to `stream` (a
 * hipStream_t, NULL = the default stream) and are asynchronous unless
 * stated otherwise.
 *
 * Threads and streams: every entry point may be called from several host
 * threads at once, each on its own stream, as the reference's frame threads
 * call its search (enc/mod.rs:207-238).  The per-call entry points
 * (pcme_search_batch, pcme_search_part_batch, pcme_b_analyse_batch,
 * pcme_b_field, pcme_skip_gate_batch, pcme_mbtree_propagate) take a
 * workspace of their own for each call (job table, node windows) from a
 * per-device pool; a workspace returns to the
 * pool when the call's last kernel has completed (a HIP event), so calls on
 * different streams run concurrently and never share a buffer a running
 * kernel reads.  Launches of one plan (pcme_plan_*, and the plan
 * pcme_search_frames caches per device) share that plan's workspace and
 * are serialised on the GPU: each eager launch waits for the previous one,
 * whatever stream it was issued on.  Results never depend on the stream,
 * the thread or the batch composition.
 *
 * Errors: 0 = success, negative = failure (PCME_E*).  The reference's search
 * is infallible and its frame threads turn a panic into a sticky
 * Error::Failed for the frame (enc/mod.rs:228-233, 630-642); a caller maps a
 * non-zero status onto that same path.  pcme_last_error() returns a
 * thread-local message for the last failure.  There is no silent fallback:
 * without a usable gfx950 device every compute entry point fails with
 * PCME_ENODEV.
 *
 * Picture layout (picture.rs:7-51, 150-180): a padded luma plane of a coded
 * picture of mbw x mbh macroblocks has stride 16*mbw + 128 and
 * 16*mbh + 128 rows; picture pixel (0,0) is at byte 64*stride + 64.  The
 * three half-pel planes (picture.rs:318-380) use the same layout.  Chroma
 * planes: stride 8*mbw + 64, 8*mbh + 64 rows, pad 32.
 */
#ifndef PCME_H
#define PCME_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCME_LUMA_PAD 64
#define PCME_CHROMA_PAD 32

#define PCME_OK 0
#define PCME_EINVAL (-1)   /* bad argument (sizes, ranges, null pointers) */
#define PCME_ENODEV (-2)   /* no usable gfx950 device */
#define PCME_EHIP (-3)     /* a HIP runtime call failed */
#define PCME_ESTALL (-4)   /* a wavefront dependency wait exceeded its bound */
#define PCME_ENOMEM (-5)   /* device allocation failed */

/* Quarter-pel motion vector; layout of `#[repr(Rust)] Mv {x: i16, y: i16}`
 * (mbinfo.rs:1-16). */
typedef struct pcme_mv { int16_t x, y; } pcme_mv;

/* Partition shapes, in the order of `enum Part` (mbinfo.rs:47-72). */
enum { PCME_P16x16 = 0, PCME_P16x8 = 1, PCME_P8x16 = 2, PCME_P8x8 = 3 };

/* Per-macroblock result of the ME-only field analysis: what analyse_p
 * commits for an inter MB (slice.rs:1921-1945, cand_inter_info
 * slice.rs:1751-1785) plus the per-mode search_part totals.  128 bytes. */
typedef struct pcme_mb_out {
  pcme_mv mv[16];       /* MbInfo.mv[0], 4x4 blocks in raster order */
  pcme_mv mvs[4];       /* best_mvs (one per partition, rest zero) */
  pcme_mv mvps[4];      /* best_mvps: the MV predictor of each partition */
  uint32_t cost[4];     /* search_part total per mode P16x16..P8x8;
                           0xFFFFFFFF = mode not searched */
  uint32_t best_cost;   /* chosen cost incl. the mode penalty (lam*k)>>8 */
  pcme_mv skip_mv;      /* pskip_mv (slice.rs:484-494), the 16x16 seed */
  pcme_mv mv16;         /* 16x16 result mv16[0], the seed of the other modes */
  uint8_t part;         /* best_part (PCME_P*) */
  uint8_t ref;          /* the MB's reference index (list 0); 0 with one
                           reference */
  uint8_t reserved[2];
} pcme_mb_out;

/* Search parameters (EncParams subset, slice.rs:21-34). */
typedef struct pcme_params {
  int32_t mbw, mbh;        /* coded size in macroblocks */
  int32_t me_range;        /* hex-iteration cap and clamp, int-pel */
  int32_t subpel_rounds;   /* extra quarter-pel rounds */
  int32_t partitions;      /* 0/1: search 16x8, 8x16, 8x8 when cost16>1024 */
  int32_t qp;              /* lambda = pcme_lambda(qp) */
} pcme_params;

/* Per-macroblock search hint of the learned pruner (include/pcme_net.h).
 * No counterpart in the reference.  When `prune` is 1 the macroblock's
 * search runs pruned: every partition search starts from the rounded
 * predictor, its two search_part candidates and the hint's candidates
 * instead of the hexagon descent, and only P16x16 plus the hinted mode are
 * searched.  When `prune` is 0 the macroblock runs the full search.
 * 40 bytes. */
typedef struct pcme_hint {
  pcme_mv cand[8];         /* ranked candidate MVs, quarter-pel units on
                              the integer grid */
  uint16_t conf_part;      /* probability of `part` x 65535 */
  uint16_t conf_mv;        /* probability mass of the candidates x 65535 */
  uint8_t part;            /* hinted partition mode (PCME_P*) */
  uint8_t ncand;           /* candidates used, 0..8 */
  uint8_t prune;           /* 1: pruned search; 0: full search */
  uint8_t reserved;
} pcme_hint;

/* The four padded planes of one reference picture (device): luma and its
 * H, V, C half-pel planes. */
typedef struct pcme_ref_planes {
  const uint8_t *luma, *h, *v, *c;
} pcme_ref_planes;

#define PCME_MAX_REFS 4

/* One P frame of the batched ME-only field search.  All pointers (device). */
typedef struct pcme_frame_job {
  const uint8_t *src;      /* padded source luma of the frame */
  const uint8_t *ref;      /* padded reference luma, pads edge-extended */
  const uint8_t *ref_h;    /* half-pel planes of the reference (H, V, C) */
  const uint8_t *ref_v;
  const uint8_t *ref_c;
  pcme_mb_out *out;        /* mbw*mbh records, raster order */
  const pcme_hint *hint;   /* NULL: full search of every MB; else mbw*mbh
                              hints, raster order (pruned search) */
  pcme_params prm;
  /* Multi-reference search (no counterpart in the reference, whose P
   * slices use one reference): nref references of list 0, `ref` being
   * index 0 (the newest) and older[k] index k + 1.  0 or 1: `ref` only.
   * Each MB's P16x16 search runs on every reference, each with the
   * predictor of its index and (lambda * te(v) bits of the index) >> 8 on
   * its cost; the least total (the lower index on ties) is the MB's
   * reference (pcme_mb_out.ref) and its partition searches run on it, each
   * partition paying the index's rate.  References other than the newest
   * are read whole (no row limit).  A pruned MB's hint candidates (motion
   * over one picture) are scaled by each reference's distance, index + 1. */
  int32_t nref;
  int32_t reserved;
  pcme_ref_planes older[PCME_MAX_REFS - 1];
} pcme_frame_job;

/* ---- library ------------------------------------------------------- */
int pcme_version(void);
const char *pcme_last_error(void);
/* Number of visible gfx950 devices (0 if none); never fails. */
int pcme_device_count(void);

/* Lambda table (slice.rs:52-68, Lambdas::new): the u32 `sad` entry. */
uint32_t pcme_lambda(int32_t qp);
/* The packed quarter-pel source table the kernels use (picture.rs:387-404),
 * 16 entries of 8 bits: [1:0] plane A, [2] dx A, [3] dy A, [5:4] plane B,
 * [6] dx B, [7] dy B; planes 0 full, 1 H, 2 V, 3 C.  Writes 4 words. */
void pcme_qpel_table(uint32_t out[4]);

/* ---- picture preparation -------------------------------------------- */
/* Import + edge extension of n planes (mod.rs:842-873 import, frame.rs:58-97
 * extend): plane i of src (device) starts at src + i*src_pitch, is w x h with
 * row stride sstride; plane i of dst (device) starts at dst + i*dst_pitch and
 * is a padded plane of stride dstride around an aw x ah aligned picture with
 * `pad` pad pixels per side (luma: aw = 16*mbw, dstride = aw + 128, pad 64;
 * chroma: aw = 8*mbw, dstride = aw + 64, pad 32).  padded(x, y) =
 * src[clamp(y, 0, h-1)][clamp(x, 0, w-1)] for x in [-pad, dstride-pad),
 * y in [-pad, ah+pad).  Requires w <= aw, h <= ah. */
int pcme_pad_planes(const uint8_t *src, int64_t src_pitch, int32_t sstride,
                    int32_t w, int32_t h, uint8_t *dst, int64_t dst_pitch,
                    int32_t dstride, int32_t aw, int32_t ah, int32_t pad,
                    int32_t n, void *stream);
/* Half-pel planes of n padded luma planes (picture.rs:318-380, run by
 * frame.rs:99-122 over the whole padded plane).  Plane i of ref (device) is
 * at ref + i*pitch; its H, V and C planes are written at h + i*pitch,
 * v + i*pitch, c + i*pitch (device).  Layout as above for mbw x mbh. */
int pcme_hpel(const uint8_t *ref, uint8_t *h, uint8_t *v, uint8_t *c,
              int64_t pitch, int32_t n, int32_t mbw, int32_t mbh,
              void *stream);
/* pcme_pad_planes (luma: pad 64, aw = 16*mbw, dstride = aw + 128) followed
 * by pcme_hpel, in one pass from the unpadded source: the same four planes
 * bit for bit (mod.rs:842-873 import, frame.rs:58-122 extend + half-pel
 * planes, picture.rs:318-380).  Plane i of src (device) is w x h at
 * src + i*src_pitch with row stride sstride; its full, H, V and C planes are
 * written at full/hp/vp/cp + i*pitch (device, 4-byte aligned). */
int pcme_luma_planes(const uint8_t *src, int64_t src_pitch, int32_t sstride,
                     int32_t w, int32_t h, uint8_t *full, uint8_t *hp,
                     uint8_t *vp, uint8_t *cp, int64_t pitch, int32_t n,
                     int32_t mbw, int32_t mbh, void *stream);
/* The arguments of one pcme_luma_planes call. */
typedef struct pcme_planes_job {
  const uint8_t *src;
  uint8_t *full, *hp, *vp, *cp;
  int64_t src_pitch, pitch;
  int32_t sstride, w, h, n, mbw, mbh;
} pcme_planes_job;
/* pcme_luma_planes of several clips, of any sizes, in one launch per 16
 * clips: the same planes bit for bit.  Each clip is the reference's import
 * + extend + half-pel planes of its frames (mod.rs:842-873, frame.rs:58-122,
 * picture.rs:318-380); the reference builds them per frame on each encoder's
 * frame threads, which a batch of clips replaces.  Every job is checked
 * before anything is launched; no device memory is allocated. */
int pcme_luma_planes_batch(const pcme_planes_job *jobs, int32_t njobs, void *stream);

/* ---- the search ----------------------------------------------------- */
/* ME-only field over a batch of P frames: for every macroblock of every
 * job, pskip_mv -> search_part(P16x16, seed = skip_mv) -> when partitions
 * and cost16 > 1024, search_part(P16x8 / P8x16 / P8x8, seed = mv16) ->
 * strict-< mode choice -> commit (slice.rs:1870-1945 without early skip,
 * intra and RD; search_part slice.rs:1810-1868; search slice.rs:1590-1723;
 * mvp slice.rs:392-482).  Each job's neighbours are its own committed MBs,
 * its vlim rule that of frame.rs:152-156 / picture.rs:261-265.  Jobs are
 * independent and may differ in size and parameters.  `jobs` is a HOST
 * array; the library copies it.  Requires 1 <= me_range <= 1024,
 * 0 <= subpel_rounds <= 8, 0 <= qp <= 51, mbw, mbh >= 1.  Blocks until the
 * job table is uploaded; the search itself runs asynchronously on stream.
 * The device's cached plan is reused while the job table is unchanged; its
 * launches run one after another on the GPU (see above).  A stall of an
 * asynchronous call is kept until pcme_search_frames_check reports it. */
int pcme_search_frames(const pcme_frame_job *jobs, int njobs, void *stream);

/* Same as pcme_search_frames, then waits for the stream and reports a
 * dependency-wait stall (PCME_ESTALL) if one happened. */
int pcme_search_frames_sync(const pcme_frame_job *jobs, int njobs,
                            void *stream);

/* pcme_search_frames_sync without any allocation: no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under stream capture set-up, tight serving loops, harnesses that forbid
 * hipMalloc).  The same results as pcme_search_frames_sync for the same
 * jobs.  Its workspace (control block, job table, tickets and 16 bytes of
 * exchange words per macroblock) lives in a 16 MiB array the library keeps
 * on each device until the process ends; a call whose jobs need more (about
 * 125 1080p frames) or more than 1024 jobs fails with PCME_EINVAL.  Every
 * call on a device shares that array: the library serialises these calls
 * and each one waits for its kernels before it returns (it is synchronous,
 * and reports PCME_ESTALL like pcme_search_frames_sync); calls from
 * several processes each use their own copy.  The caller passes the
 * inputs and outputs as for pcme_search_frames. */
int pcme_search_frames_noalloc(const pcme_frame_job *jobs, int njobs,
                               void *stream);

/* Waits for `stream` and for the last launch of the current device's
 * cached pcme_search_frames plan, and returns PCME_ESTALL if any launch of
 * that plan since the previous check (or pcme_search_frames_sync) stalled;
 * the flag is then cleared.  PCME_OK when no plan is cached.  No
 * counterpart in the reference (its search cannot stall). */
int pcme_search_frames_check(void *stream);

/* Plans: the same ME-only field search, split into a set-up step and a
 * launch step, for callers that replay one frame-step layout many times
 * (a frame pipeline, hipGraph capture).  Same computation and results as
 * pcme_search_frames (which keeps one such plan cached per device).
 *
 * pcme_plan_create copies the job table (host array) to the device and
 * allocates the plan's own workspace (synchronous).  The device buffers the
 * jobs point to must stay valid while the plan is used; their contents may
 * change between launches.
 * pcme_plan_launch enqueues kernels only -- no host synchronisation, copies
 * or memsets -- so a stream capture of it replays correctly.  Launches of one
 * plan share the plan's workspace: an eager launch makes its stream wait for
 * the plan's previous eager launch (a HIP event), so they never overlap;
 * inside a stream capture the graph's order applies.  Different plans may
 * run concurrently.
 * pcme_plan_check waits for `stream` and the plan's last launch and returns
 * PCME_ESTALL if a dependency wait of any launch since the previous check
 * exceeded its bound (a sticky flag; replays of a captured launch set it
 * too), clearing it.
 * pcme_plan_set_grid sets the persistent grid (waves; 0 = every resident
 * wave) and returns the value in effect.  pcme_plan_autotune times the plan
 * over a set of grid sizes (`reps` launches each, on `stream`) and keeps the
 * fastest; it writes the chosen wave count and its time. */
typedef struct pcme_plan pcme_plan;

/* A plan's kernel configuration (no counterpart in the reference: how the
 * search is laid out on the GPU, never what it computes -- every
 * configuration gives the same records).  waves: the persistent grid (0 =
 * every wave the configuration keeps resident); waves_per_group: row
 * walkers per workgroup (1, 2 or 4); waves_per_cu: the occupancy cap the
 * LDS each walker reserves sets (16, 12 or 8; for a plan with pruner hints
 * its kernel's own; at 16 a walker stages each MB's source block after the
 * previous MB instead of one MB ahead); reuse_regions: 1 = a walker keeps
 * staged LDS regions from MB to MB, 0 = every search stages its own; ms:
 * the time per launch the autotuner measured.  The built configurations are
 * (1, 12, 1) -- the default --, (2, 12, 1), (4, 12, 1), (1, 8, 1),
 * (2, 8, 1), (1, 12, 0), (4, 16, 0) and (4, 16, 1); a plan with pruner
 * hints or with several references runs its kernel's default or (4, 16, 1).
 * pcme_plan_get_config reads the plan's configuration; pcme_plan_set_config
 * selects one (PCME_EINVAL for a configuration that is not built);
 * pcme_plan_autotune_ex times every built configuration over the grid
 * sizes of pcme_plan_autotune (`reps` launches each, on `stream`), keeps
 * the fastest and writes it to *best. */
typedef struct pcme_tune {
  int32_t waves;
  int32_t waves_per_group;
  int32_t waves_per_cu;
  int32_t reuse_regions;
  float ms;
} pcme_tune;

int pcme_plan_create(const pcme_frame_job *jobs, int njobs, pcme_plan **out);
int pcme_plan_launch(pcme_plan *plan, void *stream);
int pcme_plan_check(pcme_plan *plan, void *stream);
int pcme_plan_set_grid(pcme_plan *plan, int32_t waves);
int pcme_plan_autotune(pcme_plan *plan, void *stream, int32_t reps,
                       int32_t *best_waves, float *best_ms);
int pcme_plan_get_config(pcme_plan *plan, pcme_tune *out);
int pcme_plan_set_config(pcme_plan *plan, const pcme_tune *cfg);
int pcme_plan_autotune_ex(pcme_plan *plan, void *stream, int32_t reps, pcme_tune *best);
void pcme_plan_destroy(pcme_plan *plan);

/* Test hook: bound of every wavefront dependency wait, in polls (default
 * 2^22; 0 restores it) for launches made after the call.  A tiny bound lets
 * a test drive the stall path (every wave exits, PCME_ESTALL) without any
 * fault.  Returns the previous bound.  No counterpart in the reference. */
uint32_t pcme_set_spin_limit(uint32_t polls);

/* Per-call form of SliceEnc::search (slice.rs:1590-1723), batched: job i
 * searches one w x h block at (bx, by) of a reference with predictor mvp and
 * candidates cands[0..ncands) (ncands <= 2, as search_part passes
 * [seed, mvs[0]]), under the vlim of the caller's current MB row.  Results
 * go to out_mv[i], out_cost[i] (device). */
typedef struct pcme_search_job {
  const uint8_t *src;      /* (device) top-left of the w x h source block */
  int32_t sstride;
  const uint8_t *ref;      /* (device) padded planes full, H, V, C */
  const uint8_t *ref_h;
  const uint8_t *ref_v;
  const uint8_t *ref_c;
  int32_t mbw, mbh;        /* coded picture size in MBs */
  int32_t bx, by, w, h;    /* block position/size in picture pixels */
  pcme_mv mvp;
  pcme_mv cands[2];
  int32_t ncands;
  uint32_t lam;
  int32_t me_range, subpel_rounds;
  int64_t vlim;            /* SliceEnc::vlim for the block's MB row */
} pcme_search_job;
int pcme_search_batch(const pcme_search_job *jobs, int njobs,
                      pcme_mv *out_mv, uint32_t *out_cost, void *stream);

/* One neighbour 4x4 block as nb_part reads it (slice.rs:392-432) for the
 * searched list: ref -2 = not available (outside the picture, or the
 * top-right MB past the right edge), -1 = available without a reference of
 * that list (an intra MB reads as ref -1, MV 0), >= 0 = reference index,
 * with its MV.  8 bytes. */
typedef struct pcme_nb_block {
  pcme_mv mv;
  int8_t ref;
  uint8_t reserved[3];
} pcme_nb_block;

/* One search_part call (slice.rs:1810-1868): the partitions of `part` of
 * the MB at (mx, my) searched in order, each with its predictor from the
 * neighbours and the partitions already searched, candidates [seed,
 * mvs[0]].  The neighbour blocks are those of the list being searched, so
 * the same call serves list 0 and list 1 (analyse_b, slice.rs:2277-2289,
 * b_partition slice.rs:2100-2102).  168 bytes. */
typedef struct pcme_part_job {
  const uint8_t *src;      /* (device) top-left of the 16x16 source MB */
  int32_t sstride;
  int32_t part;            /* PCME_P16x16 .. PCME_P8x8 */
  const uint8_t *ref;      /* (device) padded planes full, H, V, C */
  const uint8_t *ref_h;
  const uint8_t *ref_v;
  const uint8_t *ref_c;
  int32_t mbw, mbh, mx, my;
  pcme_mv seed;
  uint32_t lam;
  int32_t me_range, subpel_rounds;
  int64_t vlim;            /* SliceEnc::vlim for the MB's row */
  pcme_nb_block left[4];   /* left MB, 4x4 blocks 3, 7, 11, 15 */
  pcme_nb_block top[4];    /* top MB, blocks 12, 13, 14, 15 */
  pcme_nb_block top_right; /* top-right MB, block 12 */
  pcme_nb_block top_left;  /* top-left MB, block 15 */
} pcme_part_job;

/* search_part's result: ([Mv; 4] mvs, [Mv; 4] mvps, u32 total), unused
 * partitions zero.  36 bytes. */
typedef struct pcme_part_out {
  pcme_mv mvs[4];
  pcme_mv mvps[4];
  uint32_t total;
} pcme_part_out;

/* Independent search_part calls, one wave each; out[i] (device) gets the
 * result of jobs[i] (host array). */
int pcme_search_part_batch(const pcme_part_job *jobs, int njobs,
                           pcme_part_out *out, void *stream);

/* The motion search of one B macroblock (analyse_b, slice.rs:2189-2377,
 * and b_partition, slice.rs:2086-2187) without the parts that need the
 * residual coder or RD (BSkip, rd_cost, intra).  nb[l] holds the
 * neighbour blocks of list l in the order left[4], top[4], top_right,
 * top_left (as pcme_part_job); col_* describe the co-located MB of the
 * list-1 picture for direct_spatial (slice.rs:2057-2071).  328 bytes. */
typedef struct pcme_b_job {
  const uint8_t *src;      /* (device) top-left of the 16x16 source MB */
  int32_t sstride;
  int32_t reserved0;
  const uint8_t *ref0[4];  /* (device) list-0 picture: full, H, V, C planes */
  const uint8_t *ref1[4];  /* (device) list-1 picture */
  int32_t mbw, mbh, mx, my;
  uint32_t lam;
  int32_t me_range, subpel_rounds, reserved1;
  int64_t vlim;
  pcme_nb_block nb[2][10];
  pcme_mv col_mv[2][4];    /* co-located MB: MVs of its 4x4 blocks 0, 3, 12, 15 per list */
  int8_t col_ref[2][4];    /* its ref_idx per 8x8 per list */
  int32_t col_intra;       /* nonzero: the co-located MB is intra */
  int32_t reserved2;
} pcme_b_job;

/* Result of one B macroblock's motion search.  direct: direct_spatial's
 * refs and MVs per 8x8, whether every used MV is legal, and SATD of the
 * direct prediction + (lam*2)>>8 (all ones when not legal).  16x16: the
 * list-0/list-1 search_part results seeded with the direct MVs, the
 * direction costs (list costs + (lam*3)>>8; bi SATD + lam*(bits+5)>>8) and
 * the first least one.  part_*: b_partition of 16x8 [0] and 8x16 [1]
 * seeded with the 16x16 MVs: MVs [mode][list][partition], per-partition
 * directions and the total (lam*8)>>8 + sum of the best.  136 bytes. */
typedef struct pcme_b_out {
  pcme_mv direct_mv[2][4];
  int8_t direct_ref[2][4];
  uint32_t direct_ok, direct_cost;
  pcme_mv mv16[2], mvp16[2];
  uint32_t cost16[2];
  uint32_t cost[3];
  uint32_t dir;
  pcme_mv part_mv[2][2][2];
  uint32_t part_cost[2];
  uint8_t part_dir[2][2];
  uint8_t reserved[4];
} pcme_b_out;

/* Independent B macroblocks, one wave each; out[i] (device) gets the
 * result of jobs[i] (host array). */
int pcme_b_analyse_batch(const pcme_b_job *jobs, int njobs, pcme_b_out *out,
                         void *stream);

/* ---- the ME-only field of a whole B picture -------------------------- */
/* The co-located MB of the list-1 picture (direct_spatial, slice.rs:
 * 2057-2071).  44 bytes. */
typedef struct pcme_b_col {
  pcme_mv mv[2][4];        /* MVs of its 4x4 blocks 0, 3, 12, 15 per list */
  int8_t ref[2][4];        /* its ref_idx per 8x8 per list */
  int32_t intra;           /* nonzero: the co-located MB is intra */
} pcme_b_col;

/* One B picture: the source and both references as the frame jobs lay
 * them out (padded luma, stride 16*mbw + 128, pad 64; device). */
typedef struct pcme_b_field_job {
  const uint8_t *src;      /* padded source luma of the B picture */
  const uint8_t *ref0[4];  /* list-0 picture: full, H, V, C planes */
  const uint8_t *ref1[4];  /* list-1 picture */
  const pcme_b_col *col;   /* mbw*mbh co-located MBs, raster order */
  int32_t mbw, mbh;
  int32_t qp, me_range, subpel_rounds, partitions;
} pcme_b_field_job;

/* One MB of the B field: its search (pcme_b_out) over the MBs committed
 * before it, and what it commits -- analyse_b's decision with rd off and
 * without BSkip and intra (slice.rs:2189-2377): direct when its MVs are
 * legal, the 16x16 direction when strictly cheaper, then 16x8 and 8x16
 * when `partitions` and the best cost exceeds 1024, each when strictly
 * cheaper; committed as cand_inter_info does (slice.rs:1751-1785).
 * 280 bytes. */
typedef struct pcme_b_mb {
  pcme_b_out search;
  pcme_mv mv[2][16];       /* MbInfo.mv per list, zero where a list is unused */
  int8_t ref[2][4];        /* MbInfo.ref_idx per list and 8x8, -1 unused */
  uint32_t best_cost;
  uint8_t mode;            /* 0 direct, 1 16x16, 2 16x8, 3 8x16 */
  uint8_t dir;             /* the 16x16 direction: 0 list 0, 1 list 1, 2 bi */
  uint8_t reserved[2];
} pcme_b_mb;

/* The whole field of one B picture in raster dependency order, on the
 * device, in one launch of persistent row walkers (two waves per MB row;
 * an MB waits for the row above's MB x+1, so the rows advance as a
 * wavefront).  out: mbw*mbh records (device), raster order.  No
 * counterpart in the reference, whose encoder walks the picture on the
 * CPU.  Asynchronous on stream.  A wait that exceeds the library's spin
 * limit (which the wavefront never reaches) ends the launch with the
 * records incomplete and is kept until pcme_b_field_check reports it. */
int pcme_b_field(const pcme_b_field_job *job, pcme_b_mb *out, void *stream);

/* pcme_b_field for njobs independent B pictures (1 <= njobs <= 1024, at
 * most 2^22 MBs in all; sizes and parameters may differ) in one launch:
 * the walkers take the pictures' rows in (row, picture) order, so the
 * pictures advance side by side -- the device form of the reference's
 * frame threads, which analyse several pictures at once.  outs[i]: the
 * mbw*mbh records of picture i (device).  Results are those of one
 * pcme_b_field call per picture; stalls as for pcme_b_field. */
int pcme_b_fields(const pcme_b_field_job *jobs, int njobs, pcme_b_mb *const *outs, void *stream);

/* Waits for `stream` and returns PCME_ESTALL if any pcme_b_field launch on
 * the current device since the previous check stalled (its records are
 * incomplete); the flag is then cleared.  Launches on other streams are
 * covered once they have completed.  No counterpart in the reference (its
 * B analysis cannot stall). */
int pcme_b_field_check(void *stream);

/* ---- early P-skip gate (analyse_p, slice.rs:1884-1920) ----------------
 * Before searching an MB the encoder predicts it at its pskip_mv
 * (pred_block: mc_luma + mc_chroma), and when that MV is legal and the
 * prediction's sad16 is below 2 * 16 * v4[qp % 6][0] << (qp / 6), codes the
 * residual (code_inter, slice.rs:943-1001: 4x4 and/or 8x8 transform,
 * dead-zone quantiser, decimation; chroma DC and AC); a coded block pattern
 * of 0 makes the MB a PSkip that is never searched.  One job per MB; the
 * caller supplies the skip MV it derived from its committed neighbours
 * (pskip_mv, slice.rs:484-494).  flags: the encoder's EncParams t8x8,
 * adaptive_t8 and cabac (the level clamp of CAVLC). */
#define PCME_SKIP_T8X8 1
#define PCME_SKIP_ADAPTIVE_T8 2
#define PCME_SKIP_CABAC 4
typedef struct pcme_skip_job {
  const uint8_t *src_y;    /* (device) top-left of the 16x16 source MB */
  const uint8_t *src_u;    /* (device) top-left of its 8x8 chroma blocks */
  const uint8_t *src_v;
  int32_t ystride, cstride;
  const uint8_t *ref[4];   /* (device) padded reference luma: full, H, V, C */
  const uint8_t *ref_u;    /* (device) padded reference chroma planes, */
  const uint8_t *ref_v;    /* stride 8*mbw + 64, pad 32 */
  int32_t mbw, mbh, mx, my;
  pcme_mv skip_mv;
  int32_t qp;
  int32_t flags;           /* PCME_SKIP_* */
  int32_t chroma_qp_offset;
  int64_t vlim;            /* SliceEnc::vlim of the MB's row */
} pcme_skip_job;

/* The gate's outcome.  16 bytes. */
typedef struct pcme_skip_out {
  uint32_t sad;            /* sad16 of the prediction (0 when !mv_ok) */
  uint32_t thresh;         /* the bound sad16 must stay below */
  uint8_t mv_ok;           /* the skip MV is legal for the 16x16 block */
  uint8_t coded;           /* code_inter ran (mv_ok and sad < thresh) */
  uint8_t cbp;             /* its coded block pattern, luma bits 0-3 and
                              chroma << 4; 0xFF when not coded */
  uint8_t skip;            /* early PSkip: coded and cbp == 0 */
  uint8_t t8x8;            /* code_inter's transform_8x8 flag */
  uint8_t reserved[3];
} pcme_skip_out;

/* The gate for n MBs, one wave each; out[i] (device) gets jobs[i]'s
 * outcome (jobs is a host array).  Bit-identical to the reference,
 * including the adaptive 8x8/4x4 choice (its float rate comparison is
 * evaluated in the reference's operation order, with the C library's
 * log2 tabulated on the host). */
int pcme_skip_gate_batch(const pcme_skip_job *jobs, int njobs, pcme_skip_out *out,
                         void *stream);

/* Motion-compensated prediction of the chosen partitions (pred_block,
 * slice.rs:1725-1749 = mc_luma picture.rs:415-449 + mc_chroma
 * picture.rs:451-487) for every MB of a searched frame.  ref planes and
 * chroma planes (device, padded), mb (device) = search output.  Writes
 * pred_y (16x16 per MB, raster MB order, 256 B each) and pred_u/pred_v
 * (8x8 per MB, 64 B each). */
int pcme_pred_frame(const uint8_t *ref, const uint8_t *ref_h,
                    const uint8_t *ref_v, const uint8_t *ref_c,
                    const uint8_t *ref_u, const uint8_t *ref_vv,
                    const pcme_mb_out *mb, int32_t mbw, int32_t mbh,
                    uint8_t *pred_y, uint8_t *pred_u, uint8_t *pred_v,
                    void *stream);


/* ---- lowres lookahead: scenecut complexity and macroblock-tree costs ----
 * A lowres frame is the w x h = (16*mbw/2) x (16*mbh/2) 2x2 average of a
 * coded luma plane, rows w bytes apart, 4-byte aligned; w and h are
 * multiples of 8 (one 8x8 block per macroblock). */

/* LowRes::new (lookahead.rs:10-33): lowres frames of n padded luma planes.
 * Plane i (device) is at planes + i*plane_pitch with row stride `stride` and
 * `pad` pixels of padding (picture (0,0) at pad*stride + pad); w x h is its
 * coded area (w, h multiples of 16).  Frame i is written at out + i*out_pitch
 * (device), (w/2) x (h/2), rows w/2 apart.  planes, plane_pitch, stride and
 * pad 4-byte aligned; out and out_pitch too. */
int pcme_lowres(const uint8_t *planes, int64_t plane_pitch, int32_t stride,
                int32_t pad, int32_t w, int32_t h, uint8_t *out,
                int64_t out_pitch, int32_t n, void *stream);

/* Per 8x8 block of complexity(): intra cost (mbtree.rs:21-49), the
 * contribution to Complexity.inter (min(SATD at the diamond's best + 8,
 * intra), or intra without a previous frame) and the diamond's integer MV
 * (lowres pixels; zero without a previous frame).  12 bytes. */
typedef struct pcme_cplx_block {
  uint32_t intra, inter;
  pcme_mv mv;
} pcme_cplx_block;

/* lookahead.rs:116-139 aq_offsets (called at mod.rs:658-662 when the rate
 * control is not constant-QP): per MB, strength * (log2(var_y/256 +
 * 0.5*(var_u/64 + var_v/64) + 1) - the picture's mean), clamped to +-8,
 * with the variances of dsp.rs:104-114.  y, u, v (device, 4-byte aligned)
 * point at picture (0, 0) of the padded source planes, strides multiples
 * of 4; out (device) receives mbw*mbh floats in raster order.  Float
 * results: within 1e-5 of the reference (the log2 of the C library and of
 * the device may differ in the last place); the variances are exact. */
int pcme_aq_offsets(const uint8_t *y, int32_t ystride, const uint8_t *u,
                    const uint8_t *v, int32_t cstride, int32_t mbw,
                    int32_t mbh, float strength, float *out, void *stream);

/* One complexity() call (lookahead.rs:52-108): cur and prev (device, NULL =
 * no previous frame) are w x h lowres frames.  sums (device, 8-byte
 * aligned) receives Complexity {intra, inter}; blocks (device, may be NULL)
 * (w/8)*(h/8) records in raster order.  40 bytes. */
typedef struct pcme_cplx_job {
  const uint8_t *cur, *prev;
  pcme_cplx_block *blocks;
  uint64_t *sums;
  int32_t w, h;
} pcme_cplx_job;

/* lookahead::complexity for n frames at once (the scenecut input,
 * mod.rs:656-657).  Bit-identical to the reference. */
int pcme_complexity(const pcme_cplx_job *jobs, int32_t n, void *stream);

/* mbtree.rs:10-15 Block: intra cost, the best of intra / list 0 / list 1 /
 * bi (`list` 0..3) and its cost, and the MVs as analyse records them (a
 * list-0 win leaves mv[1] = 0).  20 bytes. */
typedef struct pcme_tree_block {
  uint32_t intra, inter;
  pcme_mv mv[2];
  uint8_t list;
  uint8_t reserved[3];
} pcme_tree_block;

/* One analyse() call (mbtree.rs:113-159): cur, ref0, ref1 (device; a NULL
 * reference skips its list) are w x h lowres frames; blocks (device)
 * receives (w/8)*(h/8) records in raster order.  40 bytes. */
typedef struct pcme_tree_job {
  const uint8_t *cur, *ref0, *ref1;
  pcme_tree_block *blocks;
  int32_t w, h;
} pcme_tree_job;

/* mbtree::analyse for n frames at once (mod.rs:760-790, the lookahead's
 * per-frame costs of the macroblock tree).  Bit-identical to the
 * reference. */
int pcme_mbtree_analyse(const pcme_tree_job *jobs, int32_t n, void *stream);

/* mbtree.rs:162-165 Node: the node's analyse() blocks (device, bw*bh
 * records) and the node index each list refers to, -1 for none (Option
 * None).  A reference at or after the node itself feeds nothing.  16 bytes. */
typedef struct pcme_tree_node {
  const pcme_tree_block *blocks;
  int32_t ref[2];
} pcme_tree_node;

/* mbtree::propagate (mbtree.rs:167-223; the encoder's tree_offsets,
 * mod.rs:819-840, runs it over the lookahead queue): prop (device, n*bw*bh
 * floats, node-major, raster order) receives every node's propagated cost.
 * Bit-identical to the reference.  The nodes run in rounds, one launch per
 * round on `stream`: a node joins a round once every node that feeds it has
 * run and every later node that feeds one of its destinations has run, so
 * each cell still receives its adds in the reference's order (nodes last to
 * first, blocks in raster order, list 0 before list 1).  One chain is one
 * node per round; independent queues in one call share rounds. */
int pcme_mbtree_propagate(const pcme_tree_node *nodes, int32_t n, int32_t bw, int32_t bh, float *prop,
                          void *stream);

/* mbtree::offsets (mbtree.rs:225-231): out[k] = -2 log2((max(intra, 1) +
 * prop[k]) / max(intra, 1)) over nblocks blocks (all device).  Within 1e-5
 * of the reference (the GPU's log2). */
int pcme_mbtree_offsets(const pcme_tree_block *blocks, const float *prop, int32_t nblocks, float *out,
                        void *stream);

#ifdef __cplusplus
}
#endif
#endif
