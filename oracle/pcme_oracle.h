/*
 * pcme_oracle.h -- CPU oracle for the H.264 P-frame motion search.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the
 * reference encoder's inter-prediction search (pito-codecs 5.3.0,
 * src/h264/enc/{dsp,picture,mbinfo,slice,frame}.rs).  Only the tests, the
 * smoke check and the benchmark's CPU-baseline leg may link or call it, and
 * only as the checker / CPU baseline -- never as the product path.
 *
 * Every function cites the reference lines it restates.  What pins each
 * function (golden vectors, closed forms, invariants) is listed in the
 * header comment of pcme_oracle.c.
 */
#ifndef PCME_ORACLE_H
#define PCME_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* picture.rs:7-8 */
#define O_LUMA_PAD 64
#define O_CHROMA_PAD 32
/* frame.rs:13 */
#define O_VMAX 96

/* mbinfo.rs:1-16 -- quarter-pel motion vector */
typedef struct { int16_t x, y; } omv;

/* mbinfo.rs:18-45 (subset used by the P-frame predictor) */
enum { OK_NONE = 0, OK_I4x4, OK_I8x8, OK_I16x16, OK_PSKIP, OK_PINTER };
/* mbinfo.rs:47-72 */
enum { OP_16x16 = 0, OP_16x8 = 1, OP_8x16 = 2, OP_8x8 = 3 };

/* mbinfo.rs:74-107, list 0 only (P slices use one list) */
typedef struct {
  int8_t kind;
  int8_t ref_idx[4]; /* per 8x8, raster */
  omv mv[16];        /* per 4x4, raster */
} ombinfo;

/* Reference view (picture.rs:116-138): four planes of identical padded
 * layout: full-pel luma, then the H ('b'), V ('h') and C ('j') half-pel
 * planes.  Coordinates are picture coordinates; (0,0) sits at
 * data + pad*stride + pad. */
typedef struct {
  const uint8_t *pl[4];
  int stride;
  int pad;
} oref;

/* Search parameters (slice.rs:21-34 EncParams subset + SliceEnc::vlim) */
typedef struct {
  int mbw, mbh;
  int me_range;
  int subpel_rounds;
  int partitions;
  long vlim;
} oparams;

/* Work counters (evaluated candidates per phase, weighted by pixels). */
typedef struct {
  uint64_t sad_px;     /* pixels through SAD */
  uint64_t satd_px;    /* pixels through MC + SATD */
  uint64_t sad_evals;  /* SAD candidate evaluations */
  uint64_t satd_evals; /* SATD candidate evaluations */
  uint64_t searches;   /* calls of search() */
  uint64_t mbs;        /* macroblocks analysed */
} ocounters;

/* Per-MB result of the ME-only field analysis. */
typedef struct {
  omv mv[16];      /* committed MbInfo.mv[0] */
  omv mvs[4];      /* best_mvs */
  omv mvps[4];     /* best_mvps */
  uint32_t cost[4];/* search_part totals per mode; 0xFFFFFFFF = not searched */
  uint32_t best_cost;
  omv skip_mv;
  omv mv16;        /* 16x16 search result, seed of the other modes */
  uint8_t part;
  uint8_t ref;     /* the MB's reference index (o_analyse_mb_mr) */
  uint8_t pad_[2];
} omb_out;

/* Learned-pruner hint (layout of pcme_hint, include/pcme.h).  No
 * counterpart in the reference: the pruned search is this build's own
 * variant, restated here so its field can be checked bit for bit. */
typedef struct {
  omv cand[8];
  uint16_t conf_part, conf_mv;
  uint8_t part, ncand, prune, reserved;
} ohint;

/* ---- dsp.rs ---- */
uint32_t o_sad(const uint8_t *a, int sa, const uint8_t *b, int sb, int w, int h);
uint32_t o_sad16(const uint8_t *a, int sa, const uint8_t *b, int sb, int h);
uint32_t o_satd(const uint8_t *a, int sa, const uint8_t *b, int sb, int w, int h);

/* ---- slice.rs:52-68, 130-147 ---- */
uint32_t o_lambda(int qp);
uint32_t o_se_bits(int32_t v);
uint32_t o_mv_bits(omv mv, omv mvp);
int16_t o_median(int16_t a, int16_t b, int16_t c);

/* ---- picture.rs / mod.rs / frame.rs: planes ---- */
/* Import (mod.rs:842-873): copy a w x h plane into the MB-aligned area
 * (aw x ah) of a padded plane, replicating the right column and bottom row.
 * dst points at the padded plane's first byte. */
void o_import_plane(const uint8_t *src, int sstride, int w, int h,
                    uint8_t *dst, int dstride, int pad, int aw, int ah);
/* Edge extension of the pad (frame.rs:58-97, picture.rs:285-316). */
void o_extend_plane(uint8_t *plane, int stride, int pad, int w, int h);
/* Half-pel planes over the whole padded plane (picture.rs:318-380). */
void o_hpel(const uint8_t *src, int stride, int rows,
            uint8_t *hout, uint8_t *vout, uint8_t *cout);
/* picture.rs:415-449 */
void o_mc_luma(const oref *r, int x, int y, omv mv, int w, int h,
               uint8_t *out, int ostride);
/* picture.rs:451-487; u/v planes padded by pad */
void o_mc_chroma(const uint8_t *u, const uint8_t *v, int cstride, int cpad,
                 int x, int y, omv mv, int w, int h,
                 uint8_t *out_u, uint8_t *out_v, int ostride);

/* ---- slice.rs: search machinery ---- */
/* vlim rule: frame.rs:13,152-156 with picture.rs:211-268 */
long o_vlim(int my, int mbh);
int o_mv_ok(const oparams *p, int x, int y, int w, int h, omv mv);
/* nb_part (slice.rs:392-432).  mbs = the current picture's committed
 * MbInfo (mbw*mbh, raster), cur = partial info of the MB being analysed.
 * Returns 1 if available and writes (ref, mv). */
int o_nb_part(const oparams *p, const ombinfo *mbs, const ombinfo *cur,
              uint16_t mask, int mx, int my, int bx, int by,
              int8_t *ref, omv *mv);
/* mvp (slice.rs:434-482); shape: 0 Normal 1 Top16x8 2 Bottom16x8
 * 3 Left8x16 4 Right8x16 */
omv o_mvp(const oparams *p, const ombinfo *mbs, const ombinfo *cur,
          uint16_t mask, int mx, int my, int x, int y, int w, int8_t rf,
          int shape);
/* pskip_mv (slice.rs:484-494) */
omv o_pskip_mv(const oparams *p, const ombinfo *mbs, int mx, int my);
/* search (slice.rs:1590-1723) */
omv o_search(const oparams *p, const oref *r, int bx, int by, int w, int h,
             const uint8_t *src, int sstride, omv mvp, const omv *cands,
             int ncands, uint32_t lam, uint32_t *cost, ocounters *ctr);
/* search_part (slice.rs:1810-1868) */
uint32_t o_search_part(const oparams *p, const ombinfo *mbs, int mx, int my,
                       const oref *r, int part, const uint8_t *src16,
                       uint32_t lam, omv seed, omv mvs[4], omv mvps[4],
                       ocounters *ctr);
/* ME-only analysis of one MB: pskip_mv, search_part(16x16), partition
 * choice (slice.rs:1884-1945) and the commit of cand_inter_info
 * (slice.rs:1751-1785, 1980-1986) into mbs[my*mbw+mx]. */
void o_analyse_mb(const oparams *p, ombinfo *mbs, int mx, int my,
                  const oref *r, const uint8_t *src_plane, int sstride,
                  int spad, uint32_t lam, omb_out *out, ocounters *ctr);
/* ME-only field over one frame, raster order.  src = padded source luma,
 * r = reference view whose planes are complete, vlim per row from o_vlim. */
void o_search_frame(int mbw, int mbh, int me_range, int subpel_rounds,
                    int partitions, uint32_t lam, const uint8_t *src,
                    int sstride, int spad, const oref *r, omb_out *out,
                    ocounters *ctr);

/* ---- pruned search (no reference counterpart) ----
 * o_search_ex: o_search when nh < 0.  With nh >= 0 the hexagon descent is
 * replaced by a scan of the hint candidates hc[0..nh) (rounded like the
 * search_part candidates, skipped when illegal) after the predictor and the
 * search_part candidates; the square and sub-pel rounds are unchanged. */
omv o_search_ex(const oparams *p, const oref *r, int bx, int by, int w, int h,
                const uint8_t *src, int sstride, omv mvp, const omv *cands,
                int ncands, const omv *hc, int nh, uint32_t lam,
                uint32_t *cost, ocounters *ctr);
/* o_analyse_mb with an optional hint: h == NULL or h->prune == 0 is the full
 * analysis.  A pruned MB searches P16x16 and, when partitions and
 * cost16 > 1024, the hinted mode only (P16x8 or P8x16 hints search both of
 * those, P8x8 hints search P8x8), all with o_search_ex(nh = h->ncand). */
void o_analyse_mb_ex(const oparams *p, ombinfo *mbs, int mx, int my,
                     const oref *r, const uint8_t *src_plane, int sstride,
                     int spad, uint32_t lam, const ohint *h, omb_out *out,
                     ocounters *ctr);
/* o_analyse_mb_ex over nref references (refs[0] the newest): see
 * pcme_oracle.c.  No counterpart in the reference (P slices there use one
 * reference); with nref = 1 it is o_analyse_mb_ex. */
void o_analyse_mb_mr(const oparams *p, ombinfo *mbs, int mx, int my,
                     const oref *refs, int nref, const uint8_t *src_plane,
                     int sstride, int spad, uint32_t lam, const ohint *h,
                     omb_out *out, ocounters *ctr);
/* te(v) length of reference index r among nref. */
uint32_t o_ref_bits(int r, int nref);
/* o_search_frame_ex over nref references. */
void o_search_frame_mr(int mbw, int mbh, int me_range, int subpel_rounds,
                       int partitions, uint32_t lam, const uint8_t *src,
                       int sstride, int spad, const oref *refs, int nref,
                       const ohint *hints, omb_out *out, ocounters *ctr);
/* o_search_frame with per-MB hints (mbw*mbh, raster; NULL = none). */
void o_search_frame_ex(int mbw, int mbh, int me_range, int subpel_rounds,
                       int partitions, uint32_t lam, const uint8_t *src,
                       int sstride, int spad, const oref *r,
                       const ohint *hints, omb_out *out, ocounters *ctr);

/* Multi-frame CPU driver (frame threads, as enc/mod.rs:189-263): for each
 * job, build the reference's half-pel planes then search the frame.  All
 * planes are padded luma of the given geometry.  Returns wall seconds. */
typedef struct {
  const uint8_t *src;   /* padded source luma */
  const uint8_t *ref;   /* padded reference luma (edges extended) */
  omb_out *out;         /* mbw*mbh */
  int mbw, mbh, me_range, subpel_rounds, partitions, qp;
  int nref;                    /* 0 or 1: `ref` only; else o_search_frame_mr */
  const uint8_t *older[3];     /* padded luma of references 1..nref-1 */
} ojob;
double o_search_jobs_mt(const ojob *jobs, int njobs, int threads,
                        ocounters *ctr);

/* Scratch size in bytes for o_hpel row buffers (none needed); kept for ABI
 * symmetry with the product library. */
int o_version(void);

#ifdef __cplusplus
}
#endif
#endif

/* ---- lookahead.rs / mbtree.rs: the scenecut complexity of lowres frames ---- */
void o_lowres(const uint8_t *plane, int stride, int pad, int pw, int ph, uint8_t *out);
uint32_t o_lowres_intra(const uint8_t *cur, int w, int x0, int y0);
void o_complexity(const uint8_t *cur, const uint8_t *prev, int w, int h, uint64_t *intra, uint64_t *inter,
                  uint32_t *blk_intra, uint32_t *blk_inter, int16_t *blk_mv);

/* mbtree.rs:10-15 (Block): lowres block costs of the macroblock tree */
typedef struct {
  uint32_t intra, inter;
  omv mv[2];
  uint8_t list; /* 0 intra, 1 list 0, 2 list 1, 3 bi */
  uint8_t reserved[3];
} otree_block;
omv o_tree_search(const uint8_t *cur, const uint8_t *r, int w, int h, int x0, int y0, omv seed);
void o_tree_analyse(const uint8_t *cur, const uint8_t *l0, const uint8_t *l1, int w, int h, otree_block *out);
double o_lookahead_mt(const uint8_t *lr, int w, int h, int n, int threads, uint64_t *sums, otree_block *blocks);

/* ---- B macroblocks: direct_spatial and the motion search of analyse_b ---- */
/* mbinfo.rs:74-107 with both lists */
typedef struct {
  int8_t kind;
  int8_t ref_idx[2][4];
  omv mv[2][16];
} ombinfo2;

/* Layout of pcme_b_out (include/pcme.h). */
typedef struct {
  omv direct_mv[2][4];
  int8_t direct_ref[2][4];
  uint32_t direct_ok, direct_cost;
  omv mv16[2], mvp16[2];
  uint32_t cost16[2];
  uint32_t cost[3];
  uint32_t dir;
  omv part_mv[2][2][2];
  uint32_t part_cost[2];
  uint8_t part_dir[2][2];
  uint8_t reserved[4];
} ob_out;

void o_direct_spatial(const oparams *p, const ombinfo2 *mbs, const ombinfo2 *col, int mx, int my,
                      int8_t out_ref[2][4], omv out_mv[2][4]);
void o_b_analyse(const oparams *p, const ombinfo2 *mbs, const ombinfo2 *col, int mx, int my, const oref *r0,
                 const oref *r1, const uint8_t *src16, uint32_t lam, ob_out *o);
/* mbtree.rs:162-231 */
void o_mbtree_propagate(const otree_block *const *blocks, const int *refs, int n, int bw, int bh, float *prop);
void o_mbtree_offsets(const otree_block *blocks, const float *prop, int nb, float *out);
/* dsp.rs:104-114, lookahead.rs:116-139 */
uint32_t o_block_var(const uint8_t *a, int sa, int w, int h);
void o_aq_offsets(const uint8_t *y, int ys, const uint8_t *u, const uint8_t *v, int cs, int mbw, int mbh,
                  float strength, float *out);
double o_b_frame_mt(const oparams *p, const ombinfo2 *mbs, const ombinfo2 *cols, const oref *r0, const oref *r1,
                    const uint8_t *src, int sstride, int pad, uint32_t lam, int threads, ob_out *out);

/* Layout of pcme_b_mb (include/pcme.h): one MB of the ME-only B field. */
typedef struct {
  ob_out search;
  omv mv[2][16];     /* committed MbInfo.mv per list (zero where unused) */
  int8_t ref[2][4];  /* committed ref_idx per list and 8x8 (-1: list unused) */
  uint32_t best_cost;
  uint8_t mode;      /* 0 direct, 1 16x16, 2 16x8, 3 8x16 */
  uint8_t dir;       /* the 16x16 direction (0 list 0, 1 list 1, 2 bi) */
  uint8_t reserved[2];
} ob_mb;
/* The ME-only field of one B picture (analyse_b's decision without BSkip,
 * RD and intra, slice.rs:2189-2377 with rd off): every MB in raster order
 * searched as o_b_analyse over the MBs already committed, then committed
 * as direct (when its MVs are legal), the 16x16 direction when cheaper
 * (strict <), then 16x8 and 8x16 when partitions and the best cost >
 * 1024, each when cheaper; cols = the list-1 picture's MbInfo. */
void o_b_field(const oparams *p, const ombinfo2 *cols, const oref *r0, const oref *r1, const uint8_t *src,
               int sstride, int pad, uint32_t lam, ob_mb *out);

/* ---- early P-skip gate (skip_oracle.c; slice.rs:1884-1920, 679-1001,
 * transform.rs, residual.rs) ---- */
typedef struct {
  int32_t mf4[6][16], v4[6][16], mf8[6][64], v8[6][64];
} oquant;
/* flags of the gate: the encoder's t8x8, adaptive_t8 and cabac params */
#define OSKIP_T8X8 1
#define OSKIP_ADAPTIVE_T8 2
#define OSKIP_CABAC 4
/* Layout of pcme_skip_out (include/pcme.h). */
typedef struct {
  uint32_t sad, thresh;
  uint8_t mv_ok, coded, cbp, skip, t8x8, reserved[3];
} oskip_out;
void o_quant_tables(oquant *q);
int o_chroma_qp(int qpi);
double o_lambda_ssd(int qp);
void o_fdct4(const int32_t d[16], int32_t o[16]);
uint32_t o_decimate_score(const int16_t *levels, int n);
int o_code_luma4(const uint8_t src[256], const uint8_t pred[256], int qp, int decimate, int cabac,
                 int16_t levels[16][16], uint8_t rec[256]);
int o_code_luma8(const uint8_t src[256], const uint8_t pred[256], int qp, int decimate, int16_t levels[4][64],
                 uint8_t rec[256]);
int o_code_chroma(const uint8_t su[64], const uint8_t sv[64], const uint8_t pu[64], const uint8_t pv[64], int qp,
                  int chroma_qp_offset, int intra, int cabac);
int o_code_inter(const uint8_t sy[256], const uint8_t su[64], const uint8_t sv[64], const uint8_t py[256],
                 const uint8_t pu[64], const uint8_t pv[64], int qp, int flags, int chroma_qp_offset,
                 int *t8x8_out);
void o_skip_gate(const oparams *p, const oref *r, const uint8_t *ru, const uint8_t *rv, int cstride, int mx, int my,
                 omv skip_mv, const uint8_t *src_y, const uint8_t *src_u, const uint8_t *src_v, int qp, int flags,
                 int chroma_qp_offset, oskip_out *out);
