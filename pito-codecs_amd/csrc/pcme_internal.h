// pcme_internal.h -- types shared by the kernels and the host side of libpcme.
#ifndef PCME_INTERNAL_H
#define PCME_INTERNAL_H

#include <stddef.h>
#include <stdint.h>

#include "../../include/pcme.h"

#define PCME_PAD 64
#define PCME_CPAD 32

namespace pcme {

// Device copy of one frame job of the ME-only field search.
struct DevJob {
  const uint8_t *src;    // padded source luma
  const uint8_t *pl[4];  // reference planes: full, H, V, C
  pcme_mb_out *out;      // mbw*mbh records
  const pcme_hint *hint; // nullptr: full search; else mbw*mbh pruner hints
  uint32_t *xchg;        // mbw*mbh*4 exchange words (bottom-row MVs)
  int mbw, mbh, stride;
  int me_range, subpel_rounds, partitions;
  uint32_t lam;
  int nref;                    // references of list 0 (1..PCME_MAX_REFS)
  const uint8_t *opl[3][4];    // planes of references 1..nref-1
};

// Ticket queues of the frame kernels (one per XCD) and the size of a
// plan's control block (see frames_body).
#define PCME_QUEUES 8
#define CTL_BYTES 128

// Lookahead jobs travel in the kernel arguments, LA_CHUNK per launch
// (64 x 40 B of the 4 KiB kernel-argument space).
#define LA_CHUNK 64
#define CPLX_BPW 8  // complexity_kernel: 8x8 blocks per wave (one per 8-lane group)
struct CplxArgs {
  pcme_cplx_job j[LA_CHUNK];
};
struct TreeArgs {
  pcme_tree_job j[LA_CHUNK];
};
struct TreeNodeArgs {
  const pcme_tree_block *b[LA_CHUNK];
};
#define TREE_WIN_EMPTY (1 << 29)  // tree_window_kernel's bound of a node no block of which feeds a cell
#define TREE_LDS 4096             // tree_prop_kernel: sources staged per 16x16 tile (16 B each)
// One (node, destination) pair of a propagation round.
struct TreePropJob {
  const pcme_tree_block *src;  // the node's blocks
  const float *mine;           // the node's propagated costs
  float *dst;                  // the destination node's propagated costs
  const int4 *win;             // the node's window (tree_window_kernel)
  int mask;                    // lists of the node that feed dst (bit l)
  int pad_;
};
#define TREE_JOBS 64
struct TreePropArgs {
  TreePropJob j[TREE_JOBS];
};
static_assert(sizeof(TreePropArgs) <= 4000, "propagation jobs exceed the kernel arguments");
static_assert(sizeof(CplxArgs) <= 4000 && sizeof(TreeArgs) <= 4000, "lookahead jobs exceed the kernel arguments");

#ifdef __HIP__
__global__ void search_frames_kernel(const DevJob *__restrict__ jobs, const int2 *__restrict__ tickets,
                                     int ntickets, uint32_t *ctl, uint32_t spin_limit);
__global__ void search_frames_hinted_kernel(const DevJob *__restrict__ jobs, const int2 *__restrict__ tickets,
                                            int ntickets, uint32_t *ctl, uint32_t spin_limit);
__global__ void search_frames_mr_kernel(const DevJob *__restrict__ jobs, const int2 *__restrict__ tickets,
                                        int ntickets, uint32_t *ctl, uint32_t spin_limit);
__global__ void search_frames_mr_hinted_kernel(const DevJob *__restrict__ jobs, const int2 *__restrict__ tickets,
                                               int ntickets, uint32_t *ctl, uint32_t spin_limit);
__global__ void search_frames_hinted16_kernel(const DevJob *__restrict__ jobs, const int2 *__restrict__ tickets,
                                              int ntickets, uint32_t *ctl, uint32_t spin_limit);
__global__ void search_frames_mr_hinted16_kernel(const DevJob *__restrict__ jobs, const int2 *__restrict__ tickets,
                                                 int ntickets, uint32_t *ctl, uint32_t spin_limit);
__global__ void search_frames_mr16_kernel(const DevJob *__restrict__ jobs, const int2 *__restrict__ tickets,
                                          int ntickets, uint32_t *ctl, uint32_t spin_limit);
template <int WPG, int WCU, bool REUSE>
__global__ void search_frames_var_kernel(const DevJob *__restrict__ jobs, const int2 *__restrict__ tickets,
                                         int ntickets, uint32_t *ctl, uint32_t spin_limit);
__global__ void search_batch_kernel(const pcme_search_job *__restrict__ jobs, int njobs,
                                    pcme_mv *out_mv, uint32_t *out_cost);
__global__ void search_part_batch_kernel(const pcme_part_job *__restrict__ jobs, int njobs, pcme_part_out *out);
__global__ void b_analyse_16_kernel(const pcme_b_job *__restrict__ jobs, int njobs, pcme_b_out *out);
template <int K>
__global__ void b_analyse_part_kernel(const pcme_b_job *__restrict__ jobs, int njobs, pcme_b_out *out);
// One picture of a B field launch: its MBs are the launch's jobs, records
// and exchange words (8 per MB) [job0, job0 + mbw*mbh), raster order.
struct BPic {
  const pcme_b_col *col;  // mbw*mbh co-located MBs
  pcme_b_mb *field;       // mbw*mbh committed records
  int job0, mbw, mbh, partitions;
};
__global__ void b_walk_kernel(const pcme_b_job *__restrict__ jobs, const BPic *__restrict__ pics,
                              const int2 *__restrict__ tickets, int ntickets, pcme_b_out *res, uint32_t *xw,
                              uint32_t *ctl, uint32_t *sticky, uint32_t spin_limit);
__global__ void b_walk_clear_kernel(uint32_t *xw, size_t nwords, uint32_t *ctl);
__global__ void pad_planes_kernel(const uint8_t *__restrict__ src, int64_t src_pitch, int sstride, int w,
                                  int h, uint8_t *__restrict__ dst, int64_t dst_pitch, int dstride,
                                  int rows, int pad);
__global__ void hpel_kernel(const uint8_t *__restrict__ ref, uint8_t *__restrict__ ho,
                            uint8_t *__restrict__ vo, uint8_t *__restrict__ co, int64_t pitch,
                            int stride, int rows);
__global__ void luma_planes_kernel(const uint8_t *__restrict__ src, int64_t src_pitch, int sstride, int w,
                                   int h, uint8_t *__restrict__ full, uint8_t *__restrict__ hp,
                                   uint8_t *__restrict__ vp, uint8_t *__restrict__ cp, int64_t pitch, int stride,
                                   int rows);
#define LP_TH 32  // rows of a luma_planes_kernel band
// One clip of luma_planes_batch_kernel: ngx 4-column groups x nbands bands
// per frame, bpf blocks per frame, its first block at `first`.
struct LpJob {
  const uint8_t *src;
  uint8_t *full, *hp, *vp, *cp;
  int64_t src_pitch, pitch;
  int sstride, w, h, stride, rows, ngx, nbands, bpf, first, unused;
};
#define LP_MAXJOBS 16  // clips per launch (the table travels in the kernel arguments)
struct LpBatch {
  LpJob j[LP_MAXJOBS];
  int n;
};
__global__ void luma_planes_batch_kernel(const LpBatch jobs);
__global__ void pred_frame_kernel(const uint8_t *__restrict__ pl0, const uint8_t *__restrict__ pl1,
                                  const uint8_t *__restrict__ pl2, const uint8_t *__restrict__ pl3,
                                  const uint8_t *__restrict__ pu, const uint8_t *__restrict__ pv,
                                  const pcme_mb_out *__restrict__ mbs, int mbw, int mbh,
                                  uint8_t *__restrict__ oy, uint8_t *__restrict__ ou,
                                  uint8_t *__restrict__ ov);
__global__ void prep_kernel(uint32_t *ctl);
__global__ void lowres_kernel(const uint8_t *__restrict__ planes, int64_t plane_pitch, int stride, int pad, int w2,
                              uint8_t *__restrict__ out, int64_t out_pitch);
__global__ void clear_kernel(const uint32_t *ctl, uint32_t *xchg, size_t nwords);
__global__ void complexity_kernel(const CplxArgs A);
__global__ void complexity_clear_kernel(const CplxArgs A, int n);
__global__ void tree_kernel(const TreeArgs A);
__global__ void aq_var_kernel(const uint8_t *__restrict__ y, int ys, const uint8_t *__restrict__ u,
                              const uint8_t *__restrict__ v, int cs, int mbw, int mbh, float *__restrict__ e);
__global__ void aq_mean_kernel(float *__restrict__ e, int n, float strength);
__global__ void tree_window_kernel(const TreeNodeArgs A, int nb, int4 *__restrict__ win);
__global__ void tree_zero_kernel(float *__restrict__ p, size_t n);
__global__ void tree_prop_kernel(const TreePropArgs A, int bw, int bh);
__global__ void tree_offsets_kernel(const pcme_tree_block *__restrict__ b, const float *__restrict__ prop, int n,
                                    float *__restrict__ out);
#endif

// Tables of the early P-skip gate, one copy per device: the quantiser
// (transform.rs QuantTables::new), Lambdas::new().ssd (slice.rs:52-68) and
// the rate estimate of a level, 3 + 2 log2|v| (0.4 for 0) for |v| <= 32768
// (slice.rs:968-974), all computed on the host with the C library.
struct SkipTables {
  int32_t mf4[6][16], v4[6][16], mf8[6][64], v8[6][64];
  double lam_ssd[52];
  double lb[32769];
};

#ifdef __HIP__
__global__ void skip_gate_kernel(const pcme_skip_job *__restrict__ jobs, int njobs, const SkipTables *__restrict__ T,
                                 pcme_skip_out *out);
#endif
// Fills the tables (host side).
void build_skip_tables(SkipTables *t);

// Message returned by pcme_last_error() on this thread (host side).
void set_error(const char *msg);

// Tile of the half-pel kernel.
#define HT_W 64
#define HT_H 32

}  // namespace pcme

#endif
