// pcme_kernels.hip -- gfx950 kernels of the H.264 P-frame motion search.
//
// Kernels:
//   pad_planes_kernel     import + edge extension (mod.rs:842-873,
//                         frame.rs:58-97); HBM-bound copy.
//   hpel_kernel           H/V/C half-pel planes over the padded plane
//                         (picture.rs:318-380); HBM-bound, LDS-tiled.
//   search_frames_kernel  persistent row walkers: one wave64 walks one MB row
//                         of one frame left to right and runs, per MB,
//                         pskip_mv -> search_part(16x16) -> partition choice
//                         (slice.rs:1870-1945) with search (slice.rs:1590-1723)
//                         evaluated 8 candidates per pass.  Rows are taken
//                         from an atomic ticket queue in (row, frame) order;
//                         row y of a frame waits for the bottom-row MVs of MB
//                         x+1 of row y-1, published as tagged 32-bit words.
//   search_batch_kernel   one wave per independent search() call.
//   pred_frame_kernel     mc_luma + mc_chroma of the chosen partitions
//                         (slice.rs:1725-1749).
//
// Candidate evaluation layout: lane = 8*slot + sub.  A slot holds one
// candidate MV; its 8 lanes each cover a 4-row x 8-column strip of the
// partition (two 4x4 blocks side by side, packed as i16 pairs for SATD).
// Costs are reduced over the 8 lanes with DPP and the strict-< scan over the
// slots runs on scalar registers, which reproduces the reference's
// sequential "first strictly smaller cost wins" order exactly.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pcme_internal.h"
#include "pcme_dev.h"

namespace pcme {



// --------------------------------------------------------- search context

struct SCtx {            // wave-uniform
  const uint8_t *pl[4];  // padded planes: full, H, V, C
  int stride;
  int pw;                // 16*mbw
  int ph;                // 16*mbh
  int vlim;
  uint32_t lam;
  int range;
  int subpel;
};


// Lane constants of one search: strip origin inside the block and whether
// the lane's strip lies inside the block.
struct Lane {
  int slot, sub, sx, sy;
  bool active;
};

template <int W, int H>
__device__ __forceinline__ Lane lane_layout() {
  Lane L;
  uint32_t l = lane_id();
  L.slot = (int)(l >> 3);
  L.sub = (int)(l & 7u);
  if (W == 16) {
    L.sx = (L.sub & 1) * 8;
    L.sy = (L.sub >> 1) * 4;
    L.active = L.sub < 2 * (H / 4);
  } else {
    L.sx = 0;
    L.sy = L.sub * 4;
    L.active = L.sub < H / 4;
  }
  if (!L.active) L.sy = 0;
  return L;
}

struct SrcRegs {
  uint32_t raw[8];  // row r: raw[2r] cols 0-3, raw[2r+1] cols 4-7
};

// u16 pairs (left block, right block) of the source strip, for SATD.
__device__ __forceinline__ void src_pairs(const SrcRegs &s, v2s *sp) {
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++)
      sp[r * 4 + c] = as_v2s(__builtin_amdgcn_perm(s.raw[2 * r + 1], s.raw[2 * r], SELC(c)));
}

// Source strip of the lane; blk = block top-left, 4-byte aligned, sstride a
// multiple of 4.
__device__ __forceinline__ void load_src(const uint8_t *blk, int sstride, const Lane &L, SrcRegs &s) {
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const uint32_t *q = (const uint32_t *)(blk + (L.sy + r) * sstride + L.sx);
    s.raw[2 * r] = q[0];
    s.raw[2 * r + 1] = q[1];
  }
}

// ---- staged windows (LDS) ------------------------------------------------
// Per wave: the full-pel window of the integer phase and the four planes
// around the integer best for the sub-pel phase.  Before every pass a
// wave-uniform check re-stages the region when the pass's candidates could
// leave it; a region centred on the current best always covers every legal
// candidate of the next pass (mv_ok keeps blocks 44 px inside the 64 px pad),
// so all hex, square and sub-pel passes read LDS.  Regions go HBM -> LDS by
// 16-byte LDS-DMA (no VGPR round trip); a search issues its first region
// before the candidate pass, so the two HBM round trips overlap.
#define WIN_MX 12  // horizontal margin of the full-pel window
#define WIN_MY 13  // vertical margin
#define WIN_W 48   // 16 + 2*WIN_MX + 8 bytes of alignment / over-read slack (3 x 16 B)
#define WIN_H 42   // 16 + 2*WIN_MY
#define SP_W 32
#define SP_H 23   // 16 + 2*3 + 1: room for a pruned search's square rounds
#define WIN_BYTES (WIN_W * WIN_H)
#define SP_PLANE (SP_W * SP_H)
#define LDS_PER_SEARCH (((WIN_BYTES + 4 * SP_PLANE) + 15) & ~15)
#define LDS_PER_WAVE (2 * LDS_PER_SEARCH)  // room for two concurrent searches
static_assert(WIN_BYTES % 16 == 0 && SP_PLANE % 16 == 0, "LDS-DMA destinations are 16-byte aligned");

typedef const __attribute__((address_space(1))) uint32_t gu32;

// 8 bytes at byte offset `off` of a 4-byte-aligned base (global or LDS),
// read as three aligned dwords and shifted into place with v_alignbyte.
template <typename P>
__device__ __forceinline__ void ld8_at(P base, int off, uint32_t &lo, uint32_t &hi) {
  const int a = off >> 2;
  const uint32_t sh = (uint32_t)off & 3u;
  uint32_t d0 = base[a], d1 = base[a + 1], d2 = base[a + 2];
  lo = __builtin_amdgcn_alignbyte(d1, d0, sh);
  hi = __builtin_amdgcn_alignbyte(d2, d1, sh);
}

__device__ __forceinline__ gu32 *gplane(const uint8_t *p) { return (gu32 *)p; }

// Regions of a padded plane (DW bytes wide, DW a multiple of 4, origin ox
// 4-aligned, picture coordinates) are staged into LDS by dma_region below.

// Integer-pel SAD cost of the lane's slot candidate (slice.rs:1605-1616):
// SAD over the block + (lam * mv_bits) >> 8.  (ux, uy) is the address MV:
// the candidate when valid, else a valid fallback (its cost is ignored).
template <bool LDS, int LPC = 8, int PITCH = WIN_W>
__device__ __forceinline__ uint32_t sad_cost(const SCtx &c, const Lane &L, const SrcRegs &s, int bx, int by,
                                             int cx, int cy, int ux, int uy, int mvpx, int mvpy,
                                             const uint32_t *lw, int wx0, int wy0) {
  const int x = bx + L.sx + (ux >> 2), y = by + L.sy + (uy >> 2);
  uint32_t acc = 0;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    uint32_t a, b;
    if (LDS)
      ld8_at(lw, (y - wy0 + r) * PITCH + (x - wx0), a, b);
    else
      ld8_at(gplane(c.pl[0]), (int)__umul24((uint32_t)(y + r + PCME_PAD), (uint32_t)c.stride) + x + PCME_PAD, a, b);
    acc = __builtin_amdgcn_sad_u8(a, s.raw[2 * r], acc);
    acc = __builtin_amdgcn_sad_u8(b, s.raw[2 * r + 1], acc);
  }
  acc = L.active ? acc : 0u;
  acc = LPC == 8 ? red8(acc) : (LPC == 4 ? red4(acc) : red42(acc));
  return acc + mv_cost(c.lam, cx - mvpx, cy - mvpy);
}

__device__ __forceinline__ const uint8_t *plane_sel(const SCtx &c, uint32_t p) {
  const uint8_t *a = (p & 1u) ? c.pl[1] : c.pl[0];
  const uint8_t *b = (p & 1u) ? c.pl[3] : c.pl[2];
  return (p & 2u) ? b : a;
}

// Sub-pel cost (slice.rs:1684-1687): mc_luma (picture.rs:415-449) as the
// rounding average of two of the staged planes, SATD, + (lam * mv_bits) >> 8.
// Full and half positions read the same plane twice, and (a+a+1)>>1 == a:
// HALF (every candidate of the pass at a full or half position) reads it once.
template <int LPC = 8, bool HALF = false>
__device__ __forceinline__ uint32_t satd_cost(const SCtx &c, const Lane &L, const v2s *sp, int bx, int by, int cx,
                                              int cy, int ux, int uy, int mvpx, int mvpy, const uint32_t *lsp,
                                              int sx0, int sy0) {
  const int xi = bx + L.sx + (ux >> 2) - sx0, yi = by + L.sy + (uy >> 2) - sy0;
  const uint32_t e = qpel_entry((uint32_t)((uy & 3) * 4 + (ux & 3)));
  const int oa = (int)(e & 3u) * SP_PLANE + (yi + (int)((e >> 3) & 1u)) * SP_W + xi + (int)((e >> 2) & 1u);
  const int ob = (int)((e >> 4) & 3u) * SP_PLANE + (yi + (int)((e >> 7) & 1u)) * SP_W + xi + (int)((e >> 6) & 1u);
  uint32_t P[8];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    if constexpr (HALF) {
      ld8_at(lsp, oa + r * SP_W, P[2 * r], P[2 * r + 1]);
    } else {
      uint32_t a0, a1, b0, b1;
      ld8_at(lsp, oa + r * SP_W, a0, a1);
      ld8_at(lsp, ob + r * SP_W, b0, b1);
      P[2 * r] = __builtin_amdgcn_lerp(a0, b0, 0x01010101u);
      P[2 * r + 1] = __builtin_amdgcn_lerp(a1, b1, 0x01010101u);
    }
  }
  uint32_t acc = satd_strip8(sp, P);
  acc = L.active ? acc : 0u;
  acc = LPC == 8 ? red8(acc) : (LPC == 4 ? red4(acc) : red42(acc));
  return acc + mv_cost(c.lam, cx - mvpx, cy - mvpy);
}

// Per-slot candidate offsets, 4-bit fields (value + bias) per slot.
// HEX: (-2,0),(2,0),(-1,-2),(1,-2),(-1,2),(1,2) (slice.rs:1634).
// SQ8: (-1,0),(1,0),(0,-1),(0,1),(-1,-1),(1,-1),(-1,1),(1,1)
// (slice.rs:1660-1669, 1693-1702).
#define HEX_DX 0x313140u  // dx + 2
#define HEX_DY 0x440022u  // dy + 2
#define SQ_DX 0x20201120u  // dx + 1
#define SQ_DY 0x22002011u  // dy + 1
__device__ __forceinline__ int hex_dx(int s) { return (int)((HEX_DX >> (4 * s)) & 15u) - 2; }
__device__ __forceinline__ int hex_dy(int s) { return (int)((HEX_DY >> (4 * s)) & 15u) - 2; }
__device__ __forceinline__ int sq_dx(int s) { return (int)((SQ_DX >> (4 * s)) & 15u) - 1; }
__device__ __forceinline__ int sq_dy(int s) { return (int)((SQ_DY >> (4 * s)) & 15u) - 1; }

struct SRes {
  int x, y;
  uint32_t cost;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }


// mv_ok (slice.rs:496-507) of a w x h block at (x, y) as a box of legal
// quarter-pel MVs: (m >> 2) >= k <=> m >= 4k and (m >> 2) <= k <=> m <= 4k+3,
// so |m| < 2000, x0 >= -44, y0 >= -44, x0 + w <= pw + 44, y0 + h < vlim
// become lo <= m <= hi per component.
struct Box {
  int xlo, xhi, ylo, yhi;
};
__device__ __forceinline__ Box legal_box(const SCtx &c, int x, int y, int w, int h) {
  Box b;
  b.xlo = max(-1999, 4 * (-44 - x));
  b.xhi = min(1999, 4 * (c.pw + 44 - w - x) + 3);
  b.ylo = max(-1999, 4 * (-44 - y));
  b.yhi = min(1999, 4 * (c.vlim - h - y - 1) + 3);
  return b;
}
__device__ __forceinline__ bool in_box(const Box &b, int mx, int my) {
  return (mx >= b.xlo) & (mx <= b.xhi) & (my >= b.ylo) & (my <= b.yhi);
}

typedef const __attribute__((address_space(1))) void *gvptr;
typedef __attribute__((address_space(3))) void *lvptr;

// The same copy with LDS-DMA loads (global_load_lds_dword: lane l of
// instruction k writes dword 64k + l of dst), issued without waiting, so a
// region can be staged while other work runs; no data passes through VGPRs.
// dma_wait() completes every load issued before it.
// Stage a DW x DH byte region (DW a multiple of 16, origin 4-aligned) into
// LDS with 16-byte LDS-DMA loads (global_load_lds_dwordx4: lane l of an
// instruction writes its 16 bytes at the instruction's LDS base + 16 l;
// masked lanes write nothing).  Chunk d = 64k + lane of instruction k is
// region row d / CPR, 16-byte column d % CPR: the lane's (row, col) of
// instruction 0 plus a compile-time step with one carry.  The empty asm
// keeps the compiler from hoisting every instruction's per-lane address
// out of the search loops (a VGPR each).
template <int DW, int DH>
__device__ __forceinline__ void dma_region(const uint8_t *plane, int stride, int ox, int oy, uint32_t *dst) {
  static_assert(DW % 16 == 0, "regions are staged in 16-byte chunks");
  constexpr int CPR = DW / 16, N = CPR * DH, IT = (N + 63) / 64;
  const int l = (int)lane_id();
  int r0 = l / CPR, c0 = l - (l / CPR) * CPR;
  asm volatile("" : "+v"(r0), "+v"(c0));
  const uint8_t *base = plane + (size_t)((oy + PCME_PAD) * stride + ox + PCME_PAD);
  const uint32_t ustride = (uint32_t)stride;
#pragma unroll
  for (int k = 0; k < IT; k++) {
    const int q = (64 * k) / CPR, m = (64 * k) % CPR;
    const int t = c0 + m;
    const bool cy = t >= CPR;
    const int row = r0 + q + (cy ? 1 : 0), col = cy ? t - CPR : t;
    if (64 * k + 63 < N || l + 64 * k < N)
      __builtin_amdgcn_global_load_lds((gvptr)(base + __umul24((uint32_t)row, ustride) + 16 * col),
                                       (lvptr)(dst + 256 * k), 16, 0, 0);
  }
}
__device__ __forceinline__ void dma_wait() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Order this wave's LDS accesses around an LDS-DMA re-stage: every LDS read
// issued before has returned (so the DMA cannot overwrite data still being
// read), and no LDS access moves across.  Each wave stages into its own LDS
// area, so no other wave is involved (workgroups may hold several walkers).
__device__ __forceinline__ void lds_fence() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Stage a 16x16 block (rows `stride` apart, 4-byte aligned) into 256 bytes
// of LDS, row r at dst + 16 r: lane r < 16 moves row r with one 16-byte
// LDS-DMA load.  The caller waits (dma_wait) before reading it.
__device__ __forceinline__ void dma_block16(const uint8_t *src, int stride, uint32_t *dst) {
  const uint32_t l = lane_id();
  if (l < 16u) __builtin_amdgcn_global_load_lds((gvptr)(src + __umul24(l, (uint32_t)stride)), (lvptr)dst, 16, 0, 0);
}

// Origin of a full-pel window centred on the block at MV (bxv, byv).
__device__ __forceinline__ void win_origin(const SCtx &c, int bx, int by, int bxv, int byv, int &wx0, int &wy0) {
  const int X = bx + (bxv >> 2), Y = by + (byv >> 2);
  wx0 = clampi((X - WIN_MX) & ~3, -PCME_PAD, c.pw + PCME_PAD - WIN_W);
  wy0 = clampi(Y - WIN_MY, -PCME_PAD, c.ph + PCME_PAD - WIN_H);
}

// Start staging the full-pel window around (bxv, byv); the caller waits.
template <int W, int H>
__device__ __forceinline__ void prefetch_win(const SCtx &c, int bx, int by, int bxv, int byv, int &wx0, int &wy0,
                                             uint32_t *lw) {
  win_origin(c, bx, by, bxv, byv, wx0, wy0);
  lds_fence();
  dma_region<WIN_W, WIN_H>(c.pl[0], c.stride, wx0, wy0, lw);
}

// Re-stage the full-pel window around the current best (bxv, byv) when a
// pass reaching +-R px around it could leave the staged window.
template <int W, int H>
__device__ __forceinline__ void ensure_win(const SCtx &c, int bx, int by, int bxv, int byv, int R, int &wx0,
                                           int &wy0, uint32_t *lw) {
  const int X = bx + (bxv >> 2), Y = by + (byv >> 2);
  const bool ok = (X - R - wx0 >= 0) & (X + R + W + 4 - wx0 <= WIN_W) & (Y - R - wy0 >= 0) & (Y + R + H - wy0 <= WIN_H);
  if (ok) return;
  win_origin(c, bx, by, bxv, byv, wx0, wy0);
  lds_fence();
  dma_region<WIN_W, WIN_H>(c.pl[0], c.stride, wx0, wy0, lw);
  dma_wait();
  lds_fence();
}

// Same for the four sub-pel planes: candidates of a round lie within +-1 px
// of the current best, reading one more row/column for the half-pel offsets.
// Start staging the four planes around the block at (bxv, byv) with RM
// pixels of margin (a search prefetches them while its integer passes run;
// a pruned search's square rounds read the full-pel plane of the region).
template <int W, int H, int RM>
__device__ __forceinline__ void prefetch_sp(const SCtx &c, int bx, int by, int bxv, int byv, int &sx0, int &sy0,
                                            uint32_t *lsp) {
  const int X = bx + (bxv >> 2), Y = by + (byv >> 2);
  sx0 = clampi((X - RM) & ~3, -PCME_PAD, c.pw + PCME_PAD - SP_W);
  sy0 = clampi(Y - RM, -PCME_PAD, c.ph + PCME_PAD - SP_H);
  lds_fence();
#pragma unroll
  for (int p = 0; p < 4; p++) dma_region<SP_W, SP_H>(c.pl[p], c.stride, sx0, sy0, lsp + p * (SP_PLANE / 4));
}

// Complete any staging in flight, then re-stage the four planes (centred
// with RM pixels of margin) when the next sub-pel round could leave them.
template <int W, int H, int RM = 2>
__device__ __forceinline__ void ensure_sp(const SCtx &c, int bx, int by, int bxv, int byv, int &sx0, int &sy0,
                                          uint32_t *lsp) {
  dma_wait();
  const int X = bx + (bxv >> 2), Y = by + (byv >> 2);
  const bool ok = (X - 1 - sx0 >= 0) & (X + W + 6 - sx0 <= SP_W) & (Y - 1 - sy0 >= 0) & (Y + H + 2 - sy0 <= SP_H);
  if (ok) return;
  prefetch_sp<W, H, RM>(c, bx, by, bxv, byv, sx0, sy0, lsp);
  dma_wait();
  lds_fence();
}

// Staged regions of one LDS search area (full-pel window origin, sub-pel
// region origin; 1 << 20 = nothing staged).  A row walker keeps them across
// the searches of its row: every pass re-checks that its reads lie inside
// the staged region (ensure_win / ensure_sp) and re-stages when they do not,
// so a region staged by an earlier search of the same reference picture is
// reused whenever it still covers, and its HBM round trip disappears.
struct Area {
  int wx0, wy0, sx0, sy0;
};
__device__ __forceinline__ Area no_area() {
  Area a;
  a.wx0 = a.wy0 = a.sx0 = a.sy0 = 1 << 20;
  return a;
}

// Whether the staged full-pel window holds every read of a W x H block at
// MV (mx, my) moved by up to R px (the reads of sad_cost<true>: 8 bytes from
// three aligned dwords per strip row, up to 4 bytes past the block).
template <int W, int H>
__device__ __forceinline__ bool win_covers(int bx, int by, int mx, int my, int R, int wx0, int wy0) {
  const int X = bx + (mx >> 2), Y = by + (my >> 2);
  return (X - R - wx0 >= 0) & (X + R + W + 4 - wx0 <= WIN_W) & (Y - R - wy0 >= 0) & (Y + R + H - wy0 <= WIN_H);
}

// Whether the staged sub-pel region serves every sub-pel round around a
// best up to R px from (mx, my) (the check of ensure_sp, widened by R).
template <int W, int H>
__device__ __forceinline__ bool sp_covers(int bx, int by, int mx, int my, int R, int sx0, int sy0) {
  const int X = bx + (mx >> 2), Y = by + (my >> 2);
  return (X - R - 1 - sx0 >= 0) & (X + R + W + 6 - sx0 <= SP_W) & (Y - R - 1 - sy0 >= 0) &
         (Y + R + H + 2 - sy0 <= SP_H);
}

// Hint candidate i (lane-varying) of the MB's hint record hv (lane j < 10
// holds dword j of its pcme_hint; cand[k] is dword k), rounded like the
// search_part candidates (slice.rs:1622); i outside [0, nh) reads as 0.
__device__ __forceinline__ void hint_cand(uint32_t hv, int i, int nh, int &x, int &y) {
  const bool in = (i >= 0) & (i < nh);
  uint32_t v = (uint32_t)__builtin_amdgcn_ds_bpermute((in ? i : 0) << 2, (int)hv);
  v = in ? v : 0u;
  x = ((int)(int16_t)(v & 0xFFFFu) + 2) & ~3;
  y = (((int)v >> 16) + 2) & ~3;
}

// SliceEnc::search (slice.rs:1590-1723) for a W x H block at picture
// position (bx, by).  mvp, cands are wave-uniform; returns (best, cost).
// lds = this wave's LDS_PER_WAVE bytes.  nh >= 0 selects the pruned search:
// the hint candidates 0..nh-1 of hv are scanned after the search_part
// candidates (slots 3-7, then slots 0-2 of a second pass) and the hexagon
// descent is skipped.
template <int W, int H, bool PR>
__device__ __forceinline__ SRes search(const SCtx &c, const SrcRegs &s, const Lane &L, int bx, int by, int mvpx,
                                       int mvpy, int c0x, int c0y, int c1x, int c1y, int ncand, uint32_t hv,
                                       int nh, uint32_t *lds, Area &ar) {
  uint32_t *lw = lds;
  uint32_t *lsp = lds + WIN_BYTES / 4;
  const Box ok = legal_box(c, bx, by, W, H);
  // slice.rs:1617-1631: rounded predictor (slot 0), then the candidates
  // (slots 1, 2), read from HBM.  A candidate equal to the running best has
  // the best's cost and cannot be taken, so least (cost, slot) reproduces
  // the scan including its "c != best" test.
  int bxv = (mvpx + 2) & ~3, byv = (mvpy + 2) & ~3;
  if (!in_box(ok, bxv, byv)) { bxv = 0; byv = 0; }
  // Staging runs beside the candidate pass (its HBM round trip): the full
  // window around the rounded predictor (the usual winner), or for a pruned
  // search the sub-pel planes its square rounds read.
  // A pruned search starts with nothing staged; a full search keeps the
  // regions its area holds (a region of an earlier search of the row).
  if constexpr (PR) ar = no_area();
  int wx0 = ar.wx0, wy0 = ar.wy0, sx0 = ar.sx0, sy0 = ar.sy0;
  bool lds_pass = false;
  if constexpr (PR) {
    prefetch_sp<W, H, 3>(c, bx, by, bxv, byv, sx0, sy0, lsp);
  } else {
    // the candidate pass reads the staged window when it holds every
    // candidate, else HBM beside a fresh window
    const int r0x = (c0x + 2) & ~3, r0y = (c0y + 2) & ~3;
    const int r1x = (c1x + 2) & ~3, r1y = (c1y + 2) & ~3;
    const bool v1 = (ncand >= 1) & in_box(ok, r0x, r0y), v2 = (ncand >= 2) & in_box(ok, r1x, r1y);
    lds_pass = win_covers<W, H>(bx, by, bxv, byv, 0, wx0, wy0) &
               (!v1 | win_covers<W, H>(bx, by, r0x, r0y, 0, wx0, wy0)) &
               (!v2 | win_covers<W, H>(bx, by, r1x, r1y, 0, wx0, wy0));
    if (!lds_pass) prefetch_win<W, H>(c, bx, by, bxv, byv, wx0, wy0, lw);
  }
  uint32_t bcost;
  if constexpr (!PR) {
    const int r0x = (c0x + 2) & ~3, r0y = (c0y + 2) & ~3;
    const int r1x = (c1x + 2) & ~3, r1y = (c1y + 2) & ~3;
    const int cx = L.slot == 1 ? r0x : (L.slot == 2 ? r1x : bxv);
    const int cy = L.slot == 1 ? r0y : (L.slot == 2 ? r1y : byv);
    const bool valid = (L.slot == 0) | ((L.slot <= ncand) & in_box(ok, cx, cy));
    const int ux = valid ? cx : bxv, uy = valid ? cy : byv;
    const uint32_t k = lds_pass ? sad_cost<true>(c, L, s, bx, by, cx, cy, ux, uy, mvpx, mvpy, lw, wx0, wy0)
                                : sad_cost<false>(c, L, s, bx, by, cx, cy, ux, uy, mvpx, mvpy, lw, 0, 0);
    const uint32_t key = min_key(k, valid, L.slot);
    bcost = key >> 3;
    const int w = (int)(key & 7u);
    bxv = w == 1 ? r0x : (w == 2 ? r1x : bxv);
    byv = w == 1 ? r0y : (w == 2 ? r1y : byv);
  } else {
    // Pruned search: the hint candidates 0..4 continue the scan in slots
    // 3-7 of the same pass.
    const int r0x = (c0x + 2) & ~3, r0y = (c0y + 2) & ~3;
    const int r1x = (c1x + 2) & ~3, r1y = (c1y + 2) & ~3;
    int hx, hy;
    hint_cand(hv, L.slot - 3, nh, hx, hy);
    const bool hs = L.slot >= 3;
    const int cx = L.slot == 1 ? r0x : (L.slot == 2 ? r1x : (hs ? hx : bxv));
    const int cy = L.slot == 1 ? r0y : (L.slot == 2 ? r1y : (hs ? hy : byv));
    const bool valid = (L.slot == 0) | (((L.slot <= ncand) | (hs & (L.slot - 3 < nh))) & in_box(ok, cx, cy));
    const uint32_t k = sad_cost<false>(c, L, s, bx, by, cx, cy, valid ? cx : bxv, valid ? cy : byv, mvpx, mvpy,
                                       lw, 0, 0);
    const uint32_t key = min_key(k, valid, L.slot);
    bcost = key >> 3;
    const int w = (int)(key & 7u);
    bxv = (int)rdlane((uint32_t)cx, 8 * w);
    byv = (int)rdlane((uint32_t)cy, 8 * w);
  }
  if (PR && nh > 5) {
    // hint candidates 5..7 continue the scan (slots 0-2)
    int hx, hy;
    hint_cand(hv, L.slot + 5, nh, hx, hy);
    const bool valid = (L.slot + 5 < nh) & in_box(ok, hx, hy);
    const uint32_t k = sad_cost<false>(c, L, s, bx, by, hx, hy, valid ? hx : bxv, valid ? hy : byv, mvpx, mvpy,
                                       lw, 0, 0);
    const uint32_t key = min_key(k, valid, L.slot);
    if ((key >> 3) < bcost) {
      bcost = key >> 3;
      const int w = (int)(key & 7u);
      bxv = rdlane((uint32_t)hx, 8 * w);
      byv = rdlane((uint32_t)hy, 8 * w);
    }
  }
  dma_wait();
  // slice.rs:1632-1656: hexagon descent clamped to +-range around centre0.
  const int c0xx = bxv, c0yy = byv;
  const int range = c.range;
  const int hdx = hex_dx(L.slot & 7) * 4, hdy = hex_dy(L.slot & 7) * 4;
  // hex candidates: legal and |m - centre0| >> 2 <= range, one box
  Box hb = ok;
  hb.xlo = max(hb.xlo, c0xx - 4 * range - 3);
  hb.xhi = min(hb.xhi, c0xx + 4 * range + 3);
  hb.ylo = max(hb.ylo, c0yy - 4 * range - 3);
  hb.yhi = min(hb.yhi, c0yy + 4 * range + 3);
  if (L.slot >= 6) hb.xlo = 1 << 20;  // slots 6, 7 carry no hex point
  for (int it = 0; it < (PR ? 0 : range); it++) {
    ensure_win<W, H>(c, bx, by, bxv, byv, 2, wx0, wy0, lw);
    const int cx = bxv + hdx, cy = byv + hdy;
    const bool valid = in_box(hb, cx, cy);
    const uint32_t k =
        sad_cost<true>(c, L, s, bx, by, cx, cy, valid ? cx : bxv, valid ? cy : byv, mvpx, mvpy, lw, wx0, wy0);
    const uint32_t key = min_key(k, valid, L.slot);
    if ((key >> 3) >= bcost) break;  // also covers "no valid candidate"
    bcost = key >> 3;
    const int j = (int)(key & 7u);
    bxv += hex_dx(j) * 4;
    byv += hex_dy(j) * 4;
  }
  // slice.rs:1657-1682: two rounds of 8-neighbour integer refinement.  The
  // full search reads the full-pel window; a pruned search stages the four
  // sub-pel planes once (margin 3) and reads the full-pel plane among them.
  // The sub-pel planes stage while the square rounds run: two rounds move
  // the best by at most 2 px, which margin 3 covers.
  if constexpr (!PR) {
    if (!sp_covers<W, H>(bx, by, bxv, byv, 2, sx0, sy0)) prefetch_sp<W, H, 3>(c, bx, by, bxv, byv, sx0, sy0, lsp);
  }
  const int qdx = sq_dx(L.slot), qdy = sq_dy(L.slot);
  for (int it = 0; it < 2; it++) {
    if constexpr (PR)
      ensure_sp<W, H, 3>(c, bx, by, bxv, byv, sx0, sy0, lsp);
    else
      ensure_win<W, H>(c, bx, by, bxv, byv, 1, wx0, wy0, lw);
    const int cx = bxv + qdx * 4, cy = byv + qdy * 4;
    const bool valid = in_box(ok, cx, cy);
    const uint32_t k =
        PR ? sad_cost<true, 8, SP_W>(c, L, s, bx, by, cx, cy, valid ? cx : bxv, valid ? cy : byv, mvpx, mvpy, lsp,
                                     sx0, sy0)
           : sad_cost<true>(c, L, s, bx, by, cx, cy, valid ? cx : bxv, valid ? cy : byv, mvpx, mvpy, lw, wx0, wy0);
    const uint32_t key = min_key(k, valid, L.slot);
    if ((key >> 3) >= bcost) break;
    bcost = key >> 3;
    const int j = (int)(key & 7u);
    bxv += sq_dx(j) * 4;
    byv += sq_dy(j) * 4;
  }
  // Stage the four planes around the integer best for the sub-pel phase.
  ensure_sp<W, H>(c, bx, by, bxv, byv, sx0, sy0, lsp);
  v2s sp[16];
  src_pairs(s, sp);
  // slice.rs:1688: SATD cost at the integer best.
  uint32_t bs = rdlane(satd_cost<8, true>(c, L, sp, bx, by, bxv, byv, bxv, byv, mvpx, mvpy, lsp, sx0, sy0), 0);
  // slice.rs:1689-1721: half-pel round, then 1 + subpel_rounds quarter-pel.
#pragma unroll
  for (int step = 2; step >= 1; step--) {
    const int rounds = step == 1 ? 1 + c.subpel : 1;
    for (int it = 0; it < rounds; it++) {
      ensure_sp<W, H>(c, bx, by, bxv, byv, sx0, sy0, lsp);
      const int cx = bxv + qdx * step, cy = byv + qdy * step;
      const bool valid = in_box(ok, cx, cy);
      // the half-pel round's candidates and fallback sit at half positions
      const uint32_t k = step == 2 ? satd_cost<8, true>(c, L, sp, bx, by, cx, cy, valid ? cx : bxv,
                                                        valid ? cy : byv, mvpx, mvpy, lsp, sx0, sy0)
                                   : satd_cost<8, false>(c, L, sp, bx, by, cx, cy, valid ? cx : bxv,
                                                         valid ? cy : byv, mvpx, mvpy, lsp, sx0, sy0);
      const uint32_t key = min_key(k, valid, L.slot);
      if ((key >> 3) >= bs) break;
      bs = key >> 3;
      const int j = (int)(key & 7u);
      bxv += sq_dx(j) * step;
      byv += sq_dy(j) * step;
    }
  }
  ar.wx0 = wx0; ar.wy0 = wy0; ar.sx0 = sx0; ar.sy0 = sy0;
  SRes r;
  r.x = bxv;
  r.y = byv;
  r.cost = bs;
  return r;
}

// ---- paired search ---------------------------------------------------------
// The 16x8, 8x16 and 8x8 modes are independent of each other (all seeded
// with mv16 and predicted from the same committed neighbours plus their own
// earlier partitions), so two partition searches of different modes run in
// one wave: lanes 0-31 search A (a 16x8 or 8x16 partition, 4 lanes per
// candidate), lanes 32-63 search B (an 8x16 partition, 4 lanes per
// candidate, or an 8x8 partition, 2 lanes per candidate on lanes 32-47).
// 8 candidates per search per pass.  Each search keeps its own scalar state,
// LDS region and strict-< scan; a search whose phase ended passes with no
// valid candidate.

struct PState {          // wave-uniform state of one search of a pair
  int bx, by;            // block position (picture pixels)
  int mvpx, mvpy;
  Box ok, hb;            // legal box, hex box
  int x, y;              // running best
  uint32_t cost;
  int wx0, wy0, sx0, sy0;
  bool act;
  int it;
};

// Per-group keys of a paired pass (lanes 0-31: A, 32-63: B): least
// cost*8 + slot of each group, read from lanes 31 and 63.  The mirror steps
// reach every slot when a slot spans 4 lanes; with 2-lane slots (TWO) a quad
// swap first gives each lane the least of its quad's two slots.
template <bool TWO = false>
__device__ __forceinline__ void min_key_pair(uint32_t cost, bool valid, int slot, uint32_t &ka, uint32_t &kb) {
  uint32_t k = valid ? ((cost << 3) | (uint32_t)slot) : 0xFFFFFFFFu;
  uint32_t t;
  if constexpr (TWO) {
    t = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)k, DPP_QP_2301, 0xF, 0xF, false);
    k = t < k ? t : k;
  }
  t = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)k, DPP_ROW_HALF_MIRROR, 0xF, 0xF, false);
  k = t < k ? t : k;
  t = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)k, DPP_ROW_MIRROR, 0xF, 0xF, false);
  k = t < k ? t : k;
  t = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)k, DPP_ROW_BCAST15, 0xA, 0xF, false);
  k = t < k ? t : k;
  ka = rdlane(k, 31);
  kb = rdlane(k, 63);
}

template <int CTL>
__device__ __forceinline__ uint32_t dpp_min(uint32_t k) {
  const uint32_t t = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)k, CTL, 0xF, 0xF, false);
  return t < k ? t : k;
}

// Least of ka over the wave and least of kb over the wave (keys of lanes
// outside a set are all ones): each row of 16 reduced with DPP, the four
// rows read out.
__device__ __forceinline__ void min_key_two(uint32_t ka, uint32_t kb, uint32_t &ma, uint32_t &mb) {
  ka = dpp_min<DPP_ROW_MIRROR>(dpp_min<DPP_ROW_HALF_MIRROR>(dpp_min<DPP_QP_2301>(dpp_min<DPP_QP_1032>(ka))));
  kb = dpp_min<DPP_ROW_MIRROR>(dpp_min<DPP_ROW_HALF_MIRROR>(dpp_min<DPP_QP_2301>(dpp_min<DPP_QP_1032>(kb))));
  ma = min(min(rdlane(ka, 0), rdlane(ka, 16)), min(rdlane(ka, 32), rdlane(ka, 48)));
  mb = min(min(rdlane(kb, 0), rdlane(kb, 16)), min(rdlane(kb, 32), rdlane(kb, 48)));
}

__device__ __forceinline__ int gsel(bool g, int a, int b) { return g ? b : a; }

// Apply a group's key: take the candidate when its cost is below the
// running best; returns whether it moved.
__device__ __forceinline__ bool take_key(PState &S, uint32_t key, int dx, int dy) {
  if ((key >> 3) >= S.cost) return false;
  S.cost = key >> 3;
  S.x += dx;
  S.y += dy;
  return true;
}

// search (slice.rs:1590-1723) for A (an AW x AH partition) and B (a BW x BH
// partition) at the same time.  src* = block top-left of each source.
// nh >= 0: pruned search with hint candidates 0..nh-1 of hv (see search()).
template <int AW, int AH, int BW, int BH, bool PR>
__device__ __forceinline__ void search_pair(const SCtx &cA, const SCtx &cB, PState &A, PState &B,
                                            const uint8_t *srcA, const uint8_t *srcB, int sstride, int c0ax,
                                            int c0ay, int c0bx, int c0by, int c1ax, int c1ay, int c1bx, int c1by,
                                            uint32_t hv, int nh, uint32_t *lds, Area &arA, Area &arB, SRes &ra,
                                            SRes &rb) {
  // cA / cB: the reference pictures of A and B (the same context for two
  // partitions of one list, two for a list-0 / list-1 pair); lam, range,
  // subpel and the plane geometry are common to both.
  const SCtx &c = cA;
  static_assert(AW * AH == 128 && (BW * BH == 128 || (BW == 8 && BH == 8)), "pair shapes");
  constexpr bool B8 = BW == 8 && BH == 8;  // B is an 8x8 partition: 2 lanes per candidate
  constexpr int LPC = B8 ? 2 : 4;         // reduction code of the cost helpers
  constexpr int BLS = B8 ? 2 : 4;         // lanes per B candidate
  const uint32_t l = lane_id();
  // With an 8x8 B, lanes 48-63 hold no B candidate: lanes 48-51 carry A's
  // data (an A candidate's 4 lanes) and lanes 52-53 B's, inactive in every
  // scan; they price the integer best of each search in the half-pel pass.
  const bool ca = B8 && (l >> 2) == 12u, cb = B8 && (l >> 1) == 26u;
  const bool g = l >= 32 && !ca;
  SCtx cl = cA;  // the lane's search's context, for reads from memory
  cl.pl[0] = g ? cB.pl[0] : cA.pl[0];
  Lane L;
  if (!g) {
    L.slot = (int)((l >> 2) & 7u);
    L.sub = (int)(l & 3u);
    L.sx = AW == 16 ? (L.sub & 1) * 8 : 0;
    L.sy = AW == 16 ? (L.sub >> 1) * 4 : L.sub * 4;
    L.active = !ca;
  } else {
    L.slot = (int)(((l - 32) / BLS) & 7u);
    L.sub = (int)((l - 32) % BLS);
    L.sx = BW == 16 ? (L.sub & 1) * 8 : 0;
    L.sy = BW == 16 ? (L.sub >> 1) * 4 : L.sub * 4;
    L.active = !B8 || l < 48;
  }
  SrcRegs s;
  load_src(g ? srcB : srcA, sstride, L, s);
  uint32_t *lw = lds + (g ? LDS_PER_SEARCH / 4 : 0);
  uint32_t *lsp = lw + WIN_BYTES / 4;
  const int bx = gsel(g, A.bx, B.bx), by = gsel(g, A.by, B.by);
  const int mvpx = gsel(g, A.mvpx, B.mvpx), mvpy = gsel(g, A.mvpy, B.mvpy);
  A.ok = legal_box(cA, A.bx, A.by, AW, AH);
  B.ok = legal_box(cB, B.bx, B.by, BW, BH);
  // slice.rs:1617-1631 (see search()).
  int a0x = (A.mvpx + 2) & ~3, a0y = (A.mvpy + 2) & ~3;
  if (!in_box(A.ok, a0x, a0y)) { a0x = 0; a0y = 0; }
  int b0x = (B.mvpx + 2) & ~3, b0y = (B.mvpy + 2) & ~3;
  if (!in_box(B.ok, b0x, b0y)) { b0x = 0; b0y = 0; }
  // Staging beside the candidate pass (see search()).  A full search keeps
  // the regions its area holds; the candidate pass reads LDS when both
  // windows hold every read of it (the candidates and, with an 8x8 B, the
  // first hexagon round around the rounded predictor).
  if constexpr (PR) {
    arA = no_area();
    arB = no_area();
  }
  A.wx0 = arA.wx0; A.wy0 = arA.wy0; A.sx0 = arA.sx0; A.sy0 = arA.sy0;
  B.wx0 = arB.wx0; B.wy0 = arB.wy0; B.sx0 = arB.sx0; B.sy0 = arB.sy0;
  bool lds_pass = false;
  if constexpr (PR) {
    prefetch_sp<AW, AH, 3>(cA, A.bx, A.by, a0x, a0y, A.sx0, A.sy0, lds + WIN_BYTES / 4);
    prefetch_sp<BW, BH, 3>(cB, B.bx, B.by, b0x, b0y, B.sx0, B.sy0, lds + (LDS_PER_SEARCH + WIN_BYTES) / 4);
  } else {
    constexpr int HR = (BW == 8 && BH == 8) ? 2 : 0;  // hexagon points priced in the pass
    const int r0ax = (c0ax + 2) & ~3, r0ay = (c0ay + 2) & ~3, r1ax = (c1ax + 2) & ~3, r1ay = (c1ay + 2) & ~3;
    const int r0bx = (c0bx + 2) & ~3, r0by = (c0by + 2) & ~3, r1bx = (c1bx + 2) & ~3, r1by = (c1by + 2) & ~3;
    const bool va1 = in_box(A.ok, r0ax, r0ay), va2 = in_box(A.ok, r1ax, r1ay);
    const bool vb1 = in_box(B.ok, r0bx, r0by), vb2 = in_box(B.ok, r1bx, r1by);
    const bool ca_ = win_covers<AW, AH>(A.bx, A.by, a0x, a0y, HR, A.wx0, A.wy0) &
                     (!va1 | win_covers<AW, AH>(A.bx, A.by, r0ax, r0ay, 0, A.wx0, A.wy0)) &
                     (!va2 | win_covers<AW, AH>(A.bx, A.by, r1ax, r1ay, 0, A.wx0, A.wy0));
    const bool cb_ = win_covers<BW, BH>(B.bx, B.by, b0x, b0y, HR, B.wx0, B.wy0) &
                     (!vb1 | win_covers<BW, BH>(B.bx, B.by, r0bx, r0by, 0, B.wx0, B.wy0)) &
                     (!vb2 | win_covers<BW, BH>(B.bx, B.by, r1bx, r1by, 0, B.wx0, B.wy0));
    lds_pass = ca_ & cb_;
    if (!lds_pass) {
      prefetch_win<AW, AH>(cA, A.bx, A.by, a0x, a0y, A.wx0, A.wy0, lds);
      prefetch_win<BW, BH>(cB, B.bx, B.by, b0x, b0y, B.wx0, B.wy0, lds + LDS_PER_SEARCH / 4);
    }
  }
  // With an 8x8 B (the full search) the candidate pass's idle slots (3-7,
  // and lanes 48-51 / 52-53) price the first hexagon round around slot 0's
  // position (the rounded predictor, which also wins ties): when the pass's
  // best lands there, that round is already done (hk*).
  constexpr bool SPEC = B8 && !PR;
  uint32_t hka = 0xFFFFFFFFu, hkb = 0xFFFFFFFFu;
  {
    // cands = [seed, mvs[0] of the search's own mode]
    const int r0ax = (c0ax + 2) & ~3, r0ay = (c0ay + 2) & ~3;
    const int r0bx = (c0bx + 2) & ~3, r0by = (c0by + 2) & ~3;
    const int r0x = gsel(g, r0ax, r0bx), r0y = gsel(g, r0ay, r0by);
    const int r1ax = (c1ax + 2) & ~3, r1ay = (c1ay + 2) & ~3;
    const int r1bx = (c1bx + 2) & ~3, r1by = (c1by + 2) & ~3;
    const int r1x = gsel(g, r1ax, r1bx), r1y = gsel(g, r1ay, r1by);
    const int px = gsel(g, a0x, b0x), py = gsel(g, a0y, b0y);
    // a pruned search scans hint candidates 0..4 in slots 3-7 of this pass
    int hx = 0, hy = 0;
    if constexpr (PR) hint_cand(hv, L.slot - 3, nh, hx, hy);
    const bool hs = PR && L.slot >= 3;
    const Box ok = g ? B.ok : A.ok;
    // speculative hex point hi of lane's search: in the hex box of centre0 = p
    const bool hl = SPEC && ((L.slot >= 3 && L.active) || ca || cb);
    const int hi = hl ? ((ca | cb) ? 5 : L.slot - 3) : 0;
    const int shx = px + hex_dx(hi) * 4, shy = py + hex_dy(hi) * 4;
    const int rr = 4 * c.range + 3;
    const bool vh = hl && c.range > 0 && (shx >= max(ok.xlo, px - rr)) && (shx <= min(ok.xhi, px + rr)) &&
                    (shy >= max(ok.ylo, py - rr)) && (shy <= min(ok.yhi, py + rr));
    const int cx = hl ? shx : (L.slot == 1 ? r0x : (L.slot == 2 ? r1x : (hs ? hx : px)));
    const int cy = hl ? shy : (L.slot == 1 ? r0y : (L.slot == 2 ? r1y : (hs ? hy : py)));
    const bool valid =
        !hl & L.active & ((L.slot == 0) | (((L.slot <= 2) | (hs & (L.slot - 3 < nh))) & in_box(ok, cx, cy)));
    Lane Lk = L;
    Lk.active = L.active | (SPEC & (ca | cb));
    const bool use = valid | vh;
    const int ux = use ? cx : px, uy = use ? cy : py;
    const uint32_t k = lds_pass ? sad_cost<true, LPC>(cl, Lk, s, bx, by, cx, cy, ux, uy, mvpx, mvpy, lw,
                                                      gsel(g, A.wx0, B.wx0), gsel(g, A.wy0, B.wy0))
                                : sad_cost<false, LPC>(cl, Lk, s, bx, by, cx, cy, ux, uy, mvpx, mvpy, lw, 0, 0);
    uint32_t ka, kb;
    min_key_pair<B8>(k, valid, L.slot, ka, kb);
    if constexpr (SPEC) {
      const uint32_t hk = vh ? ((k << 3) | (uint32_t)hi) : 0xFFFFFFFFu;
      min_key_two(g ? 0xFFFFFFFFu : hk, g ? hk : 0xFFFFFFFFu, hka, hkb);
    }
    A.cost = ka >> 3;
    B.cost = kb >> 3;
    const int wa = (int)(ka & 7u), wb = (int)(kb & 7u);
    if constexpr (PR) {
      A.x = (int)rdlane((uint32_t)cx, 4 * wa);
      A.y = (int)rdlane((uint32_t)cy, 4 * wa);
      B.x = (int)rdlane((uint32_t)cx, 32 + BLS * wb);
      B.y = (int)rdlane((uint32_t)cy, 32 + BLS * wb);
    } else {
      A.x = wa == 1 ? r0ax : (wa == 2 ? r1ax : a0x);
      A.y = wa == 1 ? r0ay : (wa == 2 ? r1ay : a0y);
      B.x = wb == 1 ? r0bx : (wb == 2 ? r1bx : b0x);
      B.y = wb == 1 ? r0by : (wb == 2 ? r1by : b0y);
    }
  }
  if (PR && nh > 5) {
    // hint candidates 5..7 continue each scan (slots 0-2)
    int hx, hy;
    hint_cand(hv, L.slot + 5, nh, hx, hy);
    const Box ok = g ? B.ok : A.ok;
    const int ox = gsel(g, A.x, B.x), oy = gsel(g, A.y, B.y);
    const bool valid = L.active & (L.slot + 5 < nh) & in_box(ok, hx, hy);
    const uint32_t k = sad_cost<false, LPC>(cl, L, s, bx, by, hx, hy, valid ? hx : ox, valid ? hy : oy, mvpx, mvpy,
                                            lw, 0, 0);
    uint32_t ka, kb;
    min_key_pair<B8>(k, valid, L.slot, ka, kb);
    if ((ka >> 3) < A.cost) {
      const int wa = (int)(ka & 7u);
      A.cost = ka >> 3;
      A.x = (int)rdlane((uint32_t)hx, 4 * wa);
      A.y = (int)rdlane((uint32_t)hy, 4 * wa);
    }
    if ((kb >> 3) < B.cost) {
      const int wb = (int)(kb & 7u);
      B.cost = kb >> 3;
      B.x = (int)rdlane((uint32_t)hx, 32 + BLS * wb);
      B.y = (int)rdlane((uint32_t)hy, 32 + BLS * wb);
    }
  }
  dma_wait();
  const int range = c.range;
  A.hb = A.ok;
  A.hb.xlo = max(A.hb.xlo, A.x - 4 * range - 3);
  A.hb.xhi = min(A.hb.xhi, A.x + 4 * range + 3);
  A.hb.ylo = max(A.hb.ylo, A.y - 4 * range - 3);
  A.hb.yhi = min(A.hb.yhi, A.y + 4 * range + 3);
  B.hb = B.ok;
  B.hb.xlo = max(B.hb.xlo, B.x - 4 * range - 3);
  B.hb.xhi = min(B.hb.xhi, B.x + 4 * range + 3);
  B.hb.ylo = max(B.hb.ylo, B.y - 4 * range - 3);
  B.hb.yhi = min(B.hb.yhi, B.y + 4 * range + 3);
  // slice.rs:1632-1656: hexagon descent of both.
  const int hdx = hex_dx(L.slot) * 4, hdy = hex_dy(L.slot) * 4;
  A.act = B.act = range > 0 && !PR;
  A.it = B.it = 0;
  if constexpr (SPEC) {
    // the first round, priced in the candidate pass, applies when centre0
    // is the position it was priced around
    if (A.act && A.x == a0x && A.y == a0y) {
      const int j = (int)(hka & 7u);
      A.act = take_key(A, hka, hex_dx(j) * 4, hex_dy(j) * 4) && ++A.it < range;
    }
    if (B.act && B.x == b0x && B.y == b0y) {
      const int j = (int)(hkb & 7u);
      B.act = take_key(B, hkb, hex_dx(j) * 4, hex_dy(j) * 4) && ++B.it < range;
    }
  }
  while (A.act | B.act) {
    if (A.act) ensure_win<AW, AH>(cA, A.bx, A.by, A.x, A.y, 2, A.wx0, A.wy0, lds);
    if (B.act) ensure_win<BW, BH>(cB, B.bx, B.by, B.x, B.y, 2, B.wx0, B.wy0, lds + LDS_PER_SEARCH / 4);
    const int ox = gsel(g, A.x, B.x), oy = gsel(g, A.y, B.y);
    const int cx = ox + hdx, cy = oy + hdy;
    Box hb = g ? B.hb : A.hb;
    const bool act = g ? B.act : A.act;
    const bool valid = L.active & (L.slot < 6) & act & in_box(hb, cx, cy);
    const uint32_t k = sad_cost<true, LPC>(c, L, s, bx, by, cx, cy, valid ? cx : ox, valid ? cy : oy, mvpx, mvpy,
                                           lw, gsel(g, A.wx0, B.wx0), gsel(g, A.wy0, B.wy0));
    uint32_t ka, kb;
    min_key_pair<B8>(k, valid, L.slot, ka, kb);
    if (A.act) {
      const int j = (int)(ka & 7u);
      A.act = take_key(A, ka, hex_dx(j) * 4, hex_dy(j) * 4) && ++A.it < range;
    }
    if (B.act) {
      const int j = (int)(kb & 7u);
      B.act = take_key(B, kb, hex_dx(j) * 4, hex_dy(j) * 4) && ++B.it < range;
    }
  }
  // slice.rs:1657-1682: two rounds of 8-neighbour refinement of both (a
  // pruned search reads the full-pel plane of its sub-pel region, staged
  // once with margin 3).
  if constexpr (!PR) {
    if (!sp_covers<AW, AH>(A.bx, A.by, A.x, A.y, 2, A.sx0, A.sy0))
      prefetch_sp<AW, AH, 3>(cA, A.bx, A.by, A.x, A.y, A.sx0, A.sy0, lds + WIN_BYTES / 4);
    if (!sp_covers<BW, BH>(B.bx, B.by, B.x, B.y, 2, B.sx0, B.sy0))
      prefetch_sp<BW, BH, 3>(cB, B.bx, B.by, B.x, B.y, B.sx0, B.sy0, lds + (LDS_PER_SEARCH + WIN_BYTES) / 4);
  }
  const int qdx = sq_dx(L.slot), qdy = sq_dy(L.slot);
  A.act = B.act = true;
  A.it = B.it = 0;
  while (A.act | B.act) {
    if constexpr (PR) {
      if (A.act) ensure_sp<AW, AH, 3>(cA, A.bx, A.by, A.x, A.y, A.sx0, A.sy0, lds + WIN_BYTES / 4);
      if (B.act)
        ensure_sp<BW, BH, 3>(cB, B.bx, B.by, B.x, B.y, B.sx0, B.sy0, lds + (LDS_PER_SEARCH + WIN_BYTES) / 4);
    } else {
      if (A.act) ensure_win<AW, AH>(cA, A.bx, A.by, A.x, A.y, 1, A.wx0, A.wy0, lds);
      if (B.act) ensure_win<BW, BH>(cB, B.bx, B.by, B.x, B.y, 1, B.wx0, B.wy0, lds + LDS_PER_SEARCH / 4);
    }
    const int ox = gsel(g, A.x, B.x), oy = gsel(g, A.y, B.y);
    const int cx = ox + qdx * 4, cy = oy + qdy * 4;
    const Box ok = g ? B.ok : A.ok;
    const bool act = g ? B.act : A.act;
    const bool valid = L.active & act & in_box(ok, cx, cy);
    const uint32_t k =
        PR ? sad_cost<true, LPC, SP_W>(c, L, s, bx, by, cx, cy, valid ? cx : ox, valid ? cy : oy, mvpx, mvpy, lsp,
                                       gsel(g, A.sx0, B.sx0), gsel(g, A.sy0, B.sy0))
           : sad_cost<true, LPC>(c, L, s, bx, by, cx, cy, valid ? cx : ox, valid ? cy : oy, mvpx, mvpy, lw,
                                 gsel(g, A.wx0, B.wx0), gsel(g, A.wy0, B.wy0));
    uint32_t ka, kb;
    min_key_pair<B8>(k, valid, L.slot, ka, kb);
    if (A.act) {
      const int j = (int)(ka & 7u);
      A.act = take_key(A, ka, sq_dx(j) * 4, sq_dy(j) * 4) && ++A.it < 2;
    }
    if (B.act) {
      const int j = (int)(kb & 7u);
      B.act = take_key(B, kb, sq_dx(j) * 4, sq_dy(j) * 4) && ++B.it < 2;
    }
  }
  // Sub-pel planes of both, then the SATD cost at each integer best.
  ensure_sp<AW, AH>(cA, A.bx, A.by, A.x, A.y, A.sx0, A.sy0, lds + WIN_BYTES / 4);
  ensure_sp<BW, BH>(cB, B.bx, B.by, B.x, B.y, B.sx0, B.sy0, lds + (LDS_PER_SEARCH + WIN_BYTES) / 4);
  v2s sp[16];
  src_pairs(s, sp);
  if constexpr (B8) {
    // slice.rs:1688-1705 in one pass: lanes 48-51 / 52-53 price the integer
    // best of A / B (the SATD the half-pel round must beat), the scan lanes
    // the half-pel candidates around it.
    const int ox = gsel(g, A.x, B.x), oy = gsel(g, A.y, B.y);
    const int cx = ox + qdx * 2, cy = oy + qdy * 2;
    const Box ok = g ? B.ok : A.ok;
    const bool valid = L.active & in_box(ok, cx, cy);
    const bool centre = ca | cb;
    Lane Lc = L;
    Lc.active = L.active | centre;
    const int px = centre ? ox : cx, py = centre ? oy : cy;
    const uint32_t k = satd_cost<LPC, true>(c, Lc, sp, bx, by, px, py, valid ? cx : ox, valid ? cy : oy, mvpx,
                                            mvpy, lsp, gsel(g, A.sx0, B.sx0), gsel(g, A.sy0, B.sy0));
    A.cost = rdlane(k, 48);
    B.cost = rdlane(k, 52);
    uint32_t ka, kb;
    min_key_pair<B8>(k, valid, L.slot, ka, kb);
    const int ja = (int)(ka & 7u), jb = (int)(kb & 7u);
    take_key(A, ka, sq_dx(ja) * 2, sq_dy(ja) * 2);
    take_key(B, kb, sq_dx(jb) * 2, sq_dy(jb) * 2);
  } else {
    const int ox = gsel(g, A.x, B.x), oy = gsel(g, A.y, B.y);
    const uint32_t k = satd_cost<LPC, true>(c, L, sp, bx, by, ox, oy, ox, oy, mvpx, mvpy, lsp,
                                            gsel(g, A.sx0, B.sx0), gsel(g, A.sy0, B.sy0));
    A.cost = rdlane(k, 0);
    B.cost = rdlane(k, 32);
  }
  // slice.rs:1689-1721: half-pel round (fused above for an 8x8 B), then
  // 1 + subpel_rounds quarter-pel.
#pragma unroll
  for (int step = B8 ? 1 : 2; step >= 1; step--) {
    const int rounds = step == 1 ? 1 + c.subpel : 1;
    A.act = B.act = true;
    A.it = B.it = 0;
    while (A.act | B.act) {
      if (A.act) ensure_sp<AW, AH>(cA, A.bx, A.by, A.x, A.y, A.sx0, A.sy0, lds + WIN_BYTES / 4);
      if (B.act) ensure_sp<BW, BH>(cB, B.bx, B.by, B.x, B.y, B.sx0, B.sy0, lds + (LDS_PER_SEARCH + WIN_BYTES) / 4);
      const int ox = gsel(g, A.x, B.x), oy = gsel(g, A.y, B.y);
      const int cx = ox + qdx * step, cy = oy + qdy * step;
      const Box ok = g ? B.ok : A.ok;
      const bool act = g ? B.act : A.act;
      const bool valid = L.active & act & in_box(ok, cx, cy);
      const uint32_t k =
          step == 2 ? satd_cost<LPC, true>(c, L, sp, bx, by, cx, cy, valid ? cx : ox, valid ? cy : oy, mvpx, mvpy,
                                           lsp, gsel(g, A.sx0, B.sx0), gsel(g, A.sy0, B.sy0))
                    : satd_cost<LPC, false>(c, L, sp, bx, by, cx, cy, valid ? cx : ox, valid ? cy : oy, mvpx, mvpy,
                                            lsp, gsel(g, A.sx0, B.sx0), gsel(g, A.sy0, B.sy0));
      uint32_t ka, kb;
      min_key_pair<B8>(k, valid, L.slot, ka, kb);
      if (A.act) {
        const int j = (int)(ka & 7u);
        A.act = take_key(A, ka, sq_dx(j) * step, sq_dy(j) * step) && ++A.it < rounds;
      }
      if (B.act) {
        const int j = (int)(kb & 7u);
        B.act = take_key(B, kb, sq_dx(j) * step, sq_dy(j) * step) && ++B.it < rounds;
      }
    }
  }
  arA.wx0 = A.wx0; arA.wy0 = A.wy0; arA.sx0 = A.sx0; arA.sy0 = A.sy0;
  arB.wx0 = B.wx0; arB.wy0 = B.wy0; arB.sx0 = B.sx0; arB.sy0 = B.sy0;
  ra.x = A.x; ra.y = A.y; ra.cost = A.cost;
  rb.x = B.x; rb.y = B.y; rb.cost = B.cost;
}

// ------------------------------------------------------------ neighbours

// MVs of the scalar state are packed like pcme_mv: x in bits 0-15, y in
// bits 16-31 (halves the scalar registers the neighbourhood needs).
__device__ __forceinline__ uint32_t pack_mv(int x, int y) { return ((uint32_t)x & 0xFFFFu) | ((uint32_t)y << 16); }
__device__ __forceinline__ int mv_x(uint32_t p) { return (int)(int16_t)(p & 0xFFFFu); }
__device__ __forceinline__ int mv_y(uint32_t p) { return (int)p >> 16; }

// Neighbour state of the MB being analysed in the ME-only field: every
// coded MB is PInter with ref 0, so an available neighbour has ref 0.
// With several references (the multi-reference field, one reference index
// per MB) the neighbours carry their MB's index and rf is the index being
// searched (the current MB's earlier partitions have it); with one they
// are all zero, constants the compiler folds away.
struct Nbr {
  bool aL, aT, aTR, aTL;
  uint32_t L[4];  // left MB, blocks 3, 7, 11, 15
  uint32_t T[4];  // top MB, blocks 12..15
  uint32_t TR;    // top-right MB, block 12
  uint32_t TL;    // top-left MB, block 15
  int rL, rT, rTR, rTL;  // the neighbours' reference indices
  int rf;                // reference index of the search
};

struct NbRes {
  bool avail;
  int ref;
  uint32_t mv;
};

// Neighbour state given explicitly, block by block (the per-call
// search_part): ref -2 = unavailable, -1 = available without a reference
// of the searched list (an intra neighbour reads as (-1, 0)), >= 0 the
// reference index, with its MV.
struct NbrG {
  int rL[4], rT[4], rTR, rTL;
  uint32_t L[4], T[4], TR, TL;
};

// Partitions already searched in the current mode (search_part's `info`).
struct Cur {
  uint32_t mv[4];
};

__host__ __device__ constexpr int part_count(int part) { return part == 0 ? 1 : (part == 3 ? 4 : 2); }
__host__ __device__ constexpr int rect_x(int part, int i) { return part == 2 ? 8 * i : (part == 3 ? 8 * (i & 1) : 0); }
__host__ __device__ constexpr int rect_y(int part, int i) { return part == 1 ? 8 * i : (part == 3 ? 8 * (i >> 1) : 0); }
__host__ __device__ constexpr int rect_w(int part) { return (part == 0 || part == 1) ? 16 : 8; }
__host__ __device__ constexpr int rect_h(int part) { return (part == 0 || part == 2) ? 16 : 8; }
// Partition of `part` covering 4x4 block r.
__host__ __device__ constexpr int part_of(int part, int r) {
  return part == 0 ? 0 : part == 1 ? (r >> 3) : part == 2 ? ((r & 3) >> 1) : (((r >> 3) << 1) | ((r & 3) >> 1));
}

// nb_part (slice.rs:392-432) for the partition P of mode PART, block offset
// (BX, BY) in 4x4 units relative to the MB.
template <int PART, int P, int BX, int BY>
__device__ __forceinline__ NbRes nb(const Nbr &n, const Cur &cur) {
  NbRes r = {false, -1, 0u};
  if constexpr (BX >= 0 && BX < 4 && BY >= 0 && BY < 4) {
    constexpr int k = part_of(PART, BY * 4 + BX);
    if constexpr (k < P) {
      r.avail = true; r.ref = n.rf; r.mv = cur.mv[k];
    }
  } else if constexpr (BY >= 4 || (BY >= 0 && BX >= 4)) {
  } else if constexpr (BY < 0) {
    if constexpr (BX < 0) {
      if (n.aTL) { r.avail = true; r.ref = n.rTL; r.mv = n.TL; }
    } else if constexpr (BX < 4) {
      if (n.aT) { r.avail = true; r.ref = n.rT; r.mv = n.T[BX]; }
    } else {
      if (n.aTR) { r.avail = true; r.ref = n.rTR; r.mv = n.TR; }
    }
  } else {
    if (n.aL) { r.avail = true; r.ref = n.rL; r.mv = n.L[BY]; }
  }
  return r;
}

__device__ __forceinline__ int search_rf(const Nbr &n) { return n.rf; }
__device__ __forceinline__ int search_rf(const NbrG &) { return 0; }

template <int PART, int P, int BX, int BY>
__device__ __forceinline__ NbRes nb(const NbrG &n, const Cur &cur) {
  NbRes r = {false, -1, 0u};
  if constexpr (BX >= 0 && BX < 4 && BY >= 0 && BY < 4) {
    constexpr int k = part_of(PART, BY * 4 + BX);
    if constexpr (k < P) {
      r.avail = true; r.ref = 0; r.mv = cur.mv[k];
    }
  } else if constexpr (BY >= 4 || (BY >= 0 && BX >= 4)) {
  } else if constexpr (BY < 0) {
    if constexpr (BX < 0) {
      r.avail = n.rTL >= -1; r.ref = n.rTL; r.mv = n.TL;
    } else if constexpr (BX < 4) {
      r.avail = n.rT[BX] >= -1; r.ref = n.rT[BX]; r.mv = n.T[BX];
    } else {
      r.avail = n.rTR >= -1; r.ref = n.rTR; r.mv = n.TR;
    }
  } else {
    r.avail = n.rL[BY] >= -1; r.ref = n.rL[BY]; r.mv = n.L[BY];
  }
  return r;
}

__device__ __forceinline__ int median3(int a, int b, int c) {
  int mx = a > b ? a : b, mn = a < b ? a : b;
  int t = mn > c ? mn : c;
  return mx < t ? mx : t;
}

// mvp (slice.rs:434-482) for reference index rf (search_rf: 0 except in
// the multi-reference field).  SHAPE: 0 Normal, 1 Top16x8,
// 2 Bottom16x8, 3 Left8x16, 4 Right8x16.  Returns a packed MV.
template <int PART, int P, int X, int Y, int WB, int SHAPE, class N>
__device__ __forceinline__ uint32_t mvp(const N &n, const Cur &cur) {
  NbRes a = nb<PART, P, X - 1, Y>(n, cur);
  NbRes b = nb<PART, P, X, Y - 1>(n, cur);
  NbRes cc = nb<PART, P, X + WB, Y - 1>(n, cur);
  if (!cc.avail) cc = nb<PART, P, X - 1, Y - 1>(n, cur);
  if (!b.avail && !cc.avail && a.avail) { b = a; cc = a; }
  if (!a.avail) { a.ref = -1; a.mv = 0u; }
  if (!b.avail) { b.ref = -1; b.mv = 0u; }
  if (!cc.avail) { cc.ref = -1; cc.mv = 0u; }
  const int rf = search_rf(n);
  if (SHAPE == 1 && b.ref == rf) return b.mv;
  if (SHAPE == 2 && a.ref == rf) return a.mv;
  if (SHAPE == 3 && a.ref == rf) return a.mv;
  if (SHAPE == 4 && cc.ref == rf) return cc.mv;
  const int cnt = (a.ref == rf) + (b.ref == rf) + (cc.ref == rf);
  if (cnt == 1) return a.ref == rf ? a.mv : (b.ref == rf ? b.mv : cc.mv);
  return pack_mv(median3(mv_x(a.mv), mv_x(b.mv), mv_x(cc.mv)), median3(mv_y(a.mv), mv_y(b.mv), mv_y(cc.mv)));
}

// pskip_mv (slice.rs:484-494): reference index 0
__device__ __forceinline__ uint32_t pskip(const Nbr &nbr) {
  Nbr n = nbr;
  n.rf = 0;
  Cur cur = {};
  NbRes a = nb<0, 0, -1, 0>(n, cur);
  NbRes b = nb<0, 0, 0, -1>(n, cur);
  if (!a.avail || !b.avail) return 0u;
  if (a.ref == 0 && a.mv == 0u) return 0u;
  if (b.ref == 0 && b.mv == 0u) return 0u;
  return mvp<0, 0, 0, 0, 4, 0>(n, cur);
}

__host__ __device__ constexpr int shape_of(int part, int p) {
  return part == 1 ? (p == 0 ? 1 : 2) : part == 2 ? (p == 0 ? 3 : 4) : 0;
}

struct PartRes {
  uint32_t mv[4];   // mvs
  uint32_t mvp[4];  // mvps
  uint32_t total;
};

// One partition of search_part (slice.rs:1826-1866), then the next.
template <int PART, int P, bool PR, class N>
__device__ __forceinline__ void part_step(const SCtx &c, const N &n, Cur &cur, const uint8_t *src,
                                          int sstride, int mbx, int mby, uint32_t seed, uint32_t hv,
                                          int nh, PartRes &res, uint32_t *lds, Area *ar) {
  if constexpr (P < part_count(PART)) {
    constexpr int X = rect_x(PART, P), Y = rect_y(PART, P);
    constexpr int W = rect_w(PART), H = rect_h(PART);
    const uint32_t pv = mvp<PART, P, X / 4, Y / 4, W / 4, shape_of(PART, P)>(n, cur);
    Lane L = lane_layout<W, H>();
    SrcRegs s;
    load_src(src + Y * sstride + X, sstride, L, s);
    const uint32_t m0 = P == 0 ? 0u : res.mv[0];
    SRes r = search<W, H, PR>(c, s, L, mbx + X, mby + Y, mv_x(pv), mv_y(pv), mv_x(seed), mv_y(seed), mv_x(m0),
                          mv_y(m0), 2, hv, nh, lds, ar[0]);
    const uint32_t m = pack_mv(r.x, r.y);
    res.mv[P] = m;
    res.mvp[P] = pv;
    res.total += r.cost;
    cur.mv[P] = m;
    part_step<PART, P + 1, PR>(c, n, cur, src, sstride, mbx, mby, seed, hv, nh, res, lds, ar);
  }
}

// search_part (slice.rs:1810-1868), list 0.  src = MB top-left.
template <int PART, bool PR, class N>
__device__ __forceinline__ void search_part(const SCtx &c, const N &n, const uint8_t *src, int sstride,
                                            int mbx, int mby, uint32_t seed, uint32_t hv, int nh,
                                            PartRes &res, uint32_t *lds, Area *ar) {
  Cur cur = {};
#pragma unroll
  for (int i = 0; i < 4; i++) { res.mv[i] = res.mvp[i] = 0u; }
  res.total = 0;
  part_step<PART, 0, PR>(c, n, cur, src, sstride, mbx, mby, seed, hv, nh, res, lds, ar);
}

// search_part (slice.rs:1810-1868) of P16x8 (ra) and P8x16 (rb) together:
// partition p of both modes in one paired search.
template <bool PR>
__device__ __forceinline__ void search_part_12(const SCtx &c, const Nbr &n, const uint8_t *src, int sstride,
                                               int mbx, int mby, uint32_t seed, uint32_t hv, int nh,
                                               PartRes &ra, PartRes &rb, uint32_t *lds, Area *ar) {
  Cur ca = {}, cb = {};
#pragma unroll
  for (int i = 0; i < 4; i++) { ra.mv[i] = ra.mvp[i] = rb.mv[i] = rb.mvp[i] = 0u; }
  PState A, B;
  SRes xa, xb;
  // partition 0: 16x8 top (x 0, y 0, w 4 blocks), 8x16 left (x 0, y 0, w 2)
  uint32_t pa = mvp<1, 0, 0, 0, 4, 1>(n, ca), pb = mvp<2, 0, 0, 0, 2, 3>(n, cb);
  A.bx = mbx; A.by = mby; A.mvpx = mv_x(pa); A.mvpy = mv_y(pa);
  B.bx = mbx; B.by = mby; B.mvpx = mv_x(pb); B.mvpy = mv_y(pb);
  search_pair<16, 8, 8, 16, PR>(c, c, A, B, src, src, sstride, mv_x(seed), mv_y(seed), mv_x(seed), mv_y(seed), 0, 0, 0, 0, hv, nh, lds, ar[0],
                                ar[1], xa, xb);
  ra.mv[0] = ca.mv[0] = pack_mv(xa.x, xa.y); ra.mvp[0] = pa; ra.total = xa.cost;
  rb.mv[0] = cb.mv[0] = pack_mv(xb.x, xb.y); rb.mvp[0] = pb; rb.total = xb.cost;
  // partition 1: 16x8 bottom (y 2 blocks), 8x16 right (x 2 blocks)
  pa = mvp<1, 1, 0, 2, 4, 2>(n, ca);
  pb = mvp<2, 1, 2, 0, 2, 4>(n, cb);
  A.bx = mbx; A.by = mby + 8; A.mvpx = mv_x(pa); A.mvpy = mv_y(pa);
  B.bx = mbx + 8; B.by = mby; B.mvpx = mv_x(pb); B.mvpy = mv_y(pb);
  search_pair<16, 8, 8, 16, PR>(c, c, A, B, src + 8 * sstride, src + 8, sstride, mv_x(seed), mv_y(seed), mv_x(seed), mv_y(seed), mv_x(ra.mv[0]),
              mv_y(ra.mv[0]), mv_x(rb.mv[0]), mv_y(rb.mv[0]), hv, nh, lds, ar[0], ar[1], xa, xb);
  ra.mv[1] = pack_mv(xa.x, xa.y); ra.mvp[1] = pa; ra.total += xa.cost;
  rb.mv[1] = pack_mv(xb.x, xb.y); rb.mvp[1] = pb; rb.total += xb.cost;
}

// search_part (slice.rs:1810-1868) of one mode (PART 1: 16x8, 2: 8x16) on
// list 0 (ra) and list 1 (rb) together: partition p of both lists in one
// paired search, each list with its own reference (c[l]), neighbours (n[l])
// and seed (seed[l]), as analyse_b searches them one after the other.
template <int PART, class N>
__device__ __forceinline__ void search_part_lists(const SCtx *c, const N *n, const uint8_t *src, int sstride,
                                                  int mbx, int mby, const uint32_t *seed, PartRes &ra,
                                                  PartRes &rb, uint32_t *lds, Area *ar) {
  constexpr int W = rect_w(PART), H = rect_h(PART);
  constexpr int X0 = rect_x(PART, 0), Y0 = rect_y(PART, 0), X1 = rect_x(PART, 1), Y1 = rect_y(PART, 1);
  Cur ca = {}, cb = {};
#pragma unroll
  for (int i = 0; i < 4; i++) { ra.mv[i] = ra.mvp[i] = rb.mv[i] = rb.mvp[i] = 0u; }
  PState A, B;
  SRes xa, xb;
  uint32_t pa = mvp<PART, 0, X0 / 4, Y0 / 4, W / 4, shape_of(PART, 0)>(n[0], ca);
  uint32_t pb = mvp<PART, 0, X0 / 4, Y0 / 4, W / 4, shape_of(PART, 0)>(n[1], cb);
  A.bx = B.bx = mbx + X0; A.by = B.by = mby + Y0;
  A.mvpx = mv_x(pa); A.mvpy = mv_y(pa);
  B.mvpx = mv_x(pb); B.mvpy = mv_y(pb);
  const uint8_t *s0 = src + Y0 * sstride + X0;
  search_pair<W, H, W, H, false>(c[0], c[1], A, B, s0, s0, sstride, mv_x(seed[0]), mv_y(seed[0]), mv_x(seed[1]),
                                 mv_y(seed[1]), 0, 0, 0, 0, 0u, -1, lds, ar[0], ar[1], xa, xb);
  ra.mv[0] = ca.mv[0] = pack_mv(xa.x, xa.y); ra.mvp[0] = pa; ra.total = xa.cost;
  rb.mv[0] = cb.mv[0] = pack_mv(xb.x, xb.y); rb.mvp[0] = pb; rb.total = xb.cost;
  pa = mvp<PART, 1, X1 / 4, Y1 / 4, W / 4, shape_of(PART, 1)>(n[0], ca);
  pb = mvp<PART, 1, X1 / 4, Y1 / 4, W / 4, shape_of(PART, 1)>(n[1], cb);
  A.bx = B.bx = mbx + X1; A.by = B.by = mby + Y1;
  A.mvpx = mv_x(pa); A.mvpy = mv_y(pa);
  B.mvpx = mv_x(pb); B.mvpy = mv_y(pb);
  const uint8_t *s1 = src + Y1 * sstride + X1;
  search_pair<W, H, W, H, false>(c[0], c[1], A, B, s1, s1, sstride, mv_x(seed[0]), mv_y(seed[0]), mv_x(seed[1]),
                                 mv_y(seed[1]), mv_x(ra.mv[0]), mv_y(ra.mv[0]), mv_x(rb.mv[0]), mv_y(rb.mv[0]), 0u,
                                 -1, lds, ar[0], ar[1], xa, xb);
  ra.mv[1] = pack_mv(xa.x, xa.y); ra.mvp[1] = pa; ra.total += xa.cost;
  rb.mv[1] = pack_mv(xb.x, xb.y); rb.mvp[1] = pb; rb.total += xb.cost;
}

// search_part (slice.rs:1810-1868) of P16x8 (r1), P8x16 (r2) and P8x8 (r3)
// in four paired steps: (16x8 p0, 8x8 p0), (8x16 p0, 8x8 p1), (16x8 p1,
// 8x8 p2), (8x16 p1, 8x8 p3).  Each partition still follows the earlier
// partitions of its own mode, with the same predictor and candidates as the
// one-at-a-time order, so the results are the same.
__device__ __forceinline__ void search_part_123(const SCtx &c, const Nbr &n, const uint8_t *src, int sstride,
                                                int mbx, int mby, uint32_t seed, PartRes &r1, PartRes &r2,
                                                PartRes &r3, uint32_t *lds, Area *ar) {
  Cur c1 = {}, c2 = {}, c3 = {};
#pragma unroll
  for (int i = 0; i < 4; i++) { r1.mv[i] = r1.mvp[i] = r2.mv[i] = r2.mvp[i] = r3.mv[i] = r3.mvp[i] = 0u; }
  PState A, B;
  SRes xa, xb;
  const uint32_t nohint = 0u;
  const int sx = mv_x(seed), sy = mv_y(seed);
  // step 1: 16x8 top (x 0, y 0, w 4 blocks) with 8x8 top-left (x 0, y 0, w 2)
  uint32_t pa = mvp<1, 0, 0, 0, 4, 1>(n, c1), pb = mvp<3, 0, 0, 0, 2, 0>(n, c3);
  A.bx = mbx; A.by = mby; A.mvpx = mv_x(pa); A.mvpy = mv_y(pa);
  B.bx = mbx; B.by = mby; B.mvpx = mv_x(pb); B.mvpy = mv_y(pb);
  search_pair<16, 8, 8, 8, false>(c, c, A, B, src, src, sstride, sx, sy, sx, sy, 0, 0, 0, 0, nohint, -1, lds, ar[0],
                                  ar[1], xa, xb);
  r1.mv[0] = c1.mv[0] = pack_mv(xa.x, xa.y); r1.mvp[0] = pa; r1.total = xa.cost;
  r3.mv[0] = c3.mv[0] = pack_mv(xb.x, xb.y); r3.mvp[0] = pb; r3.total = xb.cost;
  // step 2: 8x16 left (x 0, y 0, w 2) with 8x8 top-right (x 2, y 0)
  pa = mvp<2, 0, 0, 0, 2, 3>(n, c2);
  pb = mvp<3, 1, 2, 0, 2, 0>(n, c3);
  A.bx = mbx; A.by = mby; A.mvpx = mv_x(pa); A.mvpy = mv_y(pa);
  B.bx = mbx + 8; B.by = mby; B.mvpx = mv_x(pb); B.mvpy = mv_y(pb);
  search_pair<8, 16, 8, 8, false>(c, c, A, B, src, src + 8, sstride, sx, sy, sx, sy, 0, 0, mv_x(r3.mv[0]), mv_y(r3.mv[0]),
                                  nohint, -1, lds, ar[0], ar[1], xa, xb);
  r2.mv[0] = c2.mv[0] = pack_mv(xa.x, xa.y); r2.mvp[0] = pa; r2.total = xa.cost;
  r3.mv[1] = c3.mv[1] = pack_mv(xb.x, xb.y); r3.mvp[1] = pb; r3.total += xb.cost;
  // step 3: 16x8 bottom (y 2 blocks) with 8x8 bottom-left (x 0, y 2)
  pa = mvp<1, 1, 0, 2, 4, 2>(n, c1);
  pb = mvp<3, 2, 0, 2, 2, 0>(n, c3);
  A.bx = mbx; A.by = mby + 8; A.mvpx = mv_x(pa); A.mvpy = mv_y(pa);
  B.bx = mbx; B.by = mby + 8; B.mvpx = mv_x(pb); B.mvpy = mv_y(pb);
  search_pair<16, 8, 8, 8, false>(c, c, A, B, src + 8 * sstride, src + 8 * sstride, sstride, sx, sy, sx, sy,
                                  mv_x(r1.mv[0]), mv_y(r1.mv[0]), mv_x(r3.mv[0]), mv_y(r3.mv[0]), nohint, -1, lds,
                                  ar[0], ar[1], xa, xb);
  r1.mv[1] = pack_mv(xa.x, xa.y); r1.mvp[1] = pa; r1.total += xa.cost;
  r3.mv[2] = c3.mv[2] = pack_mv(xb.x, xb.y); r3.mvp[2] = pb; r3.total += xb.cost;
  // step 4: 8x16 right (x 2 blocks) with 8x8 bottom-right (x 2, y 2)
  pa = mvp<2, 1, 2, 0, 2, 4>(n, c2);
  pb = mvp<3, 3, 2, 2, 2, 0>(n, c3);
  A.bx = mbx + 8; A.by = mby; A.mvpx = mv_x(pa); A.mvpy = mv_y(pa);
  B.bx = mbx + 8; B.by = mby + 8; B.mvpx = mv_x(pb); B.mvpy = mv_y(pb);
  search_pair<8, 16, 8, 8, false>(c, c, A, B, src + 8, src + 8 * sstride + 8, sstride, sx, sy, sx, sy, mv_x(r2.mv[0]),
                                  mv_y(r2.mv[0]), mv_x(r3.mv[0]), mv_y(r3.mv[0]), nohint, -1, lds, ar[0], ar[1], xa,
                                  xb);
  r2.mv[1] = pack_mv(xa.x, xa.y); r2.mvp[1] = pa; r2.total += xa.cost;
  r3.mv[3] = pack_mv(xb.x, xb.y); r3.mvp[3] = pb; r3.total += xb.cost;
}

__device__ __forceinline__ uint32_t sel4(int k, uint32_t a, uint32_t b, uint32_t cc, uint32_t d) {
  return k == 0 ? a : (k == 1 ? b : (k == 2 ? cc : d));
}

// dst = take ? src : dst, element by element (keeps the records in registers).
__device__ __forceinline__ void take_if(PartRes &dst, const PartRes &src, bool take) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    dst.mv[i] = take ? src.mv[i] : dst.mv[i];
    dst.mvp[i] = take ? src.mvp[i] : dst.mvp[i];
  }
  dst.total = take ? src.total : dst.total;
}

struct MbRes {
  int part;
  int ref;  // the MB's reference index (multi-reference field; else 0)
  PartRes best;
  uint32_t cost[4];
  uint32_t best_cost;
  uint32_t skip, m16;
};

// The strict-< choice of mode `part` (total r, penalty pen) against the
// running best (slice.rs:1925-1945).
__device__ __forceinline__ void take_mode(MbRes &o, const PartRes &r, int part, uint32_t pen) {
  o.cost[part] = r.total;
  const uint32_t cc = r.total + pen;
  const bool t = cc < o.best_cost;
  o.best_cost = t ? cc : o.best_cost;
  o.part = t ? part : o.part;
  take_if(o.best, r, t);
}

// Modes P16x8 and P8x16 (searched together), then the strict-< choice
// against the running best (slice.rs:1925-1945).
template <bool PR>
__device__ __forceinline__ void mode_12(const SCtx &c, const Nbr &n, const uint8_t *src, int sstride, int mbx,
                                        int mby, uint32_t hv, int nh, MbRes &o, uint32_t *lds, Area *ar) {
  PartRes r, r2;
  search_part_12<PR>(c, n, src, sstride, mbx, mby, o.m16, hv, nh, r, r2, lds, ar);
  o.cost[1] = r.total;
  uint32_t cc = r.total + ((c.lam * 3u) >> 8);
  bool t = cc < o.best_cost;
  o.best_cost = t ? cc : o.best_cost;
  o.part = t ? 1 : o.part;
  take_if(o.best, r, t);
  o.cost[2] = r2.total;
  cc = r2.total + ((c.lam * 3u) >> 8);
  t = cc < o.best_cost;
  o.best_cost = t ? cc : o.best_cost;
  o.part = t ? 2 : o.part;
  take_if(o.best, r2, t);
}

// Mode P8x8 and its choice.
template <bool PR>
__device__ __forceinline__ void mode_3(const SCtx &c, const Nbr &n, const uint8_t *src, int sstride, int mbx,
                                       int mby, uint32_t hv, int nh, MbRes &o, uint32_t *lds, Area *ar) {
  PartRes r;
  search_part<3, PR>(c, n, src, sstride, mbx, mby, o.m16, hv, nh, r, lds, ar);
  o.cost[3] = r.total;
  const uint32_t cc = r.total + ((c.lam * 10u) >> 8);
  const bool t = cc < o.best_cost;
  o.best_cost = t ? cc : o.best_cost;
  o.part = t ? 3 : o.part;
  take_if(o.best, r, t);
}

// ME-only analysis of one MB (slice.rs:1884-1945 without early skip,
// intra and RD).  nh >= 0: pruned MB (hint candidates 0..nh-1 of hv, hinted
// mode hpart): P16x16, then only P16x8 + P8x16 (hpart 1 or 2) or only P8x8
// (hpart 3), with the same strict-< choice.
template <bool PR>
__device__ __forceinline__ void analyse_mb(const SCtx &c, const Nbr &n, const uint8_t *src, int sstride,
                                           int mbx, int mby, int partitions, uint32_t hv, int nh,
                                           int hpart, MbRes &o, uint32_t *lds, Area *ar) {
  const uint32_t skip = pskip(n);
  PartRes r16;
  search_part<0, PR>(c, n, src, sstride, mbx, mby, skip, hv, nh, r16, lds, ar);
  const uint32_t cost16 = r16.total;
  o.part = 0;
  o.ref = 0;
  take_if(o.best, r16, true);
  o.best_cost = cost16 + (c.lam >> 8);
  o.cost[0] = cost16;
  o.cost[1] = o.cost[2] = o.cost[3] = 0xFFFFFFFFu;
  o.skip = skip;
  o.m16 = r16.mv[0];
  const bool split = partitions && cost16 > 1024u;
  if (PR) {
    // A pruned MB searches the hinted mode only: P16x8 with P8x16, or P8x8.
    if (split & ((hpart == 1) | (hpart == 2)))
      mode_12<PR>(c, n, src, sstride, mbx, mby, hv, nh, o, lds, ar);
    else if (split & (hpart == 3))
      mode_3<PR>(c, n, src, sstride, mbx, mby, hv, nh, o, lds, ar);
  } else if (split) {
    PartRes r1, r2, r3;
    search_part_123(c, n, src, sstride, mbx, mby, o.m16, r1, r2, r3, lds, ar);
    take_mode(o, r1, 1, (c.lam * 3u) >> 8);
    take_mode(o, r2, 2, (c.lam * 3u) >> 8);
    take_mode(o, r3, 3, (c.lam * 10u) >> 8);
  }
}

// (lam * te(v) bits of reference index r among nref) >> 8: no bits with one
// reference, one with two, else ue(v).
__device__ __forceinline__ uint32_t ref_rate(uint32_t lam, int r, int nref) {
  const uint32_t bits = nref <= 1 ? 0u : (nref == 2 ? 1u : 2u * (31u - (uint32_t)__builtin_clz((uint32_t)r + 1u)) + 1u);
  return __umul24(lam, bits) >> 8;
}

// Search context of reference index r >= 1 of job J (its planes; read
// whole, as an older reference is complete).
__device__ __forceinline__ SCtx older_ctx(const DevJob &J, const SCtx &c0, int r) {
  SCtx c = c0;
  const uint8_t *const *pl = J.opl[r - 1];
  c.pl[0] = pl[0]; c.pl[1] = pl[1]; c.pl[2] = pl[2]; c.pl[3] = pl[3];
  c.vlim = c.ph + 44;
  return c;
}

// The hint dwords (lane i: dword i) with the candidates (dwords 0-7) scaled
// by s: motion over one picture taken over s.
__device__ __forceinline__ uint32_t scale_hint(uint32_t hv, int s) {
  return lane_id() < 8u ? pack_mv(mv_x(hv) * s, mv_y(hv) * s) : hv;
}

// Multi-reference analysis of one MB (no counterpart in the reference, whose
// P slices use one reference): the P16x16 search on each reference of the
// job, each with the predictor of its index and its index's rate added (a
// pruned MB's hint candidates scaled by the reference's distance, index +
// 1); the least total, the lower index on ties, gives the MB's reference,
// and the partition modes are searched on it, each partition paying the
// rate.  The LDS areas hold the regions of reference ar_ref and are
// restaged when the reference changes.
template <bool PR>
__device__ __forceinline__ void analyse_mb_mr(const DevJob &J, const SCtx &c0, const Nbr &nb0, const uint8_t *src,
                                              int mbx, int mby, uint32_t hv, int nh, int hpart, MbRes &o,
                                              uint32_t *lds, Area *ar, int &ar_ref) {
  Nbr n = nb0;
  const uint32_t skip = pskip(n);
  PartRes b16;
  uint32_t cost16 = 0u;
  int rs = 0;
  for (int r = 0; r < J.nref; r++) {
    const SCtx cr = r == 0 ? c0 : older_ctx(J, c0, r);
    if (ar_ref != r) {
      ar[0] = no_area();
      ar[1] = no_area();
      ar_ref = r;
    }
    n.rf = r;
    PartRes r16;
    search_part<0, PR>(cr, n, src, 16, mbx, mby, skip, PR ? scale_hint(hv, r + 1) : hv, nh, r16, lds, ar);
    const uint32_t t = r16.total + ref_rate(c0.lam, r, J.nref);
    const bool take = r == 0 || t < cost16;
    cost16 = take ? t : cost16;
    rs = take ? r : rs;
    take_if(b16, r16, take);
  }
  const SCtx cs = rs == 0 ? c0 : older_ctx(J, c0, rs);
  if (ar_ref != rs) {
    ar[0] = no_area();
    ar[1] = no_area();
    ar_ref = rs;
  }
  n.rf = rs;
  if (PR) hv = scale_hint(hv, rs + 1);
  const uint32_t rr = ref_rate(c0.lam, rs, J.nref);
  o.part = 0;
  o.ref = rs;
  take_if(o.best, b16, true);
  o.best_cost = cost16 + (c0.lam >> 8);
  o.cost[0] = cost16;
  o.cost[1] = o.cost[2] = o.cost[3] = 0xFFFFFFFFu;
  o.skip = skip;
  o.m16 = b16.mv[0];
  const bool split = J.partitions && cost16 > 1024u;
  if (PR) {
    if (split & ((hpart == 1) | (hpart == 2))) {
      PartRes a, b;
      search_part_12<PR>(cs, n, src, 16, mbx, mby, o.m16, hv, nh, a, b, lds, ar);
      a.total += 2u * rr;
      b.total += 2u * rr;
      take_mode(o, a, 1, (c0.lam * 3u) >> 8);
      take_mode(o, b, 2, (c0.lam * 3u) >> 8);
    } else if (split & (hpart == 3)) {
      PartRes a;
      search_part<3, PR>(cs, n, src, 16, mbx, mby, o.m16, hv, nh, a, lds, ar);
      a.total += 4u * rr;
      take_mode(o, a, 3, (c0.lam * 10u) >> 8);
    }
  } else if (split) {
    PartRes r1, r2, r3;
    search_part_123(cs, n, src, 16, mbx, mby, o.m16, r1, r2, r3, lds, ar);
    r1.total += 2u * rr;
    r2.total += 2u * rr;
    r3.total += 4u * rr;
    take_mode(o, r1, 1, (c0.lam * 3u) >> 8);
    take_mode(o, r2, 2, (c0.lam * 3u) >> 8);
    take_mode(o, r3, 3, (c0.lam * 10u) >> 8);
  }
}

// Packed MV of 4x4 block r after cand_inter_info (slice.rs:1751-1785).
__device__ __forceinline__ uint32_t blk_mv(const MbRes &o, int r) {
  int k = o.part == 0 ? 0 : o.part == 1 ? (r >> 3) : o.part == 2 ? ((r & 3) >> 1) : (((r >> 3) << 1) | ((r & 3) >> 1));
  return sel4(k, o.best.mv[0], o.best.mv[1], o.best.mv[2], o.best.mv[3]);
}

// 32 dwords of pcme_mb_out, lane i < 32 stores dword i.
__device__ __forceinline__ void store_out(pcme_mb_out *out, const MbRes &o) {
  // Branch-free: every value is wave-uniform, so each lane picks its dword
  // with selects (lane-range branches cost more than the whole record).
  const int l = (int)lane_id();
  const int q = l & 3;
  const int pt = o.part;
  const int kb = pt == 0 ? 0 : pt == 1 ? (l >> 3) : pt == 2 ? (q >> 1) : (((l >> 3) << 1) | (q >> 1));
  const uint32_t vm = sel4(l < 16 ? kb : q, o.best.mv[0], o.best.mv[1], o.best.mv[2], o.best.mv[3]);
  const uint32_t vp = sel4(q, o.best.mvp[0], o.best.mvp[1], o.best.mvp[2], o.best.mvp[3]);
  const uint32_t vc = sel4(q, o.cost[0], o.cost[1], o.cost[2], o.cost[3]);
  const uint32_t vx = sel4(q, o.best_cost, o.skip, o.m16, (uint32_t)pt | ((uint32_t)o.ref << 8));
  const uint32_t v = l < 20 ? vm : (l < 24 ? vp : (l < 28 ? vc : vx));
  if (l < 32) ((uint32_t *)out)[l] = v;
}

// ------------------------------------------------------ frame search kernel

// Exchange word: the launch tag in bits 26-31, the MB's reference index in
// 24-25, x in 23-12, y in 11-0 (12-bit two's complement; legal MVs lie
// within +-2000).
#define XTAG(v) ((v) >> 26)
__device__ __forceinline__ int sext12(uint32_t v) { return (int)((v & 0xFFFu) ^ 0x800u) - 0x800; }
__device__ __forceinline__ uint32_t xword(uint32_t tag, uint32_t ref, uint32_t mv) {
  return (tag << 26) | (ref << 24) | (((uint32_t)mv_x(mv) & 0xFFFu) << 12) | ((uint32_t)mv_y(mv) & 0xFFFu);
}
__device__ __forceinline__ uint32_t xword_mv(uint32_t w) { return pack_mv(sext12(w >> 12), sext12(w)); }
__device__ __forceinline__ int xword_ref(uint32_t w) { return (int)((w >> 24) & 3u); }

// Wait until the 4 exchange words of one MB carry `tag`; returns their MVs
// (packed) in w[0..3] and the MB's reference index in *ref (uniform).  False
// on stall or on a raised error flag.
__device__ __forceinline__ bool fetch_words(const uint32_t *xw, uint32_t tag, uint32_t *err,
                                            uint32_t spin_limit, uint32_t w[4], int *ref) {
  int l = (int)lane_id();
  uint32_t spins = 0;
  uint32_t v = 0;
  for (;;) {
    if (l < 4) v = __hip_atomic_load(xw + l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    bool ok = l >= 4 || XTAG(v) == tag;
    if (__ballot(!ok) == 0) break;
    if (++spins > spin_limit) {
      if (l == 0) {
        __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(err + 3, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // sticky: ctl[4]
      }
      return false;
    }
    if ((spins & 63u) == 0) {
      uint32_t e = 0;
      if (l == 0) e = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (rdlane(e, 0) != 0) return false;
    }
    __builtin_amdgcn_s_sleep(2);
  }
  *ref = xword_ref(rdlane(v, 0));
  v = xword_mv(v);
#pragma unroll
  for (int i = 0; i < 4; i++) w[i] = rdlane(v, i);
  return true;
}

static_assert(sizeof(pcme_hint) == 40 && offsetof(pcme_hint, conf_part) == 32, "pcme_hint: 10 dwords, cand[] first");
// The 10 dwords of a pcme_hint, dword i in lane i (lanes >= 10 read 0).
__device__ __forceinline__ uint32_t load_hint(const pcme_hint *h) {
  const int l = (int)lane_id();
  uint32_t v = 0u;
  if (l < 10) v = ((const uint32_t *)h)[l];
  return v;
}

// Timeline build (`make timeline`, -DPCME_TIMELINE; never the product
// library): the frame walkers stamp four moments of every MB of one job
// (g_tl_job) with the device's constant-rate clock (s_memrealtime, 100 MHz):
// [0] the MB's turn begins, [1] its top-right neighbour's words have been
// seen, [2] its analysis is done, [3] its own words are published.  The
// stamps show the per-MB analysis time, the wait for the row above and the
// hand-off latency between rows.
#ifdef PCME_TIMELINE
#define TL_MBS 32768
__device__ unsigned long long g_timeline[TL_MBS * 4];
__device__ int g_tl_job;
#define TL_STAMP(i)                                                                     \
  do {                                                                                  \
    if (tl_on && l == 0) g_timeline[tl_mb * 4 + (i)] = __builtin_amdgcn_s_memrealtime(); \
  } while (0)
#else
#define TL_STAMP(i) \
  do {              \
  } while (0)
#endif

// SRC2: the MB source blocks travel one MB ahead through two LDS buffers;
// without it one buffer is refilled at the end of each MB (the 16-waves-
// per-CU configurations, whose LDS share has room for one).
template <bool HINTS, bool REUSE = true, bool MR = false, bool SRC2 = true>
__device__ __forceinline__ void frames_body(const DevJob *__restrict__ jobs, const int2 *__restrict__ tickets,
                                            int ntickets, uint32_t *ctl, uint32_t spin_limit, uint32_t *s_lds,
                                            uint32_t *s_src) {
  uint32_t *err = ctl + 1;
  const uint32_t tag = ctl[2] & 0x3Fu;  // this launch's epoch (set by prep_kernel)
  const int l = (int)lane_id();
  // Ticket queue q holds tickets [ctl[16+q], ctl[17+q]), taken in order
  // through the counter ctl[8+q].  Workgroups are dealt to the eight XCDs
  // in turn, so a wave starts with its own XCD's queue; an exhausted queue
  // passes the wave to the next, and after all eight it ends.  Any queue's
  // ticket depends only on earlier tickets of the same queue, which running
  // waves took, so taking from another XCD's queue keeps the walk
  // deadlock-free.
  uint32_t q = blockIdx.x & (PCME_QUEUES - 1);
  int passed = 0;
  for (;;) {
    uint32_t t = 0xFFFFFFFFu;
    while (passed < PCME_QUEUES) {
      uint32_t v = 0;
      if (l == 0) v = atomicAdd(ctl + 8 + q, 1u);
      v = rdlane(v, 0);
      const uint32_t lo = ctl[16 + q], hi = ctl[17 + q];
      if (v < hi - lo) {
        t = lo + v;
        break;
      }
      q = (q + 1) & (PCME_QUEUES - 1);
      passed++;
    }
    if (passed >= PCME_QUEUES || t >= (uint32_t)ntickets) return;
    {
      uint32_t e = 0;
      if (l == 0) e = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (rdlane(e, 0) != 0) return;
    }
    const int2 tk = tickets[t];
    const DevJob &J = jobs[tk.x];
    const int my = tk.y;
    const int mbw = J.mbw, mbh = J.mbh, stride = J.stride;
    SCtx c;
    c.pl[0] = J.pl[0]; c.pl[1] = J.pl[1]; c.pl[2] = J.pl[2]; c.pl[3] = J.pl[3];
    c.stride = stride;
    c.pw = 16 * mbw;
    c.ph = 16 * mbh;
    {
      // frame.rs:152 with picture.rs:261-265
      int need = 16 * (my + 1) + 96;
      c.vlim = need >= 16 * mbh ? 16 * mbh + 44 : need;
    }
    c.lam = J.lam;
    c.range = J.me_range;
    c.subpel = J.subpel_rounds;
    uint32_t *xrow = J.xchg + (size_t)my * mbw * 4;
    const uint32_t *xup = J.xchg + (size_t)(my - 1) * mbw * 4;
    Nbr n;
    n.aL = false;
#pragma unroll
    for (int i = 0; i < 4; i++) { n.L[i] = 0u; n.T[i] = 0u; }
    n.TR = n.TL = 0u;
    n.rL = n.rT = n.rTR = n.rTL = n.rf = 0;
    // Row above: n.T = MB x, nxt = MB x+1; n.TL = block 15 of MB x-1.
    // With one reference every index read stays in rdummy and n's indices
    // are the constant 0.
    uint32_t nxt[4] = {0u, 0u, 0u, 0u};
    int rnxt = 0, rdummy = 0;
    bool ok = true;
    if (my > 0) {
      ok = fetch_words(xup, tag, err, spin_limit, n.T, MR ? &n.rT : &rdummy);
      if (ok && mbw > 1) ok = fetch_words(xup + 4, tag, err, spin_limit, nxt, MR ? &rnxt : &rdummy);
    }
    int ar_ref = 0;  // the reference whose regions the areas hold (several references)
    if (!ok) return;
    // The MB's source block travels into LDS one MB ahead of its use (two
    // buffers), so no search of the row waits for a source load; the staged
    // reference regions of the row's searches are kept from MB to MB.
    const uint8_t *srow = J.src + (size_t)(16 * my + PCME_PAD) * stride + PCME_PAD;
    dma_block16(srow, stride, s_src);
    Area ar[2] = {no_area(), no_area()};
    // Hint records travel one MB ahead of their use (a vector load whose
    // round trip hides under the previous MB's search).
    const bool hinted = HINTS && J.hint != nullptr;
    uint32_t hnext = 0u;
    if (hinted) hnext = load_hint(J.hint + (size_t)my * mbw);
    for (int mx = 0; mx < mbw; mx++) {
#ifdef PCME_TIMELINE
      const bool tl_on = tk.x == g_tl_job && my * mbw + mx < TL_MBS;
      const size_t tl_mb = (size_t)(my * mbw + mx);
#endif
      TL_STAMP(0);
      if (my > 0 && mx > 0) {
        n.TL = n.T[3];
#pragma unroll
        for (int i = 0; i < 4; i++) n.T[i] = nxt[i];
        if constexpr (MR) {
          n.rTL = n.rT;
          n.rT = rnxt;
        }
        if (mx + 1 < mbw) {
          if (!fetch_words(xup + (size_t)(mx + 1) * 4, tag, err, spin_limit, nxt, MR ? &rnxt : &rdummy)) return;
        }
      }
      TL_STAMP(1);
      n.aT = my > 0;
      n.aTL = my > 0 && mx > 0;
      n.aTR = my > 0 && mx + 1 < mbw;
      n.TR = nxt[0];
      if constexpr (MR) n.rTR = rnxt;
      dma_wait();  // this MB's source block
      if (SRC2 && mx + 1 < mbw) dma_block16(srow + 16 * (mx + 1), stride, s_src + 64 * ((mx + 1) & 1));
      const uint8_t *src = (const uint8_t *)(s_src + (SRC2 ? 64 * (mx & 1) : 0));
      // Pruner hint of this MB; a full search without one.
      uint32_t hv = 0u;
      int nh = -1, hpart = 0;
      if (hinted) {
        hv = hnext;
        if (mx + 1 < mbw) hnext = load_hint(J.hint + (size_t)my * mbw + mx + 1);
        const uint32_t w9 = rdlane(hv, 9);  // part | ncand << 8 | prune << 16
        if ((w9 >> 16) & 0xFFu) {
          nh = min((int)((w9 >> 8) & 0xFFu), 8);
          hpart = (int)(w9 & 3u);
        }
      }
      MbRes o;
      if (!REUSE) {  // stage every search's regions afresh
        ar[0] = no_area();
        ar[1] = no_area();
      }
      if constexpr (MR) {
        if (HINTS && nh >= 0)
          analyse_mb_mr<true>(J, c, n, src, 16 * mx, 16 * my, hv, nh, hpart, o, s_lds, ar, ar_ref);
        else
          analyse_mb_mr<false>(J, c, n, src, 16 * mx, 16 * my, hv, nh, hpart, o, s_lds, ar, ar_ref);
      } else {
        if (HINTS && nh >= 0)
          analyse_mb<true>(c, n, src, 16, 16 * mx, 16 * my, J.partitions, hv, nh, hpart, o, s_lds, ar);
        else
          analyse_mb<false>(c, n, src, 16, 16 * mx, 16 * my, J.partitions, hv, nh, hpart, o, s_lds, ar);
      }
      TL_STAMP(2);
      if (!SRC2 && mx + 1 < mbw) {
        lds_fence();  // every read of this MB's source block has returned
        dma_block16(srow + 16 * (mx + 1), stride, s_src);
      }
      store_out(J.out + (size_t)my * mbw + mx, o);
      // Publish the bottom row (blocks 12..15) for the row below.
      if (l < 4) {
        uint32_t wv = xword(tag, MR ? (uint32_t)o.ref : 0u, blk_mv(o, 12 + l));
        __hip_atomic_store(xrow + (size_t)mx * 4 + l, wv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      TL_STAMP(3);
      // Left column (blocks 3, 7, 11, 15) for the next MB of this row.
      n.aL = true;
#pragma unroll
      for (int i = 0; i < 4; i++) n.L[i] = blk_mv(o, 4 * i + 3);
      if constexpr (MR) n.rL = o.ref;
    }
  }
}

// The full-search kernel carries no hint code (128 VGPRs); the hinted
// kernel runs jobs with and without hints (its two MB paths together need
// 168 VGPRs, 2 waves per SIMD; forcing 3 spills and runs slower).  Each
// workgroup (one wave) holds FRAME_LDS_WORDS of LDS: the two search areas,
// the two MB source blocks and padding that keeps a CU at 12 waves, 3 per
// SIMD.  With 4 per SIMD (128 VGPRs allow it) the dispatcher packed some
// SIMDs fuller than others and the persistent walkers ran 5% slower.
#define FRAME_LDS_WORDS (13312 / 4)
static_assert(FRAME_LDS_WORDS * 4 >= LDS_PER_WAVE + 512 && 163840 / (FRAME_LDS_WORDS * 4) == 12,
              "frame kernels: room for the areas and source blocks, 12 waves per CU");
__global__ void __launch_bounds__(64) search_frames_kernel(const DevJob *__restrict__ jobs,
                                                           const int2 *__restrict__ tickets, int ntickets,
                                                           uint32_t *ctl, uint32_t spin_limit) {
  __shared__ __attribute__((aligned(16))) uint32_t s_lds[FRAME_LDS_WORDS];
  frames_body<false>(jobs, tickets, ntickets, ctl, spin_limit, s_lds, s_lds + LDS_PER_WAVE / 4);
}

__global__ void __launch_bounds__(64)
search_frames_hinted_kernel(const DevJob *__restrict__ jobs, const int2 *__restrict__ tickets, int ntickets,
                            uint32_t *ctl, uint32_t spin_limit) {
  __shared__ __attribute__((aligned(16))) uint32_t s_lds[FRAME_LDS_WORDS];
  frames_body<true>(jobs, tickets, ntickets, ctl, spin_limit, s_lds, s_lds + LDS_PER_WAVE / 4);
}

// The multi-reference field (jobs with nref > 1; the same walk, each MB
// searched on every reference): full search, and with pruner hints.
__global__ void __launch_bounds__(64) search_frames_mr_kernel(const DevJob *__restrict__ jobs,
                                                              const int2 *__restrict__ tickets, int ntickets,
                                                              uint32_t *ctl, uint32_t spin_limit) {
  __shared__ __attribute__((aligned(16))) uint32_t s_lds[FRAME_LDS_WORDS];
  frames_body<false, true, true>(jobs, tickets, ntickets, ctl, spin_limit, s_lds, s_lds + LDS_PER_WAVE / 4);
}

__global__ void __launch_bounds__(64) search_frames_mr_hinted_kernel(const DevJob *__restrict__ jobs,
                                                                     const int2 *__restrict__ tickets, int ntickets,
                                                                     uint32_t *ctl, uint32_t spin_limit) {
  __shared__ __attribute__((aligned(16))) uint32_t s_lds[FRAME_LDS_WORDS];
  frames_body<true, true, true>(jobs, tickets, ntickets, ctl, spin_limit, s_lds, s_lds + LDS_PER_WAVE / 4);
}

// The multi-reference field at 16 waves per CU: 4 walkers per workgroup,
// each with the two search areas and one source block (the 4K workload is
// bound by throughput, where the fourth wave per SIMD pays).
#define MR16_WORDS (163840 / 16 / 4)
static_assert(MR16_WORDS * 4 >= LDS_PER_WAVE + 256, "multi-reference walkers at 16 per CU");
__global__ void __launch_bounds__(256, 4) search_frames_mr16_kernel(const DevJob *__restrict__ jobs,
                                                                   const int2 *__restrict__ tickets, int ntickets,
                                                                   uint32_t *ctl, uint32_t spin_limit) {
  __shared__ __attribute__((aligned(16))) uint32_t s_lds[4 * MR16_WORDS];
  uint32_t *mine = s_lds + (threadIdx.x >> 6) * MR16_WORDS;
  frames_body<false, true, true, false>(jobs, tickets, ntickets, ctl, spin_limit, mine, mine + LDS_PER_WAVE / 4);
}

// The hinted field (pruner hints, one reference) at 16 waves per CU, for
// workloads bound by throughput: its two MB paths need 168 VGPRs, so at
// 128 some spill to scratch (96 bytes per lane).
__global__ void __launch_bounds__(256, 4) search_frames_hinted16_kernel(const DevJob *__restrict__ jobs,
                                                                       const int2 *__restrict__ tickets,
                                                                       int ntickets, uint32_t *ctl,
                                                                       uint32_t spin_limit) {
  __shared__ __attribute__((aligned(16))) uint32_t s_lds[4 * MR16_WORDS];
  uint32_t *mine = s_lds + (threadIdx.x >> 6) * MR16_WORDS;
  frames_body<true, true, false, false>(jobs, tickets, ntickets, ctl, spin_limit, mine, mine + LDS_PER_WAVE / 4);
}

// The multi-reference field with pruner hints at 16 waves per CU (166
// VGPRs at the default occupancy; some spill at 128).
__global__ void __launch_bounds__(256, 4) search_frames_mr_hinted16_kernel(const DevJob *__restrict__ jobs,
                                                                          const int2 *__restrict__ tickets,
                                                                          int ntickets, uint32_t *ctl,
                                                                          uint32_t spin_limit) {
  __shared__ __attribute__((aligned(16))) uint32_t s_lds[4 * MR16_WORDS];
  uint32_t *mine = s_lds + (threadIdx.x >> 6) * MR16_WORDS;
  frames_body<true, true, true, false>(jobs, tickets, ntickets, ctl, spin_limit, mine, mine + LDS_PER_WAVE / 4);
}

// Configurations of the full-search kernel for the autotuner: WPG row
// walkers (waves) per workgroup, each reserving the LDS that caps a CU at
// WCU waves, and with or without the staged regions kept along a row.
// search_frames_kernel is <1, 12, true>.
// With 16 waves per CU (4 per SIMD, 128 VGPRs) a wave's LDS share holds the
// two search areas and one source block.
template <int WPG, int WCU, bool REUSE>
__global__ void __launch_bounds__(64 * WPG, WCU > 12 ? 4 : 1) search_frames_var_kernel(const DevJob *__restrict__ jobs,
                                                                     const int2 *__restrict__ tickets, int ntickets,
                                                                     uint32_t *ctl, uint32_t spin_limit) {
  constexpr bool SRC2 = WCU <= 12;
  constexpr int kWords = (163840 / WCU / 4) & ~3;
  static_assert(kWords * 4 >= LDS_PER_WAVE + (SRC2 ? 512 : 256) && 163840 / (kWords * 4) == WCU,
                "LDS per wave for WCU waves per CU");
  __shared__ __attribute__((aligned(16))) uint32_t s_lds[WPG * kWords];
  uint32_t *mine = s_lds + (threadIdx.x >> 6) * kWords;
  frames_body<false, REUSE, false, SRC2>(jobs, tickets, ntickets, ctl, spin_limit, mine, mine + LDS_PER_WAVE / 4);
}
template __global__ void search_frames_var_kernel<2, 12, true>(const DevJob *__restrict__, const int2 *__restrict__,
                                                                int, uint32_t *, uint32_t);
template __global__ void search_frames_var_kernel<4, 12, true>(const DevJob *__restrict__, const int2 *__restrict__,
                                                                int, uint32_t *, uint32_t);
template __global__ void search_frames_var_kernel<1, 8, true>(const DevJob *__restrict__, const int2 *__restrict__,
                                                               int, uint32_t *, uint32_t);
template __global__ void search_frames_var_kernel<2, 8, true>(const DevJob *__restrict__, const int2 *__restrict__,
                                                               int, uint32_t *, uint32_t);
template __global__ void search_frames_var_kernel<1, 12, false>(const DevJob *__restrict__, const int2 *__restrict__,
                                                                 int, uint32_t *, uint32_t);
template __global__ void search_frames_var_kernel<4, 16, false>(const DevJob *__restrict__, const int2 *__restrict__,
                                                                 int, uint32_t *, uint32_t);
template __global__ void search_frames_var_kernel<4, 16, true>(const DevJob *__restrict__, const int2 *__restrict__,
                                                                int, uint32_t *, uint32_t);

// -------------------------------------------------- per-call search_part

static_assert(sizeof(pcme_nb_block) == 8 && sizeof(pcme_part_job) == 168 && sizeof(pcme_part_out) == 36,
              "pcme_part_job / pcme_part_out layout");

// nb_part gives MV zero with a negative reference (slice.rs:429-430).
__device__ __forceinline__ uint32_t nb_mv(const pcme_nb_block &b) { return b.ref >= 0 ? pack_mv(b.mv.x, b.mv.y) : 0u; }

// Neighbour blocks in the order left[4], top[4], top_right, top_left.
__device__ __forceinline__ NbrG nbr_of(const pcme_nb_block *b) {
  NbrG n;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    n.rL[i] = b[i].ref; n.L[i] = nb_mv(b[i]);
    n.rT[i] = b[4 + i].ref; n.T[i] = nb_mv(b[4 + i]);
  }
  n.rTR = b[8].ref; n.TR = nb_mv(b[8]);
  n.rTL = b[9].ref; n.TL = nb_mv(b[9]);
  return n;
}

__global__ void __launch_bounds__(64) search_part_batch_kernel(const pcme_part_job *__restrict__ jobs,
                                                               int njobs, pcme_part_out *out) {
  __shared__ __attribute__((aligned(16))) uint32_t s_lds[LDS_PER_WAVE / 4];
  const int j = blockIdx.x;
  if (j >= njobs) return;
  const pcme_part_job &J = jobs[j];
  SCtx c;
  c.pl[0] = J.ref; c.pl[1] = J.ref_h; c.pl[2] = J.ref_v; c.pl[3] = J.ref_c;
  c.stride = 16 * J.mbw + 2 * PCME_PAD;
  c.pw = 16 * J.mbw;
  c.ph = 16 * J.mbh;
  c.vlim = (int)J.vlim;
  c.lam = J.lam;
  c.range = J.me_range;
  c.subpel = J.subpel_rounds;
  static_assert(offsetof(pcme_part_job, top) == offsetof(pcme_part_job, left) + 32 &&
                offsetof(pcme_part_job, top_right) == offsetof(pcme_part_job, top) + 32 &&
                offsetof(pcme_part_job, top_left) == offsetof(pcme_part_job, top_right) + 8,
                "pcme_part_job: neighbour blocks contiguous");
  const NbrG n = nbr_of(J.left);
  const uint32_t seed = pack_mv(J.seed.x, J.seed.y);
  const int mbx = 16 * J.mx, mby = 16 * J.my;
  PartRes r;
  Area ar[2] = {no_area(), no_area()};
  switch (J.part) {
    case 0: search_part<0, false>(c, n, J.src, J.sstride, mbx, mby, seed, 0u, -1, r, s_lds, ar); break;
    case 1: search_part<1, false>(c, n, J.src, J.sstride, mbx, mby, seed, 0u, -1, r, s_lds, ar); break;
    case 2: search_part<2, false>(c, n, J.src, J.sstride, mbx, mby, seed, 0u, -1, r, s_lds, ar); break;
    default: search_part<3, false>(c, n, J.src, J.sstride, mbx, mby, seed, 0u, -1, r, s_lds, ar); break;
  }
  const int lane = (int)lane_id();
  if (lane < 9) {
    uint32_t v = r.total;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      v = lane == i ? r.mv[i] : v;
      v = lane == 4 + i ? r.mvp[i] : v;
    }
    ((uint32_t *)&out[j])[lane] = v;
  }
}

// ---------------------------------------------------------- B macroblocks

static_assert(sizeof(pcme_b_job) == 328 && sizeof(pcme_b_out) == 136, "pcme_b_job / pcme_b_out layout");

__device__ __forceinline__ int minp(int x, int y) { return (x >= 0 && y >= 0) ? min(x, y) : max(x, y); }

// mvp (slice.rs:434-482) of a 16x16 partition (Normal shape) for reference
// index rf, from explicit neighbours.
__device__ __forceinline__ uint32_t mvp16_rf(const NbrG &n, int rf) {
  bool aa = n.rL[0] >= -1, ab = n.rT[0] >= -1, ac = n.rTR >= -1;
  int ra = n.rL[0], rb = n.rT[0], rc = n.rTR;
  uint32_t ma = n.L[0], mb = n.T[0], mc = n.TR;
  if (!ac) {
    ac = n.rTL >= -1; rc = n.rTL; mc = n.TL;
  }
  if (!ab && !ac && aa) {
    ab = ac = true; rb = rc = ra; mb = mc = ma;
  }
  if (!aa) { ra = -1; ma = 0u; }
  if (!ab) { rb = -1; mb = 0u; }
  if (!ac) { rc = -1; mc = 0u; }
  const int cnt = (ra == rf) + (rb == rf) + (rc == rf);
  if (cnt == 1) return ra == rf ? ma : (rb == rf ? mb : mc);
  return pack_mv(median3(mv_x(ma), mv_x(mb), mv_x(mc)), median3(mv_y(ma), mv_y(mb), mv_y(mc)));
}

// One quarter-pel luma sample (mc_luma, picture.rs:415-449) of picture
// pixel (x, y) displaced by the packed MV m.
__device__ __forceinline__ uint32_t mc_px(const SCtx &c, int x, int y, uint32_t m) {
  const int mxv = mv_x(m), myv = mv_y(m);
  const int xi = x + (mxv >> 2) + PCME_PAD, yi = y + (myv >> 2) + PCME_PAD;
  const uint32_t e = qpel_entry((uint32_t)((myv & 3) * 4 + (mxv & 3)));
  const uint32_t a = plane_sel(c, e & 3u)[(size_t)(yi + ((e >> 3) & 1)) * c.stride + xi + ((e >> 2) & 1)];
  const uint32_t b = plane_sel(c, (e >> 4) & 3u)[(size_t)(yi + ((e >> 7) & 1)) * c.stride + xi + ((e >> 6) & 1)];
  return (a + b + 1u) >> 1;
}

// SATD of the W x H rect at (x, y) of the source MB against a 16x16
// prediction in LDS (row stride 16): one 4x8 strip per lane, wave-uniform.
template <int W, int H>
__device__ __forceinline__ uint32_t satd_rect(const uint8_t *src, int sstride, const uint8_t *pred, int x, int y) {
  constexpr int NS = (W / 8) * (H / 4);
  const int lane = (int)lane_id();
  const int s = lane < NS ? lane : 0;
  const int sx = x + 8 * (s % (W / 8)), sy = y + 4 * (s / (W / 8));
  uint32_t S[8], P[8];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const uint8_t *q = src + (size_t)(sy + r) * sstride + sx;
    S[2 * r] = *(const uint32_t *)q;
    S[2 * r + 1] = *(const uint32_t *)(q + 4);
    P[2 * r] = *(const uint32_t *)(pred + (sy + r) * 16 + sx);
    P[2 * r + 1] = *(const uint32_t *)(pred + (sy + r) * 16 + sx + 4);
  }
  v2s sp[16];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int cc = 0; cc < 4; cc++) sp[r * 4 + cc] = as_v2s(__builtin_amdgcn_perm(S[2 * r + 1], S[2 * r], SELC(cc)));
  const uint32_t v = lane < NS ? satd_strip8(sp, P) : 0u;
  return rdlane(red8(v), 0);
}

// Prediction of the rect (x, y, w, h) of the MB at picture (px, py) with MV
// m, into the 16x16 LDS buffer; lanes cover the rect's pixels.
__device__ __forceinline__ void mc_rect(const SCtx &c, int px, int py, int x, int y, int w, int h, uint32_t m,
                                        uint8_t *buf) {
  for (int i = (int)lane_id(); i < w * h; i += 64) {
    const int xx = x + i % w, yy = y + i / w;
    buf[yy * 16 + xx] = (uint8_t)mc_px(c, px + xx, py + yy, m);
  }
}

__device__ __forceinline__ void avg_buf(uint8_t *d, const uint8_t *a, const uint8_t *b) {
  for (int i = (int)lane_id(); i < 256; i += 64) d[i] = (uint8_t)(((uint32_t)a[i] + b[i] + 1u) >> 1);
}

// Scalar context of one B job.
__device__ __forceinline__ void b_ctx(const pcme_b_job &J, SCtx *c) {
#pragma unroll
  for (int l = 0; l < 2; l++) {
#pragma unroll
    for (int p = 0; p < 4; p++) c[l].pl[p] = l ? J.ref1[p] : J.ref0[p];
    c[l].stride = 16 * J.mbw + 2 * PCME_PAD;
    c[l].pw = 16 * J.mbw;
    c[l].ph = 16 * J.mbh;
    c[l].vlim = (int)J.vlim;
    c[l].lam = J.lam;
    c[l].range = J.me_range;
    c[l].subpel = J.subpel_rounds;
  }
}

// Orders this wave's LDS writes before its later LDS reads by other lanes
// (the B analysis runs one MB per wave).
__device__ __forceinline__ void bwave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// analyse_b's motion search (slice.rs:2189-2377 without BSkip, RD and
// intra), first half: direct_spatial and the direct prediction's cost, the
// list-0/list-1 16x16 search_part and the direction costs, written as
// dwords 0-21 of the record.  In pieces, so that the two lists' searches
// can run on two waves (b_walk_kernel) or one after the other
// (b_analyse_16).

// direct_spatial's reference indices and MVs per 8x8 and whether every
// direct MV is legal (slice.rs:2030-2084).  Wave-uniform.
struct BDirect {
  int refs[2];
  uint32_t dmv[2][4];
  bool ok;
};

__device__ __forceinline__ BDirect b_direct(const pcme_b_job &J, const SCtx *c, const NbrG *n) {
  const int px = 16 * J.mx, py = 16 * J.my;
  BDirect D;
  int *refs = D.refs;
  uint32_t dmvp[2] = {0u, 0u};
#pragma unroll
  for (int l = 0; l < 2; l++) {
    const int a = n[l].rL[0] >= -1 ? n[l].rL[0] : -1;
    const int b = n[l].rT[0] >= -1 ? n[l].rT[0] : -1;
    const int cc = n[l].rTR >= -1 ? n[l].rTR : (n[l].rTL >= -1 ? n[l].rTL : -1);
    refs[l] = minp(a, minp(b, cc));
  }
  const bool zero = refs[0] < 0 && refs[1] < 0;
  if (zero) {
    refs[0] = refs[1] = 0;
  } else {
#pragma unroll
    for (int l = 0; l < 2; l++)
      if (refs[l] >= 0) dmvp[l] = mvp16_rf(n[l], refs[l]);
  }
  uint32_t (*dmv)[4] = D.dmv;
  bool direct_ok = true;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    bool col_zero = false;
    if (!J.col_intra) {
      const int l = J.col_ref[0][q] >= 0 ? 0 : 1;
      const pcme_mv m = J.col_mv[l][q];
      col_zero = J.col_ref[l][q] == 0 && m.x >= -1 && m.x <= 1 && m.y >= -1 && m.y <= 1;
    }
#pragma unroll
    for (int l = 0; l < 2; l++) {
      dmv[l][q] = (refs[l] >= 0 && !(zero || (refs[l] == 0 && col_zero))) ? dmvp[l] : 0u;
      if (refs[l] >= 0) {
        const Box bx = legal_box(c[l], px + 8 * (q & 1), py + 8 * (q >> 1), 8, 8);
        direct_ok = direct_ok && in_box(bx, mv_x(dmv[l][q]), mv_y(dmv[l][q]));
      }
    }
  }
  D.ok = direct_ok;
  return D;
}

// The direct prediction's cost (SATD + (2 lam) >> 8), all ones when a
// direct MV is illegal.  One wave; s_pred[2] holds the prediction after.
__device__ __forceinline__ uint32_t b_direct_cost(const pcme_b_job &J, const SCtx *c, const BDirect &D,
                                                  uint8_t (*s_pred)[256]) {
  const int lane = (int)lane_id();
  const int px = 16 * J.mx, py = 16 * J.my;
  const uint32_t lam = J.lam;
  const int *refs = D.refs;
  const uint32_t (*dmv)[4] = D.dmv;
  const bool direct_ok = D.ok;
  uint32_t direct_cost = 0xFFFFFFFFu;
  if (direct_ok) {
    for (int i = lane; i < 256; i += 64) {
      const int x = i & 15, y = i >> 4, q = ((y >> 3) << 1) | (x >> 3);
      uint32_t m0 = dmv[0][0], m1 = dmv[1][0];
#pragma unroll
      for (int k = 1; k < 4; k++) {
        m0 = q == k ? dmv[0][k] : m0;
        m1 = q == k ? dmv[1][k] : m1;
      }
      uint32_t v;
      if (refs[0] >= 0 && refs[1] >= 0)
        v = (mc_px(c[0], px + x, py + y, m0) + mc_px(c[1], px + x, py + y, m1) + 1u) >> 1;
      else if (refs[0] >= 0)
        v = mc_px(c[0], px + x, py + y, m0);
      else
        v = mc_px(c[1], px + x, py + y, m1);
      s_pred[2][i] = (uint8_t)v;
    }
    bwave_sync();
    direct_cost = satd_rect<16, 16>(J.src, J.sstride, s_pred[2], 0, 0) + ((lam * 2u) >> 8);
    bwave_sync();
  }
  return direct_cost;
}

// One list's 16x16 search_part, seeded with its direct MV
// (slice.rs:2232-2244); each list's search starts with nothing staged (the
// two lists read different pictures).
__device__ __forceinline__ void b_search16(const pcme_b_job &J, const SCtx &c, const NbrG &n, uint32_t seed,
                                           uint32_t *s_lds, PartRes &r) {
  Area ar[2] = {no_area(), no_area()};
  search_part<0, false>(c, n, J.src, J.sstride, 16 * J.mx, 16 * J.my, seed, 0u, -1, r, s_lds, ar);
}

// The bi prediction's cost, the direction and dwords 0-21 of the record
// from both lists' 16x16 results (mv[0], mvp[0], total); mv16 = the two
// lists' MVs.  One wave.
__device__ __forceinline__ void b_finish16(const pcme_b_job &J, const SCtx *c, const BDirect &D, uint32_t direct_cost,
                                           const PartRes *r16, uint8_t (*s_pred)[256], pcme_b_out *rec,
                                           uint32_t mv16[2]) {
  const int lane = (int)lane_id();
  const int px = 16 * J.mx, py = 16 * J.my;
  const uint32_t lam = J.lam;
  const int *refs = D.refs;
  const uint32_t (*dmv)[4] = D.dmv;
  const bool direct_ok = D.ok;
  bwave_sync();
  mc_rect(c[0], px, py, 0, 0, 16, 16, r16[0].mv[0], s_pred[0]);
  mc_rect(c[1], px, py, 0, 0, 16, 16, r16[1].mv[0], s_pred[1]);
  bwave_sync();
  avg_buf(s_pred[2], s_pred[0], s_pred[1]);
  bwave_sync();
  uint32_t bits = 0;
#pragma unroll
  for (int l = 0; l < 2; l++)
    bits += se_bits(mv_x(r16[l].mv[0]) - mv_x(r16[l].mvp[0])) + se_bits(mv_y(r16[l].mv[0]) - mv_y(r16[l].mvp[0]));
  uint32_t cost[3];
  cost[0] = r16[0].total + ((lam * 3u) >> 8);
  cost[1] = r16[1].total + ((lam * 3u) >> 8);
  cost[2] = satd_rect<16, 16>(J.src, J.sstride, s_pred[2], 0, 0) + ((lam * (bits + 5u)) >> 8);
  const uint32_t dir = cost[1] < cost[0] ? (cost[2] < cost[1] ? 2u : 1u) : (cost[2] < cost[0] ? 2u : 0u);
  // dwords 0-21 of the record, one per lane
  uint32_t w = 0;
  if (lane < 8) {
    const int l = lane >> 2, q = lane & 3;
    uint32_t v = dmv[0][0];
#pragma unroll
    for (int ll = 0; ll < 2; ll++)
#pragma unroll
      for (int qq = 0; qq < 4; qq++) v = (l == ll && q == qq) ? dmv[ll][qq] : v;
    w = v;
  } else if (lane == 8 || lane == 9) {
    w = (uint32_t)(uint8_t)(int8_t)(lane == 8 ? refs[0] : refs[1]) * 0x01010101u;
  } else if (lane == 10) {
    w = direct_ok ? 1u : 0u;
  } else if (lane == 11) {
    w = direct_cost;
  } else if (lane == 12 || lane == 13) {
    w = lane == 12 ? r16[0].mv[0] : r16[1].mv[0];
  } else if (lane == 14 || lane == 15) {
    w = lane == 14 ? r16[0].mvp[0] : r16[1].mvp[0];
  } else if (lane == 16 || lane == 17) {
    w = lane == 16 ? r16[0].total : r16[1].total;
  } else if (lane >= 18 && lane <= 20) {
    w = lane == 18 ? cost[0] : (lane == 19 ? cost[1] : cost[2]);
  } else if (lane == 21) {
    w = dir;
  }
  if (lane < 22 || lane == 33) ((uint32_t *)rec)[lane] = w;  // dword 33: reserved, zero
  mv16[0] = r16[0].mv[0];
  mv16[1] = r16[1].mv[0];
}

// The first half on one wave.
__device__ __forceinline__ void b_analyse_16(const pcme_b_job &J, pcme_b_out *rec, uint32_t *s_lds,
                                             uint8_t (*s_pred)[256], uint32_t mv16[2]) {
  SCtx c[2];
  b_ctx(J, c);
  const NbrG n[2] = {nbr_of(J.nb[0]), nbr_of(J.nb[1])};
  const BDirect D = b_direct(J, c, n);
  const uint32_t direct_cost = b_direct_cost(J, c, D, s_pred);
  PartRes r16[2];
  b_search16(J, c[0], n[0], D.dmv[0][0], s_lds, r16[0]);
  b_search16(J, c[1], n[1], D.dmv[1][0], s_lds, r16[1]);
  b_finish16(J, c, D, direct_cost, r16, s_pred, rec, mv16);
}

__global__ void __launch_bounds__(64) b_analyse_16_kernel(const pcme_b_job *__restrict__ jobs, int njobs,
                                                          pcme_b_out *out) {
  __shared__ __attribute__((aligned(16))) uint32_t s_lds[LDS_PER_WAVE / 4];
  __shared__ __attribute__((aligned(16))) uint8_t s_pred[3][256];
  const int j = blockIdx.x;
  if (j >= njobs) return;
  uint32_t mv16[2];
  b_analyse_16(jobs[j], &out[j], s_lds, s_pred, mv16);
}

// Second half: b_partition (slice.rs:2086-2187) of 16x8 (K = 0) or 8x16
// (K = 1), seeded with the two lists' 16x16 MVs mv16 (packed; dwords 12-13
// of the record).  Writes part_mv[K], part_cost[K] and part_dir[K] of *rec.
// One wave, LDS as b_analyse_16.
template <int K>
__device__ __forceinline__ void b_analyse_part(const pcme_b_job &J, pcme_b_out *rec, uint32_t *s_lds,
                                               uint8_t (*s_pred)[256], const uint32_t mv16[2]) {
  const int lane = (int)lane_id();
  SCtx c[2];
  b_ctx(J, c);
  const NbrG n[2] = {nbr_of(J.nb[0]), nbr_of(J.nb[1])};
  const int px = 16 * J.mx, py = 16 * J.my;
  const uint32_t lam = J.lam;
  uint32_t pmv[2][2];
  PartRes rp[2];
  // both lists' searches in two paired steps (partition 0, then 1)
  Area ar[2] = {no_area(), no_area()};  // list 0 in area 0, list 1 in area 1
  search_part_lists<K + 1>(c, n, J.src, J.sstride, px, py, mv16, rp[0], rp[1], s_lds, ar);
#pragma unroll
  for (int l = 0; l < 2; l++) {
    pmv[l][0] = rp[l].mv[0];
    pmv[l][1] = rp[l].mv[1];
  }
  bwave_sync();
#pragma unroll
  for (int l = 0; l < 2; l++)
#pragma unroll
    for (int pi = 0; pi < 2; pi++) {
      if (K == 0)
        mc_rect(c[l], px, py, 0, 8 * pi, 16, 8, rp[l].mv[pi], s_pred[l]);
      else
        mc_rect(c[l], px, py, 8 * pi, 0, 8, 16, rp[l].mv[pi], s_pred[l]);
    }
  bwave_sync();
  avg_buf(s_pred[2], s_pred[0], s_pred[1]);
  bwave_sync();
  uint32_t tot = (lam * 8u) >> 8, pdir[2];
#pragma unroll
  for (int pi = 0; pi < 2; pi++) {
    uint32_t best = 0xFFFFFFFFu, bd = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const uint32_t sd = K == 0 ? satd_rect<16, 8>(J.src, J.sstride, s_pred[d], 0, 8 * pi)
                                 : satd_rect<8, 16>(J.src, J.sstride, s_pred[d], 8 * pi, 0);
      const uint32_t cd = sd + ((lam * (d == 2 ? 14u : 7u)) >> 8);
      if (cd < best) {
        best = cd;
        bd = (uint32_t)d;
      }
    }
    tot += best;
    pdir[pi] = bd;
  }
  uint32_t *o = (uint32_t *)rec;
  // part_mv[K][l][pi] = dword 22 + 4K + 2l + pi
  const uint32_t v = lane == 0 ? pmv[0][0] : (lane == 1 ? pmv[0][1] : (lane == 2 ? pmv[1][0] : pmv[1][1]));
  if (lane < 4) o[22 + 4 * K + lane] = v;
  if (lane == 4) o[30 + K] = tot;
  if (lane == 5 || lane == 6) ((uint8_t *)o)[128 + 2 * K + (lane - 5)] = (uint8_t)pdir[lane - 5];
}

// The second half over independent MBs whose first half an earlier launch
// wrote (the 16x16 MVs read from dwords 12-13 of the record).
template <int K>
__global__ void __launch_bounds__(64) b_analyse_part_kernel(const pcme_b_job *__restrict__ jobs, int njobs,
                                                            pcme_b_out *out) {
  __shared__ __attribute__((aligned(16))) uint32_t s_lds[LDS_PER_WAVE / 4];
  __shared__ __attribute__((aligned(16))) uint8_t s_pred[3][256];
  const int j = blockIdx.x;
  if (j >= njobs) return;
  const uint32_t *rec = (const uint32_t *)&out[j];
  const uint32_t mv16[2] = {rdlane(rec[12], 0), rdlane(rec[13], 0)};
  b_analyse_part<K>(jobs[j], &out[j], s_lds, s_pred, mv16);
}

template __global__ void b_analyse_part_kernel<0>(const pcme_b_job *__restrict__, int, pcme_b_out *);
template __global__ void b_analyse_part_kernel<1>(const pcme_b_job *__restrict__, int, pcme_b_out *);

// ------------------------------------------------------ the B picture field

static_assert(sizeof(pcme_b_col) == 44 && sizeof(pcme_b_mb) == 280, "pcme_b_col / pcme_b_mb layout");

// The 8x8 (raster) holding 4x4 block r.
__device__ __forceinline__ int blk8(int r) { return ((r >> 3) << 1) | ((r & 3) >> 1); }

// Before an MB's search: its neighbour blocks (both lists: left blocks 3,
// 7, 11, 15, top 12..15, top-right 12, top-left 15; -2 outside the picture)
// read from the records the earlier diagonals committed, and its co-located
// MB, into its job.  One thread.
__device__ __forceinline__ void b_gather(pcme_b_job &J, const pcme_b_mb *__restrict__ field,
                                         const pcme_b_col *__restrict__ col) {
  const int mx = J.mx, my = J.my, mbw = J.mbw;
  // (MB offset, 4x4 block) of the ten neighbour blocks
  const int ndx[10] = {-1, -1, -1, -1, 0, 0, 0, 0, 1, -1};
  const int ndy[10] = {0, 0, 0, 0, -1, -1, -1, -1, -1, -1};
  const int nbk[10] = {3, 7, 11, 15, 12, 13, 14, 15, 12, 15};
  for (int i = 0; i < 10; i++) {
    const int x = mx + ndx[i], y = my + ndy[i];
    const bool av = x >= 0 && y >= 0 && x < mbw;
    for (int l = 0; l < 2; l++) {
      pcme_nb_block b;
      b.reserved[0] = b.reserved[1] = b.reserved[2] = 0;
      if (av) {
        const pcme_b_mb &m = field[(size_t)y * mbw + x];
        b.mv = m.mv[l][nbk[i]];
        b.ref = m.ref[l][blk8(nbk[i])];
      } else {
        b.mv.x = b.mv.y = 0;
        b.ref = -2;
      }
      J.nb[l][i] = b;
    }
  }
  const pcme_b_col &c = col[(size_t)my * mbw + mx];
  for (int l = 0; l < 2; l++)
    for (int q = 0; q < 4; q++) {
      J.col_mv[l][q] = c.mv[l][q];
      J.col_ref[l][q] = c.ref[l][q];
    }
  J.col_intra = c.intra;
}

// After an MB's search: analyse_b's decision with rd off and without BSkip
// and intra (slice.rs:2189-2377) and the commit of cand_inter_info
// (slice.rs:1751-1785) into the MB's record.  One thread.
__device__ __forceinline__ void b_commit(const pcme_b_out &o, pcme_b_mb &m, int partitions) {
  uint32_t best = 0u;
  int mode = -1;
  if (o.direct_ok) {
    best = o.direct_cost;
    mode = 0;
  }
  if (mode < 0 || o.cost[o.dir] < best) {
    best = o.cost[o.dir];
    mode = 1;
  }
  if (partitions && best > 1024u)
    for (int p = 0; p < 2; p++)
      if (o.part_cost[p] < best) {
        best = o.part_cost[p];
        mode = 2 + p;
      }
  m.search = o;
  m.best_cost = best;
  m.mode = (uint8_t)mode;
  m.dir = (uint8_t)o.dir;
  m.reserved[0] = m.reserved[1] = 0;
  for (int l = 0; l < 2; l++)
    for (int r = 0; r < 16; r++) {
      const int q = blk8(r);
      int rf;
      pcme_mv mv;
      if (mode == 0) {
        rf = o.direct_ref[l][q];
        mv = o.direct_mv[l][q];
      } else if (mode == 1) {
        rf = (o.dir == 2u || o.dir == (uint32_t)l) ? 0 : -1;
        mv = o.mv16[l];
      } else {
        const int p = mode - 2;
        const int pi = p == 0 ? (r >> 3) : ((r & 3) >> 1);
        const int d = o.part_dir[p][pi];
        rf = (d == 2 || d == l) ? 0 : -1;
        mv = o.part_mv[p][l][pi];
      }
      if (rf < 0) mv.x = mv.y = 0;
      m.mv[l][r] = mv;
      m.ref[l][q] = (int8_t)rf;
    }
}

// B fields in one launch: persistent row walkers (a workgroup of two waves
// walks one MB row of one picture left to right; rows are taken through a
// ticket counter from a table of (picture, row) tickets in (row, picture)
// order, so a row's dependency, the row above of its picture, always holds
// an earlier ticket, taken by a running walker).  MB (x, y) waits for the row above's MB x+1 (x for the
// last column), publishes its bottom blocks 12-15 of both lists as eight
// exchange words (the P walker's format: tag, list used, MV) for the row
// below, and keeps its right column (blocks 3, 7, 11, 15) in LDS for the
// next MB of its row.  Per MB: its job is copied into LDS and its
// neighbour blocks filled from those two sources and its co-located MB;
// wave 0 prices direct and searches list 0's 16x16 while wave 1 searches
// list 1's; wave 0 prices bi and fills the first half of the record; the
// two waves run the 16x8 and 8x16 b_partition searches; one thread
// commits.  The job and record stay in LDS.  Picture p's MBs are jobs,
// records and exchange words [job0, job0 + mbw*mbh) of the launch (BPic).
// ctl[0]: tickets taken, ctl[1]:
// stall flag (a wait past spin_limit polls ends the launch: the walkers
// leave their rows and take no more), *sticky: the same, kept across
// launches until the host reads and clears it.
__device__ __forceinline__ uint32_t bword(uint32_t tag, int ref, pcme_mv m) {
  return xword(tag, ref >= 0 ? 1u : 0u, pack_mv(m.x, m.y));
}

// Exchange word q once it carries `tag`; past spin_limit polls raises
// *stall (this workgroup), ctl[1] (this launch) and *sticky (kept for the
// host's check).
__device__ __forceinline__ uint32_t bwait(const uint32_t *q, uint32_t tag, uint32_t spin_limit, int *stall,
                                          uint32_t *ctl, uint32_t *sticky) {
  uint32_t v = 0, spins = 0;
  for (;;) {
    v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (XTAG(v) == tag) return v;
    if (++spins > spin_limit) {
      *stall = 1;
      __hip_atomic_store(ctl + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(sticky, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return v;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

__global__ void __launch_bounds__(128) b_walk_kernel(const pcme_b_job *__restrict__ jobs,
                                                     const BPic *__restrict__ pics,
                                                     const int2 *__restrict__ tickets, int ntickets,
                                                     pcme_b_out *res, uint32_t *xw, uint32_t *ctl,
                                                     uint32_t *sticky, uint32_t spin_limit) {
  __shared__ __attribute__((aligned(16))) uint32_t s_lds[2][LDS_PER_WAVE / 4];
  __shared__ __attribute__((aligned(16))) uint8_t s_pred[2][3][256];
  __shared__ __attribute__((aligned(16))) pcme_b_job s_job;
  __shared__ __attribute__((aligned(16))) pcme_b_out s_rec;
  __shared__ __attribute__((aligned(16))) pcme_b_mb s_mb;
  __shared__ uint32_t s_mv16[3];
  __shared__ uint32_t s_above[3][8];  // exchange words of the row above: MB x-1, x, x+1
  __shared__ pcme_nb_block s_left[2][4];  // this row's previous MB: blocks 3, 7, 11, 15 per list
  __shared__ int s_ctl[2];            // the ticket taken; a stall
  static_assert(sizeof(pcme_b_job) % 4 == 0 && sizeof(pcme_b_job) / 4 <= 128 && sizeof(pcme_b_out) / 4 <= 128 &&
                    sizeof(pcme_b_mb) % 4 == 0 && sizeof(pcme_b_mb) / 4 <= 128,
                "one dword per thread");
  const uint32_t tag = 1u;
  const int t = (int)threadIdx.x, w = t >> 6;
  for (;;) {
    if (t == 0) {
      s_ctl[0] = (int)atomicAdd(ctl, 1u);
      s_ctl[1] = __hip_atomic_load(ctl + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
    }
    __syncthreads();
    if (s_ctl[0] >= ntickets || s_ctl[1]) return;  // done, or another row stalled
    // ticket = (picture, row), read-only in this launch
    const int2 tk = tickets[s_ctl[0]];
    const BPic &P = pics[tk.x];
    const int my = tk.y, mbw = P.mbw, partitions = P.partitions;
    const pcme_b_job *pj = jobs + P.job0;
    pcme_b_out *pres = res + P.job0;
    uint32_t *pxw = xw + (size_t)P.job0 * 8;
    pcme_b_mb *field = P.field;
    const pcme_b_col *col = P.col;
    for (int mx = 0; mx < mbw; mx++) {
      const int k = my * mbw + mx;
      // wait for the row above's MB x+1 (x in the last column); its words
      // and those of MB x (and x-1) read earlier give the top neighbours
      if (t < 8 && my > 0) {
        const int dx = mx + 1 < mbw ? 1 : 0;
        if (mx == 0) {
          s_above[1][t] = bwait(pxw + (size_t)(k - mbw) * 8 + t, tag, spin_limit, &s_ctl[1], ctl, sticky);  // MB x
        } else {
          s_above[0][t] = s_above[1][t];
          s_above[1][t] = s_above[2][t];
        }
        s_above[2][t] = bwait(pxw + (size_t)(k - mbw + dx) * 8 + t, tag, spin_limit, &s_ctl[1], ctl, sticky);
      }
      if (t < (int)(sizeof(pcme_b_job) / 4)) ((uint32_t *)&s_job)[t] = ((const uint32_t *)&pj[k])[t];
      if (t < (int)(sizeof(pcme_b_out) / 4)) ((uint32_t *)&s_rec)[t] = 0u;
      __syncthreads();
      if (s_ctl[1]) return;  // a stall: the launch ends (every wave reads the same flag)
      if (t < 20) {
        // neighbour block i of list l: left[4] (blocks 3, 7, 11, 15 of MB
        // x-1), top[4] (12..15 of the MB above), top-right (12 of x+1),
        // top-left (15 of x-1); -2 outside the picture
        const int l = t / 10, i = t % 10;
        pcme_nb_block b;
        b.reserved[0] = b.reserved[1] = b.reserved[2] = 0;
        b.mv.x = b.mv.y = 0;
        b.ref = -2;
        if (i < 4) {
          if (mx > 0) b = s_left[l][i];
        } else if (my > 0) {
          const bool tr = i == 8, tl = i == 9;
          const bool av = tr ? mx + 1 < mbw : (tl ? mx > 0 : true);
          if (av) {
            const int slot = tr ? 2 : (tl ? 0 : 1);
            const int blk = tr ? 0 : (tl ? 3 : i - 4);
            const uint32_t v = s_above[slot][4 * l + blk];
            b.ref = (int8_t)(xword_ref(v) ? 0 : -1);
            const uint32_t m = b.ref >= 0 ? xword_mv(v) : 0u;
            b.mv.x = (int16_t)mv_x(m);
            b.mv.y = (int16_t)mv_y(m);
          }
        }
        s_job.nb[l][i] = b;
      } else if (t >= 32 && t < 40) {
        const int l = (t - 32) >> 2, q = (t - 32) & 3;
        const pcme_b_col &c = col[k];
        s_job.col_mv[l][q] = c.mv[l][q];
        s_job.col_ref[l][q] = c.ref[l][q];
      } else if (t == 40) {
        s_job.col_intra = col[k].intra;
      }
      __syncthreads();  // the job's neighbours
      const pcme_b_job &J = s_job;
      SCtx c[2];
      b_ctx(J, c);
      const NbrG nb[2] = {nbr_of(J.nb[0]), nbr_of(J.nb[1])};
      const BDirect D = b_direct(J, c, nb);
      PartRes r16[2];
      uint32_t direct_cost = 0u;
      if (w == 0) {
        direct_cost = b_direct_cost(J, c, D, s_pred[0]);
        b_search16(J, c[0], nb[0], D.dmv[0][0], s_lds[0], r16[0]);
      } else {
        b_search16(J, c[1], nb[1], D.dmv[1][0], s_lds[1], r16[1]);
        if (lane_id() == 0) {
          s_mv16[0] = r16[1].mv[0];
          s_mv16[1] = r16[1].mvp[0];
          s_mv16[2] = r16[1].total;
        }
      }
      __syncthreads();  // list 1's 16x16 result
      if (w == 0) {
        r16[1].mv[0] = s_mv16[0];
        r16[1].mvp[0] = s_mv16[1];
        r16[1].total = s_mv16[2];
        uint32_t mv16[2];
        b_finish16(J, c, D, direct_cost, r16, s_pred[0], &s_rec, mv16);
        if (lane_id() == 0) {
          s_mv16[0] = mv16[0];
          s_mv16[1] = mv16[1];
        }
      }
      __syncthreads();  // the first half's record and MVs
      const uint32_t mv16[2] = {s_mv16[0], s_mv16[1]};
      if (w == 0)
        b_analyse_part<0>(J, &s_rec, s_lds[0], s_pred[0], mv16);
      else
        b_analyse_part<1>(J, &s_rec, s_lds[1], s_pred[1], mv16);
      __syncthreads();  // the whole record
      if (t < (int)(sizeof(pcme_b_out) / 4)) ((uint32_t *)&pres[k])[t] = ((const uint32_t *)&s_rec)[t];
      if (t == 0) b_commit(s_rec, s_mb, partitions);
      __syncthreads();  // the committed record
      if (t < (int)(sizeof(pcme_b_mb) / 4)) ((uint32_t *)&field[k])[t] = ((const uint32_t *)&s_mb)[t];
      if (t < 16) {
        // lanes 0-7: the bottom blocks 12-15 of both lists for the row below;
        // lanes 8-15: the right column 3, 7, 11, 15 for the next MB
        const int l = (t & 7) >> 2, i = t & 3;
        const pcme_b_mb &m = s_mb;
        const int blk = t < 8 ? 12 + i : 4 * i + 3;
        const pcme_mv mv = m.mv[l][blk];
        const int rf = m.ref[l][blk8(blk)];
        if (t < 8) {
          __hip_atomic_store(pxw + (size_t)k * 8 + t, bword(tag, rf, mv), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
          pcme_nb_block b;
          b.reserved[0] = b.reserved[1] = b.reserved[2] = 0;
          b.ref = (int8_t)rf;
          b.mv = rf >= 0 ? mv : pcme_mv{0, 0};
          s_left[l][i] = b;
        }
      }
      __syncthreads();  // s_left, s_above and s_job are rewritten for the next MB
    }
  }
}

// Clears the walker's exchange words and control words before a launch.
__global__ void __launch_bounds__(256) b_walk_clear_kernel(uint32_t *xw, size_t nwords, uint32_t *ctl) {
  const size_t i0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i0 < 2) ctl[i0] = 0u;
  for (size_t i = i0; i < nwords; i += (size_t)gridDim.x * blockDim.x) xw[i] = 0u;
}

// -------------------------------------------------- per-call search batch

__global__ void __launch_bounds__(64) search_batch_kernel(const pcme_search_job *__restrict__ jobs,
                                                          int njobs, pcme_mv *out_mv,
                                                          uint32_t *out_cost) {
  __shared__ __attribute__((aligned(16))) uint32_t s_lds[LDS_PER_WAVE / 4];
  const int j = blockIdx.x;
  if (j >= njobs) return;
  const pcme_search_job &J = jobs[j];
  SCtx c;
  c.pl[0] = J.ref; c.pl[1] = J.ref_h; c.pl[2] = J.ref_v; c.pl[3] = J.ref_c;
  c.stride = 16 * J.mbw + 2 * PCME_PAD;
  c.pw = 16 * J.mbw;
  c.ph = 16 * J.mbh;
  c.vlim = (int)J.vlim;
  c.lam = J.lam;
  c.range = J.me_range;
  c.subpel = J.subpel_rounds;
  int c0x = J.ncands > 0 ? J.cands[0].x : 0, c0y = J.ncands > 0 ? J.cands[0].y : 0;
  int c1x = J.ncands > 1 ? J.cands[1].x : 0, c1y = J.ncands > 1 ? J.cands[1].y : 0;
  SRes r;
  r.x = r.y = 0; r.cost = 0;
  Area ar = no_area();
  const uint32_t hv = 0u;  // full search: no hints
  if (J.w == 16 && J.h == 16) {
    Lane L = lane_layout<16, 16>(); SrcRegs s; load_src(J.src, J.sstride, L, s);
    r = search<16, 16, false>(c, s, L, J.bx, J.by, J.mvp.x, J.mvp.y, c0x, c0y, c1x, c1y, J.ncands, hv, -1, s_lds,
                           ar);
  } else if (J.w == 16 && J.h == 8) {
    Lane L = lane_layout<16, 8>(); SrcRegs s; load_src(J.src, J.sstride, L, s);
    r = search<16, 8, false>(c, s, L, J.bx, J.by, J.mvp.x, J.mvp.y, c0x, c0y, c1x, c1y, J.ncands, hv, -1, s_lds,
                           ar);
  } else if (J.w == 8 && J.h == 16) {
    Lane L = lane_layout<8, 16>(); SrcRegs s; load_src(J.src, J.sstride, L, s);
    r = search<8, 16, false>(c, s, L, J.bx, J.by, J.mvp.x, J.mvp.y, c0x, c0y, c1x, c1y, J.ncands, hv, -1, s_lds,
                           ar);
  } else {
    Lane L = lane_layout<8, 8>(); SrcRegs s; load_src(J.src, J.sstride, L, s);
    r = search<8, 8, false>(c, s, L, J.bx, J.by, J.mvp.x, J.mvp.y, c0x, c0y, c1x, c1y, J.ncands, hv, -1, s_lds,
                           ar);
  }
  if (lane_id() == 0) {
    pcme_mv m;
    m.x = (int16_t)r.x;
    m.y = (int16_t)r.y;
    out_mv[j] = m;
    out_cost[j] = r.cost;
  }
}

// ------------------------------------------------------------- pad kernel

// One thread per 4 output bytes of a padded plane row; grid.z = plane.
__global__ void __launch_bounds__(256) pad_planes_kernel(const uint8_t *__restrict__ src, int64_t src_pitch,
                                                         int sstride, int w, int h, uint8_t *__restrict__ dst,
                                                         int64_t dst_pitch, int dstride, int rows, int pad) {
  const int gx = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
  const int row = blockIdx.y;
  const int pl = blockIdx.z;
  if (gx >= dstride || row >= rows) return;
  const uint8_t *s = src + (size_t)pl * src_pitch;
  uint8_t *d = dst + (size_t)pl * dst_pitch + (size_t)row * dstride;
  int y = row - pad;
  y = y < 0 ? 0 : (y > h - 1 ? h - 1 : y);
  const uint8_t *sr = s + (size_t)y * sstride;
  uint32_t v = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    int x = gx + i - pad;
    x = x < 0 ? 0 : (x > w - 1 ? w - 1 : x);
    v |= (uint32_t)sr[x] << (8 * i);
  }
  if (gx + 4 <= dstride) {
    *(uint32_t *)(d + gx) = v;
  } else {
    for (int i = 0; gx + i < dstride; i++) d[gx + i] = (uint8_t)(v >> (8 * i));
  }
}

// ------------------------------------------------------------ hpel kernel

// Tile of HT_W columns x HT_H rows per 256-thread block: source rows and
// columns (+-2/3 taps, clamped to the plane) and the i16 horizontal
// intermediates are staged in LDS; each thread then writes 4 consecutive
// bytes of one row of each output plane per pass.
__global__ void __launch_bounds__(256) hpel_kernel(const uint8_t *__restrict__ ref, uint8_t *__restrict__ ho,
                                                   uint8_t *__restrict__ vo, uint8_t *__restrict__ co,
                                                   int64_t pitch, int stride, int rows) {
  __shared__ uint8_t s_src[HT_H + 5][HT_W + 8];
  __shared__ int16_t s_h[HT_H + 5][HT_W];
  const int x0 = blockIdx.x * HT_W, y0 = blockIdx.y * HT_H;
  const size_t poff = (size_t)blockIdx.z * pitch;
  const uint8_t *src = ref + poff;
  const int t = threadIdx.x;
  // Source rows y0-2 .. y0+HT_H+2, columns x0-2 .. x0+HT_W+5 (clamped).
  for (int i = t; i < (HT_H + 5) * (HT_W + 8); i += 256) {
    int r = i / (HT_W + 8), cx = i % (HT_W + 8);
    int yy = y0 - 2 + r, xx = x0 - 2 + cx;
    yy = yy < 0 ? 0 : (yy > rows - 1 ? rows - 1 : yy);
    xx = xx < 0 ? 0 : (xx > stride - 1 ? stride - 1 : xx);
    s_src[r][cx] = src[(size_t)yy * stride + xx];
  }
  __syncthreads();
  // Horizontal intermediates (i16) for all HT_H+5 rows.
  for (int i = t; i < (HT_H + 5) * HT_W; i += 256) {
    int r = i / HT_W, cx = i % HT_W;
    const uint8_t *e = &s_src[r][cx];  // e[0] = x-2
    s_h[r][cx] = (int16_t)((int)e[0] + (int)e[5] - 5 * ((int)e[1] + (int)e[4]) + 20 * ((int)e[2] + (int)e[3]));
  }
  __syncthreads();
  // Outputs: thread -> 4 consecutive columns of one row, 8 rows per pass.
  const int cg = (t & 15) * 4;
  for (int r = t >> 4; r < HT_H; r += 16) {
    const int y = y0 + r;
    if (y >= rows) break;
    uint32_t hv = 0, vv = 0, cv = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int cx = cg + i;
      const int k = r + 2;
      int hval = ((int)s_h[k][cx] + 16) >> 5;
      hval = hval < 0 ? 0 : (hval > 255 ? 255 : hval);
      int sv = (int)s_src[k - 2][cx + 2] + (int)s_src[k + 3][cx + 2] -
               5 * ((int)s_src[k - 1][cx + 2] + (int)s_src[k + 2][cx + 2]) +
               20 * ((int)s_src[k][cx + 2] + (int)s_src[k + 1][cx + 2]);
      int vval = (sv + 16) >> 5;
      vval = vval < 0 ? 0 : (vval > 255 ? 255 : vval);
      int sc = (int)s_h[k - 2][cx] + (int)s_h[k + 3][cx] - 5 * ((int)s_h[k - 1][cx] + (int)s_h[k + 2][cx]) +
               20 * ((int)s_h[k][cx] + (int)s_h[k + 1][cx]);
      int cval = (sc + 512) >> 10;
      cval = cval < 0 ? 0 : (cval > 255 ? 255 : cval);
      // Keep each clamped byte opaque: fused into v_ashr_pk_u8_i32 and
      // packed with v_lshl_or_b32 the upper bytes came out 0xFF on gfx950
      // (the fused op's bits 16-31 were not zero).
      asm volatile("" : "+v"(hval), "+v"(vval), "+v"(cval));
      hv |= (uint32_t)hval << (8 * i);
      vv |= (uint32_t)vval << (8 * i);
      cv |= (uint32_t)cval << (8 * i);
    }
    const int x = x0 + cg;
    const size_t o = poff + (size_t)y * stride + x;
    if (x + 4 <= stride) {
      *(uint32_t *)(ho + o) = hv;
      *(uint32_t *)(vo + o) = vv;
      *(uint32_t *)(co + o) = cv;
    } else {
      for (int i = 0; x + i < stride; i++) {
        ho[o + i] = (uint8_t)(hv >> (8 * i));
        vo[o + i] = (uint8_t)(vv >> (8 * i));
        co[o + i] = (uint8_t)(cv >> (8 * i));
      }
    }
  }
}

// ------------------------------------------------------ luma planes kernel

// Import + edge extension + the three half-pel planes of a luma frame in one
// pass from the unpadded source (pcme_pad_planes then pcme_hpel, fused).
// Clamping a tap to the padded plane and then the padded pixel to the
// picture is clamping the tap to the picture, so every output pixel reads
// the source directly: full(y, x) = S(cy(y), cx(x)), H/V = the 6-tap filters
// (1, -5, 20, 20, -5, 1) along x / y rounded (+16) >> 5, C = the same filter
// along y over the unrounded horizontal sums, rounded (+512) >> 10, each
// clipped to 0..255 (picture.rs:318-380), with cy, cx the clamps into the
// picture.  A thread owns 4 output columns of an LP_TH-row band and slides
// a 6-row window down it: per source row 8 byte pairs (packed i16), the
// horizontal sums of its 4 columns; per output row the H, V and C dwords.

// Byte pair (b[k], b[k+1]) of the 12 source bytes b[-4..7] held in D0 (b[-4..-1]),
// D1 (b[0..3]), D2 (b[4..7]) as two zero-extended 16-bit lanes.
__device__ __forceinline__ v2s lp_pair(uint32_t D0, uint32_t D1, uint32_t D2, int k) {
  switch (k) {
    case -2: return as_v2s(__builtin_amdgcn_perm(0u, D0, 0x0C030C02u));
    case -1: return as_v2s(__builtin_amdgcn_perm(D1, D0, 0x0C040C03u));
    case 0: return as_v2s(__builtin_amdgcn_perm(0u, D1, 0x0C010C00u));
    case 1: return as_v2s(__builtin_amdgcn_perm(0u, D1, 0x0C020C01u));
    case 2: return as_v2s(__builtin_amdgcn_perm(0u, D1, 0x0C030C02u));
    case 3: return as_v2s(__builtin_amdgcn_perm(D2, D1, 0x0C040C03u));
    case 4: return as_v2s(__builtin_amdgcn_perm(0u, D2, 0x0C010C00u));
    default: return as_v2s(__builtin_amdgcn_perm(0u, D2, 0x0C020C01u));
  }
}

// The 6-tap filter of packed i16 lanes: a + f - 5 (b + e) + 20 (c + d).
__device__ __forceinline__ v2s lp_tap6(v2s a, v2s b, v2s c, v2s d, v2s e, v2s f) {
  return (a + f) - (v2s)5 * (b + e) + (v2s)20 * (c + d);
}

// Four bytes of two rounded, clipped i16 pairs (lo = bytes 0-1, hi = 2-3).
__device__ __forceinline__ uint32_t lp_bytes(v2s lo, v2s hi) {
  const v2s z = (v2s)0, m = (v2s)255;
  lo = __builtin_elementwise_min(__builtin_elementwise_max(lo, z), m);
  hi = __builtin_elementwise_min(__builtin_elementwise_max(hi, z), m);
  return __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, hi), __builtin_bit_cast(uint32_t, lo), 0x06040200u);
}

// Column j (0 or 1 = low or high lane) of the 2-D filter over six rows of
// horizontal sums (i32: the sum reaches 42 * 10710).
__device__ __forceinline__ int lp_c(v2s h0, v2s h1, v2s h2, v2s h3, v2s h4, v2s h5, int j) {
  const uint32_t sel = j ? 0x07060302u : 0x05040100u;
  const auto pr = [sel](v2s a, v2s b) {
    return as_v2s(__builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, b), __builtin_bit_cast(uint32_t, a), sel));
  };
  int acc = __builtin_amdgcn_sdot2(pr(h0, h5), (v2s)1, 0, false);
  acc = __builtin_amdgcn_sdot2(pr(h1, h4), (v2s)-5, acc, false);
  return __builtin_amdgcn_sdot2(pr(h2, h3), (v2s)20, acc, false);
}

__device__ __forceinline__ uint32_t lp_cbyte(int v) {
  int b = (v + 512) >> 10;
  b = b < 0 ? 0 : (b > 255 ? 255 : b);
  asm volatile("" : "+v"(b));  // see hpel_kernel: no fused byte packing
  return (uint32_t)b;
}

// The band of LP_TH output rows from y0 and 4 output columns from x0 of one
// frame: source frame s, planes at full/hp/vp/cp + po.
__device__ __forceinline__ void lp_band(const uint8_t *__restrict__ s, int sstride, int w, int h,
                                        uint8_t *__restrict__ full, uint8_t *__restrict__ hp,
                                        uint8_t *__restrict__ vp, uint8_t *__restrict__ cp, size_t po, int stride,
                                        int rows, int x0, int y0) {
  const int sx0 = x0 - PCME_PAD;
  // Interior lanes read three aligned dwords of a row; lanes near the
  // picture's left/right edges gather clamped bytes.
  const bool fast = (sx0 >= 4) & (sx0 + 8 <= w) & ((((uintptr_t)s | (uintptr_t)sstride) & 3u) == 0u);
  v2s q01[6], q23[6], h01[6], h23[6];
#pragma unroll
  for (int i = 0; i < LP_TH + 5; i++) {
    const int yy = y0 - 2 + i;  // padded row of source row i of the window
    int sy = yy - PCME_PAD;
    sy = sy < 0 ? 0 : (sy > h - 1 ? h - 1 : sy);
    const uint8_t *row = s + (size_t)sy * sstride;
    uint32_t D0, D1, D2;
    if (fast) {
      const uint32_t *r32 = (const uint32_t *)(row + sx0);
      D0 = r32[-1];
      D1 = r32[0];
      D2 = r32[1];
    } else {
      uint32_t b[9];
#pragma unroll
      for (int k = 0; k < 9; k++) {
        int x = sx0 - 2 + k;
        x = x < 0 ? 0 : (x > w - 1 ? w - 1 : x);
        b[k] = row[x];
      }
      D0 = (b[0] << 16) | (b[1] << 24);
      D1 = b[2] | (b[3] << 8) | (b[4] << 16) | (b[5] << 24);
      D2 = b[6] | (b[7] << 8) | (b[8] << 16);
    }
    const int j = i % 6;
    const v2s pm2 = lp_pair(D0, D1, D2, -2), pm1 = lp_pair(D0, D1, D2, -1), p0 = lp_pair(D0, D1, D2, 0),
              p1 = lp_pair(D0, D1, D2, 1), p2 = lp_pair(D0, D1, D2, 2), p3 = lp_pair(D0, D1, D2, 3),
              p4 = lp_pair(D0, D1, D2, 4), p5 = lp_pair(D0, D1, D2, 5);
    q01[j] = p0;
    q23[j] = p2;
    h01[j] = lp_tap6(pm2, pm1, p0, p1, p2, p3);
    h23[j] = lp_tap6(p0, p1, p2, p3, p4, p5);
    if (i < 5) continue;
    // Output row y = yy - 3 uses window rows i-5 .. i (source rows y-2 .. y+3).
    const int y = yy - 3;
    if (y >= rows) continue;
    const int a0 = (i - 5) % 6, a1 = (i - 4) % 6, a2 = (i - 3) % 6, a3 = (i - 2) % 6, a4 = (i - 1) % 6, a5 = i % 6;
    const size_t o = po + (size_t)y * stride + x0;
    const uint32_t fv = __builtin_amdgcn_perm(__builtin_bit_cast(uint32_t, q23[a2]),
                                              __builtin_bit_cast(uint32_t, q01[a2]), 0x06040200u);
    const uint32_t hv = lp_bytes((h01[a2] + (v2s)16) >> (v2s)5, (h23[a2] + (v2s)16) >> (v2s)5);
    const v2s v01 = lp_tap6(q01[a0], q01[a1], q01[a2], q01[a3], q01[a4], q01[a5]);
    const v2s v23 = lp_tap6(q23[a0], q23[a1], q23[a2], q23[a3], q23[a4], q23[a5]);
    const uint32_t vv = lp_bytes((v01 + (v2s)16) >> (v2s)5, (v23 + (v2s)16) >> (v2s)5);
    const uint32_t c0 = lp_cbyte(lp_c(h01[a0], h01[a1], h01[a2], h01[a3], h01[a4], h01[a5], 0));
    const uint32_t c1 = lp_cbyte(lp_c(h01[a0], h01[a1], h01[a2], h01[a3], h01[a4], h01[a5], 1));
    const uint32_t c2 = lp_cbyte(lp_c(h23[a0], h23[a1], h23[a2], h23[a3], h23[a4], h23[a5], 0));
    const uint32_t c3 = lp_cbyte(lp_c(h23[a0], h23[a1], h23[a2], h23[a3], h23[a4], h23[a5], 1));
    const uint32_t cv = __builtin_amdgcn_perm(__builtin_amdgcn_perm(c3, c2, 0x0C0C0400u),
                                              __builtin_amdgcn_perm(c1, c0, 0x0C0C0400u), 0x05040100u);
    *(uint32_t *)(full + o) = fv;
    *(uint32_t *)(hp + o) = hv;
    *(uint32_t *)(vp + o) = vv;
    *(uint32_t *)(cp + o) = cv;
  }
}

__global__ void __launch_bounds__(256) luma_planes_kernel(const uint8_t *__restrict__ src, int64_t src_pitch,
                                                          int sstride, int w, int h, uint8_t *__restrict__ full,
                                                          uint8_t *__restrict__ hp, uint8_t *__restrict__ vp,
                                                          uint8_t *__restrict__ cp, int64_t pitch, int stride,
                                                          int rows) {
  const int x0 = (blockIdx.x * 256 + threadIdx.x) * 4;
  if (x0 >= stride) return;  // stride is a multiple of 16
  lp_band(src + (size_t)blockIdx.z * src_pitch, sstride, w, h, full, hp, vp, cp, (size_t)blockIdx.z * pitch,
          stride, rows, x0, blockIdx.y * LP_TH);
}

// Several clips of any sizes in one launch.  The blocks of clip k start at
// jobs.j[k].first, bpf per frame; a frame's (band, 4-column group) pairs are
// numbered band-major and dealt 256 to a block, so a narrow picture fills
// its blocks' lanes (one band per block would leave most of them idle).
__global__ void __launch_bounds__(256) luma_planes_batch_kernel(const LpBatch jobs) {
  const int b = blockIdx.x;
  int k = 0;
  while (k + 1 < jobs.n && b >= jobs.j[k + 1].first) k++;
  const LpJob &J = jobs.j[k];
  const int local = b - J.first, frame = local / J.bpf;
  const int idx = (local - frame * J.bpf) * 256 + (int)threadIdx.x;
  if (idx >= J.ngx * J.nbands) return;
  const int band = idx / J.ngx;
  lp_band(J.src + (size_t)frame * J.src_pitch, J.sstride, J.w, J.h, J.full, J.hp, J.vp, J.cp,
          (size_t)frame * J.pitch, J.stride, J.rows, (idx - band * J.ngx) * 4, band * LP_TH);
}

// ------------------------------------------------------------ pred kernel

// One 256-thread block per MB: threads make the 16x16 luma prediction
// (mc_luma of the partition covering each pixel) and threads 0..127 the two
// 8x8 chroma predictions (mc_chroma, picture.rs:451-487).
__global__ void __launch_bounds__(256) pred_frame_kernel(const uint8_t *__restrict__ pl0, const uint8_t *__restrict__ pl1,
                                                         const uint8_t *__restrict__ pl2, const uint8_t *__restrict__ pl3,
                                                         const uint8_t *__restrict__ pu, const uint8_t *__restrict__ pv,
                                                         const pcme_mb_out *__restrict__ mbs, int mbw, int mbh,
                                                         uint8_t *__restrict__ oy, uint8_t *__restrict__ ou,
                                                         uint8_t *__restrict__ ov) {
  const int mb = blockIdx.x;
  if (mb >= mbw * mbh) return;
  const int mx = mb % mbw, my = mb / mbw;
  const pcme_mb_out &m = mbs[mb];
  const int t = threadIdx.x;
  const int ys = 16 * mbw + 2 * PCME_PAD;
  {
    const int px = t & 15, py = t >> 4;
    const pcme_mv mv = m.mv[(py >> 2) * 4 + (px >> 2)];
    const int xi = 16 * mx + px + (mv.x >> 2), yi = 16 * my + py + (mv.y >> 2);
    const uint32_t e = qpel_entry((uint32_t)((mv.y & 3) * 4 + (mv.x & 3)));
    SCtx pc;
    pc.pl[0] = pl0; pc.pl[1] = pl1; pc.pl[2] = pl2; pc.pl[3] = pl3;
    const uint8_t a = plane_sel(pc, e & 3u)[(size_t)(yi + ((e >> 3) & 1) + PCME_PAD) * ys + xi + ((e >> 2) & 1) + PCME_PAD];
    const uint8_t b = plane_sel(pc, (e >> 4) & 3u)[(size_t)(yi + ((e >> 7) & 1) + PCME_PAD) * ys + xi + ((e >> 6) & 1) + PCME_PAD];
    oy[(size_t)mb * 256 + t] = (uint8_t)(((uint32_t)a + (uint32_t)b + 1u) >> 1);
  }
  if (t < 128) {
    const int cs = 8 * mbw + 2 * PCME_CPAD;
    const int k = t & 63;
    const int cx = k & 7, cy = k >> 3;
    const pcme_mv mv = m.mv[(cy >> 1) * 4 + (cx >> 1)];
    const int xi = 8 * mx + cx + (mv.x >> 3), yi = 8 * my + cy + (mv.y >> 3);
    const uint32_t fx = (uint32_t)(mv.x & 7), fy = (uint32_t)(mv.y & 7);
    const uint8_t *p = (t < 64 ? pu : pv) + (size_t)(yi + PCME_CPAD) * cs + xi + PCME_CPAD;
    const uint32_t v = ((8 - fx) * (8 - fy) * p[0] + fx * (8 - fy) * p[1] + (8 - fx) * fy * p[cs] +
                        fx * fy * p[cs + 1] + 32) >> 6;
    (t < 64 ? ou : ov)[(size_t)mb * 64 + k] = (uint8_t)v;
  }
}

// Per launch: reset the ticket counters and the stall flag ctl[1] (ctl[4],
// the sticky copy, is cleared only by the host's check) and advance the
// launch epoch ctl[2], whose low byte tags the exchange words.  Tag 0 is
// skipped; when the tag wraps to 1, ctl[3] asks clear_kernel to zero the
// exchange words so no word of 255 launches ago can pass for a fresh one.
// Everything lives in device memory, so a captured graph replays correctly.
__global__ void prep_kernel(uint32_t *ctl) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    uint32_t e = ctl[2] + 1u;
    if ((e & 0x3Fu) == 0u) e += 1u;  // tag 0 is a cleared word
    ctl[0] = 0u;
    ctl[1] = 0u;
    ctl[2] = e;
    ctl[3] = (e & 0x3Fu) == 1u ? 1u : 0u;  // the tags wrapped: clear the words
  }
  if (blockIdx.x == 0 && threadIdx.x < PCME_QUEUES) ctl[8 + threadIdx.x] = 0u;  // per-queue ticket counters
}

__global__ void __launch_bounds__(256) clear_kernel(const uint32_t *ctl, uint32_t *xchg, size_t nwords) {
  if (ctl[3] == 0u) return;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (size_t)gridDim.x * blockDim.x)
    xchg[i] = 0u;
}

}  // namespace pcme

#ifdef PCME_TIMELINE
// Timeline build only: choose the recorded job and clear its stamps; read
// them back (4 x u64 per MB, raster order, 0 = not reached).
extern "C" int pcme_timeline_set(int job) {
  static unsigned long long zero[TL_MBS * 4];
  if (hipMemcpyToSymbol(HIP_SYMBOL(pcme::g_tl_job), &job, sizeof job, 0, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpyToSymbol(HIP_SYMBOL(pcme::g_timeline), zero, sizeof zero, 0, hipMemcpyHostToDevice) != hipSuccess)
    return PCME_EHIP;
  return PCME_OK;
}
extern "C" int pcme_timeline_read(unsigned long long *dst, int mbs) {
  if (!dst || mbs < 0 || mbs > TL_MBS) return PCME_EINVAL;
  return hipMemcpyFromSymbol(dst, HIP_SYMBOL(pcme::g_timeline), (size_t)mbs * 32, 0, hipMemcpyDeviceToHost) ==
                 hipSuccess
             ? PCME_OK
             : PCME_EHIP;
}
#endif
