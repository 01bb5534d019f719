// pcme_lookahead.hip -- gfx950 kernels of the lowres lookahead: the 2x2
// downsample (lookahead.rs:10-33), the scenecut complexity
// (lookahead.rs:52-108) and the macroblock-tree block costs
// (mbtree.rs:21-159).
//
// Kernels:
//   lowres_kernel      one thread per 4 output bytes; HBM-bound (reads 1 B,
//                      writes 1/4 B per input pixel).
//   complexity_kernel  one wave per 8x8 block (blocks are independent):
//                      intra cost, diamond SAD search from (0, 0), SATD at
//                      its best; the frame sums are reduced per workgroup
//                      and added with one 64-bit atomic each (integer adds:
//                      the sums do not depend on the order).
//   tree_kernel        one wave per block row: a row is sequential in the
//                      reference (each block's search is seeded with its
//                      left neighbour's MVs), rows and frames are not.
//
// Lane layouts (wave64):
//   SAD passes   lane = 8*slot + row: slot = candidate (8 per pass), row =
//                one 8-byte row of the block; the 8 row SADs of a slot are
//                summed with DPP and the strict-< scan over the slots is a
//                min over (cost, slot) keys (min_key).
//   SATD passes  lane & 1 = strip (rows 0-3 or 4-7 of the block), bits 1-2
//                select what the lane prices (an intra mode, or list 0 /
//                list 1 / bi); the strip pair is summed with DPP and the
//                per-kind totals read from lanes 0, 2, 4.  Lanes 8-63 repeat
//                lanes 0-7: the wave stays converged and the extra lanes
//                cost no time.
//
// Memory: lowres frames are read through L2 (a 1080p lowres frame is
// 510 KiB); unaligned 8-byte rows are assembled from aligned dwords with
// v_alignbyte.  A frame's base and width are 4-byte multiples, so every
// dword that holds a needed byte lies inside the frame.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pcme_internal.h"
#include "pcme_dev.h"

namespace pcme {

// ------------------------------------------------------------------ lowres

// 2x2 sums of bytes (0,1) and (2,3) of rows x and y, rounded and halved
// twice, in bytes 0 and 2 (16-bit lanes: at most 4*255+2, no carry out).
__device__ __forceinline__ uint32_t lr_pair(uint32_t x, uint32_t y) {
  const uint32_t s = (x & 0x00FF00FFu) + ((x >> 8) & 0x00FF00FFu) + (y & 0x00FF00FFu) +
                     ((y >> 8) & 0x00FF00FFu) + 0x00020002u;
  return (s >> 2) & 0x00FF00FFu;
}

__global__ void __launch_bounds__(256) lowres_kernel(const uint8_t *__restrict__ planes, int64_t plane_pitch,
                                                     int stride, int pad, int w2, uint8_t *__restrict__ out,
                                                     int64_t out_pitch) {
  const int q = (int)(blockIdx.x * 256 + threadIdx.x);
  if (4 * q >= w2) return;
  const int y = (int)blockIdx.y, f = (int)blockIdx.z;
  const uint8_t *r0 = planes + (int64_t)f * plane_pitch + (int64_t)(2 * y + pad) * stride + pad + 8 * q;
  const uint32_t *a = (const uint32_t *)r0;
  const uint32_t *b = (const uint32_t *)(r0 + stride);
  const uint32_t lo = lr_pair(a[0], b[0]), hi = lr_pair(a[1], b[1]);
  *(uint32_t *)(out + (int64_t)f * out_pitch + (int64_t)y * w2 + 4 * q) = __builtin_amdgcn_perm(hi, lo, 0x06040200u);
}

// ------------------------------------------------------- block primitives

__device__ __forceinline__ uint32_t ld32(const uint8_t *p) { return *(const uint32_t *)p; }

// The 8 bytes of row y of a lowres frame from column x (any alignment).
__device__ __forceinline__ void row8(const uint8_t *frame, int w, int x, int y, uint32_t &lo, uint32_t &hi) {
  const uint32_t *p = (const uint32_t *)(frame + (size_t)y * w + (x & ~3));
  const uint32_t sh = (uint32_t)x & 3u;
  const uint32_t a = p[0], b = p[1], c = p[sh ? 2 : 1];
  lo = __builtin_amdgcn_alignbyte(b, a, sh);
  hi = __builtin_amdgcn_alignbyte(c, b, sh);
}

// SATD of one 4-row x 8-column strip (dsp.rs:35-88, satd_strip<8, 4>):
// S, P = source and prediction rows, two dwords (columns 0-3, 4-7) per row.
__device__ __forceinline__ uint32_t strip_satd(const uint32_t *S, const uint32_t *P) {
  v2s sp[16];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++) sp[r * 4 + c] = as_v2s(__builtin_amdgcn_perm(S[2 * r + 1], S[2 * r], SELC(c)));
  return satd_strip8(sp, P);
}

// Sum of lane pairs (2k, 2k+1): the two strips of one 8x8 SATD.
__device__ __forceinline__ uint32_t pair_sum(uint32_t v) {
  return v + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_QP_1032, 0xF, 0xF, false);
}

// Per-byte floor((a + b) / 2) and (a + b + 1) >> 1.
__device__ __forceinline__ uint32_t avg_floor(uint32_t a, uint32_t b) {
  return (a & b) + (((a ^ b) >> 1) & 0x7F7F7F7Fu);
}
__device__ __forceinline__ uint32_t avg_round(uint32_t a, uint32_t b) {
  return (a | b) - (((a ^ b) >> 1) & 0x7F7F7F7Fu);
}

// Intra costs of eight blocks at once (mbtree.rs:21-49 intra_cost): lane
// group g = lane >> 3 prices the block at (x0, y0) of its lanes, lane kind
// (lane >> 1) & 3 = mode 0 (average of top row and left column, 128 at the
// top or left edge), 1 (left column, needs x0 > 0), 2 (top row, needs
// y0 > 0), 3 (unused); lane & 1 = strip.  Loads the lane's strip of the
// source into S (kept for the batch's final SATD pass) and returns the
// block's least SATD + 16 in every lane of the group.
__device__ __forceinline__ uint32_t intra8(const uint8_t *cur, int w, int x0, int y0, uint32_t *S) {
  const int lane = (int)__lane_id();
  const int mode = (lane >> 1) & 3, strip = lane & 1;
  const bool hasL = x0 > 0, hasT = y0 > 0;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const uint8_t *p = cur + (size_t)(y0 + 4 * strip + r) * w + x0;
    S[2 * r] = ld32(p);
    S[2 * r + 1] = ld32(p + 4);
  }
  const uint8_t *tp = cur + (size_t)(hasT ? y0 - 1 : y0) * w + x0;
  const uint32_t t0 = ld32(tp), t1 = ld32(tp + 4);
  uint32_t P[8];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    // left neighbour = byte 3 of the dword ending at x0 - 1
    const uint32_t l = hasL ? (ld32(cur + (size_t)(y0 + 4 * strip + r) * w + x0 - 4) >> 24) * 0x01010101u : 0u;
    uint32_t p0, p1;
    if (mode == 0) {
      p0 = (hasL && hasT) ? avg_floor(t0, l) : 0x80808080u;
      p1 = (hasL && hasT) ? avg_floor(t1, l) : 0x80808080u;
    } else if (mode == 1) {
      p0 = p1 = l;
    } else {
      p0 = t0;
      p1 = t1;
    }
    P[2 * r] = p0;
    P[2 * r + 1] = p1;
  }
  uint32_t v = pair_sum(strip_satd(S, P));
  const bool ok = mode == 0 || (mode == 1 && hasL) || (mode == 2 && hasT);
  v = ok ? v : 0xFFFFFFFFu;
  // min over the group: lanes (0,2),(1,3),(4,6),(5,7), then i <-> 7-i
  uint32_t t = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, DPP_QP_2301, 0xF, 0xF, false);
  v = min(v, t);
  t = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, DPP_ROW_HALF_MIRROR, 0xF, 0xF, false);
  v = min(v, t);
  return v + 16u;
}

// The row (lane & 7) of the 8x8 block at (x0, y0), as the SAD passes use it.
__device__ __forceinline__ void load_row(const uint8_t *cur, int w, int x0, int y0, uint32_t &s0, uint32_t &s1) {
  const uint8_t *p = cur + (size_t)(y0 + (int)(__lane_id() & 7u)) * w + x0;
  s0 = ld32(p);
  s1 = ld32(p + 4);
}

// SAD of the block against ref at (x, y), summed over each slot's 8 lanes.
__device__ __forceinline__ uint32_t slot_sad(const uint8_t *ref, int w, int x, int y, uint32_t s0, uint32_t s1) {
  uint32_t lo, hi;
  row8(ref, w, x, y + (int)(__lane_id() & 7u), lo, hi);
  return red8(__builtin_amdgcn_sad_u8(s0, lo, __builtin_amdgcn_sad_u8(s1, hi, 0u)));
}

__device__ __forceinline__ int nib(uint32_t table, int k) { return (int)((table >> (4 * k)) & 15u) - 2; }

// lookahead.rs:76-98: (-1,0),(1,0),(0,-1),(0,1),(-2,0),(2,0),(0,-2),(0,2), +2
#define DIA_DX 0x22402231u
#define DIA_DY 0x40223122u
// mbtree.rs:79-88: (-1,0),(1,0),(0,-1),(0,1),(-2,-2),(2,-2),(-2,2),(2,2), +2
#define TREE_DX 0x40402231u
#define TREE_DY 0x44003122u

// Workgroups are handed to the 8 XCDs round-robin in launch order.  This
// remap gives XCD k the k-th contiguous eighth of the launch instead, so
// the frames one XCD works on stay few and in its own L2.  A bijection of
// [0, n) for any n.
__device__ __forceinline__ int xcd_remap(int l, int n) {
  const int xcd = l & 7, r = l >> 3, q = n >> 3, rem = n & 7;
  return xcd < rem ? xcd * (q + 1) + r : rem * (q + 1) + (xcd - rem) * q + r;
}

// ------------------------------------------------------------- complexity

// One wave = 8 consecutive blocks (raster order): one intra pass for all
// eight, the eight diamond searches in turn, one SATD pass for all eight.
__global__ void __launch_bounds__(256) complexity_kernel(const CplxArgs A) {
  const int gx = (int)gridDim.x;
  const int wg = xcd_remap((int)(blockIdx.x + blockIdx.y * gridDim.x), gx * (int)gridDim.y);
  const pcme_cplx_job &J = A.j[wg / gx];
  const int w = J.w, h = J.h, bw = w >> 3, nb = bw * (h >> 3);
  const int wave = (int)(threadIdx.x >> 6), lane = (int)__lane_id();
  const int first = (wg % gx) * 4 * CPLX_BPW;
  if (first >= nb) return;  // whole workgroup
  const uint8_t *cur = J.cur, *prev = J.prev;
  const int g = lane >> 3;
  const int b0 = first + wave * CPLX_BPW;
  const int nblk = max(0, min(CPLX_BPW, nb - b0));
  uint64_t si = 0, sp = 0;
  if (nblk > 0) {
    // the lane's block (lanes past the last block repeat block b0)
    const int lb = g < nblk ? b0 + g : b0;
    const int lx0 = 8 * (lb % bw), ly0 = 8 * (lb / bw);
    uint32_t S[8];
    const uint32_t ic = intra8(cur, w, lx0, ly0, S);
    uint32_t pc = ic;
    int lmx = 0, lmy = 0;  // the lane's block's MV
    if (prev) {
      const int gdx = nib(DIA_DX, g), gdy = nib(DIA_DY, g);
      for (int b = 0; b < nblk; b++) {
        const int x0 = 8 * ((b0 + b) % bw), y0 = 8 * ((b0 + b) / bw);
        uint32_t s0, s1;
        load_row(cur, w, x0, y0, s0, s1);
        int mx = 0, my = 0;
        uint32_t bc = rdlane(slot_sad(prev, w, x0, y0, s0, s1), 0);
        for (int it = 0; it < 8; it++) {
          const int cx = mx + gdx, cy = my + gdy;
          const bool ok = x0 + cx >= 0 && y0 + cy >= 0 && x0 + cx + 8 <= w && y0 + cy + 8 <= h;
          const uint32_t c = slot_sad(prev, w, x0 + (ok ? cx : mx), y0 + (ok ? cy : my), s0, s1);
          const uint32_t key = min_key(c, ok, g);
          if (key == 0xFFFFFFFFu || (key >> 3) >= bc) break;
          bc = key >> 3;
          mx += nib(DIA_DX, (int)(key & 7u));
          my += nib(DIA_DY, (int)(key & 7u));
        }
        lmx = g == b ? mx : lmx;
        lmy = g == b ? my : lmy;
      }
      uint32_t P[8];
      const int strip = lane & 1;
#pragma unroll
      for (int r = 0; r < 4; r++) row8(prev, w, lx0 + lmx, ly0 + lmy + 4 * strip + r, P[2 * r], P[2 * r + 1]);
      pc = min(pair_sum(strip_satd(S, P)) + 8u, ic);
    }
    for (int b = 0; b < nblk; b++) {
      si += rdlane(ic, 8 * b);
      sp += rdlane(pc, 8 * b);
    }
    if (J.blocks && (lane & 7) == 0 && g < nblk) {
      pcme_cplx_block o;
      o.intra = ic;
      o.inter = pc;
      o.mv.x = (int16_t)lmx;
      o.mv.y = (int16_t)lmy;
      J.blocks[b0 + g] = o;
    }
  }
  __shared__ unsigned long long red[2][4];
  if (lane == 0) {
    red[0][wave] = si;
    red[1][wave] = sp;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd((unsigned long long *)&J.sums[0], red[0][0] + red[0][1] + red[0][2] + red[0][3]);
    atomicAdd((unsigned long long *)&J.sums[1], red[1][0] + red[1][1] + red[1][2] + red[1][3]);
  }
}

// One thread per sum: launched with 2 * LA_CHUNK threads.
__global__ void __launch_bounds__(2 * LA_CHUNK) complexity_clear_kernel(const CplxArgs A, int n) {
  const int t = (int)threadIdx.x;
  if (t < 2 * n) A.j[t >> 1].sums[t & 1] = 0;
}

// ------------------------------------------------------------ mbtree costs

// mbtree.rs:51-102 search: SAD + |mx| + |my|; (0, 0), then the seed, then
// up to 16 rounds of the pattern.  Group 0 prices (0, 0) and group 1 the
// seed in one pass.
__device__ __forceinline__ void tree_search(const uint8_t *ref, int w, int h, int x0, int y0, uint32_t s0, uint32_t s1,
                                            int sx, int sy, int &mx, int &my) {
  const int lane = (int)__lane_id(), g = lane >> 3;
  const bool seed_ok = x0 + sx >= 0 && y0 + sy >= 0 && x0 + sx <= w - 8 && y0 + sy <= h - 8;
  const bool use_seed = g == 1 && seed_ok;
  const uint32_t c0 = slot_sad(ref, w, x0 + (use_seed ? sx : 0), y0 + (use_seed ? sy : 0), s0, s1);
  uint32_t bc = rdlane(c0, 0);
  mx = 0;
  my = 0;
  if (seed_ok) {
    const uint32_t cs = rdlane(c0, 8) + (uint32_t)(abs(sx) + abs(sy));
    if (cs < bc) {
      bc = cs;
      mx = sx;
      my = sy;
    }
  }
  const int gdx = nib(TREE_DX, g), gdy = nib(TREE_DY, g);
  for (int it = 0; it < 16; it++) {
    const int cx = mx + gdx, cy = my + gdy;
    const bool ok = x0 + cx >= 0 && y0 + cy >= 0 && x0 + cx <= w - 8 && y0 + cy <= h - 8;
    const uint32_t c = slot_sad(ref, w, x0 + (ok ? cx : mx), y0 + (ok ? cy : my), s0, s1) + (uint32_t)(abs(cx) + abs(cy));
    const uint32_t key = min_key(c, ok, g);
    if (key == 0xFFFFFFFFu || (key >> 3) >= bc) break;
    bc = key >> 3;
    mx += nib(TREE_DX, (int)(key & 7u));
    my += nib(TREE_DY, (int)(key & 7u));
  }
}

// One wave = one block row, in batches of 8 blocks: one intra pass, the
// searches block by block (each seeded with its left neighbour's MVs),
// one SATD pass pricing list 0, list 1 and bi for the 8 blocks.
__global__ void __launch_bounds__(256) tree_kernel(const TreeArgs A) {
  const int gx = (int)gridDim.x;
  const int wg = xcd_remap((int)(blockIdx.x + blockIdx.y * gridDim.x), gx * (int)gridDim.y);
  const pcme_tree_job &J = A.j[wg / gx];
  const int w = J.w, h = J.h, bw = w >> 3, bh = h >> 3;
  const int by = (wg % gx) * 4 + (int)(threadIdx.x >> 6);
  if (by >= bh) return;
  const int lane = (int)__lane_id(), g = lane >> 3;
  const uint8_t *cur = J.cur, *r0 = J.ref0, *r1 = J.ref1;
  const int kind = (lane >> 1) & 3, strip = lane & 1;
  int l0x = 0, l0y = 0, l1x = 0, l1y = 0;  // left neighbour's MVs (mbtree.rs:124, 153)
  const int y0 = 8 * by;
  for (int bx0 = 0; bx0 < bw; bx0 += 8) {
    const int nblk = min(8, bw - bx0);
    const int lx0 = 8 * (bx0 + (g < nblk ? g : 0));
    uint32_t S[8];
    const uint32_t intra = intra8(cur, w, lx0, y0, S);
    int m0x = 0, m0y = 0, m1x = 0, m1y = 0;  // the lane's block's MVs
    for (int b = 0; b < nblk; b++) {
      const int x0 = 8 * (bx0 + b);
      uint32_t s0, s1;
      load_row(cur, w, x0, y0, s0, s1);
      int ax = 0, ay = 0, cx = 0, cy = 0;
      if (r0) tree_search(r0, w, h, x0, y0, s0, s1, l0x, l0y, ax, ay);
      if (r1) tree_search(r1, w, h, x0, y0, s0, s1, l1x, l1y, cx, cy);
      l0x = ax;
      l0y = ay;
      l1x = cx;
      l1y = cy;
      const bool mine = g == b;
      m0x = mine ? ax : m0x;
      m0y = mine ? ay : m0y;
      m1x = mine ? cx : m1x;
      m1y = mine ? cy : m1y;
    }
    // kinds: 0 = list 0, 1 = list 1, 2 = bi, 3 = unused (a missing list
    // reads cur at the block, its cost is never used)
    const uint8_t *ra = (kind == 1) ? (r1 ? r1 : cur) : (r0 ? r0 : cur);
    const int ax = (kind == 1) ? m1x : m0x, ay = (kind == 1) ? m1y : m0y;
    uint32_t P[8];
#pragma unroll
    for (int r = 0; r < 4; r++) row8(ra, w, lx0 + ax, y0 + ay + 4 * strip + r, P[2 * r], P[2 * r + 1]);
    if (kind == 2) {
      const uint8_t *rb = r1 ? r1 : cur;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        uint32_t q0, q1;
        row8(rb, w, lx0 + m1x, y0 + m1y + 4 * strip + r, q0, q1);
        P[2 * r] = avg_round(P[2 * r], q0);
        P[2 * r + 1] = avg_round(P[2 * r + 1], q1);
      }
    }
    const uint32_t s = pair_sum(strip_satd(S, P));
    for (int b = 0; b < nblk; b++) {
      uint32_t best = rdlane(intra, 8 * b), list = 0;
      const uint32_t bm0 = rdlane(((uint32_t)m0x & 0xFFFFu) | ((uint32_t)m0y << 16), 8 * b);
      const uint32_t bm1 = rdlane(((uint32_t)m1x & 0xFFFFu) | ((uint32_t)m1y << 16), 8 * b);
      uint32_t o0 = 0, o1 = 0;
      if (r0) {
        const uint32_t c = rdlane(s, 8 * b) + 8u;
        if (c < best) {
          best = c;
          list = 1;
          o0 = bm0;
        }
      }
      if (r1) {
        const uint32_t c = rdlane(s, 8 * b + 2) + 8u;
        if (c < best) {
          best = c;
          list = 2;
          o0 = bm0;
          o1 = bm1;
        }
      }
      if (r0 && r1) {
        const uint32_t c = rdlane(s, 8 * b + 4) + 12u;
        if (c < best) {
          best = c;
          list = 3;
          o0 = bm0;
          o1 = bm1;
        }
      }
      if (lane == 0) {
        uint32_t *o = (uint32_t *)&J.blocks[by * bw + bx0 + b];
        o[0] = rdlane(intra, 8 * b);
        o[1] = best;
        o[2] = o0;
        o[3] = o1;
        o[4] = list;
      }
    }
  }
}


// ----------------------------------------------------------- AQ offsets

__device__ __forceinline__ uint32_t bsum4(uint32_t w) {
  return (w & 0xFFu) + ((w >> 8) & 0xFFu) + ((w >> 16) & 0xFFu) + (w >> 24);
}
__device__ __forceinline__ uint32_t bsq4(uint32_t w) {
  const uint32_t a = w & 0xFFu, b = (w >> 8) & 0xFFu, c = (w >> 16) & 0xFFu, d = w >> 24;
  return a * a + b * b + c * c + d * d;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += (uint32_t)__shfl_xor((int)v, o);
  return v;
}

// lookahead.rs:116-130: per MB, log2 of the luma and chroma block
// variances (dsp.rs:104-114 block_var).  One wave per MB: lane l holds luma
// dword (row l >> 2, bytes 4*(l & 3)..), lanes 0-15 a u dword and 16-31 a v
// dword (row (l & 15) >> 1, bytes 4*(l & 1)..).  Every f32 operation rounds
// on its own (no contraction), as in the reference.
__global__ void __launch_bounds__(256) aq_var_kernel(const uint8_t *__restrict__ y, int ys,
                                                     const uint8_t *__restrict__ u, const uint8_t *__restrict__ v,
                                                     int cs, int mbw, int mbh, float *__restrict__ e) {
#pragma clang fp contract(off)
  const int mb = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (mb >= mbw * mbh) return;
  const int lane = (int)__lane_id(), mx = mb % mbw, my = mb / mbw;
  const uint32_t wy = ld32(y + (size_t)(16 * my + (lane >> 2)) * ys + 16 * mx + 4 * (lane & 3));
  const int cl = lane & 15;
  const uint8_t *cp = lane < 16 ? u : v;
  const uint32_t wc = lane < 32 ? ld32(cp + (size_t)(8 * my + (cl >> 1)) * cs + 8 * mx + 4 * (cl & 1)) : 0u;
  const uint32_t sy = wave_sum(bsum4(wy)), qy = wave_sum(bsq4(wy));
  const uint32_t su = wave_sum(lane < 16 ? bsum4(wc) : 0u), qu = wave_sum(lane < 16 ? bsq4(wc) : 0u);
  const uint32_t sv = wave_sum(lane >= 16 ? bsum4(wc) : 0u), qv = wave_sum(lane >= 16 ? bsq4(wc) : 0u);
  if (lane == 0) {
    const uint32_t var_y = qy - (uint32_t)(((uint64_t)sy * sy) / 256u);
    const uint32_t var_u = qu - (uint32_t)(((uint64_t)su * su) / 64u);
    const uint32_t var_v = qv - (uint32_t)(((uint64_t)sv * sv) / 64u);
    const float vy = (float)var_y / 256.0f, vu = (float)var_u / 64.0f, vv = (float)var_v / 64.0f;
    e[mb] = log2f(vy + 0.5f * (vu + vv) + 1.0f);
  }
}

// lookahead.rs:134-138: the mean of e summed in MB order (one wave), then
// clamp(strength * (e - mean), -8, 8) in place.
__global__ void __launch_bounds__(64) aq_mean_kernel(float *__restrict__ e, int n, float strength) {
#pragma clang fp contract(off)
  const int lane = (int)__lane_id();
  float sum = 0.0f;
  for (int b = 0; b < n; b += 64) {
    const float x = b + lane < n ? e[b + lane] : 0.0f;
    const int m = min(64, n - b);
    for (int j = 0; j < m; j++) sum += __builtin_bit_cast(float, rdlane(__builtin_bit_cast(uint32_t, x), j));
  }
  const float mean = sum / (float)n;
  for (int i = lane; i < n; i += 64) {
    const float o = strength * (e[i] - mean);
    e[i] = o < -8.0f ? -8.0f : (o > 8.0f ? 8.0f : o);
  }
}

// ------------------------------------------------------ macroblock tree

// mbtree.rs:167-223 feeds a node's blocks into the 8x8 cells their MVs land
// on: block (bx, by) with MV m lands at cell (bx + floor(mx/8), by +
// floor(my/8)) and its right / lower neighbours.  For each node this kernel
// (one workgroup per node) finds the range of floor(m/8) over the blocks
// that can feed anything (list != 0, intra != 0, the lists list uses):
// win = {qx_min, qx_max, qy_min, qy_max}, an empty range when none can.
__global__ void __launch_bounds__(256) tree_window_kernel(const TreeNodeArgs A, int nb, int4 *__restrict__ win) {
  __shared__ int red[4][4];
  const pcme_tree_block *b = A.b[blockIdx.x];
  int v[4] = {TREE_WIN_EMPTY, -TREE_WIN_EMPTY, TREE_WIN_EMPTY, -TREE_WIN_EMPTY};
  for (int k = (int)threadIdx.x; k < nb; k += 256) {
    const pcme_tree_block q = b[k];
    if (q.list == 0 || q.intra == 0) continue;
#pragma unroll
    for (int l = 0; l < 2; l++) {
      if (!((q.list == 3) || (q.list == 1 + l))) continue;
      const int qx = (int)q.mv[l].x >> 3, qy = (int)q.mv[l].y >> 3;
      v[0] = min(v[0], qx);
      v[1] = max(v[1], qx);
      v[2] = min(v[2], qy);
      v[3] = max(v[3], qy);
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    v[0] = min(v[0], __shfl_xor(v[0], o));
    v[1] = max(v[1], __shfl_xor(v[1], o));
    v[2] = min(v[2], __shfl_xor(v[2], o));
    v[3] = max(v[3], __shfl_xor(v[3], o));
  }
  const int w = (int)(threadIdx.x >> 6);
  if (__lane_id() == 0)
    for (int c = 0; c < 4; c++) red[w][c] = v[c];
  __syncthreads();
  if (threadIdx.x == 0) {
    int4 r = make_int4(red[0][0], red[0][1], red[0][2], red[0][3]);
    for (int j = 1; j < 4; j++) {
      r.x = min(r.x, red[j][0]);
      r.y = max(r.y, red[j][1]);
      r.z = min(r.z, red[j][2]);
      r.w = max(r.w, red[j][3]);
    }
    win[blockIdx.x] = r;
  }
}

// Zero n floats (the propagated costs start at 0).
__global__ void __launch_bounds__(256) tree_zero_kernel(float *__restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

// What one source block of mbtree.rs:169-222 hands on: the lists that feed
// the destination (bit l: list l), a = amount * w / 64 (the per-unit-area
// share) and its MVs.
struct TreeSrc {
  float a;
  uint32_t lists;
  int16_t mv[4];
};

__device__ __forceinline__ TreeSrc tree_src(const pcme_tree_block &q, float m, int mask) {
#pragma clang fp contract(off)
  TreeSrc e;
  e.a = 0.0f;
  e.lists = 0;
  e.mv[0] = q.mv[0].x;
  e.mv[1] = q.mv[0].y;
  e.mv[2] = q.mv[1].x;
  e.mv[3] = q.mv[1].y;
  if (q.list == 0 || q.intra == 0) return e;
  const uint32_t lm = (q.list == 3 ? 3u : (q.list == 1 ? 1u : 2u)) & (uint32_t)mask;
  if (!lm) return e;
  const float intra = (float)q.intra;
  const float inter0 = (float)q.inter;
  const float inter = inter0 < intra ? inter0 : intra;
  const float amount = (intra + m) * (intra - inter) / intra;
  if (amount <= 0.0f) return e;
  const float wt = q.list == 3 ? 0.5f : 1.0f;
  e.a = amount * wt / 64.0f;
  e.lists = lm;
  return e;
}

// The share source e (block (sx, sy)) adds into cell (tx, ty), in the
// reference's list order.
__device__ __forceinline__ float tree_add(float acc, const TreeSrc &e, int sx, int sy, int tx, int ty) {
#pragma clang fp contract(off)
#pragma unroll
  for (int l = 0; l < 2; l++) {
    if (!((e.lists >> l) & 1u)) continue;
    const int mx = e.mv[2 * l], my = e.mv[2 * l + 1];
    const int dx = tx - (sx + (mx >> 3)), dy = ty - (sy + (my >> 3));
    if ((unsigned)dx > 1u || (unsigned)dy > 1u) continue;
    const int fx = mx & 7, fy = my & 7;
    const int area = (dx ? fx : 8 - fx) * (dy ? fy : 8 - fy);
    if (area == 0) continue;
    acc += e.a * (float)area;
  }
  return acc;
}

// One round of mbtree.rs:169-222: job j = one node feeding one destination
// (A.j[j]: the node's blocks and own costs, the destination's costs, the
// lists that feed it, the node's window); no two jobs of a round write the
// same destination or read one another's.  Workgroup (tile, j) owns a 16x16
// tile of cells of job j's destination; thread (lx, ly) owns one cell and
// gathers what the node's blocks move into it:
// the sources the window `win` allows, in raster order, list 0 before list
// 1 -- the order the reference adds them into that cell -- each `+=`
// rounded on its own, so the cell's sum is the reference's to the bit.  The
// sources the tile's cells can see are staged in LDS first (each source's
// share computed once); a rectangle larger than TREE_LDS falls back to
// reading the blocks from memory per cell.  mine holds the node's own
// propagated costs, complete because the nodes after it ran first.
__global__ void __launch_bounds__(256) tree_prop_kernel(const TreePropArgs A, int bw, int bh) {
  __shared__ TreeSrc S[TREE_LDS];
  const TreePropJob &J = A.j[blockIdx.y];
  const pcme_tree_block *__restrict__ src = J.src;
  const float *__restrict__ mine = J.mine;
  float *dst = J.dst;
  const int mask = J.mask;
  const int4 *win = J.win;
  const int tiles_x = (bw + 15) >> 4;
  const int tx0 = ((int)blockIdx.x % tiles_x) << 4, ty0 = ((int)blockIdx.x / tiles_x) << 4;
  const int tx1 = min(tx0 + 15, bw - 1), ty1 = min(ty0 + 15, bh - 1);
  const int tx = tx0 + (int)(threadIdx.x & 15), ty = ty0 + (int)(threadIdx.x >> 4);
  const bool own = tx < bw && ty < bh;
  const int4 w = *win;
  // sources that land on cell x: bx in [x - qx_max - 1, x - qx_min]
  const int sx0 = max(0, tx0 - w.y - 1), sx1 = min(bw - 1, tx1 - w.x);
  const int sy0 = max(0, ty0 - w.w - 1), sy1 = min(bh - 1, ty1 - w.z);
  const int SX = sx1 - sx0 + 1, SY = sy1 - sy0 + 1;
  if (SX <= 0 || SY <= 0) return;  // nothing lands on this tile (uniform)
  const int x0 = max(0, tx - w.y - 1), x1 = min(bw - 1, tx - w.x);
  const int y0 = max(0, ty - w.w - 1), y1 = min(bh - 1, ty - w.z);
  if (SX * SY > TREE_LDS) {
    if (!own) return;
    float acc = dst[ty * bw + tx];
    for (int sy = y0; sy <= y1; sy++)
      for (int sx = x0; sx <= x1; sx++) {
        const int k = sy * bw + sx;
        const pcme_tree_block q = src[k];
        if (q.list == 0 || q.intra == 0) continue;
        acc = tree_add(acc, tree_src(q, mine[k], mask), sx, sy, tx, ty);
      }
    dst[ty * bw + tx] = acc;
    return;
  }
  for (int j = (int)threadIdx.x; j < SX * SY; j += 256) {
    const int k = (sy0 + j / SX) * bw + sx0 + j % SX;
    const pcme_tree_block q = src[k];
    const bool feeds = q.list != 0 && q.intra != 0;
    S[j] = tree_src(q, feeds ? mine[k] : 0.0f, mask);
  }
  __syncthreads();
  if (!own) return;
  float acc = dst[ty * bw + tx];
  for (int sy = y0; sy <= y1; sy++) {
    const TreeSrc *row = S + (sy - sy0) * SX - sx0;
    for (int sx = x0; sx <= x1; sx++) {
      const TreeSrc e = row[sx];
      if (!e.lists) continue;
      acc = tree_add(acc, e, sx, sy, tx, ty);
    }
  }
  dst[ty * bw + tx] = acc;
}

// mbtree.rs:225-231: -2 log2((max(intra, 1) + prop) / max(intra, 1)).
__global__ void __launch_bounds__(256) tree_offsets_kernel(const pcme_tree_block *__restrict__ b,
                                                           const float *__restrict__ prop, int n,
                                                           float *__restrict__ out) {
#pragma clang fp contract(off)
  const int i = (int)(blockIdx.x * 256 + threadIdx.x);
  if (i >= n) return;
  const float intra = (float)(b[i].intra > 1u ? b[i].intra : 1u);
  out[i] = -2.0f * log2f((intra + prop[i]) / intra);
}

}  // namespace pcme
