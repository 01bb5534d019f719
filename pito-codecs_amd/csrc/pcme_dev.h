// pcme_dev.h -- wave64 primitives shared by the gfx950 kernels of libpcme:
// DPP reductions over lane groups, the strict-< candidate scan as a min over
// (cost, slot) keys, and the packed-i16 SATD of a 4x8 strip (dsp.rs:35-70).
#ifndef PCME_DEV_H
#define PCME_DEV_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pcme {

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// DPP controls (gfx9 encoding).
#define DPP_QP_1032 0xB1   // quad_perm [1,0,3,2]
#define DPP_QP_2301 0x4E   // quad_perm [2,3,0,1]
#define DPP_ROW_HALF_MIRROR 0x141

// Sum over each aligned group of 8 lanes; every lane of the group receives
// the sum.  Requires a full exec mask.
__device__ __forceinline__ uint32_t red8(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_QP_1032, 0xF, 0xF, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_QP_2301, 0xF, 0xF, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_ROW_HALF_MIRROR, 0xF, 0xF, false);
  return v;
}

// Sum over each aligned group of 4 lanes (quad).
__device__ __forceinline__ uint32_t red4(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_QP_1032, 0xF, 0xF, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_QP_2301, 0xF, 0xF, false);
  return v;
}

// Sums of a paired search whose lower half runs 4 lanes per candidate and
// upper half 2 (a 16x8 or 8x16 search beside an 8x8 one); lanes 48-51, idle
// in the upper half, carry one more 4-lane candidate of the lower search.
__device__ __forceinline__ uint32_t red42(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_QP_1032, 0xF, 0xF, false);
  const uint32_t w = v + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_QP_2301, 0xF, 0xF, false);
  const uint32_t l = __lane_id();
  return (l < 32 || (l >> 2) == 12) ? w : v;  // quads on lanes 0-31 and 48-51
}

// Quarter-pel source table (picture.rs:387-404), 8 bits per position q:
// [1:0] plane A, [2] dx A, [3] dy A, [5:4] plane B, [6] dx B, [7] dy B.
#define QW0 0x41111000u
#define QW1 0x61312120u
#define QW2 0x63333222u
#define QW3 0x96939282u

__device__ __forceinline__ uint32_t qpel_entry(uint32_t q) {
  uint32_t w = (q & 8u) ? ((q & 4u) ? QW3 : QW2) : ((q & 4u) ? QW1 : QW0);
  return (w >> ((q & 3u) * 8u)) & 0xFFu;
}

__device__ __forceinline__ uint32_t rdlane(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}

// Exp-Golomb se(v) length (slice.rs:130-139).
// With k = 2v-1 (v > 0) or -2v, k+1 = 2|v| + (v <= 0): branch-free.
__device__ __forceinline__ uint32_t se_bits(int v) {
  const uint32_t a = (uint32_t)(v < 0 ? -v : v);
  const uint32_t kp = 2u * a + (v <= 0 ? 1u : 0u);
  return 2u * (31u - (uint32_t)__builtin_clz(kp)) + 1u;
}

// Rate term of a candidate (slice.rs:1615, 1686): (lam * mv_bits) >> 8.
// lam <= 21362 and the bits of two components within +-4096 stay below 64,
// so the product is a full-rate 24-bit multiply.
__device__ __forceinline__ uint32_t mv_cost(uint32_t lam, int dx, int dy) {
  return __umul24(lam, se_bits(dx) + se_bits(dy)) >> 8;
}

typedef short v2s __attribute__((ext_vector_type(2)));
__device__ __forceinline__ v2s as_v2s(uint32_t u) { return __builtin_bit_cast(v2s, u); }

// v_perm selector: [lo.byte c, 0, hi.byte c, 0] -> u16 pair (lo block, hi block)
#define SELC(c) ((uint32_t)(c) | (0x0Cu << 8) | ((4u + (uint32_t)(c)) << 16) | (0x0Cu << 24))

// SATD of one 4-row x 8-column strip = two 4x4 blocks (dsp.rs:35-70 for
// one 4x4: vertical 4-point butterfly, horizontal pair butterfly, then
// sum max(|x|,|y|)), both blocks at once in packed i16.  sp = source pairs
// [r*4+c] = (left block [r][c], right block [r][c]); P[2r], P[2r+1] = the
// prediction row r, columns 0-3 and 4-7.
__device__ __forceinline__ uint32_t satd_strip8(const v2s *sp, const uint32_t *P) {
  v2s d[4][4];
#pragma unroll
  for (int r = 0; r < 4; r++)
#pragma unroll
    for (int c = 0; c < 4; c++)
      d[r][c] = sp[r * 4 + c] - as_v2s(__builtin_amdgcn_perm(P[2 * r + 1], P[2 * r], SELC(c)));
  v2s v[4][4];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    v2s a0 = d[0][c] + d[1][c], a1 = d[0][c] - d[1][c];
    v2s a2 = d[2][c] + d[3][c], a3 = d[2][c] - d[3][c];
    v[0][c] = a0 + a2;
    v[1][c] = a0 - a2;
    v[2][c] = a1 + a3;
    v[3][c] = a1 - a3;
  }
  v2s acc = {0, 0};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    v2s hs0 = v[k][0] + v[k][1], hs1 = v[k][2] + v[k][3];
    v2s hd0 = v[k][0] - v[k][1], hd1 = v[k][2] - v[k][3];
    acc += __builtin_elementwise_max(__builtin_elementwise_abs(hs0), __builtin_elementwise_abs(hs1));
    acc += __builtin_elementwise_max(__builtin_elementwise_abs(hd0), __builtin_elementwise_abs(hd1));
  }
  return (uint32_t)(uint16_t)acc.x + (uint32_t)(uint16_t)acc.y;
}

#define DPP_ROW_MIRROR 0x140
#define DPP_ROW_BCAST15 0x142
#define DPP_ROW_BCAST31 0x143

// The reference scans a pass's candidates in slot order and takes each one
// whose cost is strictly below the running best; the result is the
// candidate of least (cost, slot) among the valid ones, taken when its cost
// is below the incumbent.  Every lane of a slot group holds the group's
// cost; key = cost*8 + slot (invalid: all ones) is min-reduced over the
// eight groups (row mirror, then the row broadcasts) and read from lane 63.
// Returns the wave-uniform key.
__device__ __forceinline__ uint32_t min_key(uint32_t cost, bool valid, int slot) {
  uint32_t k = valid ? ((cost << 3) | (uint32_t)slot) : 0xFFFFFFFFu;
  // old = all ones (the identity of min) for lanes a row mask leaves out
  uint32_t t = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)k, DPP_ROW_MIRROR, 0xF, 0xF, false);
  k = t < k ? t : k;
  t = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)k, DPP_ROW_BCAST15, 0xA, 0xF, false);
  k = t < k ? t : k;
  t = (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)k, DPP_ROW_BCAST31, 0xC, 0xF, false);
  k = t < k ? t : k;
  return rdlane(k, 63);
}

}  // namespace pcme

#endif
