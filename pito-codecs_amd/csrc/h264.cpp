// h264.cpp -- the host-side H.264 codec of libpcme_h264.so: a minimal
// decoder (include/pcme_dec.h), the Baseline-profile subset that shows a
// bitstream decodes bit-exact, and the field coder (include/pcme_enc.h) that
// codes a motion field into such a bitstream with the decoder's own
// reconstruction.  Plain scalar C++: no claim to speed.
//
// The steps follow ITU-T H.264 (clause numbers below) in the structure of
// the reference's decoder:
//   NAL / RBSP, SPS, PPS, slice header    7.3, 7.4   (dec/decoder.rs, slicedec.rs)
//   CAVLC residual blocks                 9.2        (dec/cavlc.rs)
//   macroblock layer, mb_pred, sub_mb_pred 7.3.5     (dec/slicedec.rs, mbtypes.rs)
//   MV prediction, P_Skip                 8.4.1      (dec/slicedec.rs:581-681)
//   quarter-sample luma / chroma MC       8.4.2.2    (dec/inter.rs:73-260)
//   intra 4x4 / 16x16 / chroma prediction 8.3       (dec/intra.rs)
//   scaling and transforms                8.5        (dec/transform.rs)
//   deblocking                            8.7        (dec/deblock.rs)
//   reference marking, list init/modification 8.2.4, 8.2.5 (dec/picture.rs)
#include "../../include/pcme_dec.h"
#include "../../include/pcme_enc.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <deque>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pcme.h"

namespace {

thread_local char g_err[256] = "";

int fail(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return PCME_EINVAL;
}

struct Vlc4 {
  uint8_t len;
  uint16_t val;
  uint8_t t1, tc;
};
struct Vlc2 {
  uint8_t len;
  uint16_t val;
};

// ------------------------------------------------------------------ tables

// coeff_token (H.264 Table 9-5): per nC class (0-1, 2-3, 4-7, chroma DC),
// {length, value, trailing ones, total coeff}, sorted by length.
static const Vlc4 kCoeffToken[4][62] = {
  {{1,1,0,0},{2,1,1,1},{3,1,2,2},{5,3,3,3},{6,3,3,4},{6,4,1,2},{6,5,0,1},{7,4,3,5},{7,5,2,3},{8,4,3,6},{8,5,2,4},{8,6,1,3},{8,7,0,2},{9,4,3,7},{9,5,2,5},{9,6,1,4},{9,7,0,3},{10,4,3,8},{10,5,2,6},{10,6,1,5},{10,7,0,4},{11,4,3,9},{11,5,2,7},{11,6,1,6},{11,7,0,5},{13,8,0,8},{13,9,2,9},{13,10,1,8},{13,11,0,7},{13,12,3,10},{13,13,2,8},{13,14,1,7},{13,15,0,6},{14,8,3,12},{14,9,2,11},{14,10,1,10},{14,11,0,10},{14,12,3,11},{14,13,2,10},{14,14,1,9},{14,15,0,9},{15,1,1,13},{15,8,3,14},{15,9,2,13},{15,10,1,12},{15,11,0,12},{15,12,3,13},{15,13,2,12},{15,14,1,11},{15,15,0,11},{16,4,0,16},{16,5,2,16},{16,6,1,16},{16,7,0,15},{16,8,3,16},{16,9,2,15},{16,10,1,15},{16,11,0,14},{16,12,3,15},{16,13,2,14},{16,14,1,14},{16,15,0,13}},
  {{2,2,1,1},{2,3,0,0},{3,3,2,2},{4,4,3,4},{4,5,3,3},{5,6,3,5},{5,7,1,2},{6,4,3,7},{6,5,2,4},{6,6,1,4},{6,7,0,2},{6,8,3,6},{6,9,2,3},{6,10,1,3},{6,11,0,1},{7,4,3,8},{7,5,2,5},{7,6,1,5},{7,7,0,3},{8,4,0,5},{8,5,2,6},{8,6,1,6},{8,7,0,4},{9,4,3,9},{9,5,2,7},{9,6,1,7},{9,7,0,6},{11,8,3,11},{11,9,2,9},{11,10,1,9},{11,11,0,8},{11,12,3,10},{11,13,2,8},{11,14,1,8},{11,15,0,7},{12,8,0,11},{12,9,2,11},{12,10,1,11},{12,11,0,10},{12,12,3,12},{12,13,2,10},{12,14,1,10},{12,15,0,9},{13,1,3,15},{13,6,2,14},{13,7,0,14},{13,8,3,14},{13,9,2,13},{13,10,1,13},{13,11,0,13},{13,12,3,13},{13,13,2,12},{13,14,1,12},{13,15,0,12},{14,4,3,16},{14,5,2,16},{14,6,1,16},{14,7,0,16},{14,8,1,15},{14,9,0,15},{14,10,2,15},{14,11,1,14}},
  {{4,8,3,7},{4,9,3,6},{4,10,3,5},{4,11,3,4},{4,12,3,3},{4,13,2,2},{4,14,1,1},{4,15,0,0},{5,8,1,5},{5,9,2,5},{5,10,1,4},{5,11,2,4},{5,12,1,3},{5,13,3,8},{5,14,2,3},{5,15,1,2},{6,8,0,3},{6,9,2,7},{6,10,1,7},{6,11,0,2},{6,12,3,9},{6,13,2,6},{6,14,1,6},{6,15,0,1},{7,8,0,7},{7,9,0,6},{7,10,2,9},{7,11,0,5},{7,12,3,10},{7,13,2,8},{7,14,1,8},{7,15,0,4},{8,8,3,12},{8,9,2,11},{8,10,1,10},{8,11,0,9},{8,12,3,11},{8,13,2,10},{8,14,1,9},{8,15,0,8},{9,7,1,13},{9,8,0,12},{9,9,2,13},{9,10,1,12},{9,11,0,11},{9,12,3,13},{9,13,2,12},{9,14,1,11},{9,15,0,10},{10,1,0,16},{10,2,3,16},{10,3,2,16},{10,4,1,16},{10,5,0,15},{10,6,3,15},{10,7,2,15},{10,8,1,15},{10,9,0,14},{10,10,3,14},{10,11,2,14},{10,12,1,14},{10,13,0,13}},
  {{1,1,1,1},{2,1,0,0},{3,1,2,2},{6,2,0,4},{6,3,0,3},{6,4,0,2},{6,5,3,3},{6,6,1,2},{6,7,0,1},{7,0,3,4},{7,2,2,3},{7,3,1,3},{8,2,2,4},{8,3,1,4}},
};
static const int kCoeffTokenN[4] = {62,62,62,14};
// total_zeros, 4x4 blocks (Tables 9-7, 9-8): [TotalCoeff-1][total_zeros], {length, value}; length 0: no such code
static const Vlc2 kTotalZeros4[15][16] = {
  {{1,1},{3,3},{3,2},{4,3},{4,2},{5,3},{5,2},{6,3},{6,2},{7,3},{7,2},{8,3},{8,2},{9,3},{9,2},{9,1}},
  {{3,7},{3,6},{3,5},{3,4},{3,3},{4,5},{4,4},{4,3},{4,2},{5,3},{5,2},{6,3},{6,2},{6,1},{6,0},{0,0}},
  {{4,5},{3,7},{3,6},{3,5},{4,4},{4,3},{3,4},{3,3},{4,2},{5,3},{5,2},{6,1},{5,1},{6,0},{0,0},{0,0}},
  {{5,3},{3,7},{4,5},{4,4},{3,6},{3,5},{3,4},{4,3},{3,3},{4,2},{5,2},{5,1},{5,0},{0,0},{0,0},{0,0}},
  {{4,5},{4,4},{4,3},{3,7},{3,6},{3,5},{3,4},{3,3},{4,2},{5,1},{4,1},{5,0},{0,0},{0,0},{0,0},{0,0}},
  {{6,1},{5,1},{3,7},{3,6},{3,5},{3,4},{3,3},{3,2},{4,1},{3,1},{6,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
  {{6,1},{5,1},{3,5},{3,4},{3,3},{2,3},{3,2},{4,1},{3,1},{6,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
  {{6,1},{4,1},{5,1},{3,3},{2,3},{2,2},{3,2},{3,1},{6,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
  {{6,1},{6,0},{4,1},{2,3},{2,2},{3,1},{2,1},{5,1},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
  {{5,1},{5,0},{3,1},{2,3},{2,2},{2,1},{4,1},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
  {{4,0},{4,1},{3,1},{3,2},{1,1},{3,3},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
  {{4,0},{4,1},{2,1},{1,1},{3,1},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
  {{3,0},{3,1},{1,1},{2,1},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
  {{2,0},{2,1},{1,1},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
  {{1,0},{1,1},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
};
// total_zeros, chroma DC 2x2 (Table 9-9a): [TotalCoeff-1][total_zeros], {length, value}; length 0: no such code
static const Vlc2 kTotalZerosDc[3][4] = {
  {{1,1},{2,1},{3,1},{3,0}},
  {{1,1},{2,1},{2,0},{0,0}},
  {{1,1},{1,0},{0,0},{0,0}},
};
// run_before (Table 9-10): [min(zerosLeft, 7)-1][run_before], {length, value}; length 0: no such code
static const Vlc2 kRunBefore[7][15] = {
  {{1,1},{1,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
  {{1,1},{2,1},{2,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
  {{2,3},{2,2},{2,1},{2,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
  {{2,3},{2,2},{2,1},{3,1},{3,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
  {{2,3},{2,2},{3,3},{3,2},{3,1},{3,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
  {{2,3},{3,0},{3,1},{3,3},{3,2},{3,5},{3,4},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0},{0,0}},
  {{3,7},{3,6},{3,5},{3,4},{3,3},{3,2},{3,1},{4,1},{5,1},{6,1},{7,1},{8,1},{9,1},{10,1},{11,1}},
};

// Inverse 4x4 zigzag (frame scan): the raster position of coefficient k.
const uint8_t kZigzag4[16] = {0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15};
// coded_block_pattern for me(v) (Table 9-4, ChromaArrayType 1): {Intra_4x4, Inter}
const uint8_t kCbpIntra[48] = {47, 31, 15, 0,  23, 27, 29, 30, 7,  11, 13, 14, 39, 43, 45, 46,
                               16, 3,  5,  10, 12, 19, 21, 26, 28, 35, 37, 42, 44, 1,  2,  4,
                               8,  17, 18, 20, 24, 6,  9,  22, 25, 32, 33, 34, 36, 40, 38, 41};
const uint8_t kCbpInter[48] = {0,  16, 1,  2,  4,  8,  32, 3,  5,  10, 12, 15, 47, 7,  11, 13,
                               14, 6,  9,  31, 35, 37, 42, 44, 33, 34, 36, 40, 39, 43, 45, 46,
                               17, 18, 20, 24, 19, 21, 26, 28, 23, 27, 29, 30, 22, 25, 38, 41};
// normAdjust4x4 (8.5.9): v[m][class] for positions (even, even), (odd, odd), mixed
const int kNorm4[6][3] = {{10, 16, 13}, {11, 18, 14}, {13, 20, 16}, {14, 23, 18}, {16, 25, 20}, {18, 29, 23}};
// QPc as a function of qPI (Table 8-15)
const uint8_t kQpc[52] = {0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 16, 17,
                          18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 29, 30, 31, 32, 32, 33,
                          34, 34, 35, 35, 36, 36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39};
// Deblocking thresholds (Tables 8-16, 8-17): alpha', beta' and tC0' by indexA/B
const uint8_t kAlpha[52] = {0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,   0,   0,   0,   0,   0,   4,   4,
                            5,  6,  7,  8,  9,  10, 12, 13, 15, 17, 20,  22,  25,  28,  32,  36,  40,  45,
                            50, 56, 63, 71, 80, 90, 101, 113, 127, 144, 162, 182, 203, 226, 255, 255};
const uint8_t kBeta[52] = {0, 0, 0, 0, 0, 0, 0, 0, 0,  0,  0,  0,  0,  0,  0,  0,  2,  2,  2,  3,  3,  3,  3,  4,  4,  4,
                           6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13, 14, 14, 15, 15, 16, 16, 17, 17, 18, 18};
const uint8_t kTc0[52][3] = {
    {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0},
    {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 1},
    {0, 0, 1}, {0, 0, 1}, {0, 0, 1}, {0, 1, 1}, {0, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1},
    {1, 1, 2}, {1, 1, 2}, {1, 1, 2}, {1, 1, 2}, {1, 2, 3}, {1, 2, 3}, {2, 2, 3}, {2, 2, 4}, {2, 3, 4},
    {2, 3, 4}, {3, 3, 5}, {3, 4, 6}, {3, 4, 6}, {4, 5, 7}, {4, 5, 8}, {4, 6, 9}, {5, 7, 10}, {6, 8, 11},
    {6, 8, 13}, {7, 10, 14}, {8, 11, 16}, {9, 12, 18}, {10, 13, 20}, {11, 15, 23}, {13, 17, 25}};

// luma4x4BlkIdx -> (x, y) in 4x4 units (6.4.3)
inline int blk_x(int b) { return (b & 1) | ((b >> 1) & 2); }
inline int blk_y(int b) { return ((b >> 1) & 1) | ((b >> 2) & 2); }

inline int clip3(int lo, int hi, int v) { return v < lo ? lo : (v > hi ? hi : v); }
inline uint8_t clip1(int v) { return (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v)); }
inline int median3(int a, int b, int c) { return std::max(std::min(a, b), std::min(std::max(a, b), c)); }
inline int norm4(int m, int pos) {
  const int i = pos >> 2, j = pos & 3;
  return kNorm4[m][(i & 1) == 0 && (j & 1) == 0 ? 0 : ((i & 1) && (j & 1) ? 1 : 2)];
}

// ------------------------------------------------------------ bit reader

class Bits {
 public:
  Bits(const uint8_t *p, size_t n) : p_(p), n_(n) {
    // the rbsp_stop_one_bit: the last 1 bit of the payload
    size_t k = n;
    while (k > 0 && p[k - 1] == 0) k--;
    stop_ = 0;
    if (k > 0) {
      int b = 0;
      while (((p[k - 1] >> b) & 1) == 0) b++;
      stop_ = 8 * (k - 1) + (7 - b);
    }
  }
  uint32_t peek(int k) const {
    uint32_t v = 0;
    for (int i = 0; i < k; i++) {
      const size_t q = pos_ + i;
      const uint32_t bit = q < 8 * n_ ? (p_[q >> 3] >> (7 - (q & 7))) & 1u : 0u;
      v = (v << 1) | bit;
    }
    return v;
  }
  uint32_t u(int k) {
    const uint32_t v = peek(k);
    pos_ += k;
    return v;
  }
  bool flag() { return u(1) != 0; }
  uint32_t ue() {
    int z = 0;
    while (z < 32 && u(1) == 0) z++;
    if (z >= 32) {
      bad_ = true;
      return 0;
    }
    return ((1u << z) - 1u) + (z ? u(z) : 0u);
  }
  int32_t se() {
    const uint32_t k = ue();
    return (k & 1u) ? (int32_t)((k + 1) >> 1) : -(int32_t)(k >> 1);
  }
  // te(v) with the range cmax (7.4.2)
  uint32_t te(int cmax) { return cmax > 1 ? ue() : (u(1) ? 0u : 1u); }
  bool more_rbsp() const { return pos_ < stop_; }
  void align() { pos_ = (pos_ + 7) & ~(size_t)7; }
  bool bad() const { return bad_ || pos_ > 8 * n_; }
  size_t pos() const { return pos_; }
  void skip(size_t k) { pos_ += k; }

  template <int N>
  bool vlc(const Vlc2 (&t)[N], int &out) {
    const uint32_t w = peek(16);
    for (int i = 0; i < N; i++)
      if (t[i].len && (w >> (16 - t[i].len)) == t[i].val) {
        pos_ += t[i].len;
        out = i;
        return true;
      }
    return false;
  }

 private:
  const uint8_t *p_;
  size_t n_;
  size_t pos_ = 0, stop_ = 0;
  bool bad_ = false;
};

// ------------------------------------------------------ parameter sets

struct Sps {
  bool valid = false;
  int profile = 0, level = 0;
  int log2_max_frame_num = 4, poc_type = 0, log2_max_poc_lsb = 4;
  bool delta_pic_order_always_zero = false;
  int offset_for_non_ref_pic = 0, offset_for_top_to_bottom = 0;
  std::vector<int> offset_for_ref_frame;
  int max_num_ref_frames = 1;
  bool gaps = false;
  int mbw = 0, mbh = 0;
  int crop[4] = {0, 0, 0, 0};  // left, right, top, bottom (in 2-sample units)
};

struct Pps {
  bool valid = false;
  int sps_id = 0;
  bool bottom_field_pic_order_present = false;
  int num_ref_idx_default = 1;
  int pic_init_qp = 26;
  int chroma_qp_offset = 0;
  bool deblock_ctrl = false, constrained_intra = false, redundant_pic_cnt = false;
};

// ------------------------------------------------------------- pictures

// A picture plane with clamp-to-picture reads (8.4.2.2.1's Clip3 on coordinates).
struct PlaneRef {
  const uint8_t *p;
  int stride, w, h;
  int at(int x, int y) const { return p[(size_t)clip3(0, h - 1, y) * stride + clip3(0, w - 1, x)]; }
};

struct Mb;

struct Picture {
  int id = 0;
  int w = 0, h = 0;  // coded luma size
  std::vector<uint8_t> y, u, v;
  int frame_num = 0, frame_num_wrap = 0, poc = 0;
  bool short_ref = false, long_ref = false;
  int long_idx = -1;
  int crop[4] = {0, 0, 0, 0};  // frame cropping of its SPS (2-sample units)
  std::vector<Mb> mbs;  // the decoded macroblocks (output copies only)
  PlaneRef luma() const { return PlaneRef{y.data(), w, w, h}; }
  uint8_t C(const std::vector<uint8_t> &p, int x, int yy) const {
    return p[(size_t)clip3(0, h / 2 - 1, yy) * (w / 2) + clip3(0, w / 2 - 1, x)];
  }
};

enum MbKind : int8_t { kNone = -1, kI4x4, kI16x16, kIPCM, kPInter, kPSkip };

struct Mb {
  int8_t kind = kNone;
  int slice = -1;
  int qp = 0;       // QPY (the deblocking filter reads 0 for I_PCM)
  int qpc = 0;      // QPc of the MB (chroma_qp_index_offset of its slice)
  int8_t i4[16];    // Intra4x4PredMode per 4x4 block (raster)
  uint8_t tc[16];   // TotalCoeff per luma 4x4 block (raster)
  uint8_t tcc[2][4];  // TotalCoeff per chroma AC block (raster)
  int8_t ref[4];    // refIdxL0 per 8x8 (raster); -1 intra
  int refpic[4];    // the reference picture's id per 8x8 (-1 none)
  int16_t mv[16][2];
  int16_t mvp[16][2];  // predictor of the partition covering each 4x4 block
  int8_t part = -1, sub[4] = {-1, -1, -1, -1};
  bool intra() const { return kind == kI4x4 || kind == kI16x16 || kind == kIPCM; }
};

struct SliceInfo {
  int disable_deblock = 0, alpha_off = 0, beta_off = 0;
};

// The header fields of a picture that its reference marking needs.
struct PicHeader {
  bool idr = false;
  int nal_ref_idc = 0;
  int frame_num = 0;
  bool long_term_reference = false;
  bool adaptive_marking = false;
  std::vector<std::pair<int, int>> mmco;  // (op, value) ; op 3 carries two values packed
  std::vector<int> mmco_b;
};

struct SliceHeader {
  int first_mb = 0, type = 0, pps_id = 0, frame_num = 0, idr_pic_id = 0;
  int poc_lsb = 0, delta_poc_bottom = 0, delta_poc[2] = {0, 0};
  int num_ref_idx_active = 1;
  std::vector<std::pair<int, int>> reorder;  // (idc, value)
  int qp = 26;
  SliceInfo dbk;
};

}  // namespace

struct pcme_h264dec {
  Sps sps[32];
  Pps pps[256];
  std::vector<uint8_t> rbsp;
  // the picture in progress
  std::unique_ptr<Picture> cur;
  const Sps *csps = nullptr;
  PicHeader ph;
  std::vector<Mb> mbs;
  std::vector<SliceInfo> slices;
  int mbw = 0, mbh = 0;
  // reference pictures and output
  std::vector<std::unique_ptr<Picture>> dpb;
  std::deque<std::unique_ptr<Picture>> out;
  int next_id = 1;
  int prev_ref_frame_num = 0;
  int frame_num_offset = 0, prev_frame_num = 0, prev_frame_num_offset = 0;
  int prev_poc_msb = 0, prev_poc_lsb = 0;
  // the slice being decoded
  SliceHeader sh;
  const Pps *cpps = nullptr;
  std::vector<Picture *> list0;
};

namespace {

using Dec = pcme_h264dec;

// Remove emulation prevention bytes (7.4.1).
void unescape(const uint8_t *p, size_t n, std::vector<uint8_t> &o) {
  o.clear();
  o.reserve(n);
  int zeros = 0;
  for (size_t i = 0; i < n; i++) {
    if (zeros >= 2 && p[i] == 3) {
      zeros = 0;
      continue;
    }
    o.push_back(p[i]);
    zeros = p[i] == 0 ? zeros + 1 : 0;
  }
}

int parse_sps(Dec *d, Bits &b) {
  Sps s;
  s.profile = (int)b.u(8);
  b.u(8);  // constraint flags, reserved
  s.level = (int)b.u(8);
  const uint32_t id = b.ue();
  if (id >= 32) return fail("SPS id %u", id);
  if (s.profile == 100 || s.profile == 110 || s.profile == 122 || s.profile == 244 || s.profile == 44 ||
      s.profile == 83 || s.profile == 86 || s.profile == 118 || s.profile == 128)
    return fail("profile %d is outside the Baseline subset", s.profile);
  s.log2_max_frame_num = (int)b.ue() + 4;
  s.poc_type = (int)b.ue();
  if (s.poc_type == 0) {
    s.log2_max_poc_lsb = (int)b.ue() + 4;
  } else if (s.poc_type == 1) {
    s.delta_pic_order_always_zero = b.flag();
    s.offset_for_non_ref_pic = b.se();
    s.offset_for_top_to_bottom = b.se();
    const uint32_t n = b.ue();
    if (n > 255) return fail("num_ref_frames_in_pic_order_cnt_cycle %u", n);
    for (uint32_t i = 0; i < n; i++) s.offset_for_ref_frame.push_back(b.se());
  } else if (s.poc_type != 2) {
    return fail("pic_order_cnt_type %d", s.poc_type);
  }
  s.max_num_ref_frames = (int)b.ue();
  s.gaps = b.flag();
  s.mbw = (int)b.ue() + 1;
  s.mbh = (int)b.ue() + 1;
  if (!b.flag()) return fail("field coding (frame_mbs_only_flag 0) is outside the subset");
  b.u(1);  // direct_8x8_inference_flag
  if (b.flag())
    for (int i = 0; i < 4; i++) s.crop[i] = (int)b.ue();
  // vui_parameters: nothing in them changes the decoded samples
  if (b.bad() || s.mbw > 1024 || s.mbh > 1024) return fail("malformed SPS");
  s.valid = true;
  d->sps[id] = s;
  return PCME_OK;
}

int parse_pps(Dec *d, Bits &b) {
  Pps p;
  const uint32_t id = b.ue();
  if (id >= 256) return fail("PPS id %u", id);
  p.sps_id = (int)b.ue();
  if (p.sps_id >= 32) return fail("PPS refers to SPS %d", p.sps_id);
  if (b.flag()) return fail("CABAC is outside the subset");
  p.bottom_field_pic_order_present = b.flag();
  if (b.ue() != 0) return fail("slice groups (FMO) are outside the subset");
  p.num_ref_idx_default = (int)b.ue() + 1;
  b.ue();  // num_ref_idx_l1_default_active_minus1
  if (b.flag()) return fail("weighted prediction is outside the subset");
  b.u(2);  // weighted_bipred_idc
  p.pic_init_qp = 26 + b.se();
  b.se();  // pic_init_qs
  p.chroma_qp_offset = b.se();
  p.deblock_ctrl = b.flag();
  p.constrained_intra = b.flag();
  p.redundant_pic_cnt = b.flag();
  if (b.more_rbsp()) {
    if (b.flag()) return fail("the 8x8 transform is outside the subset");
    if (b.flag()) return fail("scaling matrices are outside the subset");
    const int second = b.se();
    if (second != p.chroma_qp_offset) return fail("a second chroma QP offset is outside the subset");
  }
  if (b.bad() || p.num_ref_idx_default > 32) return fail("malformed PPS");
  p.valid = true;
  d->pps[id] = p;
  return PCME_OK;
}

// ------------------------------------------------------ picture lifecycle

// 8.2.1, the picture order count (types 0, 1 and 2).
int picture_poc(Dec *d, const Sps &s, int nal_ref_idc, bool idr) {
  const int max_frame_num = 1 << s.log2_max_frame_num;
  const SliceHeader &h = d->sh;
  if (s.poc_type == 0) {
    const int max_lsb = 1 << s.log2_max_poc_lsb;
    int pmsb = idr ? 0 : d->prev_poc_msb, plsb = idr ? 0 : d->prev_poc_lsb;
    int msb;
    if (h.poc_lsb < plsb && plsb - h.poc_lsb >= max_lsb / 2) msb = pmsb + max_lsb;
    else if (h.poc_lsb > plsb && h.poc_lsb - plsb > max_lsb / 2) msb = pmsb - max_lsb;
    else msb = pmsb;
    if (nal_ref_idc) {
      d->prev_poc_msb = msb;
      d->prev_poc_lsb = h.poc_lsb;
    }
    return msb + h.poc_lsb;
  }
  int offset;
  if (idr) offset = 0;
  else if (d->prev_frame_num > h.frame_num) offset = d->prev_frame_num_offset + max_frame_num;
  else offset = d->prev_frame_num_offset;
  d->prev_frame_num_offset = offset;
  d->prev_frame_num = h.frame_num;
  if (s.poc_type == 2) {
    if (idr) return 0;
    return nal_ref_idc ? 2 * (offset + h.frame_num) : 2 * (offset + h.frame_num) - 1;
  }
  // type 1
  const int n = (int)s.offset_for_ref_frame.size();
  int abs_frame_num = n ? offset + h.frame_num : 0;
  if (!nal_ref_idc && abs_frame_num > 0) abs_frame_num--;
  int expected = 0;
  if (abs_frame_num > 0) {
    int delta_cycle = 0;
    for (int i = 0; i < n; i++) delta_cycle += s.offset_for_ref_frame[i];
    const int cycle = (abs_frame_num - 1) / n, in_cycle = (abs_frame_num - 1) % n;
    expected = cycle * delta_cycle;
    for (int i = 0; i <= in_cycle; i++) expected += s.offset_for_ref_frame[i];
  }
  if (!nal_ref_idc) expected += s.offset_for_non_ref_pic;
  return expected + h.delta_poc[0];
}

void start_picture(Dec *d, const Sps &s, int nal_ref_idc, bool idr) {
  d->csps = &s;
  d->mbw = s.mbw;
  d->mbh = s.mbh;
  d->cur.reset(new Picture());
  Picture &p = *d->cur;
  p.id = d->next_id++;
  p.w = 16 * s.mbw;
  p.h = 16 * s.mbh;
  p.y.assign((size_t)p.w * p.h, 0);
  p.u.assign((size_t)p.w * p.h / 4, 0);
  p.v.assign((size_t)p.w * p.h / 4, 0);
  p.frame_num = d->sh.frame_num;
  p.poc = picture_poc(d, s, nal_ref_idc, idr);
  for (int i = 0; i < 4; i++) p.crop[i] = s.crop[i];
  d->mbs.assign((size_t)s.mbw * s.mbh, Mb());
  d->slices.clear();
}

void deblock_picture(Dec *d);

// 8.2.5: reference marking of the finished picture.
int mark_references(Dec *d) {
  Picture *cur = d->cur.get();
  const Sps &s = *d->csps;
  const int max_frame_num = 1 << s.log2_max_frame_num;
  auto &dpb = d->dpb;
  auto drop_unused = [&]() {
    dpb.erase(std::remove_if(dpb.begin(), dpb.end(),
                             [](const std::unique_ptr<Picture> &p) { return !p->short_ref && !p->long_ref; }),
              dpb.end());
  };
  if (d->ph.idr) {
    for (auto &p : dpb) p->short_ref = p->long_ref = false;
    drop_unused();
    if (d->ph.long_term_reference) {
      cur->long_ref = true;
      cur->long_idx = 0;
    } else {
      cur->short_ref = true;
    }
  } else {
    for (auto &p : dpb) p->frame_num_wrap = p->frame_num > cur->frame_num ? p->frame_num - max_frame_num : p->frame_num;
    if (d->ph.adaptive_marking) {
      for (size_t k = 0; k < d->ph.mmco.size(); k++) {
        const int op = d->ph.mmco[k].first, a = d->ph.mmco[k].second, bb = d->ph.mmco_b[k];
        if (op == 1) {
          const int pic_num = cur->frame_num - (a + 1);
          for (auto &p : dpb)
            if (p->short_ref && p->frame_num_wrap == pic_num) p->short_ref = false;
        } else if (op == 2) {
          for (auto &p : dpb)
            if (p->long_ref && p->long_idx == a) p->long_ref = false;
        } else if (op == 3) {
          const int pic_num = cur->frame_num - (a + 1);
          for (auto &p : dpb)
            if (p->long_ref && p->long_idx == bb) p->long_ref = false;
          for (auto &p : dpb)
            if (p->short_ref && p->frame_num_wrap == pic_num) {
              p->short_ref = false;
              p->long_ref = true;
              p->long_idx = bb;
            }
        } else if (op == 4) {
          for (auto &p : dpb)
            if (p->long_ref && p->long_idx > a - 1) p->long_ref = false;
        } else if (op == 5) {
          for (auto &p : dpb) p->short_ref = p->long_ref = false;
        } else if (op == 6) {
          for (auto &p : dpb)
            if (p->long_ref && p->long_idx == a) p->long_ref = false;
          cur->long_ref = true;
          cur->long_idx = a;
        }
      }
    } else {
      // sliding window (8.2.5.3)
      int nshort = 0, nlong = 0;
      for (auto &p : dpb) {
        nshort += p->short_ref;
        nlong += p->long_ref;
      }
      if (nshort + nlong >= std::max(s.max_num_ref_frames, 1) && nshort > 0) {
        Picture *old = nullptr;
        for (auto &p : dpb)
          if (p->short_ref && (!old || p->frame_num_wrap < old->frame_num_wrap)) old = p.get();
        old->short_ref = false;
      }
    }
    drop_unused();
    if (!cur->long_ref) cur->short_ref = true;
  }
  d->prev_ref_frame_num = cur->frame_num;
  return PCME_OK;
}

// Finish the picture in progress: deblock, queue for output, and keep it
// as a reference when it is one.
int finish_picture(Dec *d) {
  if (!d->cur) return PCME_OK;
  for (const Mb &m : d->mbs)
    if (m.kind == kNone) return fail("picture %d: a macroblock was never decoded", d->cur->frame_num);
  deblock_picture(d);
  std::unique_ptr<Picture> pic = std::move(d->cur);
  std::unique_ptr<Picture> copy(new Picture(*pic));
  copy->short_ref = copy->long_ref = false;
  copy->mbs = d->mbs;
  d->out.push_back(std::move(copy));
  if (d->ph.nal_ref_idc) {
    d->cur = std::move(pic);
    int rc = mark_references(d);
    if (rc) return rc;
    d->dpb.push_back(std::move(d->cur));
  }
  d->cur.reset();
  return PCME_OK;
}

// 8.2.4: RefPicList0 of a P slice, then its modification.
int build_list0(Dec *d) {
  const Sps &s = *d->csps;
  const int max_frame_num = 1 << s.log2_max_frame_num;
  const int cur_fn = d->sh.frame_num;
  std::vector<Picture *> st, lt;
  for (auto &p : d->dpb) {
    if (p->short_ref) {
      p->frame_num_wrap = p->frame_num > cur_fn ? p->frame_num - max_frame_num : p->frame_num;
      st.push_back(p.get());
    }
    if (p->long_ref) lt.push_back(p.get());
  }
  std::sort(st.begin(), st.end(), [](Picture *a, Picture *b) { return a->frame_num_wrap > b->frame_num_wrap; });
  std::sort(lt.begin(), lt.end(), [](Picture *a, Picture *b) { return a->long_idx < b->long_idx; });
  std::vector<Picture *> init = st;
  init.insert(init.end(), lt.begin(), lt.end());
  const int n = d->sh.num_ref_idx_active;
  std::vector<Picture *> list(n, nullptr);
  for (int i = 0; i < n && i < (int)init.size(); i++) list[i] = init[i];
  // 8.2.4.3: modification
  int pred = cur_fn, idx = 0;
  for (auto &m : d->sh.reorder) {
    Picture *pick = nullptr;
    if (m.first == 0 || m.first == 1) {
      const int diff = m.second + 1;
      int no_wrap = m.first == 0 ? pred - diff : pred + diff;
      if (no_wrap < 0) no_wrap += max_frame_num;
      if (no_wrap >= max_frame_num) no_wrap -= max_frame_num;
      pred = no_wrap;
      const int pic_num = no_wrap > cur_fn ? no_wrap - max_frame_num : no_wrap;
      for (Picture *p : st)
        if (p->frame_num_wrap == pic_num) pick = p;
    } else if (m.first == 2) {
      for (Picture *p : lt)
        if (p->long_idx == m.second) pick = p;
    }
    if (!pick) return fail("reference list modification names a picture that is not a reference");
    if (idx >= n) return fail("too many reference list modifications");
    list.insert(list.begin() + idx, pick);
    int w = idx + 1;
    for (int r = idx + 1; r < (int)list.size(); r++)
      if (list[r] != pick) list[w++] = list[r];
    list.resize(n);
    idx++;
  }
  d->list0 = list;
  return PCME_OK;
}

}  // namespace

namespace {

// ------------------------------------------------------------- neighbours

// The MB at (mx, my) is available to the MB `cur` (6.4.8): inside the
// picture, decoded, and in the same slice.
inline const Mb *mb_at(const Dec *d, int mx, int my, const Mb &cur) {
  if (mx < 0 || my < 0 || mx >= d->mbw || my >= d->mbh) return nullptr;
  const Mb &m = d->mbs[(size_t)my * d->mbw + mx];
  if (m.kind == kNone || m.slice != cur.slice) return nullptr;
  return &m;
}

// Motion data of the 4x4 block holding luma sample (x, y) relative to the
// current MB (x in -1..16, y in -1..15) for MV prediction (6.4.11.7):
// `avail` says whether the partition exists and is decoded; ref -1 with
// zero MV for intra or unavailable ones.
struct NbMv {
  bool avail;
  int ref;
  int refpic;
  int mvx, mvy;
};

NbMv nb_mv(const Dec *d, int mx, int my, const Mb &cur, uint32_t done, int x, int y) {
  NbMv r = {false, -1, -1, 0, 0};
  const Mb *m;
  int bx, by;
  if (x >= 0 && x < 16 && y >= 0) {
    bx = x >> 2;
    by = y >> 2;
    if (!((done >> (by * 4 + bx)) & 1u)) return r;
    m = &cur;
  } else {
    if (x >= 16 && y >= 0) return r;  // right MB: not yet decoded
    const int dx = x < 0 ? -1 : (x >= 16 ? 1 : 0), dy = y < 0 ? -1 : 0;
    m = mb_at(d, mx + dx, my + dy, cur);
    if (!m) return r;
    bx = ((x + 16) & 15) >> 2;
    by = ((y + 16) & 15) >> 2;
  }
  r.avail = true;
  if (m->intra()) return r;
  const int b8 = (by >> 1) * 2 + (bx >> 1);
  r.ref = m->ref[b8];
  r.refpic = m->refpic[b8];
  r.mvx = m->mv[by * 4 + bx][0];
  r.mvy = m->mv[by * 4 + bx][1];
  return r;
}

// 8.4.1.3: the MV predictor of a partition at (x, y), w x h luma samples,
// with reference index ref; shape 1 = 16x8, 2 = 8x16 (directional rules).
void mv_pred(const Dec *d, int mx, int my, const Mb &cur, uint32_t done, int x, int y, int w, int h, int ref,
             int shape, int part, int &px, int &py) {
  (void)h;  // the neighbours depend on the width only
  NbMv A = nb_mv(d, mx, my, cur, done, x - 1, y);
  NbMv B = nb_mv(d, mx, my, cur, done, x, y - 1);
  NbMv C = nb_mv(d, mx, my, cur, done, x + w, y - 1);
  if (!C.avail) C = nb_mv(d, mx, my, cur, done, x - 1, y - 1);
  if (shape == 1) {
    if (part == 0 && B.ref == ref) { px = B.mvx; py = B.mvy; return; }
    if (part == 1 && A.ref == ref) { px = A.mvx; py = A.mvy; return; }
  } else if (shape == 2) {
    if (part == 0 && A.ref == ref) { px = A.mvx; py = A.mvy; return; }
    if (part == 1 && C.ref == ref) { px = C.mvx; py = C.mvy; return; }
  }
  if (!B.avail && !C.avail && A.avail) {
    B = A;
    C = A;
  }
  const int n = (A.ref == ref) + (B.ref == ref) + (C.ref == ref);
  if (n == 1) {
    const NbMv &s = A.ref == ref ? A : (B.ref == ref ? B : C);
    px = s.mvx;
    py = s.mvy;
    return;
  }
  px = median3(A.mvx, B.mvx, C.mvx);
  py = median3(A.mvy, B.mvy, C.mvy);
}

// 8.4.1.1: the P_Skip MV.
void pskip_mv(const Dec *d, int mx, int my, const Mb &cur, int &px, int &py) {
  const NbMv A = nb_mv(d, mx, my, cur, 0u, -1, 0), B = nb_mv(d, mx, my, cur, 0u, 0, -1);
  if (!A.avail || !B.avail || (A.ref == 0 && A.mvx == 0 && A.mvy == 0) || (B.ref == 0 && B.mvx == 0 && B.mvy == 0)) {
    px = py = 0;
    return;
  }
  mv_pred(d, mx, my, cur, 0u, 0, 0, 16, 16, 0, 0, 0, px, py);
}

// nC of a luma (comp < 0) or chroma AC block at (bx, by) in 4x4 units (9.2.1).
int coeff_nc(const Dec *d, int mx, int my, const Mb &cur, int comp, int bx, int by) {
  const int lim = comp < 0 ? 4 : 2;
  auto total = [&](const Mb *m, int x, int y) -> int {
    if (m->kind == kIPCM) return 16;
    if (m->kind == kPSkip) return 0;
    return comp < 0 ? m->tc[y * 4 + x] : m->tcc[comp][y * 2 + x];
  };
  int na = 0, nb = 0;
  bool aa = true, ab = true;
  if (bx > 0) na = total(&cur, bx - 1, by);
  else {
    const Mb *m = mb_at(d, mx - 1, my, cur);
    if (m) na = total(m, lim - 1, by);
    else aa = false;
  }
  if (by > 0) nb = total(&cur, bx, by - 1);
  else {
    const Mb *m = mb_at(d, mx, my - 1, cur);
    if (m) nb = total(m, bx, lim - 1);
    else ab = false;
  }
  if (aa && ab) return (na + nb + 1) >> 1;
  if (aa) return na;
  if (ab) return nb;
  return 0;
}

// -------------------------------------------------------------- CAVLC

// residual_block_cavlc (7.3.5.3.2 / 9.2): coefficients [start, end] of a
// block of maxn, into lv[] (lv[0] is coefficient `start`).  Returns
// TotalCoeff, or -1 on a syntax error.
int residual_block(Bits &b, int nc, int start, int end, int maxn, int *lv) {
  for (int i = 0; i < maxn; i++) lv[i] = 0;
  int tc, t1;
  if (nc >= 8) {
    const uint32_t c = b.u(6);
    if (c == 3) {
      tc = 0;
      t1 = 0;
    } else {
      tc = (int)(c >> 2) + 1;
      t1 = (int)(c & 3u);
    }
  } else {
    const int cls = nc == -1 ? 3 : (nc < 2 ? 0 : (nc < 4 ? 1 : 2));
    const uint32_t w = b.peek(16);
    int hit = -1;
    for (int i = 0; i < kCoeffTokenN[cls]; i++) {
      const Vlc4 &e = kCoeffToken[cls][i];
      if ((w >> (16 - e.len)) == e.val) {
        hit = i;
        break;
      }
    }
    if (hit < 0) return -1;
    b.skip(kCoeffToken[cls][hit].len);
    tc = kCoeffToken[cls][hit].tc;
    t1 = kCoeffToken[cls][hit].t1;
  }
  if (tc == 0) return 0;
  if (tc > end - start + 1 || t1 > tc) return -1;
  int level[16], run[16];
  int suffix_len = (tc > 10 && t1 < 3) ? 1 : 0;
  for (int i = 0; i < tc; i++) {
    if (i < t1) {
      level[i] = b.u(1) ? -1 : 1;
      continue;
    }
    int prefix = 0;
    while (prefix < 32 && b.u(1) == 0) prefix++;
    if (prefix >= 32) return -1;
    int code = (std::min(15, prefix) << suffix_len);
    if (suffix_len > 0 || prefix >= 14) {
      int size = suffix_len;
      if (prefix == 14 && suffix_len == 0) size = 4;
      if (prefix >= 15) size = prefix - 3;
      if (size > 0) code += (int)b.u(size);
    }
    if (prefix >= 15 && suffix_len == 0) code += 15;
    if (prefix >= 16) code += (1 << (prefix - 3)) - 4096;
    if (i == t1 && t1 < 3) code += 2;
    level[i] = (code & 1) == 0 ? (code + 2) >> 1 : (-code - 1) >> 1;
    if (suffix_len == 0) suffix_len = 1;
    if (std::abs(level[i]) > (3 << (suffix_len - 1)) && suffix_len < 6) suffix_len++;
  }
  int zeros = 0;
  if (tc < end - start + 1) {
    int z = -1;
    bool ok = maxn == 4 ? b.vlc(kTotalZerosDc[tc - 1], z) : b.vlc(kTotalZeros4[tc - 1], z);
    if (!ok) return -1;
    zeros = z;
  }
  for (int i = 0; i < tc - 1; i++) {
    if (zeros > 0) {
      int r = -1;
      if (!b.vlc(kRunBefore[std::min(zeros, 7) - 1], r) || r > zeros) return -1;
      run[i] = r;
      zeros -= r;
    } else {
      run[i] = 0;
    }
  }
  run[tc - 1] = zeros;
  int k = -1;
  for (int i = tc - 1; i >= 0; i--) {
    k += run[i] + 1;
    if (k >= maxn) return -1;
    lv[k] = level[i];
  }
  return tc;
}

// --------------------------------------------------------- transforms

// 8.5.12: inverse 4x4 transform of d (raster), result r = (x + 32) >> 6.
void idct4(const int *d, int *r) {
  int t[16];
  for (int i = 0; i < 4; i++) {
    const int *s = d + 4 * i;
    const int e = s[0] + s[2], f = s[0] - s[2], g = (s[1] >> 1) - s[3], h = s[1] + (s[3] >> 1);
    t[4 * i] = e + h;
    t[4 * i + 1] = f + g;
    t[4 * i + 2] = f - g;
    t[4 * i + 3] = e - h;
  }
  for (int j = 0; j < 4; j++) {
    const int e = t[j] + t[8 + j], f = t[j] - t[8 + j], g = (t[4 + j] >> 1) - t[12 + j], h = t[4 + j] + (t[12 + j] >> 1);
    r[j] = (e + h + 32) >> 6;
    r[4 + j] = (f + g + 32) >> 6;
    r[8 + j] = (f - g + 32) >> 6;
    r[12 + j] = (e - h + 32) >> 6;
  }
}

// 8.5.12.1: scale one coefficient at raster position pos (flat weights 16).
inline int scale4(int c, int qp, int pos) {
  const int ls = 16 * norm4(qp % 6, pos);
  if (qp >= 24) return (c * ls) << (qp / 6 - 4);
  return (c * ls + (1 << (3 - qp / 6))) >> (4 - qp / 6);
}

// Add residual r (raster 4x4) to the prediction already in the plane.
inline void add4(uint8_t *p, int stride, const int *r) {
  for (int y = 0; y < 4; y++)
    for (int x = 0; x < 4; x++) p[y * stride + x] = clip1(p[y * stride + x] + r[y * 4 + x]);
}

// ------------------------------------------------------ intra prediction

struct Edge {  // neighbouring samples of a block: top[-1..15] (top[-1] = corner), left[0..15]
  int t[17], l[16];
  bool ht, hl, htl, htr;
  int top(int i) const { return t[i + 1]; }
};

// Intra sample availability (8.3.1.2 with constrained_intra_pred).
inline bool intra_ok(const Dec *d, const Mb *m) {
  return m && !(d->cpps->constrained_intra && !m->intra());
}

// 8.3.1.2: Intra_4x4 prediction of block (bx, by) of the MB at (mx, my).
void pred4x4(const Dec *d, int mx, int my, const Mb &cur, int bx, int by, uint32_t done, int mode, uint8_t *dst,
             int stride) {
  const Picture &P = *d->cur;
  const int x0 = 16 * mx + 4 * bx, y0 = 16 * my + 4 * by;
  const Mb *L = bx > 0 ? &cur : mb_at(d, mx - 1, my, cur);
  const Mb *T = by > 0 ? &cur : mb_at(d, mx, my - 1, cur);
  const Mb *TL = (bx > 0 && by > 0) ? &cur : (bx > 0 ? mb_at(d, mx, my - 1, cur) : (by > 0 ? mb_at(d, mx - 1, my, cur) : mb_at(d, mx - 1, my - 1, cur)));
  const Mb *TR = nullptr;
  if (by == 0) TR = bx < 3 ? mb_at(d, mx, my - 1, cur) : mb_at(d, mx + 1, my - 1, cur);
  else if (bx < 3 && ((done >> ((by - 1) * 4 + bx + 1)) & 1u)) TR = &cur;
  const bool hl = intra_ok(d, L), ht = intra_ok(d, T), htl = intra_ok(d, TL), htr = intra_ok(d, TR);
  int top[9], left[4];  // top[0] = corner, top[1..8] = p[0..7, -1]
  for (int i = 0; i < 4; i++) {
    left[i] = hl ? P.y[(size_t)(y0 + i) * P.w + x0 - 1] : 0;
    top[1 + i] = ht ? P.y[(size_t)(y0 - 1) * P.w + x0 + i] : 0;
  }
  for (int i = 4; i < 8; i++) top[1 + i] = htr ? P.y[(size_t)(y0 - 1) * P.w + x0 + i] : top[4];
  top[0] = htl ? P.y[(size_t)(y0 - 1) * P.w + x0 - 1] : 0;
  auto T_ = [&](int i) { return top[1 + i]; };   // p[i, -1], i = -1..7
  auto L_ = [&](int i) { return i < 0 ? top[0] : left[i]; };  // p[-1, i]
  int pr[16];
  for (int y = 0; y < 4; y++)
    for (int x = 0; x < 4; x++) {
      int v = 128;
      switch (mode) {
        case 0: v = T_(x); break;
        case 1: v = L_(y); break;
        case 2: {
          if (ht && hl) v = (T_(0) + T_(1) + T_(2) + T_(3) + L_(0) + L_(1) + L_(2) + L_(3) + 4) >> 3;
          else if (hl) v = (L_(0) + L_(1) + L_(2) + L_(3) + 2) >> 2;
          else if (ht) v = (T_(0) + T_(1) + T_(2) + T_(3) + 2) >> 2;
          else v = 128;
          break;
        }
        case 3:
          v = (x == 3 && y == 3) ? (T_(6) + 3 * T_(7) + 2) >> 2 : (T_(x + y) + 2 * T_(x + y + 1) + T_(x + y + 2) + 2) >> 2;
          break;
        case 4:
          if (x > y) v = (T_(x - y - 2) + 2 * T_(x - y - 1) + T_(x - y) + 2) >> 2;
          else if (x < y) v = (L_(y - x - 2) + 2 * L_(y - x - 1) + L_(y - x) + 2) >> 2;
          else v = (T_(0) + 2 * L_(-1) + L_(0) + 2) >> 2;
          break;
        case 5: {
          const int z = 2 * x - y;
          if (z >= 0 && (z & 1) == 0) v = (T_(x - (y >> 1) - 1) + T_(x - (y >> 1)) + 1) >> 1;
          else if (z >= 0) v = (T_(x - (y >> 1) - 2) + 2 * T_(x - (y >> 1) - 1) + T_(x - (y >> 1)) + 2) >> 2;
          else if (z == -1) v = (L_(0) + 2 * L_(-1) + T_(0) + 2) >> 2;
          else v = (L_(y - 1) + 2 * L_(y - 2) + L_(y - 3) + 2) >> 2;
          break;
        }
        case 6: {
          const int z = 2 * y - x;
          if (z >= 0 && (z & 1) == 0) v = (L_(y - (x >> 1) - 1) + L_(y - (x >> 1)) + 1) >> 1;
          else if (z >= 0) v = (L_(y - (x >> 1) - 2) + 2 * L_(y - (x >> 1) - 1) + L_(y - (x >> 1)) + 2) >> 2;
          else if (z == -1) v = (L_(0) + 2 * L_(-1) + T_(0) + 2) >> 2;
          else v = (T_(x - 1) + 2 * T_(x - 2) + T_(x - 3) + 2) >> 2;
          break;
        }
        case 7:
          if ((y & 1) == 0) v = (T_(x + (y >> 1)) + T_(x + (y >> 1) + 1) + 1) >> 1;
          else v = (T_(x + (y >> 1)) + 2 * T_(x + (y >> 1) + 1) + T_(x + (y >> 1) + 2) + 2) >> 2;
          break;
        case 8: {
          const int z = x + 2 * y;
          if (z > 5) v = L_(3);
          else if (z == 5) v = (L_(2) + 3 * L_(3) + 2) >> 2;
          else if ((z & 1) == 0) v = (L_(y + (x >> 1)) + L_(y + (x >> 1) + 1) + 1) >> 1;
          else v = (L_(y + (x >> 1)) + 2 * L_(y + (x >> 1) + 1) + L_(y + (x >> 1) + 2) + 2) >> 2;
          break;
        }
      }
      pr[y * 4 + x] = v;
    }
  for (int y = 0; y < 4; y++)
    for (int x = 0; x < 4; x++) dst[y * stride + x] = (uint8_t)pr[y * 4 + x];
}

// 8.3.3: Intra_16x16 prediction into the picture.
int pred16x16(Dec *d, int mx, int my, const Mb &cur, int mode) {
  Picture &P = *d->cur;
  const int x0 = 16 * mx, y0 = 16 * my;
  const bool hl = intra_ok(d, mb_at(d, mx - 1, my, cur)), ht = intra_ok(d, mb_at(d, mx, my - 1, cur));
  const bool htl = intra_ok(d, mb_at(d, mx - 1, my - 1, cur));
  auto T_ = [&](int i) -> int { return i < 0 ? P.y[(size_t)(y0 - 1) * P.w + x0 - 1] : P.y[(size_t)(y0 - 1) * P.w + x0 + i]; };
  auto L_ = [&](int i) -> int { return i < 0 ? P.y[(size_t)(y0 - 1) * P.w + x0 - 1] : P.y[(size_t)(y0 + i) * P.w + x0 - 1]; };
  if ((mode == 0 && !ht) || (mode == 1 && !hl) || (mode == 3 && !(ht && hl && htl)))
    return fail("Intra_16x16 mode %d without its neighbours", mode);
  int dc = 128;
  if (mode == 2) {
    int s = 0;
    if (ht && hl) {
      for (int i = 0; i < 16; i++) s += T_(i) + L_(i);
      dc = (s + 16) >> 5;
    } else if (hl) {
      for (int i = 0; i < 16; i++) s += L_(i);
      dc = (s + 8) >> 4;
    } else if (ht) {
      for (int i = 0; i < 16; i++) s += T_(i);
      dc = (s + 8) >> 4;
    }
  }
  int a = 0, bb = 0, c = 0;
  if (mode == 3) {
    int H = 0, V = 0;
    for (int i = 0; i < 8; i++) {
      H += (i + 1) * (T_(8 + i) - T_(6 - i));
      V += (i + 1) * (L_(8 + i) - L_(6 - i));
    }
    a = 16 * (L_(15) + T_(15));
    bb = (5 * H + 32) >> 6;
    c = (5 * V + 32) >> 6;
  }
  uint8_t pr[256];
  for (int y = 0; y < 16; y++)
    for (int x = 0; x < 16; x++) {
      int v;
      if (mode == 0) v = T_(x);
      else if (mode == 1) v = L_(y);
      else if (mode == 2) v = dc;
      else v = clip1((a + bb * (x - 7) + c * (y - 7) + 16) >> 5);
      pr[y * 16 + x] = (uint8_t)v;
    }
  for (int y = 0; y < 16; y++) memcpy(&P.y[(size_t)(y0 + y) * P.w + x0], pr + 16 * y, 16);
  return PCME_OK;
}

// 8.3.4: intra chroma prediction (4:2:0) of one component into the picture.
int pred_chroma(Dec *d, int mx, int my, const Mb &cur, int mode, std::vector<uint8_t> &pl) {
  const int w = d->cur->w / 2;
  const int x0 = 8 * mx, y0 = 8 * my;
  const bool hl = intra_ok(d, mb_at(d, mx - 1, my, cur)), ht = intra_ok(d, mb_at(d, mx, my - 1, cur));
  const bool htl = intra_ok(d, mb_at(d, mx - 1, my - 1, cur));
  auto T_ = [&](int i) -> int { return i < 0 ? pl[(size_t)(y0 - 1) * w + x0 - 1] : pl[(size_t)(y0 - 1) * w + x0 + i]; };
  auto L_ = [&](int i) -> int { return i < 0 ? pl[(size_t)(y0 - 1) * w + x0 - 1] : pl[(size_t)(y0 + i) * w + x0 - 1]; };
  if ((mode == 2 && !ht) || (mode == 1 && !hl) || (mode == 3 && !(ht && hl && htl)))
    return fail("intra chroma mode %d without its neighbours", mode);
  uint8_t pr[64];
  if (mode == 0) {
    for (int blk = 0; blk < 4; blk++) {
      const int xo = 4 * (blk & 1), yo = 4 * (blk >> 1);
      int st = 0, sl = 0;
      for (int i = 0; i < 4; i++) {
        if (ht) st += T_(xo + i);
        if (hl) sl += L_(yo + i);
      }
      int v = 128;
      const bool corner = (xo == 0 && yo == 0) || (xo > 0 && yo > 0);
      if (corner) {
        if (ht && hl) v = (st + sl + 4) >> 3;
        else if (hl) v = (sl + 2) >> 2;
        else if (ht) v = (st + 2) >> 2;
      } else if (xo > 0) {  // top-right block: the top row first
        if (ht) v = (st + 2) >> 2;
        else if (hl) v = (sl + 2) >> 2;
      } else {  // bottom-left block: the left column first
        if (hl) v = (sl + 2) >> 2;
        else if (ht) v = (st + 2) >> 2;
      }
      for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) pr[(yo + y) * 8 + xo + x] = (uint8_t)v;
    }
  } else {
    int a = 0, bb = 0, c = 0;
    if (mode == 3) {
      int H = 0, V = 0;
      for (int i = 0; i < 4; i++) {
        H += (i + 1) * (T_(4 + i) - T_(2 - i));
        V += (i + 1) * (L_(4 + i) - L_(2 - i));
      }
      a = 16 * (L_(7) + T_(7));
      bb = (34 * H + 32) >> 6;
      c = (34 * V + 32) >> 6;
    }
    for (int y = 0; y < 8; y++)
      for (int x = 0; x < 8; x++) {
        int v;
        if (mode == 1) v = L_(y);
        else if (mode == 2) v = T_(x);
        else v = clip1((a + bb * (x - 3) + c * (y - 3) + 16) >> 5);
        pr[y * 8 + x] = (uint8_t)v;
      }
  }
  for (int y = 0; y < 8; y++) memcpy(&pl[(size_t)(y0 + y) * w + x0], pr + 8 * y, 8);
  return PCME_OK;
}

// ------------------------------------------------------ inter prediction

// 8.4.2.2.1: the luma sample at quarter position (4 x + fx, 4 y + fy).
int luma_qpel(const PlaneRef &R, int x, int y, int fx, int fy) {
  auto G = [&](int xx, int yy) { return R.at(xx, yy); };
  auto tap = [](int a, int b, int c, int dd, int e, int f) { return a - 5 * b + 20 * c + 20 * dd - 5 * e + f; };
  auto b1 = [&](int xx, int yy) { return tap(G(xx - 2, yy), G(xx - 1, yy), G(xx, yy), G(xx + 1, yy), G(xx + 2, yy), G(xx + 3, yy)); };
  auto h1 = [&](int xx, int yy) { return tap(G(xx, yy - 2), G(xx, yy - 1), G(xx, yy), G(xx, yy + 1), G(xx, yy + 2), G(xx, yy + 3)); };
  auto bh = [&](int xx, int yy) { return (int)clip1((b1(xx, yy) + 16) >> 5); };
  auto hh = [&](int xx, int yy) { return (int)clip1((h1(xx, yy) + 16) >> 5); };
  auto jj = [&](int xx, int yy) {
    const int j1 = tap(b1(xx, yy - 2), b1(xx, yy - 1), b1(xx, yy), b1(xx, yy + 1), b1(xx, yy + 2), b1(xx, yy + 3));
    return (int)clip1((j1 + 512) >> 10);
  };
  const int g = G(x, y);
  switch (fy * 4 + fx) {
    case 0: return g;
    case 1: return (g + bh(x, y) + 1) >> 1;                 // a
    case 2: return bh(x, y);                                 // b
    case 3: return (bh(x, y) + G(x + 1, y) + 1) >> 1;        // c
    case 4: return (g + hh(x, y) + 1) >> 1;                  // d
    case 5: return (bh(x, y) + hh(x, y) + 1) >> 1;           // e
    case 6: return (bh(x, y) + jj(x, y) + 1) >> 1;           // f
    case 7: return (bh(x, y) + hh(x + 1, y) + 1) >> 1;       // g
    case 8: return hh(x, y);                                 // h
    case 9: return (hh(x, y) + jj(x, y) + 1) >> 1;           // i
    case 10: return jj(x, y);                                // j
    case 11: return (jj(x, y) + hh(x + 1, y) + 1) >> 1;      // k
    case 12: return (hh(x, y) + G(x, y + 1) + 1) >> 1;       // n
    case 13: return (hh(x, y) + bh(x, y + 1) + 1) >> 1;      // p
    case 14: return (jj(x, y) + bh(x, y + 1) + 1) >> 1;      // q
    default: return (hh(x + 1, y) + bh(x, y + 1) + 1) >> 1;  // r
  }
}

// MC of one 4x4 luma block and its 2x2 chroma blocks (8.4.2.2).
void mc_block(Dec *d, const Picture &R, int mx, int my, int bx, int by, int mvx, int mvy) {
  Picture &P = *d->cur;
  const int xl = 16 * mx + 4 * bx, yl = 16 * my + 4 * by;
  const PlaneRef L = R.luma();
  for (int y = 0; y < 4; y++)
    for (int x = 0; x < 4; x++) {
      const int qx = 4 * (xl + x) + mvx, qy = 4 * (yl + y) + mvy;
      P.y[(size_t)(yl + y) * P.w + xl + x] = (uint8_t)luma_qpel(L, qx >> 2, qy >> 2, qx & 3, qy & 3);
    }
  const int cw = P.w / 2, xc = xl / 2, yc = yl / 2;
  for (int c = 0; c < 2; c++) {
    const std::vector<uint8_t> &src = c ? R.v : R.u;
    std::vector<uint8_t> &dst = c ? P.v : P.u;
    for (int y = 0; y < 2; y++)
      for (int x = 0; x < 2; x++) {
        const int ex = 8 * (xc + x) + mvx, ey = 8 * (yc + y) + mvy;
        const int xi = ex >> 3, yi = ey >> 3, fx = ex & 7, fy = ey & 7;
        const int A = R.C(src, xi, yi), B = R.C(src, xi + 1, yi), C = R.C(src, xi, yi + 1), D = R.C(src, xi + 1, yi + 1);
        dst[(size_t)(yc + y) * cw + xc + x] =
            (uint8_t)(((8 - fx) * (8 - fy) * A + fx * (8 - fy) * B + (8 - fx) * fy * C + fx * fy * D + 32) >> 6);
      }
  }
}

}  // namespace

namespace {

// ------------------------------------------------------ macroblock layer

// Dequantise and add the luma residual of block (bx, by) (raster coefficients lv, with dc given or not).
void luma_residual(Dec *d, int mx, int my, int bx, int by, const int *zz, int qp, bool has_dc, int dc) {
  int c[16], r[16];
  for (int k = 0; k < 16; k++) c[k] = 0;
  for (int k = 0; k < 16; k++) {
    const int pos = kZigzag4[k];
    c[pos] = zz[k] ? scale4(zz[k], qp, pos) : 0;
  }
  if (has_dc) c[0] = dc;
  idct4(c, r);
  Picture &P = *d->cur;
  add4(&P.y[(size_t)(16 * my + 4 * by) * P.w + 16 * mx + 4 * bx], P.w, r);
}

// An inter MB's prediction (all 16 4x4 blocks, luma and chroma) plus its
// luma residual (levels in zigzag order per block, blocks of the 8x8s whose
// cbp bit is set).
void recon_inter(Dec *d, int mx, int my, const Mb &m, const int (*acl)[16], int cbp_l, int qp) {
  for (int blk = 0; blk < 16; blk++) {
    const int bx = blk & 3, by = blk >> 2, b8 = (by >> 1) * 2 + (bx >> 1);
    mc_block(d, *d->list0[m.ref[b8]], mx, my, bx, by, m.mv[blk][0], m.mv[blk][1]);
  }
  for (int blk = 0; blk < 16; blk++) {
    const int bx = blk_x(blk), by = blk_y(blk);
    if ((cbp_l >> (blk >> 2)) & 1) luma_residual(d, mx, my, bx, by, acl[blk], qp, false, 0);
  }
}

// 8.5.11: the chroma residual of one component (DC levels dc[4], AC levels
// ac[4][1..15] in zigzag order) added to the prediction in pl.
void chroma_residual(Dec *d, int mx, int my, int qpc, const int *dc, const int (*ac)[16], std::vector<uint8_t> &pl) {
  const int c0 = dc[0], c1 = dc[1], c2 = dc[2], c3 = dc[3];
  const int f[4] = {c0 + c1 + c2 + c3, c0 - c1 + c2 - c3, c0 + c1 - c2 - c3, c0 - c1 - c2 + c3};
  const int ls = 16 * kNorm4[qpc % 6][0];
  const int cw = d->cur->w / 2;
  for (int i = 0; i < 4; i++) {
    int cc[16], r[16];
    for (int k = 0; k < 16; k++) cc[k] = 0;
    for (int k = 1; k < 16; k++) {
      const int pos = kZigzag4[k];
      cc[pos] = ac[i][k] ? scale4(ac[i][k], qpc, pos) : 0;
    }
    cc[0] = ((f[i] * ls) << (qpc / 6)) >> 5;
    idct4(cc, r);
    add4(&pl[(size_t)(8 * my + 4 * (i >> 1)) * cw + 8 * mx + 4 * (i & 1)], cw, r);
  }
}

int decode_mb(Dec *d, Bits &b, int addr, int slice_type, int &qp) {
  const int mx = addr % d->mbw, my = addr / d->mbw;
  Mb &m = d->mbs[addr];
  m.slice = (int)d->slices.size() - 1;
  for (int i = 0; i < 16; i++) {
    m.i4[i] = 2;
    m.tc[i] = 0;
    m.mv[i][0] = m.mv[i][1] = 0;
    m.mvp[i][0] = m.mvp[i][1] = 0;
  }
  m.part = -1;
  for (int i = 0; i < 4; i++) {
    m.sub[i] = -1;
    m.ref[i] = -1;
    m.refpic[i] = -1;
    m.tcc[0][i] = m.tcc[1][i] = 0;
  }
  const int cqo = d->cpps->chroma_qp_offset;
  Picture &P = *d->cur;
  const bool pslice = slice_type == 0;
  uint32_t mb_type = b.ue();
  if (pslice) {
    if (mb_type >= 5) mb_type -= 5;
    else mb_type += 100;  // inter: 100..104
  }
  if (mb_type > 25 && mb_type < 100) return fail("mb_type %u", mb_type);
  // ---- I_PCM
  if (mb_type == 25) {
    m.kind = kIPCM;
    b.align();
    for (int y = 0; y < 16; y++)
      for (int x = 0; x < 16; x++) P.y[(size_t)(16 * my + y) * P.w + 16 * mx + x] = (uint8_t)b.u(8);
    for (int c = 0; c < 2; c++)
      for (int y = 0; y < 8; y++)
        for (int x = 0; x < 8; x++) (c ? P.v : P.u)[(size_t)(8 * my + y) * (P.w / 2) + 8 * mx + x] = (uint8_t)b.u(8);
    for (int i = 0; i < 16; i++) m.tc[i] = 16;
    for (int i = 0; i < 4; i++) m.tcc[0][i] = m.tcc[1][i] = 16;
    m.qp = 0;  // the deblocking filter's qP for I_PCM
    m.qpc = kQpc[clip3(0, 51, 0 + cqo)];
    return b.bad() ? fail("I_PCM past the end of the slice") : PCME_OK;
  }
  const bool intra = mb_type < 100;
  const bool i16 = intra && mb_type >= 1;
  int i16_mode = 0, cbp = 0, chroma_mode = 0;
  uint32_t done = 0;
  if (intra) {
    m.kind = i16 ? kI16x16 : kI4x4;
    if (i16) {
      i16_mode = (int)(mb_type - 1) % 4;
      cbp = (((int)(mb_type - 1) / 4) % 3) << 4 | (mb_type >= 13 ? 15 : 0);
    } else {
      // prev_intra4x4_pred_mode_flag / rem_intra4x4_pred_mode in block order
      for (int blk = 0; blk < 16; blk++) {
        const bool prev = b.flag();
        const int rem = prev ? 0 : (int)b.u(3);
        const int bx = blk_x(blk), by = blk_y(blk);
        // 8.3.1.1: the predicted mode from blocks A (left) and B (above)
        auto nb_mode = [&](int x, int y, bool &dcpred) -> int {
          const Mb *n;
          int nx = x, ny = y;
          if (x < 0) { n = mb_at(d, mx - 1, my, m); nx = 3; }
          else if (y < 0) { n = mb_at(d, mx, my - 1, m); ny = 3; }
          else return m.i4[y * 4 + x];
          if (!n || (d->cpps->constrained_intra && !n->intra())) { dcpred = true; return 2; }
          return n->kind == kI4x4 ? n->i4[ny * 4 + nx] : 2;
        };
        bool dcpred = false;
        const int ma = nb_mode(bx - 1, by, dcpred), mbm = nb_mode(bx, by - 1, dcpred);
        const int pred = dcpred ? 2 : std::min(ma, mbm);
        const int mode = prev ? pred : (rem < pred ? rem : rem + 1);
        m.i4[by * 4 + bx] = (int8_t)mode;
      }
    }
    chroma_mode = (int)b.ue();
    if (chroma_mode > 3) return fail("intra_chroma_pred_mode %d", chroma_mode);
  } else {
    m.kind = kPInter;
    const int t = (int)mb_type - 100;
    m.part = (int8_t)std::min(t, 3);
    const int nref = d->sh.num_ref_idx_active;
    int ref[4] = {0, 0, 0, 0};
    int mvd[16][2];
    if (t <= 2) {
      const int np = t == 0 ? 1 : 2;
      for (int p = 0; p < np; p++) ref[p] = nref > 1 ? (int)b.te(nref - 1) : 0;
      for (int p = 0; p < np; p++) {
        mvd[p][0] = b.se();
        mvd[p][1] = b.se();
      }
      for (int p = 0; p < np; p++) {
        if (ref[p] >= nref || !d->list0[ref[p]]) return fail("ref_idx %d has no picture", ref[p]);
        const int x = t == 2 ? 8 * p : 0, y = t == 1 ? 8 * p : 0, w = t == 2 ? 8 : 16, h = t == 1 ? 8 : 16;
        int px, py;
        mv_pred(d, mx, my, m, done, x, y, w, h, ref[p], t, p, px, py);
        const int vx = px + mvd[p][0], vy = py + mvd[p][1];
        for (int by = y / 4; by < (y + h) / 4; by++)
          for (int bx = x / 4; bx < (x + w) / 4; bx++) {
            m.mv[by * 4 + bx][0] = (int16_t)vx;
            m.mv[by * 4 + bx][1] = (int16_t)vy;
            m.mvp[by * 4 + bx][0] = (int16_t)px;
            m.mvp[by * 4 + bx][1] = (int16_t)py;
            done |= 1u << (by * 4 + bx);
          }
        for (int b8 = 0; b8 < 4; b8++) {
          const int bx8 = (b8 & 1) * 8, by8 = (b8 >> 1) * 8;
          if (bx8 >= x && bx8 < x + w && by8 >= y && by8 < y + h) {
            m.ref[b8] = (int8_t)ref[p];
            m.refpic[b8] = d->list0[ref[p]]->id;
          }
        }
      }
    } else {
      int sub[4];
      for (int s = 0; s < 4; s++) {
        sub[s] = (int)b.ue();
        if (sub[s] > 3) return fail("sub_mb_type %d", sub[s]);
        m.sub[s] = (int8_t)sub[s];
      }
      for (int s = 0; s < 4; s++) ref[s] = (nref > 1 && t == 3) ? (int)b.te(nref - 1) : 0;
      int nsub[4], k = 0;
      for (int s = 0; s < 4; s++) {
        nsub[s] = sub[s] == 0 ? 1 : (sub[s] == 3 ? 4 : 2);
        for (int j = 0; j < nsub[s]; j++, k++) {
          mvd[k][0] = b.se();
          mvd[k][1] = b.se();
        }
      }
      k = 0;
      for (int s = 0; s < 4; s++) {
        if (ref[s] >= nref || !d->list0[ref[s]]) return fail("ref_idx %d has no picture", ref[s]);
        m.ref[s] = (int8_t)ref[s];
        m.refpic[s] = d->list0[ref[s]]->id;
        const int x8 = (s & 1) * 8, y8 = (s >> 1) * 8;
        const int w = (sub[s] == 0 || sub[s] == 1) ? 8 : 4, h = (sub[s] == 0 || sub[s] == 2) ? 8 : 4;
        for (int j = 0; j < nsub[s]; j++, k++) {
          const int x = x8 + (w == 8 ? 0 : 4 * (j & 1)), y = y8 + (h == 8 ? 0 : 4 * (w == 8 ? j : (j >> 1)));
          int px, py;
          mv_pred(d, mx, my, m, done, x, y, w, h, ref[s], 0, 0, px, py);
          const int vx = px + mvd[k][0], vy = py + mvd[k][1];
          for (int by = y / 4; by < (y + h) / 4; by++)
            for (int bx = x / 4; bx < (x + w) / 4; bx++) {
              m.mv[by * 4 + bx][0] = (int16_t)vx;
              m.mv[by * 4 + bx][1] = (int16_t)vy;
              m.mvp[by * 4 + bx][0] = (int16_t)px;
              m.mvp[by * 4 + bx][1] = (int16_t)py;
              done |= 1u << (by * 4 + bx);
            }
        }
      }
    }
  }
  if (!i16) {
    const uint32_t cn = b.ue();
    if (cn > 47) return fail("coded_block_pattern code %u", cn);
    cbp = intra ? kCbpIntra[cn] : kCbpInter[cn];
  }
  const int cbp_l = cbp & 15, cbp_c = cbp >> 4;
  if (cbp_l || cbp_c || i16) {
    const int dq = b.se();
    if (dq < -26 || dq > 25) return fail("mb_qp_delta %d", dq);
    qp = (qp + dq + 52) % 52;
  }
  m.qp = qp;
  m.qpc = kQpc[clip3(0, 51, qp + cqo)];
  // ---- residual (7.3.5.3), then reconstruction
  int dcl[16] = {0}, acl[16][16], cdc[2][4], cac[2][4][16];
  for (int i = 0; i < 16; i++)
    for (int k = 0; k < 16; k++) acl[i][k] = 0;
  bool has_dc16 = false;
  if (i16) {
    int lv[16];
    const int tc = residual_block(b, coeff_nc(d, mx, my, m, -1, 0, 0), 0, 15, 16, lv);
    if (tc < 0) return fail("MB %d: bad Intra16x16DCLevel", addr);
    for (int k = 0; k < 16; k++) dcl[k] = lv[k];
    has_dc16 = true;
  }
  for (int b8 = 0; b8 < 4; b8++)
    for (int i = 0; i < 4; i++) {
      const int blk = b8 * 4 + i, bx = blk_x(blk), by = blk_y(blk);
      if (!((cbp_l >> b8) & 1)) continue;
      int lv[16];
      const int nc = coeff_nc(d, mx, my, m, -1, bx, by);
      int tc;
      if (i16) {
        tc = residual_block(b, nc, 0, 14, 15, lv);
        if (tc < 0) return fail("MB %d: bad Intra16x16ACLevel", addr);
        acl[blk][0] = 0;
        for (int k = 0; k < 15; k++) acl[blk][k + 1] = lv[k];
      } else {
        tc = residual_block(b, nc, 0, 15, 16, lv);
        if (tc < 0) return fail("MB %d: bad LumaLevel4x4", addr);
        for (int k = 0; k < 16; k++) acl[blk][k] = lv[k];
      }
      m.tc[by * 4 + bx] = (uint8_t)tc;
    }
  for (int c = 0; c < 2; c++)
    for (int k = 0; k < 4; k++) cdc[c][k] = 0;
  for (int c = 0; c < 2; c++)
    for (int i = 0; i < 4; i++)
      for (int k = 0; k < 16; k++) cac[c][i][k] = 0;
  if (cbp_c & 3)
    for (int c = 0; c < 2; c++) {
      int lv[16];
      if (residual_block(b, -1, 0, 3, 4, lv) < 0) return fail("MB %d: bad ChromaDCLevel", addr);
      for (int k = 0; k < 4; k++) cdc[c][k] = lv[k];
    }
  if (cbp_c & 2)
    for (int c = 0; c < 2; c++)
      for (int i = 0; i < 4; i++) {
        int lv[16];
        const int tc = residual_block(b, coeff_nc(d, mx, my, m, c, i & 1, i >> 1), 0, 14, 15, lv);
        if (tc < 0) return fail("MB %d: bad ChromaACLevel", addr);
        for (int k = 0; k < 15; k++) cac[c][i][k + 1] = lv[k];
        m.tcc[c][i] = (uint8_t)tc;
      }
  if (b.bad()) return fail("MB %d: past the end of the slice", addr);

  // luma prediction + residual
  if (m.kind == kPInter) {
    recon_inter(d, mx, my, m, acl, cbp_l, qp);
  } else if (m.kind == kI4x4) {
    uint32_t dn = 0;
    for (int blk = 0; blk < 16; blk++) {
      const int bx = blk_x(blk), by = blk_y(blk);
      pred4x4(d, mx, my, m, bx, by, dn, m.i4[by * 4 + bx], &P.y[(size_t)(16 * my + 4 * by) * P.w + 16 * mx + 4 * bx],
              P.w);
      if ((cbp_l >> (blk >> 2)) & 1) luma_residual(d, mx, my, bx, by, acl[blk], qp, false, 0);
      dn |= 1u << (by * 4 + bx);
    }
  } else {
    int rc = pred16x16(d, mx, my, m, i16_mode);
    if (rc) return rc;
    // 8.5.10: the Intra16x16 DC levels (zigzag -> raster of blocks), Hadamard, scale
    int c[16], f[16], t[16];
    for (int k = 0; k < 16; k++) c[kZigzag4[k]] = dcl[k];
    for (int i = 0; i < 4; i++) {
      const int *s = c + 4 * i;
      t[4 * i] = s[0] + s[1] + s[2] + s[3];
      t[4 * i + 1] = s[0] + s[1] - s[2] - s[3];
      t[4 * i + 2] = s[0] - s[1] - s[2] + s[3];
      t[4 * i + 3] = s[0] - s[1] + s[2] - s[3];
    }
    for (int j = 0; j < 4; j++) {
      f[j] = t[j] + t[4 + j] + t[8 + j] + t[12 + j];
      f[4 + j] = t[j] + t[4 + j] - t[8 + j] - t[12 + j];
      f[8 + j] = t[j] - t[4 + j] - t[8 + j] + t[12 + j];
      f[12 + j] = t[j] - t[4 + j] + t[8 + j] - t[12 + j];
    }
    const int ls = 16 * kNorm4[qp % 6][0];
    for (int blk = 0; blk < 16; blk++) {
      const int bx = blk_x(blk), by = blk_y(blk);
      const int fv = f[by * 4 + bx];
      const int dc = qp >= 36 ? (fv * ls) << (qp / 6 - 6) : (fv * ls + (1 << (5 - qp / 6))) >> (6 - qp / 6);
      luma_residual(d, mx, my, bx, by, acl[blk], qp, has_dc16, dc);
    }
  }
  // chroma prediction + residual
  for (int c = 0; c < 2; c++) {
    std::vector<uint8_t> &pl = c ? P.v : P.u;
    if (intra) {
      int rc = pred_chroma(d, mx, my, m, chroma_mode, pl);
      if (rc) return rc;
    }
    if (cbp_c) chroma_residual(d, mx, my, m.qpc, cdc[c], cac[c], pl);
  }
  return PCME_OK;
}

void decode_skip(Dec *d, int addr, int qp) {
  const int mx = addr % d->mbw, my = addr / d->mbw;
  Mb &m = d->mbs[addr];
  m.slice = (int)d->slices.size() - 1;
  m.kind = kPSkip;
  for (int i = 0; i < 16; i++) {
    m.i4[i] = 2;
    m.tc[i] = 0;
  }
  for (int i = 0; i < 4; i++) m.tcc[0][i] = m.tcc[1][i] = 0;
  int px, py;
  pskip_mv(d, mx, my, m, px, py);
  for (int i = 0; i < 16; i++) {
    m.mv[i][0] = m.mvp[i][0] = (int16_t)px;
    m.mv[i][1] = m.mvp[i][1] = (int16_t)py;
  }
  m.part = -1;
  for (int i = 0; i < 4; i++) {
    m.sub[i] = -1;
    m.ref[i] = 0;
    m.refpic[i] = d->list0[0]->id;
  }
  m.qp = qp;
  m.qpc = kQpc[clip3(0, 51, qp + d->cpps->chroma_qp_offset)];
  for (int blk = 0; blk < 16; blk++) mc_block(d, *d->list0[0], mx, my, blk & 3, blk >> 2, px, py);
}

// ------------------------------------------------------------ deblocking

// 8.7.2: filter one line of samples across an edge (p[i] = s[-(i+1)*step], q[i] = s[i*step]).
void filter_line(uint8_t *s, int step, int bs, int alpha, int beta, int tc0, bool chroma) {
  const int p0 = s[-step], p1 = s[-2 * step], q0 = s[0], q1 = s[step];
  if (!(std::abs(p0 - q0) < alpha && std::abs(p1 - p0) < beta && std::abs(q1 - q0) < beta)) return;
  if (chroma) {
    if (bs < 4) {
      const int tc = tc0 + 1;
      const int delta = clip3(-tc, tc, (((q0 - p0) << 2) + (p1 - q1) + 4) >> 3);
      s[-step] = clip1(p0 + delta);
      s[0] = clip1(q0 - delta);
    } else {
      s[-step] = (uint8_t)((2 * p1 + p0 + q1 + 2) >> 2);
      s[0] = (uint8_t)((2 * q1 + q0 + p1 + 2) >> 2);
    }
    return;
  }
  const int p2 = s[-3 * step], q2 = s[2 * step];
  const int ap = std::abs(p2 - p0), aq = std::abs(q2 - q0);
  if (bs < 4) {
    const int tc = tc0 + (ap < beta) + (aq < beta);
    const int delta = clip3(-tc, tc, (((q0 - p0) << 2) + (p1 - q1) + 4) >> 3);
    s[-step] = clip1(p0 + delta);
    s[0] = clip1(q0 - delta);
    if (ap < beta) s[-2 * step] = (uint8_t)(p1 + clip3(-tc0, tc0, (p2 + ((p0 + q0 + 1) >> 1) - (p1 << 1)) >> 1));
    if (aq < beta) s[step] = (uint8_t)(q1 + clip3(-tc0, tc0, (q2 + ((p0 + q0 + 1) >> 1) - (q1 << 1)) >> 1));
    return;
  }
  const int p3 = s[-4 * step], q3 = s[3 * step];
  const bool strong = std::abs(p0 - q0) < ((alpha >> 2) + 2);
  if (ap < beta && strong) {
    s[-step] = (uint8_t)((p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3);
    s[-2 * step] = (uint8_t)((p2 + p1 + p0 + q0 + 2) >> 2);
    s[-3 * step] = (uint8_t)((2 * p3 + 3 * p2 + p1 + p0 + q0 + 4) >> 3);
  } else {
    s[-step] = (uint8_t)((2 * p1 + p0 + q1 + 2) >> 2);
  }
  if (aq < beta && strong) {
    s[0] = (uint8_t)((p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3);
    s[step] = (uint8_t)((p0 + q0 + q1 + q2 + 2) >> 2);
    s[2 * step] = (uint8_t)((2 * q3 + 3 * q2 + q1 + q0 + p0 + 4) >> 3);
  } else {
    s[0] = (uint8_t)((2 * q1 + q0 + p1 + 2) >> 2);
  }
}

// 8.7.2.1: bS between 4x4 block bp of MB P and bq of MB Q (raster indices).
int boundary_strength(const Mb &P, int bp, const Mb &Q, int bq, bool mb_edge) {
  if (P.intra() || Q.intra()) return mb_edge ? 4 : 3;
  if (P.tc[bp] || Q.tc[bq]) return 2;
  const int rp = P.refpic[((bp >> 3) << 1) | ((bp & 3) >> 1)], rq = Q.refpic[((bq >> 3) << 1) | ((bq & 3) >> 1)];
  if (rp != rq) return 1;
  if (std::abs(P.mv[bp][0] - Q.mv[bq][0]) >= 4 || std::abs(P.mv[bp][1] - Q.mv[bq][1]) >= 4) return 1;
  return 0;
}

void deblock_picture(Dec *d) {
  Picture &P = *d->cur;
  const int W = P.w, cw = W / 2;
  for (int my = 0; my < d->mbh; my++)
    for (int mx = 0; mx < d->mbw; mx++) {
      const Mb &Q = d->mbs[(size_t)my * d->mbw + mx];
      const SliceInfo &si = d->slices[Q.slice];
      if (si.disable_deblock == 1) continue;
      for (int dir = 0; dir < 2; dir++) {  // 0: vertical edges, 1: horizontal edges
        for (int e = 0; e < 4; e++) {
          const bool mb_edge = e == 0;
          const Mb *Pm = &Q;
          if (mb_edge) {
            const int nx = dir ? mx : mx - 1, ny = dir ? my - 1 : my;
            if (nx < 0 || ny < 0) continue;
            Pm = &d->mbs[(size_t)ny * d->mbw + nx];
            if (si.disable_deblock == 2 && Pm->slice != Q.slice) continue;
          }
          int bs[4];
          for (int k = 0; k < 4; k++) {
            const int bq = dir ? (e * 4 + k) : (k * 4 + e);
            const int bp = mb_edge ? (dir ? (12 + k) : (k * 4 + 3)) : (dir ? ((e - 1) * 4 + k) : (k * 4 + e - 1));
            bs[k] = boundary_strength(*Pm, bp, Q, bq, mb_edge);
          }
          // luma
          {
            const int qpav = (Pm->qp + Q.qp + 1) >> 1;
            const int ia = clip3(0, 51, qpav + si.alpha_off), ib = clip3(0, 51, qpav + si.beta_off);
            const int alpha = kAlpha[ia], beta = kBeta[ib];
            for (int i = 0; i < 16; i++) {
              const int k = i >> 2;
              if (!bs[k]) continue;
              const int x = 16 * mx + (dir ? i : 4 * e), y = 16 * my + (dir ? 4 * e : i);
              uint8_t *s = &P.y[(size_t)y * W + x];
              filter_line(s, dir ? W : 1, bs[k], alpha, beta, bs[k] < 4 ? kTc0[ia][bs[k] - 1] : 0, false);
            }
          }
          // chroma: edges 0 and 2 of luma are chroma edges 0 and 4
          if (e == 0 || e == 2) {
            const int qpav = (Pm->qpc + Q.qpc + 1) >> 1;
            const int ia = clip3(0, 51, qpav + si.alpha_off), ib = clip3(0, 51, qpav + si.beta_off);
            const int alpha = kAlpha[ia], beta = kBeta[ib];
            for (int c = 0; c < 2; c++) {
              std::vector<uint8_t> &pl = c ? P.v : P.u;
              for (int i = 0; i < 8; i++) {
                const int k = i >> 1;
                if (!bs[k]) continue;
                const int x = 8 * mx + (dir ? i : 2 * e), y = 8 * my + (dir ? 2 * e : i);
                uint8_t *s = &pl[(size_t)y * cw + x];
                filter_line(s, dir ? cw : 1, bs[k], alpha, beta, bs[k] < 4 ? kTc0[ia][bs[k] - 1] : 0, true);
              }
            }
          }
        }
      }
    }
}

// ------------------------------------------------------------------ slices

// 7.3.3: slice_header() and slice_data() of one slice NAL unit.
int decode_slice(Dec *d, Bits &b, int nal_type, int nal_ref_idc) {
  SliceHeader h;
  h.first_mb = (int)b.ue();
  h.type = (int)(b.ue() % 5);
  if (h.type != 0 && h.type != 2) return fail("slice type %d (only I and P are in the subset)", h.type);
  h.pps_id = (int)b.ue();
  if (h.pps_id >= 256 || !d->pps[h.pps_id].valid) return fail("slice refers to a missing PPS %d", h.pps_id);
  const Pps &pps = d->pps[h.pps_id];
  const Sps &sps = d->sps[pps.sps_id];
  if (!sps.valid) return fail("PPS %d refers to a missing SPS", h.pps_id);
  const bool idr = nal_type == 5;
  h.frame_num = (int)b.u(sps.log2_max_frame_num);
  if (idr) h.idr_pic_id = (int)b.ue();
  if (sps.poc_type == 0) {
    h.poc_lsb = (int)b.u(sps.log2_max_poc_lsb);
    if (pps.bottom_field_pic_order_present) h.delta_poc_bottom = b.se();
  }
  if (sps.poc_type == 1 && !sps.delta_pic_order_always_zero) {
    h.delta_poc[0] = b.se();
    if (pps.bottom_field_pic_order_present) h.delta_poc[1] = b.se();
  }
  if (pps.redundant_pic_cnt) {
    if (b.ue() != 0) return fail("redundant pictures are outside the subset");
  }
  h.num_ref_idx_active = pps.num_ref_idx_default;
  if (h.type == 0) {
    if (b.flag()) h.num_ref_idx_active = (int)b.ue() + 1;
    if (h.num_ref_idx_active > 32) return fail("num_ref_idx_active %d", h.num_ref_idx_active);
    if (b.flag()) {  // ref_pic_list_modification_flag_l0
      for (;;) {
        const int idc = (int)b.ue();
        if (idc == 3) break;
        if (idc > 2 || h.reorder.size() > 32 || b.bad()) return fail("bad ref_pic_list_modification");
        h.reorder.push_back({idc, (int)b.ue()});
      }
    }
  }
  PicHeader ph;
  ph.idr = idr;
  ph.nal_ref_idc = nal_ref_idc;
  ph.frame_num = h.frame_num;
  if (nal_ref_idc) {
    if (idr) {
      b.u(1);  // no_output_of_prior_pics_flag
      ph.long_term_reference = b.flag();
    } else {
      ph.adaptive_marking = b.flag();
      if (ph.adaptive_marking)
        for (;;) {
          const int op = (int)b.ue();
          if (op == 0) break;
          if (op > 6 || ph.mmco.size() > 64 || b.bad()) return fail("bad memory_management_control_operation");
          int a = 0, c = 0;
          if (op == 1 || op == 3) a = (int)b.ue();
          if (op == 2) a = (int)b.ue();
          if (op == 3 || op == 6) c = (int)b.ue();
          if (op == 4) a = (int)b.ue();
          if (op == 6) a = c;
          ph.mmco.push_back({op, a});
          ph.mmco_b.push_back(c);
        }
    }
  }
  h.qp = pps.pic_init_qp + b.se();
  if (h.qp < 0 || h.qp > 51) return fail("slice QP %d", h.qp);
  if (pps.deblock_ctrl) {
    h.dbk.disable_deblock = (int)b.ue();
    if (h.dbk.disable_deblock > 2) return fail("disable_deblocking_filter_idc %d", h.dbk.disable_deblock);
    if (h.dbk.disable_deblock != 1) {
      h.dbk.alpha_off = 2 * b.se();
      h.dbk.beta_off = 2 * b.se();
    }
  }
  if (b.bad()) return fail("malformed slice header");
  // a new picture begins with its first macroblock
  if (h.first_mb == 0) {
    int rc = finish_picture(d);
    if (rc) return rc;
  }
  if (!d->cur) {
    if (h.first_mb != 0) return fail("a picture starts with a slice at MB %d", h.first_mb);
    if (!idr && d->dpb.empty()) return fail("a P picture without a reference");
    d->sh = h;
    start_picture(d, sps, nal_ref_idc, idr);
    d->ph = ph;
  } else if (d->csps != &sps || h.frame_num != d->cur->frame_num) {
    return fail("a slice of another picture before the first MB of a new one");
  }
  d->sh = h;
  d->cpps = &pps;
  if (h.first_mb >= d->mbw * d->mbh) return fail("first_mb_in_slice %d", h.first_mb);
  if (h.type == 0) {
    int rc = build_list0(d);
    if (rc) return rc;
    if (!d->list0[0]) return fail("RefPicList0[0] is empty");
  }
  d->slices.push_back(h.dbk);
  // 7.3.4: slice_data()
  int qp = h.qp, addr = h.first_mb;
  const int n = d->mbw * d->mbh;
  bool more = true;
  while (more) {
    if (h.type == 0) {
      const uint32_t run = b.ue();
      if (b.bad() || run > (uint32_t)(n - addr)) return fail("mb_skip_run %u past the picture", run);
      for (uint32_t i = 0; i < run; i++) {
        if (d->mbs[addr].kind != kNone) return fail("MB %d decoded twice", addr);
        decode_skip(d, addr++, qp);
      }
      if (run > 0) {
        more = b.more_rbsp();
        if (!more) break;
      }
    }
    if (addr >= n) return fail("slice runs past the picture");
    if (d->mbs[addr].kind != kNone) return fail("MB %d decoded twice", addr);
    int rc = decode_mb(d, b, addr, h.type, qp);
    if (rc) return rc;
    more = b.more_rbsp();
    addr++;
  }
  if (b.bad()) return fail("slice data past the end of the NAL unit");
  return PCME_OK;
}

int decode_nal(Dec *d, const uint8_t *p, size_t n) {
  if (n < 1) return PCME_OK;
  if (p[0] & 0x80) return fail("forbidden_zero_bit set");
  const int ref_idc = (p[0] >> 5) & 3, type = p[0] & 31;
  unescape(p + 1, n - 1, d->rbsp);
  Bits b(d->rbsp.data(), d->rbsp.size());
  switch (type) {
    case 7: return parse_sps(d, b);
    case 8: return parse_pps(d, b);
    case 1:
    case 5: return decode_slice(d, b, type, ref_idc);
    case 2:
    case 3:
    case 4: return fail("data partitioning is outside the subset");
    case 20: return fail("NAL unit type %d is outside the subset", type);
    default: return PCME_OK;  // SEI, AUD, end of sequence/stream, filler: no samples
  }
}

}  // namespace

// ----------------------------------------------------------------- C ABI

extern "C" {

int pcme_h264dec_create(pcme_h264dec **out) {
  if (!out) return fail("pcme_h264dec_create: null out");
  *out = new pcme_h264dec();
  return PCME_OK;
}

void pcme_h264dec_destroy(pcme_h264dec *dec) { delete dec; }

int pcme_h264dec_decode(pcme_h264dec *d, const uint8_t *data, size_t size) {
  if (!d || (!data && size)) return fail("pcme_h264dec_decode: bad arguments");
  // Annex B: NAL units after 0x000001 start codes
  size_t i = 0, start = (size_t)-1;
  while (i + 2 < size) {
    if (data[i] == 0 && data[i + 1] == 0 && data[i + 2] == 1) {
      if (start != (size_t)-1) {
        size_t end = i;
        while (end > start && data[end - 1] == 0) end--;
        int rc = decode_nal(d, data + start, end - start);
        if (rc) return rc;
      }
      i += 3;
      start = i;
    } else {
      i++;
    }
  }
  if (start != (size_t)-1) {
    size_t end = size;
    while (end > start && data[end - 1] == 0) end--;
    int rc = decode_nal(d, data + start, end - start);
    if (rc) return rc;
  }
  return PCME_OK;
}

int pcme_h264dec_flush(pcme_h264dec *d) {
  if (!d) return fail("pcme_h264dec_flush: null decoder");
  return finish_picture(d);
}

int pcme_h264dec_frame(pcme_h264dec *d, pcme_h264_frame_info *info, uint8_t *y, int32_t ystride, uint8_t *u,
                       uint8_t *v, int32_t cstride) {
  if (!d || !info) return fail("pcme_h264dec_frame: bad arguments");
  if (d->out.empty()) return 0;
  const Picture &p = *d->out.front();
  const int cl = 2 * p.crop[0], cr = 2 * p.crop[1], ct = 2 * p.crop[2], cb = 2 * p.crop[3];
  info->width = p.w - cl - cr;
  info->height = p.h - ct - cb;
  info->poc = p.poc;
  info->frame_num = p.frame_num;
  if (!y || !u || !v) return 1;
  if (ystride < info->width || cstride < info->width / 2) return fail("pcme_h264dec_frame: strides too small");
  for (int r = 0; r < info->height; r++) memcpy(y + (size_t)r * ystride, &p.y[(size_t)(r + ct) * p.w + cl], info->width);
  for (int r = 0; r < info->height / 2; r++) {
    memcpy(u + (size_t)r * cstride, &p.u[(size_t)(r + ct / 2) * (p.w / 2) + cl / 2], info->width / 2);
    memcpy(v + (size_t)r * cstride, &p.v[(size_t)(r + ct / 2) * (p.w / 2) + cl / 2], info->width / 2);
  }
  d->out.pop_front();
  return 1;
}

static_assert(sizeof(pcme_h264_mb) == 144, "pcme_h264_mb layout");

int pcme_h264dec_frame_mbs(pcme_h264dec *d, pcme_h264_mb *out, int32_t n) {
  if (!d || n < 0 || (n > 0 && !out)) return fail("pcme_h264dec_frame_mbs: bad arguments");
  if (d->out.empty()) return 0;
  const std::vector<Mb> &mbs = d->out.front()->mbs;
  for (int i = 0; i < n && i < (int)mbs.size(); i++) {
    const Mb &m = mbs[i];
    pcme_h264_mb &o = out[i];
    memset(&o, 0, sizeof o);
    static const int8_t kinds[] = {0, 1, 2, 3, 4};
    o.kind = kinds[m.kind];
    o.part = m.part;
    for (int k = 0; k < 4; k++) {
      o.sub[k] = m.sub[k];
      o.ref[k] = m.ref[k];
    }
    o.qp = (uint8_t)m.qp;
    memcpy(o.mv, m.mv, sizeof o.mv);
    memcpy(o.mvp, m.mvp, sizeof o.mvp);
    o.slice = m.slice;
  }
  return (int)mbs.size();
}

int pcme_h264_mc_luma(const uint8_t *plane, int32_t stride, int32_t w, int32_t h, int32_t x, int32_t y, int32_t mvx,
                      int32_t mvy, int32_t bw, int32_t bh, uint8_t *out, int32_t ostride) {
  if (!plane || !out || w < 1 || h < 1 || stride < w || bw < 1 || bh < 1 || ostride < bw)
    return fail("pcme_h264_mc_luma: bad arguments");
  const PlaneRef R{plane, stride, w, h};
  for (int r = 0; r < bh; r++)
    for (int c = 0; c < bw; c++) {
      const int qx = 4 * (x + c) + mvx, qy = 4 * (y + r) + mvy;
      out[(size_t)r * ostride + c] = (uint8_t)luma_qpel(R, qx >> 2, qy >> 2, qx & 3, qy & 3);
    }
  return PCME_OK;
}

const char *pcme_h264dec_last_error(void) { return g_err; }

}  // extern "C"

// ================================================================ encoder
//
// The field coder of include/pcme_enc.h.  It writes the syntax the decoder
// above reads and reconstructs every MB with the decoder's own functions
// (mc_block, luma_residual, chroma_residual, deblock_picture), so its
// reconstruction is the one any decoder of the stream produces.

namespace {

class BitWriter {
 public:
  void u(int k, uint32_t v) {
    for (int i = k - 1; i >= 0; i--) bit((v >> i) & 1u);
  }
  void ue(uint32_t v) {
    const uint32_t x = v + 1;
    int n = 0;
    while ((x >> (n + 1)) != 0) n++;
    u(n, 0);
    u(n + 1, x);
  }
  void se(int32_t v) { ue(v > 0 ? (uint32_t)(2 * v - 1) : (uint32_t)(-2 * v)); }
  void bit(uint32_t b) {
    cur_ = (uint8_t)((cur_ << 1) | b);
    if (++n_ == 8) {
      buf_.push_back(cur_);
      cur_ = 0;
      n_ = 0;
    }
  }
  bool aligned() const { return n_ == 0; }
  void align_zero() {
    while (n_) bit(0);
  }
  void trailing() {  // rbsp_trailing_bits
    bit(1);
    align_zero();
  }
  const std::vector<uint8_t> &bytes() const { return buf_; }

 private:
  std::vector<uint8_t> buf_;
  uint8_t cur_ = 0;
  int n_ = 0;
};

// A NAL unit: start code, header, the RBSP with emulation prevention (7.4.1).
void put_nal(std::vector<uint8_t> &out, int ref_idc, int type, const std::vector<uint8_t> &rbsp) {
  out.insert(out.end(), {0, 0, 0, 1, (uint8_t)((ref_idc << 5) | type)});
  int zeros = 0;
  for (uint8_t b : rbsp) {
    if (zeros >= 2 && b <= 3) {
      out.push_back(3);
      zeros = 0;
    }
    out.push_back(b);
    zeros = b == 0 ? zeros + 1 : 0;
  }
}

// QUANT4 (the multipliers that invert normAdjust4x4): by position class
const int kQuant4[6][3] = {{13107, 5243, 8066}, {11916, 4660, 7490}, {10082, 4194, 6554},
                           {9362, 3647, 5825},  {8192, 3355, 5243},  {7282, 2893, 4559}};
inline int mf4(int m, int pos) {
  const int i = pos >> 2, j = pos & 3;
  return kQuant4[m][(i & 1) == 0 && (j & 1) == 0 ? 0 : ((i & 1) && (j & 1) ? 1 : 2)];
}

// Forward 4x4 core transform (the inverse of 8.5.12's up to scaling).
void fdct4(const int *d, int *o) {
  int t[16];
  for (int i = 0; i < 4; i++) {
    const int *s = d + 4 * i;
    const int s03 = s[0] + s[3], d03 = s[0] - s[3], s12 = s[1] + s[2], d12 = s[1] - s[2];
    t[4 * i] = s03 + s12;
    t[4 * i + 1] = 2 * d03 + d12;
    t[4 * i + 2] = s03 - s12;
    t[4 * i + 3] = d03 - 2 * d12;
  }
  for (int j = 0; j < 4; j++) {
    const int s03 = t[j] + t[12 + j], d03 = t[j] - t[12 + j], s12 = t[4 + j] + t[8 + j], d12 = t[4 + j] - t[8 + j];
    o[j] = s03 + s12;
    o[4 + j] = 2 * d03 + d12;
    o[8 + j] = s03 - s12;
    o[12 + j] = d03 - 2 * d12;
  }
}

// Dead-zone quantiser: |level| = (|w| * mf + f) >> qbits, f = 2^qbits / 6.
inline int quant(int w, int mf, int qbits) {
  const int64_t f = ((int64_t)1 << qbits) / 6;
  const int a = (int)(((int64_t)(w < 0 ? -w : w) * mf + f) >> qbits);
  return w < 0 ? -a : a;
}

// residual_block_cavlc (the writer of 9.2): lv[0..n-1] in scan order.
void write_block(BitWriter &bw, int nc, const int *lv, int n) {
  int pos[16], tc = 0;
  for (int k = 0; k < n; k++)
    if (lv[k]) pos[tc++] = k;
  int t1 = 0;
  for (int i = tc - 1; i >= 0 && t1 < 3 && (lv[pos[i]] == 1 || lv[pos[i]] == -1); i--) t1++;
  if (nc >= 8) {
    bw.u(6, tc == 0 ? 3u : (uint32_t)(((tc - 1) << 2) | t1));
  } else {
    const int cls = nc == -1 ? 3 : (nc < 2 ? 0 : (nc < 4 ? 1 : 2));
    for (int i = 0; i < kCoeffTokenN[cls]; i++) {
      const Vlc4 &e = kCoeffToken[cls][i];
      if (e.tc == tc && e.t1 == t1) {
        bw.u(e.len, e.val);
        break;
      }
    }
  }
  if (tc == 0) return;
  // levels from the highest frequency down
  int suffix_len = (tc > 10 && t1 < 3) ? 1 : 0;
  for (int i = 0; i < tc; i++) {
    const int level = lv[pos[tc - 1 - i]];
    if (i < t1) {
      bw.u(1, level < 0 ? 1u : 0u);
      continue;
    }
    int code = level > 0 ? 2 * level - 2 : -2 * level - 1;
    if (i == t1 && t1 < 3) code -= 2;
    int prefix, ssize = 0, suffix = 0;
    if (suffix_len == 0) {
      if (code < 14) prefix = code;
      else if (code < 30) { prefix = 14; ssize = 4; suffix = code - 14; }
      else { prefix = 15; ssize = 12; suffix = code - 30; }
    } else if (code < (15 << suffix_len)) {
      prefix = code >> suffix_len;
      ssize = suffix_len;
      suffix = code & ((1 << suffix_len) - 1);
    } else {
      prefix = 15;
      ssize = 12;
      suffix = code - (15 << suffix_len);
    }
    bw.u(prefix, 0);
    bw.u(1, 1);
    if (ssize) bw.u(ssize, (uint32_t)suffix);
    if (suffix_len == 0) suffix_len = 1;
    if (std::abs(level) > (3 << (suffix_len - 1)) && suffix_len < 6) suffix_len++;
  }
  int zeros = pos[tc - 1] + 1 - tc;
  if (tc < n) {
    const Vlc2 &e = n == 4 ? kTotalZerosDc[tc - 1][zeros] : kTotalZeros4[tc - 1][zeros];
    bw.u(e.len, e.val);
  }
  for (int i = tc - 1; i > 0 && zeros > 0; i--) {
    const int run = pos[i] - pos[i - 1] - 1;
    const Vlc2 &e = kRunBefore[std::min(zeros, 7) - 1][run];
    bw.u(e.len, e.val);
    zeros -= run;
  }
}

}  // namespace

struct pcme_h264enc {
  int w = 0, h = 0, mbw = 0, mbh = 0;
  Sps sps;
  Pps pps;
  pcme_h264dec d;                 // picture under construction, MB records, reference list
  std::deque<std::unique_ptr<Picture>> dpb;  // reconstructions of the reference pictures, newest first
  int nref = 1;                   // reference indices of the following P pictures (set_refs)
  int frame_num = 0;
  bool started = false;
  int n_skip = 0, n_inter = 0, n_pcm = 0;
  std::vector<uint8_t> sy, su, sv;  // the coded-size source (edges replicated)
};

namespace {

using Enc = pcme_h264enc;

void load_source(Enc *e, const uint8_t *y, int ys, const uint8_t *u, const uint8_t *v, int cs) {
  const int W = 16 * e->mbw, H = 16 * e->mbh;
  e->sy.resize((size_t)W * H);
  e->su.resize((size_t)W * H / 4);
  e->sv.resize((size_t)W * H / 4);
  for (int r = 0; r < H; r++)
    for (int c = 0; c < W; c++) e->sy[(size_t)r * W + c] = y[(size_t)std::min(r, e->h - 1) * ys + std::min(c, e->w - 1)];
  for (int r = 0; r < H / 2; r++)
    for (int c = 0; c < W / 2; c++) {
      const size_t o = (size_t)std::min(r, e->h / 2 - 1) * cs + std::min(c, e->w / 2 - 1);
      e->su[(size_t)r * (W / 2) + c] = u[o];
      e->sv[(size_t)r * (W / 2) + c] = v[o];
    }
}

void write_sps_pps(Enc *e, std::vector<uint8_t> &out) {
  BitWriter b;
  b.u(8, 66);    // profile_idc: Baseline
  b.u(8, 0xC0);  // constraint_set0_flag, constraint_set1_flag
  b.u(8, 51);    // level_idc
  b.ue(0);       // seq_parameter_set_id
  b.ue(12);      // log2_max_frame_num_minus4: 16 bits
  b.ue(2);       // pic_order_cnt_type
  b.ue((uint32_t)e->sps.max_num_ref_frames);
  b.u(1, 0);     // gaps_in_frame_num_value_allowed_flag
  b.ue((uint32_t)e->mbw - 1);
  b.ue((uint32_t)e->mbh - 1);
  b.u(1, 1);  // frame_mbs_only_flag
  b.u(1, 1);  // direct_8x8_inference_flag
  const int cr = (16 * e->mbw - e->w) / 2, cb = (16 * e->mbh - e->h) / 2;
  b.u(1, (cr || cb) ? 1u : 0u);
  if (cr || cb) {
    b.ue(0);
    b.ue((uint32_t)cr);
    b.ue(0);
    b.ue((uint32_t)cb);
  }
  b.u(1, 0);  // vui_parameters_present_flag
  b.trailing();
  put_nal(out, 3, 7, b.bytes());
  BitWriter p;
  p.ue(0);    // pic_parameter_set_id
  p.ue(0);    // seq_parameter_set_id
  p.u(1, 0);  // entropy_coding_mode_flag: CAVLC
  p.u(1, 0);  // bottom_field_pic_order_in_frame_present_flag
  p.ue(0);    // num_slice_groups_minus1
  p.ue(0);    // num_ref_idx_l0_default_active_minus1
  p.ue(0);    // num_ref_idx_l1_default_active_minus1
  p.u(1, 0);  // weighted_pred_flag
  p.u(2, 0);  // weighted_bipred_idc
  p.se(0);    // pic_init_qp_minus26
  p.se(0);    // pic_init_qs_minus26
  p.se(0);    // chroma_qp_index_offset
  p.u(1, 0);  // deblocking_filter_control_present_flag
  p.u(1, 0);  // constrained_intra_pred_flag
  p.u(1, 0);  // redundant_pic_cnt_present_flag
  p.trailing();
  put_nal(out, 3, 8, p.bytes());
}

void begin_picture(Enc *e) {
  pcme_h264dec &d = e->d;
  d.csps = &e->sps;
  d.cpps = &e->pps;
  d.mbw = e->mbw;
  d.mbh = e->mbh;
  d.cur.reset(new Picture());
  Picture &p = *d.cur;
  p.id = d.next_id++;
  p.w = 16 * e->mbw;
  p.h = 16 * e->mbh;
  p.y.assign((size_t)p.w * p.h, 0);
  p.u.assign((size_t)p.w * p.h / 4, 0);
  p.v.assign((size_t)p.w * p.h / 4, 0);
  p.frame_num = e->frame_num;
  for (int i = 0; i < 4; i++) p.crop[i] = e->sps.crop[i];
  d.mbs.assign((size_t)e->mbw * e->mbh, Mb());
  d.slices.assign(1, SliceInfo());
  d.list0.clear();
  for (const auto &r : e->dpb) d.list0.push_back(r.get());  // short-term, descending frame_num
  e->n_skip = e->n_inter = e->n_pcm = 0;
}

void end_picture(Enc *e) {
  deblock_picture(&e->d);
  // sliding window (8.2.5.3) over max_num_ref_frames pictures
  e->dpb.push_front(std::move(e->d.cur));
  while ((int)e->dpb.size() > e->sps.max_num_ref_frames) e->dpb.pop_back();
  e->frame_num = (e->frame_num + 1) & 0xFFFF;
}

int code_idr(Enc *e, int qp, BitWriter &b) {
  b.ue(0);       // first_mb_in_slice
  b.ue(7);       // slice_type: I (all slices)
  b.ue(0);       // pic_parameter_set_id
  b.u(16, 0);    // frame_num
  b.ue(0);       // idr_pic_id
  b.u(1, 0);     // no_output_of_prior_pics_flag
  b.u(1, 0);     // long_term_reference_flag
  b.se(qp - 26);  // slice_qp_delta
  Picture &P = *e->d.cur;
  for (int addr = 0; addr < e->mbw * e->mbh; addr++) {
    const int mx = addr % e->mbw, my = addr / e->mbw;
    b.ue(25);  // I_PCM
    b.align_zero();
    Mb &m = e->d.mbs[addr];
    m.kind = kIPCM;
    m.slice = 0;
    m.qp = 0;
    m.qpc = kQpc[0];
    for (int i = 0; i < 16; i++) {
      m.tc[i] = 16;
      m.i4[i] = 2;
      m.mv[i][0] = m.mv[i][1] = m.mvp[i][0] = m.mvp[i][1] = 0;
    }
    for (int i = 0; i < 4; i++) {
      m.tcc[0][i] = m.tcc[1][i] = 16;
      m.ref[i] = -1;
      m.refpic[i] = -1;
    }
    for (int r = 0; r < 16; r++)
      for (int c = 0; c < 16; c++) {
        const uint8_t s = e->sy[(size_t)(16 * my + r) * P.w + 16 * mx + c];
        b.u(8, s);
        P.y[(size_t)(16 * my + r) * P.w + 16 * mx + c] = s;
      }
    for (int k = 0; k < 2; k++)
      for (int r = 0; r < 8; r++)
        for (int c = 0; c < 8; c++) {
          const size_t o = (size_t)(8 * my + r) * (P.w / 2) + 8 * mx + c;
          const uint8_t s = (k ? e->sv : e->su)[o];
          b.u(8, s);
          (k ? P.v : P.u)[o] = s;
        }
    e->n_pcm++;
  }
  b.trailing();
  return PCME_OK;
}

// te(v) (9.1): one inverted bit when the range is 1, else ue(v).
void te(BitWriter &b, int range, uint32_t v) {
  if (range == 1)
    b.u(1, v ? 0u : 1u);
  else
    b.ue(v);
}

int code_p(Enc *e, const pcme_mb_out *field, int qp, BitWriter &b) {
  pcme_h264dec *d = &e->d;
  const int nra = std::min(e->nref, (int)d->list0.size());  // num_ref_idx_l0_active
  for (int i = 0; i < e->mbw * e->mbh; i++)
    if (field[i].ref >= nra) return fail("MB %d: reference index %d beyond the %d in use", i, field[i].ref, nra);
  b.ue(0);        // first_mb_in_slice
  b.ue(5);        // slice_type: P (all slices)
  b.ue(0);        // pic_parameter_set_id
  b.u(16, (uint32_t)e->frame_num);
  b.u(1, nra > 1 ? 1u : 0u);  // num_ref_idx_active_override_flag
  if (nra > 1) b.ue((uint32_t)nra - 1);
  b.u(1, 0);      // ref_pic_list_modification_flag_l0
  b.u(1, 0);      // adaptive_ref_pic_marking_mode_flag
  b.se(qp - 26);  // slice_qp_delta
  Picture &P = *d->cur;
  const int W = P.w, CW = W / 2;
  const int qpc = kQpc[clip3(0, 51, qp)];
  int skip_run = 0;
  for (int addr = 0; addr < e->mbw * e->mbh; addr++) {
    const int mx = addr % e->mbw, my = addr / e->mbw;
    const pcme_mb_out &f = field[addr];
    Mb &m = d->mbs[addr];
    m.slice = 0;
    m.kind = kPInter;
    m.part = (int8_t)f.part;
    m.qp = qp;
    m.qpc = qpc;
    const Picture &R = *d->list0[f.ref];
    for (int i = 0; i < 4; i++) {
      m.ref[i] = (int8_t)f.ref;
      m.refpic[i] = R.id;
      m.sub[i] = f.part == 3 ? 0 : -1;
      m.tcc[0][i] = m.tcc[1][i] = 0;
    }
    for (int i = 0; i < 16; i++) {
      m.i4[i] = 2;
      m.tc[i] = 0;
      m.mv[i][0] = f.mv[i].x;
      m.mv[i][1] = f.mv[i].y;
    }
    // the partitions (x, y, w, h in samples) and their predictors in decoding order
    static const int kParts[4][4][4] = {{{0, 0, 16, 16}},
                                        {{0, 0, 16, 8}, {0, 8, 16, 8}},
                                        {{0, 0, 8, 16}, {8, 0, 8, 16}},
                                        {{0, 0, 8, 8}, {8, 0, 8, 8}, {0, 8, 8, 8}, {8, 8, 8, 8}}};
    const int np = f.part == 0 ? 1 : (f.part == 3 ? 4 : 2);
    int mvd[4][2];
    uint32_t done = 0;
    for (int p = 0; p < np; p++) {
      const int *q = kParts[f.part][p];
      const int blk = (q[1] / 4) * 4 + q[0] / 4;
      int px, py;
      mv_pred(d, mx, my, m, done, q[0], q[1], q[2], q[3], f.ref, f.part == 3 ? 0 : f.part, p, px, py);
      mvd[p][0] = m.mv[blk][0] - px;
      mvd[p][1] = m.mv[blk][1] - py;
      for (int by = q[1] / 4; by < (q[1] + q[3]) / 4; by++)
        for (int bx = q[0] / 4; bx < (q[0] + q[2]) / 4; bx++) {
          if (m.mv[by * 4 + bx][0] != m.mv[blk][0] || m.mv[by * 4 + bx][1] != m.mv[blk][1])
            return fail("MB %d: the field's MVs differ inside partition %d", addr, p);
          m.mvp[by * 4 + bx][0] = (int16_t)px;
          m.mvp[by * 4 + bx][1] = (int16_t)py;
          done |= 1u << (by * 4 + bx);
        }
    }
    // prediction (into the picture), then the residual levels
    for (int blk = 0; blk < 16; blk++)
      mc_block(d, R, mx, my, blk & 3, blk >> 2, m.mv[blk][0], m.mv[blk][1]);
    int acl[16][16], cdc[2][4], cac[2][4][16];
    int cbp_l = 0;
    const int qbits = 15 + qp / 6;
    for (int blk = 0; blk < 16; blk++) {
      const int bx = blk_x(blk), by = blk_y(blk);
      int r[16], w[16];
      for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) {
          const size_t o = (size_t)(16 * my + 4 * by + y) * W + 16 * mx + 4 * bx + x;
          r[y * 4 + x] = (int)e->sy[o] - (int)P.y[o];
        }
      fdct4(r, w);
      int nz = 0;
      for (int k = 0; k < 16; k++) {
        const int pos = kZigzag4[k];
        acl[blk][k] = quant(w[pos], mf4(qp % 6, pos), qbits);
        nz |= acl[blk][k] != 0;
      }
      if (nz) cbp_l |= 1 << (blk >> 2);
    }
    int cbp_c = 0;
    const int qbc = 15 + qpc / 6;
    for (int c = 0; c < 2; c++) {
      const std::vector<uint8_t> &src = c ? e->sv : e->su;
      const std::vector<uint8_t> &pr = c ? P.v : P.u;
      int dc[4];
      for (int i = 0; i < 4; i++) {
        int r[16], w[16];
        for (int y = 0; y < 4; y++)
          for (int x = 0; x < 4; x++) {
            const size_t o = (size_t)(8 * my + 4 * (i >> 1) + y) * CW + 8 * mx + 4 * (i & 1) + x;
            r[y * 4 + x] = (int)src[o] - (int)pr[o];
          }
        fdct4(r, w);
        dc[i] = w[0];
        cac[c][i][0] = 0;
        for (int k = 1; k < 16; k++) {
          const int pos = kZigzag4[k];
          cac[c][i][k] = quant(w[pos], mf4(qpc % 6, pos), qbc);
          if (cac[c][i][k]) cbp_c = 2;
        }
      }
      const int hd[4] = {dc[0] + dc[1] + dc[2] + dc[3], dc[0] - dc[1] + dc[2] - dc[3], dc[0] + dc[1] - dc[2] - dc[3],
                         dc[0] - dc[1] - dc[2] + dc[3]};
      for (int i = 0; i < 4; i++) {
        cdc[c][i] = quant(hd[i], mf4(qpc % 6, 0), qbc + 1);
        if (cdc[c][i] && cbp_c == 0) cbp_c = 1;
      }
    }
    // P_Skip: reference 0, the 16x16 MV is the skip MV and nothing is coded
    if (f.part == 0 && f.ref == 0 && cbp_l == 0 && cbp_c == 0) {
      int sx, sy;
      pskip_mv(d, mx, my, m, sx, sy);
      if (sx == m.mv[0][0] && sy == m.mv[0][1]) {
        m.kind = kPSkip;
        m.part = -1;
        skip_run++;
        e->n_skip++;
        continue;  // the prediction is the reconstruction
      }
    }
    b.ue((uint32_t)skip_run);
    skip_run = 0;
    b.ue((uint32_t)f.part);  // P_L0_16x16, P_L0_L0_16x8, P_L0_L0_8x16, P_8x8
    if (f.part == 3)
      for (int s = 0; s < 4; s++) b.ue(0);  // sub_mb_type P_L0_8x8
    if (nra > 1)
      for (int p = 0; p < np; p++) te(b, nra - 1, f.ref);  // ref_idx_l0 of each (sub-)partition
    for (int p = 0; p < np; p++) {
      b.se(mvd[p][0]);
      b.se(mvd[p][1]);
    }
    const int cbp = cbp_l | (cbp_c << 4);
    int code = 0;
    while (kCbpInter[code] != cbp) code++;
    b.ue((uint32_t)code);
    if (cbp) b.se(0);  // mb_qp_delta
    for (int b8 = 0; b8 < 4; b8++) {
      if (!((cbp_l >> b8) & 1)) continue;
      for (int i = 0; i < 4; i++) {
        const int blk = b8 * 4 + i, bx = blk_x(blk), by = blk_y(blk);
        write_block(b, coeff_nc(d, mx, my, m, -1, bx, by), acl[blk], 16);
        int tc = 0;
        for (int k = 0; k < 16; k++) tc += acl[blk][k] != 0;
        m.tc[by * 4 + bx] = (uint8_t)tc;
      }
    }
    if (cbp_c)
      for (int c = 0; c < 2; c++) write_block(b, -1, cdc[c], 4);
    if (cbp_c == 2)
      for (int c = 0; c < 2; c++)
        for (int i = 0; i < 4; i++) {
          write_block(b, coeff_nc(d, mx, my, m, c, i & 1, i >> 1), cac[c][i] + 1, 15);
          int tc = 0;
          for (int k = 1; k < 16; k++) tc += cac[c][i][k] != 0;
          m.tcc[c][i] = (uint8_t)tc;
        }
    // reconstruction: the decoder's own steps on the same levels
    for (int blk = 0; blk < 16; blk++) {
      const int bx = blk_x(blk), by = blk_y(blk);
      if ((cbp_l >> (blk >> 2)) & 1) luma_residual(d, mx, my, bx, by, acl[blk], qp, false, 0);
    }
    if (cbp_c)
      for (int c = 0; c < 2; c++) chroma_residual(d, mx, my, qpc, cdc[c], cac[c], c ? P.v : P.u);
    e->n_inter++;
  }
  if (skip_run) b.ue((uint32_t)skip_run);
  b.trailing();
  return PCME_OK;
}

}  // namespace

extern "C" {

int pcme_h264enc_create(int32_t width, int32_t height, pcme_h264enc **out) {
  if (!out || width < 16 || height < 16 || (width & 1) || (height & 1) || width > 8192 || height > 8192)
    return fail("pcme_h264enc_create: bad arguments");
  Enc *e = new Enc();
  e->w = width;
  e->h = height;
  e->mbw = (width + 15) / 16;
  e->mbh = (height + 15) / 16;
  Sps &s = e->sps;
  s.valid = true;
  s.profile = 66;
  s.level = 51;
  s.log2_max_frame_num = 16;
  s.poc_type = 2;
  s.max_num_ref_frames = PCME_MAX_REFS;
  s.mbw = e->mbw;
  s.mbh = e->mbh;
  s.crop[1] = (16 * e->mbw - width) / 2;
  s.crop[3] = (16 * e->mbh - height) / 2;
  e->pps.valid = true;
  *out = e;
  return PCME_OK;
}

void pcme_h264enc_destroy(pcme_h264enc *enc) { delete enc; }

int pcme_h264enc_picture(pcme_h264enc *e, const uint8_t *y, int32_t ystride, const uint8_t *u, const uint8_t *v,
                         int32_t cstride, const pcme_mb_out *mbs, int32_t qp, uint8_t *out, size_t cap, size_t *len) {
  if (!e || !y || !u || !v || !out || !len || ystride < e->w || cstride < e->w / 2 || qp < 0 || qp > 51)
    return fail("pcme_h264enc_picture: bad arguments");
  if (mbs && e->dpb.empty()) return fail("pcme_h264enc_picture: a P picture before the first IDR");
  if (mbs)
    for (int i = 0; i < e->mbw * e->mbh; i++) {
      if (mbs[i].part > 3) return fail("MB %d: partition %d", i, mbs[i].part);
      if (mbs[i].ref >= PCME_MAX_REFS) return fail("MB %d: reference index %d", i, mbs[i].ref);
      for (int k = 0; k < 16; k++)
        if (std::abs(mbs[i].mv[k].x) > 2048 || std::abs(mbs[i].mv[k].y) > 2048)
          return fail("MB %d: MV (%d, %d) beyond +-2048", i, mbs[i].mv[k].x, mbs[i].mv[k].y);
    }
  load_source(e, y, ystride, u, v, cstride);
  std::vector<uint8_t> nals;
  if (!mbs) {
    if (!e->started) write_sps_pps(e, nals);
    e->started = true;
    e->frame_num = 0;
    e->dpb.clear();  // an IDR picture empties the reference list (8.2.5.1)
  }
  begin_picture(e);
  BitWriter b;
  const int rc = mbs ? code_p(e, mbs, qp, b) : code_idr(e, qp, b);
  if (rc) {
    e->d.cur.reset();
    return rc;
  }
  end_picture(e);
  put_nal(nals, mbs ? 2 : 3, mbs ? 1 : 5, b.bytes());
  if (nals.size() > cap) return fail("pcme_h264enc_picture: %zu bytes do not fit in %zu", nals.size(), cap);
  memcpy(out, nals.data(), nals.size());
  *len = nals.size();
  return PCME_OK;
}

int pcme_h264enc_recon(pcme_h264enc *e, uint8_t *y, int32_t ystride, uint8_t *u, uint8_t *v, int32_t cstride) {
  if (!e || !y || !u || !v || e->dpb.empty() || ystride < e->w || cstride < e->w / 2)
    return fail("pcme_h264enc_recon: bad arguments");
  const Picture &p = *e->dpb.front();
  for (int r = 0; r < e->h; r++) memcpy(y + (size_t)r * ystride, &p.y[(size_t)r * p.w], e->w);
  for (int r = 0; r < e->h / 2; r++) {
    memcpy(u + (size_t)r * cstride, &p.u[(size_t)r * (p.w / 2)], e->w / 2);
    memcpy(v + (size_t)r * cstride, &p.v[(size_t)r * (p.w / 2)], e->w / 2);
  }
  return PCME_OK;
}

int pcme_h264enc_set_refs(pcme_h264enc *e, int32_t nref) {
  if (!e || nref < 1 || nref > PCME_MAX_REFS) return fail("pcme_h264enc_set_refs: nref %d not in [1, %d]", nref,
                                                         PCME_MAX_REFS);
  e->nref = nref;
  return PCME_OK;
}

int pcme_h264enc_stats(pcme_h264enc *e, int32_t *skipped, int32_t *inter, int32_t *pcm) {
  if (!e) return fail("pcme_h264enc_stats: null coder");
  if (skipped) *skipped = e->n_skip;
  if (inter) *inter = e->n_inter;
  if (pcm) *pcm = e->n_pcm;
  return PCME_OK;
}

}  // extern "C"
