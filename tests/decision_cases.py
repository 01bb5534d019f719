"""Closed-form decision fixtures of the motion search.

Each case is a small picture whose cost at every candidate MV can be written
down by hand, and an expected result derived by hand from the reference's
rules, in the case's comment.  They pin the decisions of `search`
(src/h264/enc/slice.rs:1617-1721) and of the partition choice
(slice.rs:1921-1945) that random clips cannot: which MV wins a tie, where
the hexagon stops, what the sub-pel phase starts from.  Nothing here runs
an implementation of the search; the expected values are data.

Picture: 4x4 macroblocks (64x64), padded plane 192x192, pad 64.

Cost surfaces used (src = a w x h block, ref = the padded reference plane):

* flat: ref = 100 everywhere, src = 100 + delta.  Every full-, half- and
  quarter-pel sample is 100 (the 6-tap filter sums to 32 and its
  rounding is exact on a constant), so SAD = w*h*|delta|, SATD =
  w*h*|delta|/2 (a constant 4x4 residual d has one Hadamard coefficient,
  16d, and SATD = sum|coeff|/2 = 8|d|).
* x-ramp: ref(X, Y) = 128 + 2*(X - XC), unclipped over every column the
  search and the filter taps touch (asserted by the tests); src(x, y) =
  ref(bx + x + T, by + y), i.e. the block shifted by T px.  On a linear
  ramp of slope 2 the half-pel samples are exact midpoints (H = G + 1,
  V = G, C = G + 1), and the quarter-pel averages (a+b+1)>>1 of H.264
  8.4.2.2.1 give, for horizontal phase f = mv.x & 3, the offsets
  OFF = {0, 1, 1, 2} over the full sample, whatever the vertical phase
  (the twelve averaging rules a, c, d, n, e, f, g, i, k, p, q, r
  evaluated on G, b = G+1, h = G, j = G+1, m = G+2, s = G+1).  So the
  residual is the same at every pixel of the block:
      D(mv) = 2*T - 2*(mv.x >> 2) - OFF[mv.x & 3]
  and SAD = w*h*|D| (integer MVs), SATD = w*h*|D|/2.
* y-ramp: the same with rows (ref(X, Y) = 128 + 2*(Y - YC)); by the
  symmetry of the averaging rules D(mv) = 2*T - 2*(mv.y >> 2) -
  OFF[mv.y & 3], whatever the horizontal phase.

MV bits: se(v) = 2*floor(log2(k+1)) + 1 with k = 2v-1 (v > 0), -2v (v <= 0)
(slice.rs:130-143); cost term (lam * (se(dx) + se(dy))) >> 8.  Values
used: se(0)=1, se(+-1)=3, se(+-2)=se(+-3)=5, se(+-4..7)=7, se(+-8..15)=9,
se(v)=17 for 128 <= |v| <= 255, se(v)=21 for -1023 <= v <= -512.
With mvp = (640, 640) every MV the search reaches has v in
[-1023, -512] in both components, so the bits are the constant 42: a
"plateau" on which only distortion decides (and ties stay put).

Legality (slice.rs:496-507) for a w x h block at (bx, by) of this picture:
x0 = bx + (mv.x >> 2) >= -44, x0 + w <= 108, y0 = by + (mv.y >> 2) >= -44,
y0 + h < vlim.  Note mv.x >> 2 floors: mv.x = -177 gives x0 = bx - 45.
"""
from __future__ import annotations

from dataclasses import dataclass, field

MBW = MBH = 4
PAD = 64
STRIDE = 16 * MBW + 2 * PAD  # 192
ROWS = 16 * MBH + 2 * PAD
PLATEAU = (640, 640)


@dataclass
class SearchCase:
    """One call of search(): expected (mv, cost) derived by hand."""
    name: str
    surface: str            # "flat", "xramp", "yramp"
    bx: int
    by: int
    w: int
    h: int
    mvp: tuple
    lam: int
    me_range: int
    subpel: int
    vlim: int
    want_mv: tuple
    want_cost: int
    shift: int = 0          # T: the source block is the reference shifted by T px
    delta: int = 0          # flat surface: src = ref + delta
    centre: int = 0         # ramp centre column/row (padded coordinates)
    touch: tuple = (0, 0)   # padded columns/rows the search and its taps may read
    cands: tuple = ()
    want_evals: tuple | None = None  # (sad_evals, satd_evals) of the oracle's counters


def _v(my: int) -> int:
    # vlim of MB row `my` of this 4-row picture (frame.rs:152, picture.rs:261-265):
    # 16*(my+1)+96 >= 64 for every row, so vlim = 16*mbh + 44.
    return 16 * MBH + 44


SEARCH_CASES = [
    # --- (mvp + 2) & !3 on a negative odd predictor, x (slice.rs:1617-1620) -----------------
    # flat, delta 0: SAD = SATD = 0, cost = bits only; block 16x16 at (0, 16), lam 256.
    # start: ((-177+2)&~3, (-3+2)&~3) = (-176, -4): x0 = -44, legal.  v = mv - mvp = (1, -1): 6.
    #   (floor rounding would give -180, x0 = -45, illegal -> start (0, 0): another result.)
    # hexagon: (-184,-4) (-180,-12) (-180,4) illegal; (-168,-4) v=(9,-1) 12; (-172,-12) v=(5,-9)
    #   16; (-172,4) v=(5,7) 14 -> no move.  squares: (-172,-4) 10, (-176,-8) 10, (-176,0) 8,
    #   (-172,-8) 14, (-172,0) 12, three illegal -> no move.  SATD restart: 6.
    # half: (-174,-4) 8, (-176,-6) 8, (-176,-2) v=(1,1) 6 (tie, kept), (-174,-6) 10,
    #   (-174,-2) 8, three illegal -> no move.
    # quarter 1: (-175,-4) 8, (-176,-5) 8, (-176,-3) v=(1,0) 4 -> taken; (-175,-5) 10,
    #   (-175,-3) 6.  quarter 2 around (-176,-3): (-175,-3) 6, (-176,-4) 6, (-176,-2) 6,
    #   (-175,-4) 8, (-175,-2) 8, three illegal -> break.  Result (-176,-3), cost 4.
    # evals: SAD 1 + 3 (hex) + 5 (square); SATD 1 + 5 + 5 + 5.
    SearchCase("round_neg_odd_x", "flat", 0, 16, 16, 16, (-177, -3), 256, 16, 1, _v(1),
               (-176, -3), 4, want_evals=(9, 16)),
    # --- the same rounding in y at the top edge, where the rounded start is illegal ------------
    # flat, block 16x16 at (16, 0), mvp (-3, -179): start (-4, (-177)&~3 = -180): y0 = -45,
    #   illegal -> (0, 0) (truncating toward zero would give -176, legal, and another result).
    #   v = (3, 179): se 5 + 17 = 22.
    # hex 1 around (0,0): (-8,0) v=(-5,179) 24; (8,0) 26; (-4,-8) v=(-1,171) 20 -> taken;
    #   (4,-8) 24; (-4,8) 20 tie; (4,8) 24.  hex 2 around (-4,-8): (-12,-8) 26, (4,-8) 24,
    #   (-8,-16) 24, (0,-16) 22, (-8,0) 24, (0,0) 22 -> stop.
    # squares around (-4,-8): (-8,-8) 24, (0,-8) 22, (-4,-12) 20, (-4,-4) 20, corners 22-24.
    # SATD 20.  half: (-6,-8) 22, (-2,-8) v=(1,171) 20 tie, (-4,-10)/(-4,-6) 20 ties.
    # quarter 1: (-5,-8) 22, (-3,-8) v=(0,171) 18 -> taken; (-3,-9), (-3,-7) 18 ties.
    # quarter 2 around (-3,-8): (-4,-8) 20, (-2,-8) 20, (-3,-9)/(-3,-7) 18 ties -> break.
    # Result (-3,-8), cost 18.
    SearchCase("round_neg_odd_y_top_edge", "flat", 16, 0, 16, 16, (-3, -179), 256, 16, 1, _v(0),
               (-3, -8), 18),
    # --- strict < between two equal hexagon points, HEX order (slice.rs:1634-1656) ----------
    # x-ramp, T = -1: D(0) = -2, D(-4) = 0, D(-8) = 2, D(4) = -4.  Plateau bits 42, lam 256.
    # start (640,640) illegal -> (0,0): 256*2 + 42 = 554.
    # hex 1: (-8,0) D=2 554 tie; (8,0) D=-6; (-4,-8) D=0 -> 42 taken; (4,-8) D=-4; (-4,8) 42 tie
    #   (kept: the point (-1,-2) comes before (-1,2)); (4,8).  hex 2 around (-4,-8): no point
    #   on the column x = -4, every |D| >= 2 -> stop.
    # squares: (-4,-12), (-4,-4) 42 ties, the rest |D| = 2 -> no move.  SATD 128*0 + 42.
    # half: (-6,*) D = -2-2(-2)-1 = 1, (-2,*) D = -1 -> 170; (-4,-10), (-4,-6) ties.
    # quarter: (-5,*) D = -2-2(-2)-2 = 0 ties, (-3,*) D = -1.  Result (-4,-8), cost 42.
    # (HEX order (-1,2) first would give (-4,8); "<=" would take (-8,0) first.)
    SearchCase("hex_tie_first_in_order", "xramp", 16, 16, 16, 16, PLATEAU, 256, 16, 1, _v(1),
               (-4, -8), 42, shift=-1, centre=PAD + 16 + 8, touch=(PAD + 16 - 12, PAD + 16 + 16 + 12)),
    # --- the hexagon's box |m - centre0| >> 2 <= range (slice.rs:1640-1641) -------------------
    # x-ramp, T = -3: D(0) = -6, D(-4) = -4, D(-8) = -2, D(-12) = 0, D(-16) = 2.
    # me_range 1: every hexagon point has a component of 8 qpel, 8>>2 = 2 > 1: none is priced.
    #   squares: (-4,0) D=-4 taken, then (-8,0) D=-2 (1066 -> 554).  SATD 128*2+42 = 298.
    #   half: (-10,0) D = -6-2(-3)-1 = -1 -> 170 taken; (-6,0) D=-3; (-8,+-2) 298; (-10,+-2) 170
    #   ties.  quarter: (-11,0) D=-1 tie, (-9,0) D=-2, (-10,+-1) ties -> (-10,0), cost 170.
    SearchCase("hex_box_range1", "xramp", 16, 16, 16, 16, PLATEAU, 256, 1, 1, _v(1),
               (-10, 0), 170, shift=-3, centre=PAD + 16 + 8, touch=(PAD + 16 - 12, PAD + 16 + 16 + 12)),
    # me_range 2: hex 1 takes (-8,0) (554; (-4,+-8) D=-4).  hex 2 around (-8,0): (-16,0) has
    #   16>>2 = 4 > 2 and (-12,+-8) 12>>2 = 3 > 2: skipped; (-4,+-8), (0,0) worse -> stop.
    #   squares (no box): (-12,0) D=0 -> 42 taken; then around (-12,0) no move.
    #   sub-pel: (-14,0) D=1, (-10,0) D=-1, (-13,0) D=0 tie, (-12,+-k) ties.  (-12,0), cost 42.
    SearchCase("hex_box_range2_square_beyond", "xramp", 16, 16, 16, 16, PLATEAU, 256, 2, 1, _v(1),
               (-12, 0), 42, shift=-3, centre=PAD + 16 + 8, touch=(PAD + 16 - 12, PAD + 16 + 16 + 12)),
    # me_range 3: hex 2 around (-8,0): (-16,0) 4 > 3 skipped, (-12,-8) 3 <= 3, D=0 -> 42 taken
    #   ((-12,8) ties).  hex 3 around (-12,-8): (-20,*), (-16,*), (*,-16) outside the box,
    #   (-4,-8), (-8,0) worse -> stop.  squares, sub-pel: only ties.  (-12,-8), cost 42.
    SearchCase("hex_box_range3", "xramp", 16, 16, 16, 16, PLATEAU, 256, 3, 1, _v(1),
               (-12, -8), 42, shift=-3, centre=PAD + 16 + 8, touch=(PAD + 16 - 12, PAD + 16 + 16 + 12)),
    # --- the SATD restart at the integer best (slice.rs:1688) --------------------------------
    # x-ramp, T = -50, block 16x16 at (0,16), me_range 64: D(4p) = -100 - 2p.  Each hexagon
    #   round takes (-8,0) first (D + 4; (-4,+-8) give only D + 2) until x0 = -44 (mv.x = -176,
    #   D = -12): 22 moves, then (-184,0), (-180,+-8) illegal, the rest worse -> stop.
    #   squares: (-176,+-4) ties, (-172,*) D=-14, (-180,*) illegal.
    # SATD at (-176,0): 128*12 + 42 = 1578.  half: (-174,0) D = -100+88-1 = -13 -> 1706, not
    #   below 1578 (it would be below the SAD cost 256*12+42 = 3114 without the restart);
    #   (-178,*) illegal; ties.  quarter: (-175,*) D=-13, (-177,*) illegal.
    # Result (-176,0), cost 1578.
    SearchCase("satd_restart_at_left_edge", "xramp", 0, 16, 16, 16, PLATEAU, 256, 64, 1, _v(1),
               (-176, 0), 1578, shift=-50, centre=PAD + 0 - 14, touch=(PAD - 50, PAD + 16 + 16)),
    # --- 1 + subpel_rounds quarter-pel rounds (slice.rs:1689-1721) ---------------------------
    # x-ramp, T = -2, 8x8 block at (24,24): SATD = 32|D|, SAD = 64|D|; mvp (-4,0), lam 5120
    #   (cost = 20 * bits).  D(-8) = 0, D(-9) = 0, D(-10) = 1, D(-7) = D(-6) = -1,
    #   D(-5) = D(-4) = -2, D(-3) = -3, D(-12) = 2, D(0) = -4.
    # start (-4,0): 128 + 20*2 = 168.  hex (me_range 4): (-12,0) 128+200, (4,0), (-8,+-8)
    #   20*(7+9) = 320, (0,+-8) -> no move.  squares: (-8,0) 0 + 20*8 = 160 taken; around it
    #   (-12,0) 328, (-4,0) 168, (-8,+-4) 280 -> stop.
    # SATD at (-8,0): 0 + 20*(7+1) = 160.  half: (-10,0) 32 + 20*(7+1) = 192, (-6,0) 32 +
    #   20*(5+1) = 152 taken, (*,+-2) add 20*4.  quarter 1 around (-6,0): (-7,0) 32 + 20*(5+1)
    #   = 152 tie, (-5,0) 64 + 20*(3+1) = 144 taken.  quarter 2 around (-5,0): (-6,0) 152,
    #   (-4,0) 64 + 20*2 = 104 taken.  quarter 3 around (-4,0): (-5,0) 144, (-3,0) 96 + 20*4
    #   = 176, (-4,+-1) 64 + 80 = 144 -> no move.
    # subpel_rounds 0 (one quarter round): (-5,0), 144; 1: (-4,0), 104; 2: the same (round 3
    #   moves nothing).
    SearchCase("quarter_rounds_sub0", "xramp", 24, 24, 8, 8, (-4, 0), 5120, 4, 0, _v(1),
               (-5, 0), 144, shift=-2, centre=PAD + 24 + 4, touch=(PAD + 24 - 12, PAD + 24 + 8 + 12)),
    SearchCase("quarter_rounds_sub1", "xramp", 24, 24, 8, 8, (-4, 0), 5120, 4, 1, _v(1),
               (-4, 0), 104, shift=-2, centre=PAD + 24 + 4, touch=(PAD + 24 - 12, PAD + 24 + 8 + 12)),
    SearchCase("quarter_rounds_sub2", "xramp", 24, 24, 8, 8, (-4, 0), 5120, 4, 2, _v(1),
               (-4, 0), 104, shift=-2, centre=PAD + 24 + 4, touch=(PAD + 24 - 12, PAD + 24 + 8 + 12)),
    # --- y0 + h < vlim (slice.rs:503-506) --------------------------------------------------------
    # y-ramp, T = +4 (the match lies below): D(0) = 8, D(4) = 6, D(6) = 5, D(7) = 4, D(8) = 4,
    #   D(9) = D(10) = 3, D(11) = 2, D(-8) = 12; block 16x16 at (16,16), plateau, me_range 4.
    # vlim 35: legal iff 16 + (mv.y >> 2) + 16 < 35, i.e. mv.y <= 11.
    #   hex 1 around (0,0) (2090): (+-8,0) ties, (*,-8) D=12, (-4,8) D=4 -> 1066 taken, (4,8)
    #   tie.  hex 2: (*,16) illegal, the rest worse or ties.  squares: (-4,12) illegal, ties.
    #   SATD 128*4+42 = 554.  half: (-4,10) D=3 -> 426 taken.  quarter 1: (-4,9) tie,
    #   (-4,11) D=2 -> 298 taken.  quarter 2: (-4,12) illegal -> stop.  (-4,11), cost 298.
    # vlim 34: legal iff mv.y <= 7.  hex: (*,8) illegal, ties -> no move.  squares: (0,4)
    #   D=6 1578 taken; around it (0,8) illegal.  SATD 128*6+42 = 810.  half (0,6) D=5 -> 682.
    #   quarter 1: (0,5) tie, (0,7) D=4 -> 554.  quarter 2: (0,8) illegal.  (0,7), cost 554.
    SearchCase("vlim_last_legal_row", "yramp", 16, 16, 16, 16, PLATEAU, 256, 4, 1, 35,
               (-4, 11), 298, shift=4, centre=PAD + 16 + 8, touch=(PAD + 16 - 12, PAD + 16 + 16 + 12)),
    SearchCase("vlim_one_row_less", "yramp", 16, 16, 16, 16, PLATEAU, 256, 4, 1, 34,
               (0, 7), 554, shift=4, centre=PAD + 16 + 8, touch=(PAD + 16 - 12, PAD + 16 + 16 + 12)),
]


@dataclass
class FrameCase:
    """A one-macroblock P frame of the ME-only field (16x16 picture):
    reference flat 100, source flat 100 + delta.  Expected record fields."""
    name: str
    delta: int
    qp: int
    want_cost: tuple
    want_best: int
    want_part: int = 0
    want_evals: tuple | None = None  # (searches, sad_evals, satd_evals)
    extra: dict = field(default_factory=dict)


NOT_SEARCHED = 0xFFFFFFFF

FRAME_CASES = [
    # --- cost16 > 1024 gates the partition modes (slice.rs:1930) ----------------------------------
    # No neighbours: mvp = skip_mv = (0,0).  Every search of a flat picture stays at (0,0): a
    # move changes no distortion and costs bits ((lam*4)>>8 >= (lam*2)>>8 at a quarter step,
    # equal when lam*4 < 256, and a tie keeps the incumbent).
    # qp 0, lam 59: cost16 = SATD 128*8 + (59*2)>>8 = 1024 -> not > 1024: no partition search.
    #   best_cost = 1024 + (59>>8) = 1024.
    #   evals: start 1, both candidates equal best (skipped), 6 hexagon points (no move),
    #   8 square points (no move): 15 SAD; SATD restart 1, half 8, quarter 8 (no move): 17.
    FrameCase("cost16_1024_no_partitions", 8, 0, (1024, NOT_SEARCHED, NOT_SEARCHED, NOT_SEARCHED), 1024,
              want_evals=(1, 15, 17)),
    # qp 8, lam 149: cost16 = 1024 + (149*2)>>8 = 1025 > 1024: 16x8, 8x16, 8x8 searched.
    #   16x8 / 8x16: two partitions of 64*8 + 1 = 513 -> 1026; 8x8: four of 32*8 + 1 -> 1028.
    #   totals with penalties: 1025 + (149>>8 = 0); 1026 + (447>>8 = 1) = 1027; 1027;
    #   1028 + (1490>>8 = 5) = 1033 -> P16x16 (part 0), best_cost 1025.
    #   evals: 9 searches of 15 SAD and 17 SATD each.
    FrameCase("cost16_1025_partitions", 8, 8, (1025, 1026, 1026, 1028), 1025, want_evals=(9, 135, 153)),
    # --- strict < between equal mode totals (slice.rs:1937-1944) ------------------------------
    # qp 0, delta 9: cost16 = 128*9 = 1152; 16x8 = 8x16 = 2*(64*9) = 1152, penalty (59*3)>>8
    #   = 0: equal to 1152 + (59>>8) = 1152, not below -> 16x16 kept; 8x8 = 4*288 + (590>>8)
    #   = 1154.  ("<=" would end on P8x16.)
    FrameCase("mode_tie_keeps_16x16", 9, 0, (1152, 1152, 1152, 1152), 1152),
]


def ref_plane(c: SearchCase):
    """The padded reference plane of a search case (uint8 [ROWS, STRIDE])."""
    import numpy as np
    if c.surface == "flat":
        return np.full((ROWS, STRIDE), 100, np.uint8)
    idx = np.arange(STRIDE if c.surface == "xramp" else ROWS)
    lin = 128 + 2 * (idx - c.centre)
    lo, hi = c.touch
    assert lin[lo:hi + 1].min() >= 0 and lin[lo:hi + 1].max() <= 255, "ramp clipped where the search reads"
    line = np.clip(lin, 0, 255).astype(np.uint8)
    if c.surface == "xramp":
        return np.ascontiguousarray(np.broadcast_to(line[None, :], (ROWS, STRIDE)))
    return np.ascontiguousarray(np.broadcast_to(line[:, None], (ROWS, STRIDE)))


def src_block(c: SearchCase, plane):
    """The w x h source block (stride w): the reference shifted by `shift`
    px (ramps) or ref + delta (flat)."""
    import numpy as np
    if c.surface == "flat":
        return np.full((c.h, c.w), 100 + c.delta, np.uint8)
    x0, y0 = PAD + c.bx, PAD + c.by
    if c.surface == "xramp":
        x0 += c.shift
    else:
        y0 += c.shift
    return np.ascontiguousarray(plane[y0:y0 + c.h, x0:x0 + c.w])


def frame_pair(c: FrameCase):
    """[reference frame, source frame] of a one-MB frame case (16x16)."""
    import numpy as np
    return np.stack([np.full((16, 16), 100, np.uint8), np.full((16, 16), 100 + c.delta, np.uint8)])
