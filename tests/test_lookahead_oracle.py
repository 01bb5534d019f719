"""Pins of the oracle's lookahead restatement (lookahead.rs LowRes::new and
complexity, mbtree.rs intra_cost, search and analyse): closed forms the
definitions fix, exact recoveries of planted motion, and a second,
independently written Python restatement on small frames that prices with a
textbook Hadamard SATD instead of the oracle's butterflies."""
from __future__ import annotations

import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

from oracle import oracle as O

H4 = np.array([[1, 1, 1, 1], [1, -1, 1, -1], [1, 1, -1, -1], [1, -1, -1, 1]], np.int64)


def satd_textbook(a: np.ndarray, b: np.ndarray) -> int:
    """sum |H D H^T| / 2 over the 4x4 blocks of a - b (the reference's
    max(|x|,|y|) butterflies compute this half sum exactly)."""
    d = a.astype(np.int64) - b.astype(np.int64)
    t = 0
    for y in range(0, d.shape[0], 4):
        for x in range(0, d.shape[1], 4):
            t += int(np.abs(H4 @ d[y:y + 4, x:x + 4] @ H4.T).sum())
    return t // 2


def smooth(shape, seed, sigma=2.0, lo=20, hi=235):
    """A textured but smooth frame: SAD has one clear minimum at the true
    displacement for shifts of a few pixels."""
    rng = np.random.default_rng(seed)
    f = gaussian_filter(rng.normal(size=shape), sigma)
    f = (f - f.min()) / (f.max() - f.min())
    return (lo + f * (hi - lo)).round().astype(np.uint8)


# ------------------------------------------------------------------ LowRes


def test_lowres_of_constant_and_2x2_constant_frames():
    pw, ph = 64, 48
    plane = np.full((ph + 2 * O.LUMA_PAD, pw + 2 * O.LUMA_PAD), 7, np.uint8)
    assert np.all(O.lowres(plane, pw, ph) == 7)
    rng = np.random.default_rng(3)
    small = rng.integers(0, 256, (ph // 2, pw // 2)).astype(np.uint8)
    plane[O.LUMA_PAD:O.LUMA_PAD + ph, O.LUMA_PAD:O.LUMA_PAD + pw] = np.kron(small, np.ones((2, 2), np.uint8))
    assert np.array_equal(O.lowres(plane, pw, ph), small)


def test_lowres_rounds_half_up_and_reads_the_coded_area_only():
    pw, ph = 32, 16
    plane = np.full((ph + 2 * O.LUMA_PAD, pw + 2 * O.LUMA_PAD), 255, np.uint8)  # padding must not leak in
    area = plane[O.LUMA_PAD:O.LUMA_PAD + ph, O.LUMA_PAD:O.LUMA_PAD + pw]
    area[:] = 0
    area[0, 1] = 1          # sum 1 -> (1 + 2) >> 2 = 0
    area[0, 2:4] = 1        # sum 2 -> 1 (half rounds up)
    area[2, 4:6] = 3        # sum 6 -> 2
    area[3, 4] = 1          # (sum 7) -> 2
    area[4:6, 6:8] = 255    # 1020 -> 255
    lr = O.lowres(plane, pw, ph)
    assert lr.shape == (ph // 2, pw // 2)
    assert lr[0, 0] == 0 and lr[0, 1] == 1 and lr[1, 2] == 2 and lr[2, 3] == 255
    assert lr.sum() == 0 + 1 + 2 + 255


# -------------------------------------------------------------- intra_cost


def test_intra_cost_is_16_for_exactly_predicted_blocks():
    f = smooth((32, 40), 1)
    # top-left block flat 128: mode 0's 128 prediction is exact
    f[0:8, 0:8] = 128
    assert O.lowres_intra(f, 0, 0) == 16
    # a block whose rows repeat the column to its left: mode 1 exact
    f[8:16, 16:24] = f[8:16, 15:16]
    assert O.lowres_intra(f, 16, 8) == 16
    # a block whose columns repeat the row above it: mode 2 exact
    f[16:24, 8:16] = f[15:16, 8:16]
    assert O.lowres_intra(f, 8, 16) == 16


@pytest.mark.parametrize("d", [1, 3, -5, 12])
def test_intra_cost_of_a_constant_residual(d):
    """A constant residual d over the left-column prediction costs 8|d| per
    4x4 (one DC coefficient 16d, halved) = 32|d| over 8x8, plus 16, when the
    other two predictions are far off."""
    f = np.zeros((24, 24), np.uint8)
    left = 100 + 10 * np.arange(8)
    f[8:16, 7] = left
    f[7, 8:16] = 250  # top row far from the block: modes 0 and 2 lose
    f[8:16, 8:16] = (left + d)[:, None]
    assert O.lowres_intra(f, 8, 8) == 32 * abs(d) + 16


def _intra_py(f: np.ndarray, x0: int, y0: int) -> int:
    blk = f[y0:y0 + 8, x0:x0 + 8]
    preds = []
    if x0 > 0 and y0 > 0:
        top = f[y0 - 1, x0:x0 + 8].astype(np.int64)
        left = f[y0:y0 + 8, x0 - 1].astype(np.int64)
        preds.append(((top[None, :] + left[:, None]) // 2).astype(np.uint8))
    else:
        preds.append(np.full((8, 8), 128, np.uint8))
    if x0 > 0:
        preds.append(np.repeat(f[y0:y0 + 8, x0 - 1:x0], 8, axis=1))
    if y0 > 0:
        preds.append(np.repeat(f[y0 - 1:y0, x0:x0 + 8], 8, axis=0))
    return min(satd_textbook(blk, p) for p in preds) + 16


def test_intra_cost_matches_textbook_satd_over_the_three_predictions():
    rng = np.random.default_rng(11)
    for trial in range(4):
        f = rng.integers(0, 256, (32, 48)).astype(np.uint8) if trial % 2 else smooth((32, 48), trial)
        for y0 in range(0, 32, 8):
            for x0 in range(0, 48, 8):
                assert O.lowres_intra(f, x0, y0) == _intra_py(f, x0, y0), (trial, x0, y0)


# -------------------------------------------------------------- complexity


def test_complexity_without_a_previous_frame_is_all_intra():
    f = smooth((48, 64), 4)
    intra, inter, bi, bp, mv = O.complexity(f, None)
    assert intra == inter == int(bi.sum()) and np.array_equal(bi, bp) and not mv.any()


def test_complexity_of_identical_frames_costs_8_per_block():
    f = smooth((48, 64), 5)
    intra, inter, bi, bp, mv = O.complexity(f, f.copy())
    assert inter == 8 * bi.size and np.all(bp == 8) and not mv.any()
    assert intra == int(bi.sum()) and intra > inter


@pytest.mark.parametrize("dx,dy", [(2, 0), (-1, 0), (0, -2), (0, 1)])
def test_complexity_recovers_a_pure_axis_shift(dx, dy):
    """prev shifted so that cur's block at (x, y) sits at (x + dx, y + dy):
    the first diamond round holds the exact match (SAD 0, unique on a
    textured frame), so every block whose match lies inside the frame gets
    MV (dx, dy) and inter cost 8."""
    prev = smooth((64, 96), 6)
    cur = np.roll(prev, (-dy, -dx), axis=(0, 1))
    _, _, _, bp, mv = O.complexity(cur, prev)
    bh, bw = mv.shape[:2]
    n = 0
    for by in range(bh):
        for bx in range(bw):
            x, y = 8 * bx + dx, 8 * by + dy
            # blocks whose rolled content wrapped around are not pure shifts
            wrapped = (dx < 0 and bx == 0) or (dx > 0 and bx == bw - 1) or (dy < 0 and by == 0) or \
                      (dy > 0 and by == bh - 1)
            if 0 <= x <= 96 - 8 and 0 <= y <= 64 - 8 and not wrapped:
                assert tuple(mv[by, bx]) == (dx, dy), (bx, by)
                assert bp[by * bw + bx] == 8
                n += 1
    assert n >= (bh - 1) * (bw - 1)


def _complexity_py(cur, prev):
    h, w = cur.shape
    ds = [(-1, 0), (1, 0), (0, -1), (0, 1), (-2, 0), (2, 0), (0, -2), (0, 2)]
    ci = cp = 0
    mvs = []
    for y0 in range(0, h - h % 8, 8):
        for x0 in range(0, w - w % 8, 8):
            blk = cur[y0:y0 + 8, x0:x0 + 8].astype(np.int64)
            ic = _intra_py(cur, x0, y0)
            ci += ic

            def sad(mx, my):
                x, y = x0 + mx, y0 + my
                if x < 0 or y < 0 or x + 8 > w or y + 8 > h:
                    return None
                return int(np.abs(blk - prev[y:y + 8, x:x + 8]).sum())

            best, bc = (0, 0), sad(0, 0)
            for _ in range(8):
                cands = [(best[0] + a, best[1] + b) for a, b in ds]
                costs = [(sad(*m), i, m) for i, m in enumerate(cands)]
                costs = [c for c in costs if c[0] is not None and c[0] < bc]
                if not costs:
                    break
                bc, _, best = min(costs)
            sc = satd_textbook(blk, prev[y0 + best[1]:y0 + best[1] + 8, x0 + best[0]:x0 + best[0] + 8]) + 8
            cp += min(sc, ic)
            mvs.append(best)
    return ci, cp, mvs


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_complexity_matches_an_independent_restatement(seed):
    rng = np.random.default_rng(seed)
    prev = smooth((40, 56), seed + 20, sigma=1.5)
    cur = np.roll(prev, (rng.integers(-3, 4), rng.integers(-3, 4)), axis=(0, 1))
    cur = np.clip(cur.astype(np.int64) + rng.integers(-6, 7, cur.shape), 0, 255).astype(np.uint8)
    if seed == 2:
        cur[:, :] = rng.integers(0, 256, cur.shape)  # no motion to find
    intra, inter, _, _, mv = O.complexity(cur, prev)
    ci, cp, mvs = _complexity_py(cur, prev)
    assert (intra, inter) == (ci, cp)
    assert [tuple(m) for m in mv.reshape(-1, 2)] == mvs


# ------------------------------------------------------------ mbtree search


def test_tree_analyse_of_identical_frames():
    f = smooth((48, 64), 7)
    blocks = O.tree_analyse(f, f.copy(), None)
    assert np.all(blocks["list"] == 1) and np.all(blocks["inter"] == 8) and not blocks["mv"].any()
    # both lists exact: list 1 wins at cost 8 and the others tie or lose
    blocks = O.tree_analyse(f, f.copy(), f.copy())
    assert np.all(blocks["list"] == 1) and np.all(blocks["inter"] == 8)
    assert np.array_equal(blocks["intra"], O.complexity(f, None)[2])


def test_tree_search_recovers_a_diagonal_shift():
    """(-2,-2) is one of the pattern's steps: SAD 0 plus MV cost 4 beats
    every misaligned candidate on a textured frame; later blocks of the row
    start from it through the left seed."""
    ref = smooth((64, 96), 8, sigma=1.2)
    cur = np.roll(ref, (2, 2), axis=(0, 1))  # cur(x, y) = ref(x - 2, y - 2)
    blocks = O.tree_analyse(cur, ref, None).reshape(8, 12)
    inner = blocks[1:, 1:]
    assert np.all(inner["mv"][..., 0, 0] == -2) and np.all(inner["mv"][..., 0, 1] == -2)
    assert np.all(inner["inter"] == 8) and np.all(inner["list"] == 1)
    assert np.all(inner["mv"][..., 1, :] == 0)


def test_tree_bi_average_wins_when_it_is_exact():
    """l0 = cur + 6 and l1 = cur - 6: each list alone leaves a constant
    residual (SATD 4 * 8 * 6 = 192), the rounded average reproduces cur
    (SATD 0, cost 12)."""
    rng = np.random.default_rng(9)
    cur = rng.integers(6, 250, (32, 48)).astype(np.uint8)
    l0 = (cur + 6).astype(np.uint8)
    l1 = (cur - 6).astype(np.uint8)
    b = O.tree_analyse(cur, l0, l1)
    assert np.all(b["list"] == 3) and np.all(b["inter"] == 12) and not b["mv"].any()
    b0 = O.tree_analyse(cur, l0, None)
    assert np.all(b0["list"] == 1) and np.all(b0["inter"] == 192 + 8)


def _tree_py(cur, l0, l1):
    h, w = cur.shape
    pat = [(-1, 0), (1, 0), (0, -1), (0, 1), (-2, -2), (2, -2), (-2, 2), (2, 2)]
    out = []
    for y0 in range(0, h - h % 8, 8):
        left = [(0, 0), (0, 0)]
        for x0 in range(0, w - w % 8, 8):
            blk = cur[y0:y0 + 8, x0:x0 + 8].astype(np.int64)
            intra = _intra_py(cur, x0, y0)
            best = (intra, 0, [(0, 0), (0, 0)])
            mvs = [(0, 0), (0, 0)]
            preds = [None, None]
            for l, r in enumerate((l0, l1)):
                if r is None:
                    continue

                def cost(m):
                    x, y = x0 + m[0], y0 + m[1]
                    if x < 0 or y < 0 or x > w - 8 or y > h - 8:
                        return None
                    return int(np.abs(blk - r[y:y + 8, x:x + 8]).sum()) + abs(m[0]) + abs(m[1])

                bm, bc = (0, 0), cost((0, 0))
                sc = cost(left[l])
                if sc is not None and sc < bc:
                    bm, bc = left[l], sc
                for _ in range(16):
                    cands = [(bm[0] + a, bm[1] + b) for a, b in pat]
                    better = [(cost(m), i, m) for i, m in enumerate(cands)]
                    better = [c for c in better if c[0] is not None and c[0] < bc]
                    if not better:
                        break
                    bc, _, bm = min(better)
                mvs[l] = bm
                preds[l] = r[y0 + bm[1]:y0 + bm[1] + 8, x0 + bm[0]:x0 + bm[0] + 8].astype(np.int64)
                c = satd_textbook(blk, preds[l]) + 8
                if c < best[0]:
                    best = (c, 1 + l, list(mvs))
            if l0 is not None and l1 is not None:
                c = satd_textbook(blk, (preds[0] + preds[1] + 1) >> 1) + 12
                if c < best[0]:
                    best = (c, 3, list(mvs))
            left = list(mvs)
            out.append((intra, best[0], best[1], best[2]))
    return out


@pytest.mark.parametrize("seed,bi", [(0, False), (1, True), (2, True)])
def test_tree_analyse_matches_an_independent_restatement(seed, bi):
    rng = np.random.default_rng(seed + 40)
    base = smooth((40, 64), seed + 30, sigma=1.3)
    cur = np.clip(base.astype(np.int64) + rng.integers(-4, 5, base.shape), 0, 255).astype(np.uint8)
    l0 = np.roll(base, (rng.integers(-5, 6), rng.integers(-5, 6)), axis=(0, 1))
    l1 = np.roll(base, (rng.integers(-5, 6), rng.integers(-5, 6)), axis=(0, 1)) if bi else None
    got = O.tree_analyse(cur, l0, l1)
    want = _tree_py(cur, l0, l1)
    assert len(got) == len(want)
    for g, (intra, inter, lst, mv) in zip(got, want):
        assert (int(g["intra"]), int(g["inter"]), int(g["list"])) == (intra, inter, lst)
        assert [tuple(int(v) for v in m) for m in g["mv"]] == mv
