"""Pins of the oracle's ME-only B field (o_b_field: analyse_b's searches and
decision with rd off, without BSkip and intra, MB by MB in raster order,
slice.rs:2189-2377, committed as cand_inter_info does, slice.rs:1751-1785):

* every MB's search record equals o_b_analyse (pinned in test_b_oracle.py)
  over the field the earlier MBs committed, so the field feeds itself;
* the decision and commit, restated here from the reference's rule: direct
  when its MVs are legal, the 16x16 direction when strictly cheaper, then
  16x8 and 8x16 when partitions and the best cost exceeds 1024, each when
  strictly cheaper; ref_idx per 8x8 and MVs per 4x4, zero where a list is
  unused;
* closed forms: a B picture equal to its list-0 (list-1) picture, the other
  list noise: the first MB (no neighbours: spatial direct is bi-predicted
  at the zero MV) takes that list's 16x16 at the zero MV, cost
  (2 lam >> 8) + (3 lam >> 8); every later MB inherits the list through
  direct_spatial, whose prediction is then exact: direct, cost
  (2 lam) >> 8."""
from __future__ import annotations

import numpy as np
import pytest

LAM_QP = 27


def _setup(O, w, h, seed, cur=1):
    import pcme
    mbw, mbh = w // 16, h // 16
    y, _, _ = pcme.synth_clip(pcme.ClipSpec(w, h, 3, seed=seed, nrect=3, rect_lo=16, rect_hi=40, vmax=4.0))
    rs = np.random.default_rng(seed & 0xFFFF)
    cols = (O.ombinfo2 * (mbw * mbh))()
    for i in range(mbw * mbh):  # the list-1 picture's MBs: inter, some near-still
        cols[i].kind = O.OK_PINTER if rs.random() > 0.1 else O.OK_I16x16
        for l in range(2):
            for q in range(4):
                cols[i].ref_idx[l][q] = int(rs.choice([-1, 0, 0]))
            for b in range(16):
                cols[i].mv[l][b] = O.omv(int(rs.integers(-6, 7)), int(rs.integers(-6, 7)))
    return y, mbw, mbh, cols


def _decide(rec, partitions):
    s = rec["search"]
    mode, best = -1, 0
    if s["direct_ok"]:
        mode, best = 0, int(s["direct_cost"])
    c16 = int(s["cost"][int(s["dir"])])
    if mode < 0 or c16 < best:
        mode, best = 1, c16
    if partitions and best > 1024:
        for k in range(2):
            if int(s["part_cost"][k]) < best:
                mode, best = 2 + k, int(s["part_cost"][k])
    ref = np.zeros((2, 4), np.int8)
    mv = np.zeros((2, 16, 2), np.int16)
    for l in range(2):
        for r in range(16):
            q = (r >> 3) * 2 + ((r & 3) >> 1)
            if mode == 0:
                rf, m = int(s["direct_ref"][l][q]), s["direct_mv"][l][q]
            elif mode == 1:
                d = int(s["dir"])
                rf, m = (0 if d in (2, l) else -1), s["mv16"][l]
            else:
                k = mode - 2
                pi = (r >> 3) if k == 0 else ((r & 3) >> 1)
                d = int(s["part_dir"][k][pi])
                rf, m = (0 if d in (2, l) else -1), s["part_mv"][k][l][pi]
            ref[l][q] = rf
            mv[l][r] = m if rf >= 0 else (0, 0)
    return mode, best, ref, mv


@pytest.mark.parametrize("seed,partitions", [(0x5EED1800, 1), (0x5EED1801, 1), (0x5EED1802, 0)])
def test_b_field_feeds_itself_and_decides_by_the_rule(oracle, seed, partitions):
    O = oracle
    w, h = 96, 64
    y, mbw, mbh, cols = _setup(O, w, h, seed)
    src = O.pad_luma(y[1], mbw, mbh)
    r0, r1 = O.Ref(O.pad_luma(y[0], mbw, mbh)), O.Ref(O.pad_luma(y[2], mbw, mbh))
    lam = O.load().o_lambda(LAM_QP)
    prm = O.oparams(mbw, mbh, 16, 1, partitions, 0)
    got = O.b_field(prm, cols, r0, r1, src, lam)
    field = (O.ombinfo2 * (mbw * mbh))()
    for i, r in enumerate(got):
        field[i].kind = O.OK_PINTER
        for l in range(2):
            for q in range(4):
                field[i].ref_idx[l][q] = int(r["ref"][l][q])
            for b in range(16):
                field[i].mv[l][b] = O.omv(int(r["mv"][l][b][0]), int(r["mv"][l][b][1]))
    modes = set()
    for my in range(mbh):
        p = O.oparams(mbw, mbh, 16, 1, partitions, O.load().o_vlim(my, mbh))
        for mx in range(mbw):
            i = my * mbw + mx
            s16 = src[64 + 16 * my:80 + 16 * my, 64 + 16 * mx:80 + 16 * mx]
            want = O.b_analyse(p, field, cols[i], mx, my, r0, r1, s16, lam)
            assert got[i]["search"].tobytes() == want.tobytes(), (mx, my)
            mode, best, ref, mv = _decide(got[i], partitions)
            assert (int(got[i]["mode"]), int(got[i]["best_cost"])) == (mode, best), (mx, my)
            assert np.array_equal(got[i]["ref"], ref) and np.array_equal(got[i]["mv"], mv), (mx, my)
            modes.add(mode)
    assert len(modes) >= 2


@pytest.mark.parametrize("lst", [0, 1])
def test_b_picture_equal_to_a_reference(oracle, lst):
    O = oracle
    w, h = 64, 48
    y, mbw, mbh, cols = _setup(O, w, h, 0x5EED1810)
    noise = np.random.default_rng(3).integers(0, 256, (h, w)).astype(np.uint8)
    pics = [y[0], noise] if lst == 0 else [noise, y[0]]
    src = O.pad_luma(y[0], mbw, mbh)
    r0, r1 = (O.Ref(O.pad_luma(p, mbw, mbh)) for p in pics)
    lam = int(O.load().o_lambda(LAM_QP))
    got = O.b_field(O.oparams(mbw, mbh, 16, 1, 1, 0), cols, r0, r1, src, lam)
    assert got[0]["mode"] == 1 and got[0]["dir"] == lst
    assert got[0]["best_cost"] == ((2 * lam) >> 8) + ((3 * lam) >> 8)
    assert np.all(got["mode"][1:] == 0) and np.all(got["best_cost"][1:] == (2 * lam) >> 8)
    assert np.all(got["ref"][:, lst] == 0) and np.all(got["ref"][:, 1 - lst] == -1)
    assert not got["mv"].any()
