"""Pins of the oracle's B-macroblock restatement (slice.rs direct_spatial
2030-2084, b_partition 2086-2187, analyse_b 2189-2377):

* direct_spatial against spatial direct restated in the reference decoder's
  structure (dec/slicedec.rs:718-760: nb_abc, min_pos, median_pred, the
  co-located block of direct_8x8_inference) over random neighbourhoods;
* every cost o_b_analyse reports, recomputed from its MVs with the H.264
  quarter-sample equations and the textbook Hadamard (tests/pysearch.py),
  the rounded bi-average of picture.rs:491-495, the Exp-Golomb MV bits and
  the fixed direction bits; the direction choices as first minima.

The searches themselves are o_search_part on a list's projection, pinned
with the P path."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import pysearch
from test_oracle_search import spec_mvp

LAM_QP = 27


def _rand_mbinfo2(O, m, rs, intra_p=0.15):
    m.kind = O.OK_I16x16 if rs.random() < intra_p else int(rs.choice([O.OK_PINTER, O.OK_PSKIP]))
    for l in range(2):
        for q in range(4):
            m.ref_idx[l][q] = int(rs.choice([-1, 0, 0, 1]))
        for b in range(16):
            m.mv[l][b] = O.omv(int(rs.integers(-20, 21)), int(rs.integers(-20, 21)))


def _list_view(O, m, l):
    """The neighbour description spec_mvp expects, for list l."""
    return {"intra": m.kind in (O.OK_I4x4, O.OK_I8x8, O.OK_I16x16), "ref": [m.ref_idx[l][q] for q in range(4)],
            "mv": [(m.mv[l][b].x, m.mv[l][b].y) for b in range(16)]}


def spatial_direct_decoder(O, mbs, col, mx, my, mbw):
    """dec/slicedec.rs:733-752 (spatial_direct_base) and 754-800 (zeroed,
    co-located with direct_8x8_inference, col_short_term true)."""
    refs, base = [-1, -1], [(0, 0), (0, 0)]
    nbs = [{}, {}]
    for l in range(2):
        for dx in (-1, 0, 1):
            for dy in (-1, 0):
                nx, ny = mx + dx, my + dy
                if 0 <= nx < mbw and ny >= 0 and (dx, dy) != (0, 0):
                    nbs[l][(nx, ny)] = _list_view(O, mbs[ny * mbw + nx], l)
        cur = {"ref": [-1] * 4, "mv": [(0, 0)] * 16}

        def nb(bx, by):
            if bx < 0 and by < 0:
                key, b = (mx - 1, my - 1), 15
            elif bx < 0:
                key, b = (mx - 1, my), by * 4 + 3
            elif bx > 3:
                key, b = (mx + 1, my - 1), 12
            else:
                key, b = (mx, my - 1), 12 + bx
            if key not in nbs[l]:
                return (False, -1)
            m = nbs[l][key]
            if m["intra"]:
                return (True, -1)
            return (True, m["ref"][((b >> 2) >> 1) * 2 + ((b & 3) >> 1)])

        a, b, c = nb(-1, 0), nb(0, -1), nb(4, -1)
        if not c[0]:
            c = nb(-1, -1)
        mp = lambda x, y: min(x, y) if x >= 0 and y >= 0 else max(x, y)  # noqa: E731
        refs[l] = mp(a[1], mp(b[1], c[1]))
        if refs[l] >= 0:
            base[l] = spec_mvp(nbs[l], cur, set(), 0, 0, 4, refs[l], 0, 0, mx, my, mbw)
    zero = refs[0] < 0 and refs[1] < 0
    out_ref = np.zeros((2, 4), np.int8)
    out_mv = np.zeros((2, 4, 2), np.int16)
    for b8 in range(4):
        raster = (b8 >> 1) * 8 + (b8 & 1) * 2
        bx, by = raster & 3, raster >> 2
        raster = (12 if by >= 2 else 0) + (3 if bx >= 2 else 0)  # direct_8x8_inference
        col_zero = False
        if col.kind not in (O.OK_I4x4, O.OK_I8x8, O.OK_I16x16):
            q8 = (raster >> 3) * 2 + ((raster & 3) >> 1)
            lc = 0 if col.ref_idx[0][q8] >= 0 else 1
            mv = col.mv[lc][raster]
            col_zero = col.ref_idx[lc][q8] == 0 and -1 <= mv.x <= 1 and -1 <= mv.y <= 1
        for l in range(2):
            out_ref[l][b8] = 0 if zero else refs[l]
            if zero or refs[l] < 0 or (refs[l] == 0 and col_zero):
                out_mv[l][b8] = (0, 0)
            else:
                out_mv[l][b8] = base[l]
    return out_ref, out_mv


def test_direct_spatial_matches_the_decoders_spatial_direct(oracle):
    O = oracle
    rs = np.random.default_rng(21)
    mbw, mbh = 4, 3
    prm = O.oparams(mbw, mbh, 16, 1, 1, 200)
    seen_zero = seen_colzero = seen_mixed = 0
    for trial in range(500):
        mbs = (O.ombinfo2 * (mbw * mbh))()
        for i in range(mbw * mbh):
            _rand_mbinfo2(O, mbs[i], rs, intra_p=0.3 if trial % 3 == 0 else 0.1)
        col = O.ombinfo2()
        _rand_mbinfo2(O, col, rs)
        if trial % 4 == 0:  # a near-still co-located MB: exercises col_zero
            for l in range(2):
                for b in range(16):
                    col.mv[l][b] = O.omv(int(rs.integers(-1, 2)), int(rs.integers(-1, 2)))
        mx, my = int(rs.integers(0, mbw)), int(rs.integers(0, mbh))
        refs, mvs = O.direct_spatial(prm, mbs, col, mx, my)
        wr, wm = spatial_direct_decoder(O, mbs, col, mx, my, mbw)
        assert np.array_equal(refs, wr), (trial, refs, wr)
        assert np.array_equal(mvs, wm), (trial, mvs, wm)
        seen_zero += int(np.all(refs == 0) and not mvs.any())
        seen_colzero += int((refs == 0).any() and not mvs[refs == 0].any())
        seen_mixed += int((refs[0] >= 0).all() != (refs[1] >= 0).all())
    assert seen_zero > 10 and seen_colzero > 10 and seen_mixed > 10


def _mc(planes, x, y, w, h, mv):
    (pa, dxa, dya), (pb, dxb, dyb) = pysearch.QSPEC[(mv[0] & 3, mv[1] & 3)]
    X, Y = x + (mv[0] >> 2) + pysearch.PAD, y + (mv[1] >> 2) + pysearch.PAD
    a = planes[pa][Y + dya:Y + dya + h, X + dxa:X + dxa + w]
    b = planes[pb][Y + dyb:Y + dyb + h, X + dxb:X + dxb + w]
    return (a + b + 1) >> 1


@pytest.mark.parametrize("seed", range(4))
def test_b_analyse_costs_match_the_definitions(oracle, seed):
    O = oracle
    rs = np.random.default_rng(100 + seed)
    w, h = 96, 64
    mbw, mbh = w // 16, h // 16
    import pcme
    y, _, _ = pcme.synth_clip(pcme.ClipSpec(w, h, 3, seed=0x5EED1700 + seed, nrect=3, rect_lo=16, rect_hi=40,
                                            vmax=4.0))
    src = pysearch.pad_plane(y[1], mbw, mbh)
    planes = []
    refs = []
    for t in (0, 2):
        G = pysearch.pad_plane(y[t], mbw, mbh)
        planes.append([G] + pysearch.half_planes(G))
        refs.append(O.Ref(O.pad_luma(y[t], mbw, mbh)))
    lam = O.load().o_lambda(LAM_QP)
    mbs = (O.ombinfo2 * (mbw * mbh))()
    for i in range(mbw * mbh):
        _rand_mbinfo2(O, mbs[i], rs)
    n_direct = 0
    for my in range(mbh):
        prm = O.oparams(mbw, mbh, 16, 1, 1, O.load().o_vlim(my, mbh))
        for mx in range(mbw):
            col = O.ombinfo2()
            _rand_mbinfo2(O, col, rs)
            s16 = src[64 + 16 * my:80 + 16 * my, 64 + 16 * mx:80 + 16 * mx]
            r = O.b_analyse(prm, mbs, col, mx, my, refs[0], refs[1], s16.astype(np.uint8), lam)
            px, py = 16 * mx, 16 * my
            # direct prediction: per 8x8, list 0 / list 1 / rounded average
            if r["direct_ok"]:
                n_direct += 1
                pred = np.zeros((16, 16), np.int64)
                for q in range(4):
                    qx, qy = 8 * (q & 1), 8 * (q >> 1)
                    use = [r["direct_ref"][l][q] >= 0 for l in range(2)]
                    ps = [_mc(planes[l], px + qx, py + qy, 8, 8, tuple(r["direct_mv"][l][q])) for l in range(2)]
                    pred[qy:qy + 8, qx:qx + 8] = ((ps[0] + ps[1] + 1) >> 1) if all(use) else ps[0 if use[0] else 1]
                assert r["direct_cost"] == pysearch.satd(s16 - pred) + ((lam * 2) >> 8)
            else:
                assert r["direct_cost"] == 0xFFFFFFFF
            # 16x16: list costs = search_part totals + (lam*3)>>8, bi by definition
            p0 = _mc(planes[0], px, py, 16, 16, tuple(r["mv16"][0]))
            p1 = _mc(planes[1], px, py, 16, 16, tuple(r["mv16"][1]))
            bits = sum(pysearch.se_len(int(r["mv16"][l][k]) - int(r["mvp16"][l][k])) for l in range(2) for k in range(2))
            want = [int(r["cost16"][0]) + ((lam * 3) >> 8), int(r["cost16"][1]) + ((lam * 3) >> 8),
                    pysearch.satd(s16 - ((p0 + p1 + 1) >> 1)) + ((lam * (bits + 5)) >> 8)]
            assert list(r["cost"]) == want
            assert r["dir"] == int(np.argmin(want))
            # b_partition 16x8 / 8x16
            for k, rects in enumerate((pysearch.PARTS[1], pysearch.PARTS[2])):
                pr = [np.zeros((16, 16), np.int64), np.zeros((16, 16), np.int64)]
                for l in range(2):
                    for pi, (x, yy, ww, hh) in enumerate(rects):
                        pr[l][yy:yy + hh, x:x + ww] = _mc(planes[l], px + x, py + yy, ww, hh,
                                                          tuple(r["part_mv"][k][l][pi]))
                bi = (pr[0] + pr[1] + 1) >> 1
                cost = (lam * 8) >> 8
                for pi, (x, yy, ww, hh) in enumerate(rects):
                    cs = [pysearch.satd(s16[yy:yy + hh, x:x + ww] - p[yy:yy + hh, x:x + ww]) + ((lam * b) >> 8)
                          for p, b in ((pr[0], 7), (pr[1], 7), (bi, 14))]
                    assert r["part_dir"][k][pi] == int(np.argmin(cs))
                    cost += min(cs)
                assert r["part_cost"][k] == cost
    assert n_direct > 0


def test_b_frame_driver_equals_per_mb_calls(oracle):
    """o_b_frame_mt (one projection per frame, rows over threads) gives the
    per-MB o_b_analyse records."""
    O = oracle
    rs = np.random.default_rng(5)
    w, h = 64, 48
    mbw, mbh = w // 16, h // 16
    import pcme
    y, _, _ = pcme.synth_clip(pcme.ClipSpec(w, h, 3, seed=0x5EED1900, nrect=2, rect_lo=16, rect_hi=32, vmax=3.0))
    src = O.pad_luma(y[1], mbw, mbh)
    refs = [O.Ref(O.pad_luma(y[0], mbw, mbh)), O.Ref(O.pad_luma(y[2], mbw, mbh))]
    lam = O.load().o_lambda(LAM_QP)
    mbs = (O.ombinfo2 * (mbw * mbh))()
    cols = (O.ombinfo2 * (mbw * mbh))()
    for i in range(mbw * mbh):
        _rand_mbinfo2(O, mbs[i], rs)
        _rand_mbinfo2(O, cols[i], rs)
    prm = O.oparams(mbw, mbh, 16, 1, 1, 0)
    _, got = O.b_frame_mt(prm, mbs, cols, refs[0], refs[1], src, lam, 3)
    for my in range(mbh):
        p1 = O.oparams(mbw, mbh, 16, 1, 1, O.load().o_vlim(my, mbh))
        for mx in range(mbw):
            s16 = np.ascontiguousarray(src[64 + 16 * my:80 + 16 * my, 64 + 16 * mx:80 + 16 * mx])
            want = O.b_analyse(p1, mbs, cols[my * mbw + mx], mx, my, refs[0], refs[1], s16, lam)
            assert got[my * mbw + mx].tobytes() == want.tobytes(), (mx, my)
