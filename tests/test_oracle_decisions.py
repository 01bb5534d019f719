"""The oracle's search against the closed-form decision fixtures.

The fixtures (tests/decision_cases.py) fix by hand which MV each rule of
search (slice.rs:1617-1721) and of the partition choice (slice.rs:1921-1945)
selects: predictor rounding, strict-< ties in candidate order, the hexagon's
box, square rounds beyond it, the SATD restart, the number of quarter-pel
rounds, the vlim row, the cost16 > 1024 gate and equal mode totals.  The
oracle's evaluation counters pin the rules that leave no trace in the
result (candidates equal to the best are not priced, a round that moves
nothing ends the phase)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import decision_cases as D


def _oracle_search(O, c: D.SearchCase):
    lib = O.load()
    plane = D.ref_plane(c)
    ref = O.Ref(plane)
    blk = D.src_block(c, plane)
    p = O.oparams(D.MBW, D.MBH, c.me_range, c.subpel, 1, c.vlim)
    cands = (O.omv * 2)(*[O.omv(*m) for m in c.cands], *[O.omv(0, 0)] * (2 - len(c.cands)))
    cost = ctypes.c_uint32()
    ctr = O.ocounters()
    m = lib.o_search(ctypes.byref(p), ctypes.byref(ref.view), c.bx, c.by, c.w, c.h, blk.ctypes.data, c.w,
                     O.omv(*c.mvp), cands, len(c.cands), c.lam, ctypes.byref(cost), ctypes.byref(ctr))
    return (m.x, m.y), cost.value, ctr


@pytest.mark.parametrize("c", D.SEARCH_CASES, ids=lambda c: c.name)
def test_search_decision(oracle, c):
    mv, cost, ctr = _oracle_search(oracle, c)
    assert (mv, cost) == (tuple(c.want_mv), c.want_cost)
    if c.want_evals is not None:
        assert (ctr.sad_evals, ctr.satd_evals) == c.want_evals


@pytest.mark.parametrize("c", D.FRAME_CASES, ids=lambda c: c.name)
def test_frame_decision(oracle, c):
    frames = D.frame_pair(c)
    ctr = oracle.ocounters()
    res = oracle.search_clip_me(frames, 16, 1, True, c.qp, ctr)
    r = res[0, 0, 0]
    assert tuple(int(v) for v in r["cost"]) == c.want_cost
    assert int(r["best_cost"]) == c.want_best
    assert int(r["part"]) == c.want_part
    assert not r["mv"].any() and not r["mvs"].any() and not r["mvps"].any()
    assert not r["skip_mv"].any() and not r["mv16"].any()
    if c.want_evals is not None:
        assert (ctr.searches, ctr.sad_evals, ctr.satd_evals) == c.want_evals


def test_ramp_samples_follow_the_offsets(oracle):
    """The surfaces' premise, on the oracle's own interpolation: on an x-ramp
    of slope 2 every quarter-pel sample is the full sample plus OFF[fx] =
    {0, 1, 1, 2}, whatever fy; on a y-ramp plus OFF[fy], whatever fx."""
    off = [0, 1, 1, 2]
    for surface in ("xramp", "yramp"):
        c = D.SearchCase("probe", surface, 16, 16, 16, 16, (0, 0), 256, 1, 1, 108, (0, 0), 0, centre=D.PAD + 24,
                         touch=(D.PAD + 4, D.PAD + 44))
        plane = D.ref_plane(c)
        ref = oracle.Ref(plane)
        lib = oracle.load()
        for q in range(16):
            fx, fy = q & 3, q >> 2
            buf = np.zeros((16, 16), np.uint8)
            lib.o_mc_luma(ctypes.byref(ref.view), 16, 16, oracle.omv(fx, fy), 16, 16, buf.ctypes.data, 16)
            full = plane[D.PAD + 16:D.PAD + 32, D.PAD + 16:D.PAD + 32].astype(int)
            want = full + (off[fx] if surface == "xramp" else off[fy])
            np.testing.assert_array_equal(buf.astype(int), want, err_msg=f"{surface} q={q}")
