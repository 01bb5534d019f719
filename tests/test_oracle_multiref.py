"""Pin the oracle's multi-reference P analysis (o_analyse_mb_mr,
o_search_frame_mr; no counterpart in the reference, whose P slices use one
reference): against the independent restatement extended to several
references (tests/pysearch.py), against closed forms on clips whose
pictures repeat, and against textbook SATD + Exp-Golomb rates per MB."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import pysearch
from _util import assert_same_field
from test_oracle_kat import satd_textbook
from test_oracle_search import _frames


def _ue(k: int) -> int:
    return 2 * (int(k + 1).bit_length() - 1) + 1


def _se(v: int) -> int:
    return _ue(2 * v - 1 if v > 0 else -2 * v)


def test_ref_bits_is_te(oracle):
    """te(v) with num_ref_idx_active 1 (not coded), 2 (one inverted bit) and
    more (ue(v)) -- H.264 9.1."""
    lib = oracle.load()
    assert [lib.o_ref_bits(0, 1)] == [0]
    assert [lib.o_ref_bits(r, 2) for r in range(2)] == [1, 1]
    assert [lib.o_ref_bits(r, 4) for r in range(4)] == [1, 3, 3, 5]
    assert [lib.o_ref_bits(r, 16) for r in range(16)] == [_ue(r) for r in range(16)]


CASES = [
    # (seed, w, h, frames, nref, me_range, subpel, partitions, qp)
    (31, 48, 48, 4, 3, 16, 1, True, 27),
    (32, 64, 32, 5, 4, 8, 0, True, 22),
    (33, 40, 40, 3, 2, 16, 2, False, 37),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"s{c[0]}_{c[1]}x{c[2]}_r{c[4]}")
def test_oracle_matches_restatement(oracle, case):
    seed, w, h, n, nref, rng, sub, parts, qp = case
    y = _frames(seed, w, h, n, shift=(2, 1))
    # make older pictures worth choosing: frame t also resembles frame t-2
    y[2:] = ((y[2:].astype(int) + y[:-2].astype(int)) // 2).astype(np.uint8)
    got = oracle.search_clip_mr(y, nref, rng, sub, parts, qp)
    for t in range(1, n):
        refs = [y[t - 1 - k] for k in range(min(t, nref))]
        fr = pysearch.Frame(y[t], refs[0], rng, sub, parts, qp, older_refs=refs[1:])
        want = pysearch.to_records(fr.run(), oracle.OMB_DTYPE)
        assert_same_field(got[t - 1], want, f"frame {t}")
    assert (got["ref"] > 0).any()


def test_one_reference_is_the_plain_field(oracle):
    y = _frames(34, 64, 48, 3)
    a = oracle.search_clip_mr(y, 1, 16, 1, True, 28)
    b = oracle.search_clip_me(y, 16, 1, True, 28)
    assert a.tobytes() == b.tobytes()


def _noise(seed, w, h):
    return np.random.default_rng(seed).integers(0, 256, (h, w)).astype(np.uint8)


def test_repeating_pictures_choose_the_exact_reference(oracle):
    """Pictures A, B, A, B, A of independent noise: picture 4's references
    are B (index 0), A (1), B (2), A (3).  Index 1 and 3 match exactly at
    the zero MV; their totals are SATD 0 + (lam * 2) >> 8 (the zero MV
    against a zero predictor) + the index's te(v) rate, 3 and 5 bits, so
    every MB takes index 1 with P16x16 and cost (2 lam >> 8) + (3 lam >> 8).
    With the two references B, A (nref 2) index 1 costs one bit."""
    A, B = _noise(1, 64, 48), _noise(2, 64, 48)
    y = np.stack([A, B, A, B, A])
    qp = 27
    lam = int(oracle.load().o_lambda(qp))
    got = oracle.search_clip_mr(y, 4, 16, 1, True, qp)[3]
    assert np.all(got["ref"] == 1) and np.all(got["part"] == 0)
    assert np.all(got["mv16"] == 0) and np.all(got["mv"] == 0)
    assert np.all(got["cost"][..., 0] == ((2 * lam) >> 8) + ((3 * lam) >> 8))
    assert np.all(got["best_cost"] == got["cost"][..., 0] + (lam >> 8))
    got2 = oracle.search_clip_mr(y[2:5], 2, 16, 1, True, qp)[1]  # refs B, A for picture A
    assert np.all(got2["ref"] == 1)
    assert np.all(got2["cost"][..., 0] == ((2 * lam) >> 8) + ((1 * lam) >> 8))


def test_equal_totals_keep_the_newest(oracle):
    """A static clip: every reference matches at the zero MV.  With two
    references both indices cost one bit, so the totals tie and the lower
    index stays (strict <); with three, index 0 costs 1 bit and 1 costs 3."""
    A = _noise(3, 48, 32)
    y = np.stack([A, A, A, A])
    for nref in (2, 3):
        got = oracle.search_clip_mr(y, nref, 16, 1, True, 30)
        assert np.all(got["ref"] == 0) and np.all(got["mv16"] == 0)


def test_costs_are_satd_plus_rates(oracle):
    """Each P16x16 MB: cost[0] = textbook SATD of the source against the
    chosen reference's quarter-pel prediction at mv16 + (lam * se(v) bits
    of mv16 - mvps[0]) >> 8 + (lam * te(v) bits of ref) >> 8, and mvps[0]
    is o_mvp over the final field with the MB's reference index."""
    seed, w, h, n, nref, qp = 35, 64, 48, 4, 3, 27
    y = _frames(seed, w, h, n, shift=(3, -1))
    y[2:] = ((y[2:].astype(int) + y[:-2].astype(int)) // 2).astype(np.uint8)
    got = oracle.search_clip_mr(y, nref, 16, 1, True, qp)[2]  # picture 3: three references
    mbw, mbh = (w + 15) // 16, (h + 15) // 16
    lib = oracle.load()
    lam = int(lib.o_lambda(qp))
    refs = [y[2], y[1], y[0]]
    fr = pysearch.Frame(y[3], refs[0], 16, 1, True, qp, older_refs=refs[1:])
    field = (oracle.ombinfo * (mbw * mbh))()
    for i, r in enumerate(got.reshape(-1)):
        field[i].kind = oracle.OK_PINTER
        for k in range(4):
            field[i].ref_idx[k] = int(r["ref"])
        for k in range(16):
            field[i].mv[k] = oracle.omv(int(r["mv"][k][0]), int(r["mv"][k][1]))
    prm = oracle.oparams(mbw, mbh, 16, 1, 1, 0)
    cur = oracle.ombinfo()
    for k in range(4):
        cur.ref_idx[k] = -1
    checked = 0
    for i, r in enumerate(got.reshape(-1)):
        if r["part"] != 0:
            continue
        mx, my, rf = i % mbw, i // mbw, int(r["ref"])
        m16 = (int(r["mv16"][0]), int(r["mv16"][1]))
        mvp = lib.o_mvp(ctypes.byref(prm), field, ctypes.byref(cur), 0, mx, my, 0, 0, 4, rf, 0)
        assert tuple(int(v) for v in r["mvps"][0]) == (mvp.x, mvp.y), (mx, my)
        fr.use_ref(rf)
        pred = fr.pred(16 * mx, 16 * my, 16, 16, m16)
        sblk = fr.src[16 * my + pysearch.PAD:16 * my + pysearch.PAD + 16,
                      16 * mx + pysearch.PAD:16 * mx + pysearch.PAD + 16]
        rate = (lam * (_se(m16[0] - mvp.x) + _se(m16[1] - mvp.y))) >> 8
        assert int(r["cost"][0]) == satd_textbook(sblk, pred) + rate + ((lam * _ue(rf)) >> 8), (mx, my)
        checked += 1
    assert checked >= 4


def test_pruned_mbs_search_every_reference(oracle):
    """Pictures A, B, A, B, A with every MB pruned (hints of one zero
    candidate): the pruned search of reference 1 (distance 2) finds the
    identical picture, so the pruned field takes index 1 like the plain one;
    and on a clip moving (5, 3) px per picture, hints of the one-picture
    motion (20, 12) quarter samples, scaled to distance 2, (40, 24), find
    picture t-2 at that MV where the source repeats it."""
    A, B = _noise(11, 64, 48), _noise(12, 64, 48)
    y = np.stack([A, B, A, B, A])
    mbw, mbh = 4, 3
    h = np.zeros((4, mbh * mbw), oracle.OHINT_DTYPE)
    h["prune"] = 1
    h["ncand"] = 1
    got = oracle.search_clip_mr(y, 2, 16, 1, True, 27, hints=h.reshape(4, mbh, mbw))[3]
    assert np.all(got["ref"] == 1) and np.all(got["mv16"] == 0)
    # a moving picture and noise in turn: picture 4 repeats picture 2 moved
    # twice.  MB (0, 0) has no neighbour, so only the scaled hint can reach
    # the 10 px motion: from the unscaled (5, 3) px the two square rounds and
    # the sub-pel rounds reach 8 px at most.
    base = _frames(13, 96, 64, 5, shift=(5, 3))
    y2 = base.copy()
    y2[1] = _noise(14, 96, 64)
    y2[3] = _noise(15, 96, 64)
    mbw, mbh = 6, 4
    h2 = np.zeros((4, mbh * mbw), oracle.OHINT_DTYPE)
    h2["prune"] = 1
    h2["ncand"] = 1
    h2["cand"][:, :, 0, 0] = 20  # picture t at x shows picture t-1 at x + (5, 3) px
    h2["cand"][:, :, 0, 1] = 12
    got2 = oracle.search_clip_mr(y2, 2, 8, 1, True, 27, hints=h2.reshape(4, mbh, mbw))[3]
    assert got2[0, 0]["ref"] == 1 and tuple(got2[0, 0]["mv16"]) == (40, 24)
    inner = got2[:-1, :-1]
    assert np.mean(inner["ref"] == 1) > 0.9
    assert np.mean((inner["mv16"][..., 0] == 40) & (inner["mv16"][..., 1] == 24)) > 0.9
