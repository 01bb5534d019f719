"""Host logic of pcme.integrated, no GPU: the ten neighbour blocks and
pskip_mv computed from a committed picture state agree with the oracle's
nb_part / pskip_mv (slice.rs:392-432, 484-494) over random states (intra,
skip, references -1/0/1), and verify() flags exactly the MBs whose blocks
changed."""
from __future__ import annotations

import ctypes

import numpy as np

from pcme import integrated as I


def _state(O, rs, mbw, mbh):
    kind = rs.choice([O.OK_PINTER, O.OK_PINTER, O.OK_PSKIP, O.OK_I16x16, O.OK_I4x4], (mbh, mbw))
    ref = rs.choice([-1, 0, 0, 0, 1], (mbh, mbw, 4))
    mv = rs.integers(-40, 41, (mbh, mbw, 16, 2))
    return kind, ref, mv


def test_nb_blocks_and_pskip_match_the_oracle(oracle):
    O = oracle
    rs = np.random.default_rng(3)
    mbw, mbh = 5, 4
    for trial in range(60):
        kind, ref, mv = _state(O, rs, mbw, mbh)
        mbs = (O.ombinfo * (mbw * mbh))()
        for my in range(mbh):
            for mx in range(mbw):
                m = mbs[my * mbw + mx]
                m.kind = int(kind[my, mx])
                for q in range(4):
                    m.ref_idx[q] = int(ref[my, mx, q])
                for b in range(16):
                    m.mv[b] = O.omv(int(mv[my, mx, b, 0]), int(mv[my, mx, b, 1]))
        nbs = I.nb_blocks(kind, ref, mv)
        prm = O.oparams(mbw, mbh, 16, 1, 1, 200)
        cur = O.ombinfo()
        for my in range(mbh):
            for mx in range(mbw):
                for i, ((dx, dy), blk) in enumerate([((-1, 0), 3), ((0, -1), 12), ((1, -1), 12), ((-1, -1), 15)]):
                    bx = -1 if dx < 0 else (4 if dx > 0 else 0)
                    by = -1 if dy < 0 else 0
                    rr, mm = ctypes.c_int8(), O.omv()
                    has = O.load().o_nb_part(ctypes.byref(prm), ctypes.cast(mbs, ctypes.c_void_p), ctypes.byref(cur),
                                             0, mx, my, bx, by, ctypes.byref(rr), ctypes.byref(mm))
                    k = (0, 4, 8, 9)[i]
                    want = (rr.value, mm.x, mm.y) if has else (-2, 0, 0)
                    assert tuple(int(v) for v in nbs[my, mx, k]) == want, (trial, mx, my, k)
                s = O.load().o_pskip_mv(ctypes.byref(prm), ctypes.cast(mbs, ctypes.c_void_p), mx, my)
                assert I.pskip_mv(nbs[my, mx]) == (s.x, s.y), (trial, mx, my)


def test_verify_flags_exactly_the_changed_neighbourhoods():
    import pcme
    mbw, mbh = 6, 5
    rs = np.random.default_rng(4)
    field = np.zeros((mbh, mbw), pcme.MB_DTYPE)
    field["mv"] = rs.integers(-30, 31, (mbh, mbw, 16, 2))
    kind, ref, mv = I.field_state(field)
    assert I.verify(field, kind, ref, mv).all()
    k2 = kind.copy()
    k2[2, 3] = 3  # one intra MB: its left-, top- and both diagonal readers change
    bad = ~I.verify(field, k2, ref, mv)
    assert {(int(x), int(y)) for y, x in zip(*np.nonzero(bad))} == {(4, 2), (3, 3), (2, 3), (4, 3)}
