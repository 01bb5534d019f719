"""Encoder-integrated use of the ME-only field (pcme.integrated): on a
picture whose committed state differs from the field's assumption (intra
MBs, skipped MBs with their own MVs, reference index 1 in some 8x8s),
every MB that verify() passes has the field's record, every MB it flags is
re-analysed on the GPU, and both equal the oracle's analyse_mb on the
committed state, record for record.

Each case prints what it runs before the GPU step so a failing run names
its case."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("seed", [0, 1])
def test_field_verification_and_reanalysis_match_the_oracle(gpu, oracle, seed):
    import torch
    from pcme import integrated as I
    pcme, O = gpu, oracle
    w, h, qp, rng_ = 320, 240, 27, 24
    mbw, mbh = pcme.mb_dims(w, h)
    say("integrated field", seed)
    y, _, _ = pcme.synth_clip(pcme.ClipSpec(w, h, 2, seed=0x5EED1C00 + seed, nrect=5, vmax=5.0))
    dc = pcme.DeviceClip(y, mbw, mbh)
    field = pcme.search_clip_me(dc, rng_, 1, True, qp)[0]
    torch.cuda.synchronize()
    # the encoder's committed state: the field, then intra / skip / ref-1 MBs
    rs = np.random.default_rng(seed)
    kind, ref, mv = I.field_state(field)
    kind, ref, mv = kind.copy(), ref.copy(), mv.copy()
    u = rs.random((mbh, mbw))
    intra = u < 0.12
    skip = (u >= 0.12) & (u < 0.22)
    ref1 = (u >= 0.22) & (u < 0.3)
    kind[intra] = O.OK_I16x16
    ref[intra] = -1
    mv[intra] = 0
    kind[skip] = O.OK_PSKIP
    mv[skip] = rs.integers(-12, 13, (int(skip.sum()), 1, 2))
    ref[ref1, 1] = 1
    ok = I.verify(field, kind, ref, mv)
    redo = [(mx, my) for my in range(mbh) for mx in range(mbw) if not ok[my, mx]]
    assert 0 < len(redo) < mbw * mbh
    got = I.reanalyse(dc, 1, 0, redo, kind, ref, mv, qp, rng_, 1, True)
    # oracle: analyse_mb of each MB with the committed neighbours
    mbs = (O.ombinfo * (mbw * mbh))()
    for my in range(mbh):
        for mx in range(mbw):
            m = mbs[my * mbw + mx]
            m.kind = int(kind[my, mx])
            for q in range(4):
                m.ref_idx[q] = int(ref[my, mx, q])
            for b in range(16):
                m.mv[b] = O.omv(int(mv[my, mx, b, 0]), int(mv[my, mx, b, 1]))
    src = O.pad_luma(y[1], mbw, mbh)
    r = O.Ref(O.pad_luma(y[0], mbw, mbh))
    lam = O.load().o_lambda(qp)
    redo_at = {p: i for i, p in enumerate(redo)}
    for my in range(mbh):
        prm = O.oparams(mbw, mbh, rng_, 1, 1, O.load().o_vlim(my, mbh))
        for mx in range(mbw):
            saved = O.ombinfo.from_buffer_copy(mbs[my * mbw + mx])
            want = np.zeros(1, O.OMB_DTYPE)
            O.load().o_analyse_mb(ctypes.byref(prm), ctypes.cast(mbs, ctypes.c_void_p), mx, my, ctypes.byref(r.view),
                                  src.ctypes.data, src.shape[1], O.LUMA_PAD, lam, want.ctypes.data, None)
            mbs[my * mbw + mx] = saved
            rec = got[redo_at[(mx, my)]] if (mx, my) in redo_at else field[my, mx]
            assert rec.tobytes() == want[0].tobytes(), ((mx, my), (mx, my) in redo_at)
