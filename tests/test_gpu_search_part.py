"""GPU parity of the per-call search_part (pcme_search_part_batch) against
the oracle's o_search_part, bit-exact, over neighbourhoods the ME-only field
never produces: intra neighbours (ref -1, MV 0), neighbours without a
reference of the list, reference index 1, MB positions on every picture
edge, all four partition modes, random seeds.

Each case prints what it runs before the GPU step so a failing run names
its case."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _random_mbs(O, mbw, mbh, rng):
    """A picture of coded MbInfo: 15% intra, the rest P inter with per-8x8
    ref in {-1, 0, 1} (mostly 0) and MVs within +-24 px."""
    mbs = (O.ombinfo * (mbw * mbh))()
    for i in range(mbw * mbh):
        m = mbs[i]
        m.kind = O.OK_I16x16 if rng.random() < 0.15 else O.OK_PINTER
        for q in range(4):
            m.ref_idx[q] = int(rng.choice([-1, 0, 1], p=[0.15, 0.7, 0.15]))
        for b in range(16):
            m.mv[b] = O.omv(int(rng.integers(-96, 97)), int(rng.integers(-96, 97)))
    return mbs


def _nb_block(pcme, O, mbs, mbw, mx, my, dx, dy, r):
    """nb_part's answer for 4x4 block r of the neighbour MB (mx+dx, my+dy)
    (slice.rs:409-431)."""
    b = pcme.NbBlock()
    nx, ny = mx + dx, my + dy
    if nx < 0 or nx >= mbw or ny < 0:
        b.ref = -2
        return b
    m = mbs[ny * mbw + nx]
    if m.kind == O.OK_I16x16:
        b.ref = -1
        return b
    rf = m.ref_idx[((r >> 2) >> 1) * 2 + ((r & 3) >> 1)]
    b.ref = rf
    if rf >= 0:
        b.mv = pcme.Mv(m.mv[r].x, m.mv[r].y)
    return b


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_search_part_batch_matches_oracle(gpu, oracle, seed):
    import torch
    pcme, O = gpu, oracle
    w, h, qp = 320, 240, 27
    mbw, mbh = pcme.mb_dims(w, h)
    say("search_part batch", seed)
    y, _, _ = pcme.synth_clip(pcme.ClipSpec(w, h, 2, seed=0x5EED1600 + seed, nrect=5, vmax=5.0))
    dc = pcme.DeviceClip(y, mbw, mbh)
    dc.pad()
    dc.hpel()
    torch.cuda.synchronize()
    rng = np.random.default_rng(seed)
    mbs = _random_mbs(O, mbw, mbh, rng)
    src = O.pad_luma(y[1], mbw, mbh)
    ref = O.Ref(O.pad_luma(y[0], mbw, mbh))
    lam = O.load().o_lambda(qp)
    stride = src.shape[1]
    positions = [(0, 0), (mbw - 1, 0), (0, mbh - 1), (mbw - 1, mbh - 1)] + \
                [(int(rng.integers(0, mbw)), int(rng.integers(0, mbh))) for _ in range(28)]
    jobs, want = [], []
    for k, (mx, my) in enumerate(positions):
        part = k % 4
        seedmv = (int(rng.integers(-40, 41)), int(rng.integers(-40, 41)))
        vlim = O.load().o_vlim(my, mbh)
        prm = O.oparams(mbw, mbh, 32, 1, 1, vlim)
        src16 = np.ascontiguousarray(src[O.LUMA_PAD + 16 * my:O.LUMA_PAD + 16 * my + 16,
                                         O.LUMA_PAD + 16 * mx:O.LUMA_PAD + 16 * mx + 16])
        mvs = (O.omv * 4)()
        mvps = (O.omv * 4)()
        total = O.load().o_search_part(ctypes.byref(prm), ctypes.cast(mbs, ctypes.c_void_p), mx, my,
                                       ctypes.byref(ref.view), part, src16.ctypes.data, lam, O.omv(*seedmv),
                                       mvs, mvps, None)
        want.append(([(m.x, m.y) for m in mvs], [(m.x, m.y) for m in mvps], int(total)))
        j = pcme.PartJob()
        j.src = dc.plane_ptr(0, 1) + (pcme.LUMA_PAD + 16 * my) * dc.stride + pcme.LUMA_PAD + 16 * mx
        j.sstride = dc.stride
        j.part = part
        j.ref, j.ref_h, j.ref_v, j.ref_c = (dc.plane_ptr(i, 0) for i in range(4))
        j.mbw, j.mbh, j.mx, j.my = mbw, mbh, mx, my
        j.seed = pcme.Mv(*seedmv)
        j.lam = lam
        j.me_range, j.subpel_rounds, j.vlim = 32, 1, vlim
        for i in range(4):
            j.left[i] = _nb_block(pcme, O, mbs, mbw, mx, my, -1, 0, 4 * i + 3)
            j.top[i] = _nb_block(pcme, O, mbs, mbw, mx, my, 0, -1, 12 + i)
        j.top_right = _nb_block(pcme, O, mbs, mbw, mx, my, 1, -1, 12)
        j.top_left = _nb_block(pcme, O, mbs, mbw, mx, my, -1, -1, 15)
        jobs.append(j)
    assert stride == dc.stride
    out = torch.zeros(len(jobs) * pcme.PART_OUT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    pcme.search_part_batch(jobs, out.data_ptr())
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(pcme.PART_OUT_DTYPE)
    for k, ((mvs, mvps, total), g) in enumerate(zip(want, got)):
        assert [tuple(int(v) for v in m) for m in g["mvs"]] == mvs, (k, positions[k], k % 4)
        assert [tuple(int(v) for v in m) for m in g["mvps"]] == mvps, (k, positions[k], k % 4)
        assert int(g["total"]) == total, (k, positions[k])
