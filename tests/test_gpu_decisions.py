"""The GPU search against the closed-form decision fixtures.

The same hand-derived cases as test_oracle_decisions.py
(tests/decision_cases.py), through the C ABI: every search case in one
pcme_search_batch call (half-pel planes made by pcme_hpel on the device),
every one-MB frame case through pcme_search_frames_sync.  The expected
values are the fixtures' own; the oracle is not consulted."""
from __future__ import annotations

import numpy as np
import pytest

import decision_cases as D

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def test_search_batch_decisions(gpu):
    import torch
    lib = gpu.load()
    say("decision fixtures: search_batch", len(D.SEARCH_CASES))
    keep, jobs = [], []
    pitch = D.ROWS * D.STRIDE
    for c in D.SEARCH_CASES:
        plane = D.ref_plane(c)
        planes = torch.zeros((4, D.ROWS, D.STRIDE), dtype=torch.uint8, device="cuda")
        planes[0].copy_(torch.from_numpy(plane))
        gpu._check(lib.pcme_hpel(planes[0].data_ptr(), planes[1].data_ptr(), planes[2].data_ptr(),
                                 planes[3].data_ptr(), pitch, 1, D.MBW, D.MBH, None))
        src = torch.from_numpy(D.src_block(c, plane)).cuda()
        keep += [planes, src]
        sj = gpu.SearchJob()
        sj.src = src.data_ptr()
        sj.sstride = c.w
        sj.ref, sj.ref_h, sj.ref_v, sj.ref_c = (planes[k].data_ptr() for k in range(4))
        sj.mbw, sj.mbh, sj.bx, sj.by, sj.w, sj.h = D.MBW, D.MBH, c.bx, c.by, c.w, c.h
        sj.mvp = gpu.Mv(*c.mvp)
        for i, m in enumerate(c.cands):
            sj.cands[i] = gpu.Mv(*m)
        sj.ncands = len(c.cands)
        sj.lam = c.lam
        sj.me_range = c.me_range
        sj.subpel_rounds = c.subpel
        sj.vlim = c.vlim
        jobs.append(sj)
    torch.cuda.synchronize()
    omv = torch.zeros((len(jobs), 2), dtype=torch.int16, device="cuda")
    ocost = torch.zeros(len(jobs), dtype=torch.int32, device="cuda")
    gpu.search_batch(jobs, omv.data_ptr(), ocost.data_ptr(), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    gm = omv.cpu().numpy()
    gc = ocost.cpu().numpy().view(np.uint32)
    bad = [f"{c.name}: got ({gm[i, 0]}, {gm[i, 1]}) cost {gc[i]}, want {c.want_mv} cost {c.want_cost}"
           for i, c in enumerate(D.SEARCH_CASES)
           if (int(gm[i, 0]), int(gm[i, 1]), int(gc[i])) != (*c.want_mv, c.want_cost)]
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("c", D.FRAME_CASES, ids=lambda c: c.name)
def test_frame_decisions(gpu, c):
    say("decision fixtures: frame", c.name)
    frames = D.frame_pair(c)
    dc = gpu.DeviceClip(frames, 1, 1)
    r = gpu.search_clip_me(dc, 16, 1, True, c.qp)[0, 0, 0]
    assert tuple(int(v) for v in r["cost"]) == c.want_cost
    assert int(r["best_cost"]) == c.want_best
    assert int(r["part"]) == c.want_part
    assert not r["mv"].any() and not r["mvs"].any() and not r["mvps"].any()
    assert not r["skip_mv"].any() and not r["mv16"].any()
