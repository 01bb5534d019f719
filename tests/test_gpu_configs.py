"""GPU parity on the benchmark configurations themselves.

* Batched clips: one persistent launch over one clip of each size of the
  64-clip batch workload (352x288, 640x360 coded 640x368, 1280x720,
  1920x1080), each at its own QP from {22, 27, 32, 37}, me_range 32,
  2 P frames each, seeded like the workload; every clip bit-exact against
  the oracle.
* CNN-pruned 1080p: the committed per-QP networks (QP 22/27/32/37), 2 P
  frames each of the seeded 1080p clip, default confidence rule; the
  pruned field bit-exact against the oracle's pruned search on the GPU's
  own hints.

Each case prints what it runs before the GPU step."""
from __future__ import annotations

import os

import numpy as np
import pytest

from _util import assert_same_field

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


BATCH_SIZES = [(352, 288, 8, 120, 6.0), (640, 360, 8, 180, 6.0), (1280, 720, 10, 300, 8.0),
               (1920, 1080, 12, 400, 8.0)]
BATCH_QPS = (22, 27, 32, 37)


def test_batch_workload_sizes_one_launch(gpu, oracle):
    import torch
    dcs, outs, jobs, ys, prm = [], [], [], [], []
    for k, (w, h, nrect, hi, v) in enumerate(BATCH_SIZES):
        clip = 17 * k  # clip index in the 64-clip workload (sizes in blocks of 16): QPs 37, 32, 27, 22
        qp = BATCH_QPS[(clip * 7 + 3) % 4]
        spec = gpu.ClipSpec(w, h, 3, nrect=nrect, rect_lo=24 if w < 1000 else 64, rect_hi=hi, vmax=v,
                            pan_max=1.0, seed=0x5EED0000 + clip)
        y, _, _ = gpu.synth_clip(spec)
        mbw, mbh = gpu.mb_dims(w, h)
        dc = gpu.DeviceClip(y, mbw, mbh)
        dc.pad()
        dc.hpel(0, 2)
        o = torch.zeros((2, mbw * mbh * gpu.MB_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
        jobs += [gpu.frame_job(dc, t, t - 1, o[t - 1].data_ptr(), 32, 1, True, qp) for t in (1, 2)]
        dcs.append(dc)
        outs.append(o)
        ys.append(y)
        prm.append((w, h, mbw, mbh, qp))
    assert sorted(p[4] for p in prm) == [22, 27, 32, 37]
    torch.cuda.synchronize()
    say("batch workload sizes, one plan, 8 P frames")
    plan = gpu.Plan(jobs)
    plan.launch()
    plan.check()
    for k, (w, h, mbw, mbh, qp) in enumerate(prm):
        got = outs[k].cpu().numpy().view(gpu.MB_DTYPE).reshape(2, mbh, mbw)
        srcs = [oracle.pad_luma(ys[k][t], mbw, mbh) for t in (1, 2)]
        refs = [oracle.pad_luma(ys[k][t - 1], mbw, mbh) for t in (1, 2)]
        want, _, _ = oracle.search_jobs_mt(srcs, refs, mbw, mbh, 32, 1, True, qp, min(8, os.cpu_count() or 1))
        assert_same_field(got, np.stack(want), f"clip {w}x{h} qp {qp}")
    plan.close()


@pytest.fixture(scope="module")
def clip1080(gpu):
    spec = gpu.ClipSpec(1920, 1080, 3, nrect=12, rect_lo=64, rect_hi=400, vmax=8.0, pan_max=1.0,
                        seed=0x5EED0000)
    y, _, _ = gpu.synth_clip(spec)
    mbw, mbh = gpu.mb_dims(1920, 1080)
    dc = gpu.DeviceClip(y, mbw, mbh)
    dc.pad()
    dc.hpel(0, 2)
    return y, dc


@pytest.mark.parametrize("qp", [22, 27, 32, 37])
def test_pruned_1080p_committed_weights(gpu, oracle, clip1080, qp):
    import pcme.pruner as P
    y, dc = clip1080
    say("pruned 1080p, committed weights, 2 P frames, qp", qp)
    net = P.Net(seed=1, max_batch=1024, weights=P.WEIGHTS, qp=qp)
    field, hints = P.search_clip(dc, 32, qp, net=net, ncand=5, thr_part=P.thr_part_for(qp), thr_mv=0.0,
                                 prepared=True)
    net.close()
    assert 0.0 < float(np.mean(hints["prune"])) < 1.0  # both the pruned and the fallback path run
    mbw, mbh = dc.mbw, dc.mbh
    for t in (1, 2):
        src = oracle.pad_luma(y[t], mbw, mbh)
        ref = oracle.Ref(oracle.pad_luma(y[t - 1], mbw, mbh))
        want = oracle.search_frame(src, ref, mbw, mbh, 32, 1, True, qp, hints=hints[t - 1].reshape(-1))
        assert_same_field(field[t - 1], want, f"qp {qp} frame {t}")
