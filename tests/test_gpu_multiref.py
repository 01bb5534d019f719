"""The multi-reference field (pcme_frame_job.nref > 1; no counterpart in the
reference, whose P slices use one reference) on the GPU, bit for bit
against the oracle's o_search_frame_mr: up to four references, pictures
that repeat (the closed-form case), pruner hints (a pruned MB searches
reference 0 only), the per-call and no-allocation entry points, and plan
launches past the exchange words' tag wrap.

Each case prints what it runs before the GPU step so a failing run names
its case."""
from __future__ import annotations

import numpy as np
import pytest

from _util import assert_same_field

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _dclip(pcme, y):
    n, h, w = y.shape
    mbw, mbh = pcme.mb_dims(w, h)
    dc = pcme.DeviceClip(np.ascontiguousarray(y), mbw, mbh)
    dc.prepare()
    return dc


def _synth(pcme, w, h, n, seed, **kw):
    y, _, _ = pcme.synth_clip(pcme.ClipSpec(w, h, n, seed=seed, **kw))
    # an older picture is worth choosing: frame t also resembles frame t-2
    y = y.copy()
    y[2:] = ((y[2:].astype(int) + y[:-2].astype(int)) // 2).astype(np.uint8)
    return y


@pytest.mark.parametrize("w,h,nref,rng,qp", [(320, 240, 4, 16, 27), (640, 368, 3, 32, 22), (176, 144, 2, 24, 37)])
def test_multiref_matches_oracle(gpu, oracle, w, h, nref, rng, qp):
    say("multiref", w, h, nref, rng, qp)
    y = _synth(gpu, w, h, nref + 2, 0x5EED5000 + w, nrect=5, vmax=5.0)
    dc = _dclip(gpu, y)
    got = gpu.search_clip_mr(dc, nref, rng, 1, True, qp, prepared=True)
    want = oracle.search_clip_mr(y, nref, rng, 1, True, qp)
    for t in range(got.shape[0]):
        assert_same_field(got[t], want[t], f"frame {t + 1}")
    assert (got["ref"] > 0).mean() > 0.05 and set(np.unique(got["part"])) == {0, 1, 2, 3}


def test_repeating_pictures_closed_form(gpu, oracle):
    """Pictures A, B, A, B, A of independent noise, four references: every
    MB of picture 4 takes index 1 at the zero MV with P16x16, cost
    (2 lam >> 8) + (3 lam >> 8)."""
    say("multiref repeating pictures")
    rs = np.random.default_rng(5)
    A, B = (rs.integers(0, 256, (96, 128)).astype(np.uint8) for _ in range(2))
    y = np.stack([A, B, A, B, A])
    dc = _dclip(gpu, y)
    got = gpu.search_clip_mr(dc, 4, 16, 1, True, 27, prepared=True)[3]
    lam = int(oracle.load().o_lambda(27))
    assert np.all(got["ref"] == 1) and np.all(got["part"] == 0) and np.all(got["mv"] == 0)
    assert np.all(got["cost"][..., 0] == ((2 * lam) >> 8) + ((3 * lam) >> 8))


def test_multiref_with_pruner_hints(gpu, oracle):
    """Pruned MBs search every reference with their hint candidates scaled
    to its distance."""
    import torch
    import pcme.pruner as P
    say("multiref + network hints 320x240, 3 refs")
    y = _synth(gpu, 320, 240, 5, 0x5EED5100, nrect=5, vmax=5.0)
    dc = _dclip(gpu, y)
    mbw, mbh = dc.mbw, dc.mbh
    net = P.Net(seed=1, max_batch=256, weights=P.WEIGHTS, qp=27)
    nb = 4
    out = torch.empty((nb, mbw * mbh * gpu.MB_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    hints = torch.zeros((nb, mbw * mbh * gpu.HINT_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    jobs = [gpu.frame_job(dc, t, t - 1, out[t - 1].data_ptr(), 16, 1, True, 27, hint_ptr=hints[t - 1].data_ptr(),
                          older=[t - 1 - k for k in range(1, min(t, 3))]) for t in range(1, 5)]
    net.hints(jobs, 5, 0.7, 0.0)
    gpu.search_frames(jobs, sync=True)
    net.close()
    field = out.cpu().numpy().view(gpu.MB_DTYPE).reshape(nb, mbh, mbw)
    hs = hints.cpu().numpy().view(gpu.HINT_DTYPE).reshape(nb, mbh, mbw)
    assert 0.0 < float(np.mean(hs["prune"])) < 1.0
    want = oracle.search_clip_mr(y, 3, 16, 1, True, 27, hints=hs)
    for t in range(nb):
        assert_same_field(field[t], want[t], f"frame {t + 1}")
    # the same table through a plan at 16 waves per CU (4 walkers per group)
    plan = gpu.Plan(jobs)
    for waves in (0, 6):
        say("multiref + hints, configuration 4/16/1, waves", waves)
        plan.set_config(waves, 4, 16, 1)
        out.zero_()
        plan.launch()
        plan.check()
        got = out.cpu().numpy().view(gpu.MB_DTYPE).reshape(nb, mbh, mbw)
        for t in range(nb):
            assert_same_field(got[t], want[t], f"4/16 waves {waves} frame {t + 1}")


def test_multiref_entry_points_and_tag_wrap(gpu):
    """One job table through a plan launched 130 times (past two wraps of
    the 6-bit launch tag), the per-call search and the no-allocation
    search: one field."""
    import torch
    lib = gpu.load()
    say("multiref plan x130 vs per-call vs noalloc 256x128")
    y = _synth(gpu, 256, 128, 5, 0x5EED5200, nrect=4, vmax=4.0)
    dc = _dclip(gpu, y)
    mbw, mbh = dc.mbw, dc.mbh
    nbytes = mbw * mbh * gpu.MB_DTYPE.itemsize
    outs = [torch.full((4, nbytes), v, dtype=torch.uint8, device="cuda") for v in (0x11, 0x22, 0x33, 0x44)]
    jobs = [[gpu.frame_job(dc, t, t - 1, o[t - 1].data_ptr(), 16, 1, True, 30,
                           older=[t - 1 - k for k in range(1, min(t, 4))]) for t in range(1, 5)] for o in outs]
    s = torch.cuda.current_stream().cuda_stream
    plan = gpu.Plan(jobs[0])
    plan.launch(s)
    plan.check(s)
    first = outs[0].cpu().numpy().copy()
    plan2 = gpu.Plan(jobs[1])
    for _ in range(130):
        plan2.launch(s)
    plan2.check(s)
    plan.close()
    plan2.close()
    gpu.search_frames(jobs[2], sync=True)
    arr = (gpu.FrameJob * 4)(*jobs[3])
    gpu._check(lib.pcme_search_frames_noalloc(arr, 4, s))
    torch.cuda.synchronize()
    f0 = first.view(gpu.MB_DTYPE).reshape(4, mbh, mbw)
    for k, name in ((1, "plan x130"), (2, "per-call"), (3, "noalloc")):
        assert_same_field(outs[k].cpu().numpy().view(gpu.MB_DTYPE).reshape(4, mbh, mbw), f0, name)


def test_multiref_plan_configurations(gpu, oracle):
    """A multi-reference plan runs the default configuration and (4, 16, 1)
    (four walkers per workgroup at 16 per CU), at its resident grid and a
    small one, bit for bit the oracle's field; other configurations are
    refused, and autotune_ex keeps one of the two."""
    import torch
    w, h, nref, rng, qp = 320, 240, 3, 16, 27
    say("multiref plan configurations", w, h, nref)
    y = _synth(gpu, w, h, nref + 2, 0x5EED5A00, nrect=5, vmax=5.0)
    dc = _dclip(gpu, y)
    nb = dc.n - 1
    out = torch.zeros((nb, dc.mbh * dc.mbw * gpu.MB_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    jobs = [gpu.frame_job(dc, t, t - 1, out[t - 1].data_ptr(), rng, 1, True, qp,
                          older=[t - 1 - k for k in range(1, min(t, nref))]) for t in range(1, dc.n)]
    plan = gpu.Plan(jobs)
    want = oracle.search_clip_mr(y, nref, rng, 1, True, qp)
    for wpg, wcu, reuse in ((1, 12, 1), (4, 16, 1)):
        for waves in (0, 6):
            say("configuration", wpg, wcu, reuse, "waves", waves)
            t = plan.set_config(waves, wpg, wcu, reuse)
            assert (t.waves_per_group, t.waves_per_cu) == (wpg, wcu)
            out.zero_()
            plan.launch()
            plan.check()
            got = out.cpu().numpy().view(gpu.MB_DTYPE).reshape(nb, dc.mbh, dc.mbw)
            for k in range(nb):
                assert_same_field(got[k], want[k], f"config {wpg}/{wcu} waves {waves} frame {k + 1}")
    with pytest.raises(gpu.PcmeError):
        plan.set_config(0, 1, 12, 0)
    best = plan.autotune_ex(reps=1)
    assert (best.waves_per_group, best.waves_per_cu) in ((1, 12), (4, 16))
