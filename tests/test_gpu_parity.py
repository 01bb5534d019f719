"""GPU parity: every kernel of libpcme against the CPU oracle, bit-exact.

Each case prints what it runs before the GPU step so a failing run names
its case."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from _util import assert_same_field, scene_clip

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _clip(pcme, w, h, n, seed=0x5EED0000, **kw):
    spec = pcme.ClipSpec(w, h, n, seed=seed, **kw)
    return pcme.synth_clip(spec)


def test_library_reports_device(gpu):
    assert gpu.device_count() >= 1


@pytest.mark.parametrize("w,h", [(320, 240), (100, 70), (16, 16), (34, 18)])
def test_pad_planes_match_oracle(gpu, oracle, w, h):
    import torch
    say("pad", w, h)
    y, u, v = _clip(gpu, w, h, 3)
    mbw, mbh = gpu.mb_dims(w, h)
    dc = gpu.DeviceClip(y, mbw, mbh, u, v)
    dc.pad()
    torch.cuda.synchronize()
    got = dc.planes[0].cpu().numpy()
    gu = dc.cplanes.cpu().numpy()
    for i in range(3):
        np.testing.assert_array_equal(got[i], oracle.pad_luma(y[i], mbw, mbh))
        np.testing.assert_array_equal(gu[0, i], oracle.pad_chroma(u[i], mbw, mbh))
        np.testing.assert_array_equal(gu[1, i], oracle.pad_chroma(v[i], mbw, mbh))


@pytest.mark.parametrize("w,h", [(320, 240), (1920, 1080), (48, 32)])
def test_hpel_matches_oracle(gpu, oracle, w, h):
    import torch
    say("hpel", w, h)
    y, _, _ = _clip(gpu, w, h, 2)
    mbw, mbh = gpu.mb_dims(w, h)
    dc = gpu.DeviceClip(y, mbw, mbh)
    dc.pad()
    dc.hpel()
    torch.cuda.synchronize()
    planes = dc.planes.cpu().numpy()
    for i in range(2):
        want = oracle.hpel(planes[0, i])
        for k in range(3):
            np.testing.assert_array_equal(planes[1 + k, i], want[k], err_msg=f"plane {'HVC'[k]} frame {i}")


@pytest.mark.parametrize("w,h", [(1920, 1080), (320, 240), (100, 70), (16, 16), (34, 18), (52, 200)])
def test_fused_luma_planes_match_pad_and_hpel(gpu, oracle, w, h):
    """pcme_luma_planes (one pass from the unpadded source) writes the four
    planes of pcme_pad_planes + pcme_hpel bit for bit, on widths that are and
    are not multiples of 4 (the clamped-byte path) and tiny pictures where
    every lane is an edge lane."""
    import torch
    say("luma planes", w, h)
    y, u, v = _clip(gpu, w, h, 3)
    mbw, mbh = gpu.mb_dims(w, h)
    a = gpu.DeviceClip(y, mbw, mbh, u, v)
    a.pad()
    a.hpel()
    b = gpu.DeviceClip(y, mbw, mbh, u, v)
    b.planes.fill_(0xA5)
    b.prepare()
    torch.cuda.synchronize()
    pa, pb = a.planes.cpu().numpy(), b.planes.cpu().numpy()
    for k in range(4):
        np.testing.assert_array_equal(pb[k], pa[k], err_msg=f"plane {'FHVC'[k]}")
    np.testing.assert_array_equal(b.cplanes.cpu().numpy(), a.cplanes.cpu().numpy())
    want = oracle.hpel(oracle.pad_luma(y[1], mbw, mbh))
    for k in range(3):
        np.testing.assert_array_equal(pb[1 + k, 1], want[k], err_msg=f"oracle plane {'HVC'[k]}")


def test_luma_planes_batch_matches_one_clip_at_a_time(gpu):
    """pcme_luma_planes_batch over clips of mixed sizes (more than one launch's
    16, narrow pictures whose blocks hold several bands, an empty job array)
    writes every clip's planes as pcme_luma_planes of that clip alone."""
    import torch
    sizes = [(352, 288, 3), (640, 360, 2), (1280, 720, 2), (1920, 1080, 2), (100, 70, 3), (16, 16, 2),
             (34, 18, 2), (52, 200, 2), (200, 16, 3)] * 2
    say("luma planes batch", len(sizes), "clips")
    want, got = [], []
    for i, (w, h, n) in enumerate(sizes):
        y, _, _ = _clip(gpu, w, h, n)
        mbw, mbh = gpu.mb_dims(w, h)
        a = gpu.DeviceClip(y, mbw, mbh)
        a.planes.fill_(0x5A)
        a.prepare()
        b = gpu.DeviceClip(y, mbw, mbh)
        b.planes.fill_(0xA5)
        want.append(a)
        got.append(b)
    gpu.prepare_clips(got)
    torch.cuda.synchronize()
    for (w, h, n), a, b in zip(sizes, want, got):
        assert torch.equal(a.planes, b.planes), f"{w}x{h} x{n}"
    assert gpu.load().pcme_luma_planes_batch(None, 0, None) == 0


CASES = [
    # (w, h, frames, me_range, subpel, partitions, qp, generator kwargs)
    (320, 240, 8, 16, 1, True, 28, dict(nrect=4, rect_lo=24, rect_hi=80, vmax=3.0, pan_max=1.0)),
    (352, 288, 4, 32, 1, True, 22, dict(nrect=6, rect_lo=24, rect_hi=120, vmax=8.0, pan_max=1.0)),
    (320, 240, 4, 8, 0, False, 37, dict(nrect=4, rect_lo=24, rect_hi=80, vmax=3.0, pan_max=1.0)),
    (320, 240, 4, 24, 2, True, 32, dict(nrect=4, rect_lo=24, rect_hi=80, vmax=6.0, pan_max=1.0)),
    (176, 144, 4, 64, 1, True, 27, dict(nrect=3, rect_lo=16, rect_hi=64, vmax=12.0, pan_max=2.0)),
    (16, 16, 3, 16, 1, True, 28, dict(nrect=1, rect_lo=4, rect_hi=8, vmax=2.0, pan_max=1.0)),
    (200, 16, 3, 16, 1, True, 28, dict(nrect=2, rect_lo=4, rect_hi=12, vmax=2.0, pan_max=1.0)),
    (16, 200, 3, 16, 1, True, 28, dict(nrect=2, rect_lo=4, rect_hi=12, vmax=2.0, pan_max=1.0)),
    (100, 70, 3, 16, 1, True, 28, dict(nrect=2, rect_lo=8, rect_hi=30, vmax=3.0, pan_max=1.0)),
    # tiny ranges with partitions: hexagon boxes of +-1 and +-2 px around each centre (the first
    # round priced in the candidate pass must clamp like the descent), motion beyond the range
    (320, 240, 4, 1, 1, True, 27, dict(nrect=4, rect_lo=24, rect_hi=80, vmax=6.0, pan_max=2.0)),
    (320, 240, 4, 2, 1, True, 32, dict(nrect=4, rect_lo=24, rect_hi=80, vmax=6.0, pan_max=2.0)),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}r{c[3]}s{c[4]}p{int(c[5])}q{c[6]}")
def test_search_frames_match_oracle(gpu, oracle, case):
    w, h, n, rng, sub, parts, qp, kw = case
    say("search_frames", case[:7])
    y, _, _ = _clip(gpu, w, h, n, **kw)
    mbw, mbh = gpu.mb_dims(w, h)
    dc = gpu.DeviceClip(y, mbw, mbh)
    got = gpu.search_clip_me(dc, rng, sub, parts, qp)
    want = oracle.search_clip_me(y, rng, sub, parts, qp)
    assert_same_field(got, want, f"{w}x{h}")


def test_search_frames_golden(gpu):
    """The committed golden vectors (tests/golden/), through the GPU path."""
    import os
    gdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    for f in sorted(x for x in os.listdir(gdir) if x.endswith(".npz")):
        say("golden", f)
        z = np.load(os.path.join(gdir, f))
        frames, prm = z["frames"], z["params"]
        n, h, w = frames.shape
        mbw, mbh = gpu.mb_dims(w, h)
        dc = gpu.DeviceClip(frames, mbw, mbh)
        got = gpu.search_clip_me(dc, int(prm[0]), int(prm[1]), bool(prm[2]), int(prm[3]))
        want = z["expected"].view(gpu.MB_DTYPE).reshape(got.shape)
        assert_same_field(got, want, f)


@pytest.mark.parametrize("scene", ["flat", "gradient", "noise", "motion"])
def test_search_frames_reference_scenes(gpu, oracle, scene):
    say("scene", scene)
    y = scene_clip(scene, 96, 64, 4)
    mbw, mbh = gpu.mb_dims(96, 64)
    dc = gpu.DeviceClip(y, mbw, mbh)
    got = gpu.search_clip_me(dc, 16, 1, True, 26)
    want = oracle.search_clip_me(y, 16, 1, True, 26)
    assert_same_field(got, want, scene)


def test_search_frames_1080p(gpu, oracle):
    """1920x1080 (coded 1920x1088), me_range 32, QP 27, 12 rects: the full
    frame size of the benchmark, two P frames against the oracle."""
    say("search_frames 1080p")
    y, _, _ = _clip(gpu, 1920, 1080, 3, nrect=12, rect_lo=64, rect_hi=400, vmax=8.0, pan_max=1.0)
    mbw, mbh = gpu.mb_dims(1920, 1080)
    dc = gpu.DeviceClip(y, mbw, mbh)
    got = gpu.search_clip_me(dc, 32, 1, True, 27)
    want = oracle.search_clip_me(y, 32, 1, True, 27)
    assert_same_field(got, want, "1080p")


def test_search_frames_repeat_is_deterministic(gpu):
    """Repeat launches (reused workspace, advancing exchange tags) and a
    different batch composition give identical results."""
    import torch
    say("determinism")
    y, _, _ = _clip(gpu, 320, 240, 6)
    mbw, mbh = gpu.mb_dims(320, 240)
    dc = gpu.DeviceClip(y, mbw, mbh)
    a = gpu.search_clip_me(dc, 16, 1, True, 28)
    for _ in range(3):
        b = gpu.search_clip_me(dc, 16, 1, True, 28)
        assert_same_field(b, a, "repeat")
    # one frame alone
    out = torch.empty(mbw * mbh * gpu.MB_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    gpu.search_frames([gpu.frame_job(dc, 3, 2, out.data_ptr(), 16, 1, True, 28)])
    single = out.cpu().numpy().view(gpu.MB_DTYPE).reshape(mbh, mbw)
    assert_same_field(single, a[2], "single job")


def test_search_batch_matches_oracle(gpu, oracle):
    """Per-call search() (slice.rs:1590-1723) on random blocks, predictors
    and candidates, against o_search."""
    import torch
    say("search_batch")
    rs = np.random.default_rng(7)
    y, _, _ = _clip(gpu, 160, 128, 2, nrect=3, rect_lo=16, rect_hi=48, vmax=4.0)
    mbw, mbh = gpu.mb_dims(160, 128)
    dc = gpu.DeviceClip(y, mbw, mbh)
    dc.pad()
    dc.hpel(0, 1)
    torch.cuda.synchronize()
    planes = dc.planes.cpu().numpy()
    ref = oracle.Ref(planes[0, 0])
    src_plane = planes[0, 1]
    stride = src_plane.shape[1]
    jobs, want = [], []
    lib = oracle.load()
    for i in range(300):
        w, h = [(16, 16), (16, 8), (8, 16), (8, 8)][i % 4]
        bx = int(rs.integers(0, 16 * mbw - w + 1)) // 4 * 4
        by = int(rs.integers(0, 16 * mbh - h + 1))
        my = by // 16
        vlim = int(lib.o_vlim(my, mbh))
        if vlim <= by + h:
            vlim = by + h + 1
        mvp = (int(rs.integers(-120, 121)), int(rs.integers(-120, 121)))
        c0 = (int(rs.integers(-200, 201)), int(rs.integers(-200, 201)))
        c1 = (int(rs.integers(-60, 61)), int(rs.integers(-60, 61)))
        rng_ = int(rs.choice([4, 16, 32]))
        sub = int(rs.integers(0, 3))
        lam = int(lib.o_lambda(int(rs.integers(0, 52))))
        nc = int(rs.integers(0, 3))
        p = oracle.oparams(mbw, mbh, rng_, sub, 1, vlim)
        cands = (oracle.omv * 2)(oracle.omv(*c0), oracle.omv(*c1))
        cost = ctypes.c_uint32()
        blk = np.ascontiguousarray(src_plane[64 + by:64 + by + h, 64 + bx:64 + bx + w])
        m = lib.o_search(ctypes.byref(p), ctypes.byref(ref.view), bx, by, w, h, blk.ctypes.data, w,
                         oracle.omv(*mvp), cands, nc, lam, ctypes.byref(cost), None)
        want.append((m.x, m.y, cost.value))
        sj = gpu.SearchJob()
        sj.src = dc.plane_ptr(0, 1) + (64 + by) * stride + 64 + bx
        sj.sstride = stride
        sj.ref, sj.ref_h, sj.ref_v, sj.ref_c = (dc.plane_ptr(k, 0) for k in range(4))
        sj.mbw, sj.mbh, sj.bx, sj.by, sj.w, sj.h = mbw, mbh, bx, by, w, h
        sj.mvp = gpu.Mv(*mvp)
        sj.cands[0] = gpu.Mv(*c0)
        sj.cands[1] = gpu.Mv(*c1)
        sj.ncands = nc
        sj.lam = lam
        sj.me_range = rng_
        sj.subpel_rounds = sub
        sj.vlim = vlim
        jobs.append(sj)
    omv = torch.zeros((len(jobs), 2), dtype=torch.int16, device="cuda")
    ocost = torch.zeros(len(jobs), dtype=torch.int32, device="cuda")
    gpu.search_batch(jobs, omv.data_ptr(), ocost.data_ptr())
    torch.cuda.synchronize()
    gm = omv.cpu().numpy()
    gc = ocost.cpu().numpy().view(np.uint32)
    for i, (x, yy, c) in enumerate(want):
        assert (gm[i, 0], gm[i, 1], gc[i]) == (x, yy, c), f"job {i}"


def test_pred_frame_matches_oracle(gpu, oracle):
    """mc_luma + mc_chroma of the chosen partitions (slice.rs:1725-1749)."""
    import torch
    say("pred_frame")
    w, h = 160, 96
    y, u, v = _clip(gpu, w, h, 2, nrect=3, rect_lo=16, rect_hi=48, vmax=4.0)
    mbw, mbh = gpu.mb_dims(w, h)
    dc = gpu.DeviceClip(y, mbw, mbh, u, v)
    res = gpu.search_clip_me(dc, 16, 1, True, 26)
    mb = torch.as_tensor(np.ascontiguousarray(res[0]).view(np.uint8).reshape(-1)).cuda()
    py = torch.zeros(mbw * mbh * 256, dtype=torch.uint8, device="cuda")
    pu = torch.zeros(mbw * mbh * 64, dtype=torch.uint8, device="cuda")
    pv = torch.zeros(mbw * mbh * 64, dtype=torch.uint8, device="cuda")
    gpu.pred_frame(dc, 0, mb.data_ptr(), py.data_ptr(), pu.data_ptr(), pv.data_ptr())
    torch.cuda.synchronize()
    gy = py.cpu().numpy().reshape(mbh, mbw, 16, 16)
    guv = (pu.cpu().numpy().reshape(mbh, mbw, 8, 8), pv.cpu().numpy().reshape(mbh, mbw, 8, 8))
    ref = oracle.Ref(dc.planes[0, 0].cpu().numpy())
    cu = oracle.pad_chroma(u[0], mbw, mbh)
    cv = oracle.pad_chroma(v[0], mbw, mbh)
    lib = oracle.load()
    rects = {0: [(0, 0, 16, 16)], 1: [(0, 0, 16, 8), (0, 8, 16, 8)], 2: [(0, 0, 8, 16), (8, 0, 8, 16)],
             3: [(0, 0, 8, 8), (8, 0, 8, 8), (0, 8, 8, 8), (8, 8, 8, 8)]}
    for my in range(mbh):
        for mx in range(mbw):
            r = res[0, my, mx]
            wy = np.zeros((16, 16), np.uint8)
            wu = np.zeros((8, 8), np.uint8)
            wv = np.zeros((8, 8), np.uint8)
            for p, (x, yy, bw, bh) in enumerate(rects[int(r["part"])]):
                mv = oracle.omv(int(r["mvs"][p][0]), int(r["mvs"][p][1]))
                tmp = np.zeros((16, 16), np.uint8)
                lib.o_mc_luma(ctypes.byref(ref.view), 16 * mx + x, 16 * my + yy, mv, bw, bh, tmp.ctypes.data, 16)
                wy[yy:yy + bh, x:x + bw] = tmp[:bh, :bw]
                tu = np.zeros((8, 8), np.uint8)
                tv = np.zeros((8, 8), np.uint8)
                lib.o_mc_chroma(cu.ctypes.data, cv.ctypes.data, cu.shape[1], 32, (16 * mx + x) // 2,
                                (16 * my + yy) // 2, mv, bw // 2, bh // 2, tu.ctypes.data, tv.ctypes.data, 8)
                wu[yy // 2:(yy + bh) // 2, x // 2:(x + bw) // 2] = tu[:bh // 2, :bw // 2]
                wv[yy // 2:(yy + bh) // 2, x // 2:(x + bw) // 2] = tv[:bh // 2, :bw // 2]
            np.testing.assert_array_equal(gy[my, mx], wy, err_msg=f"luma MB {mx},{my}")
            np.testing.assert_array_equal(guv[0][my, mx], wu, err_msg=f"U MB {mx},{my}")
            np.testing.assert_array_equal(guv[1][my, mx], wv, err_msg=f"V MB {mx},{my}")


# ------------------------------------------------------------------ plans

def _jobs_for(gpu, dc, outs, rng, sub, parts, qp):
    return [gpu.frame_job(dc, t, t - 1, outs[t - 1].data_ptr(), rng, sub, parts, qp) for t in range(1, dc.n)]


def _field(gpu, outs, n, mbh, mbw):
    return outs[:n].cpu().numpy().view(gpu.MB_DTYPE).reshape(n, mbh, mbw)


def test_plan_graph_replay_matches_oracle(gpu, oracle):
    """A captured hipGraph of a plan launch, replayed over changed inputs
    (eager pad/hpel of a new clip between replays), equals the oracle on
    every replay: the launch tag lives in device memory."""
    import torch
    say("plan graph replay")
    w, h, n = 160, 96, 4
    mbw, mbh = gpu.mb_dims(w, h)
    clips = [_clip(gpu, w, h, n, seed=0x5EED0000 + k, nrect=3, rect_lo=16, rect_hi=48, vmax=4.0)[0]
             for k in range(3)]
    dc = gpu.DeviceClip(clips[0], mbw, mbh)
    outs = torch.zeros((n - 1, mbw * mbh * gpu.MB_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    plan = gpu.Plan(_jobs_for(gpu, dc, outs, 16, 1, True, 28))
    dc.pad()
    dc.hpel(0, n - 1)
    torch.cuda.synchronize()
    plan.launch()  # eager launch before capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
            plan.launch(s.cuda_stream)
    for k in (1, 2, 0, 1):
        dc.raw.copy_(torch.as_tensor(clips[k]).cuda())
        dc.pad()
        dc.hpel(0, n - 1)
        outs.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        plan.check()
        want = oracle.search_clip_me(clips[k], 16, 1, True, 28)
        assert_same_field(_field(gpu, outs, n - 1, mbh, mbw), want, f"replay of clip {k}")
    plan.close()


def test_batch_of_uneven_clips_in_one_launch(gpu, oracle):
    """Several clips of different sizes and parameters searched by one plan
    (one persistent launch), each equal to the oracle."""
    import torch
    say("uneven batch")
    specs = [(176, 144, 3, 32, 1, True, 22), (320, 240, 3, 16, 1, True, 37), (96, 64, 4, 8, 0, False, 27),
             (352, 288, 2, 32, 2, True, 32)]
    dcs, outs, jobs, ys = [], [], [], []
    for k, (w, h, n, rng, sub, parts, qp) in enumerate(specs):
        y, _, _ = _clip(gpu, w, h, n, seed=0x5EED0000 + 17 * k, nrect=4, rect_lo=16, rect_hi=80, vmax=6.0)
        mbw, mbh = gpu.mb_dims(w, h)
        dc = gpu.DeviceClip(y, mbw, mbh)
        dc.pad()
        dc.hpel(0, n - 1)
        o = torch.zeros((n - 1, mbw * mbh * gpu.MB_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
        jobs += _jobs_for(gpu, dc, o, rng, sub, parts, qp)
        dcs.append(dc)
        outs.append(o)
        ys.append(y)
    torch.cuda.synchronize()
    plan = gpu.Plan(jobs)
    plan.launch()
    plan.check()
    for k, (w, h, n, rng, sub, parts, qp) in enumerate(specs):
        mbw, mbh = gpu.mb_dims(w, h)
        want = oracle.search_clip_me(ys[k], rng, sub, parts, qp)
        assert_same_field(_field(gpu, outs[k], n - 1, mbh, mbw), want, f"clip {k} {w}x{h}")
    plan.close()


def test_autotune_keeps_results(gpu, oracle):
    import torch
    say("autotune")
    w, h, n = 320, 240, 5
    y, _, _ = _clip(gpu, w, h, n)
    mbw, mbh = gpu.mb_dims(w, h)
    dc = gpu.DeviceClip(y, mbw, mbh)
    dc.pad()
    dc.hpel(0, n - 1)
    outs = torch.zeros((n - 1, mbw * mbh * gpu.MB_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    plan = gpu.Plan(_jobs_for(gpu, dc, outs, 16, 1, True, 28))
    waves, ms = plan.autotune(reps=2)
    assert waves >= 1 and ms > 0
    for grid in (1, 7, 0):
        plan.set_grid(grid)
        outs.zero_()
        plan.launch()
        plan.check()
        assert_same_field(_field(gpu, outs, n - 1, mbh, mbw), oracle.search_clip_me(y, 16, 1, True, 28),
                          f"grid {grid}")
    plan.close()


def test_every_kernel_configuration_gives_the_same_field(gpu, oracle):
    """Each built configuration (row walkers per workgroup, occupancy cap,
    LDS region reuse) searches two 640x368 P frames bit-exact with the
    oracle, at its resident grid and at a small one; autotune_ex then keeps
    a built configuration and its field is exact too."""
    import torch
    w, h, n = 640, 360, 3
    y, _, _ = _clip(gpu, w, h, n)
    mbw, mbh = gpu.mb_dims(w, h)
    dc = gpu.DeviceClip(y, mbw, mbh)
    dc.pad()
    dc.hpel(0, n - 1)
    outs = torch.zeros((n - 1, mbw * mbh * gpu.MB_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    plan = gpu.Plan(_jobs_for(gpu, dc, outs, 32, 1, True, 27))
    want = oracle.search_clip_me(y, 32, 1, True, 27)
    for wpg, wcu, reuse in gpu.CONFIGS:
        for waves in (0, 9):
            say("configuration", wpg, wcu, reuse, "waves", waves)
            t = plan.set_config(waves, wpg, wcu, reuse)
            assert (t.waves_per_group, t.waves_per_cu, t.reuse_regions) == (wpg, wcu, reuse)
            assert t.waves >= 1 and (waves == 0 or t.waves == waves)
            outs.zero_()
            plan.launch()
            plan.check()
            assert_same_field(_field(gpu, outs, n - 1, mbh, mbw), want, f"config {wpg}/{wcu}/{reuse} waves {waves}")
    with pytest.raises(gpu.PcmeError):
        plan.set_config(0, 3, 12, 1)
    say("autotune_ex")
    best = plan.autotune_ex(reps=1)
    assert (best.waves_per_group, best.waves_per_cu, best.reuse_regions) in gpu.CONFIGS and best.ms > 0
    outs.zero_()
    plan.launch()
    plan.check()
    assert_same_field(_field(gpu, outs, n - 1, mbh, mbw), want, "autotuned")
    plan.close()


def test_stall_path_reports_estall(gpu):
    """With the dependency-wait bound forced to one poll, rows that must wait
    give up, raise the flag, every wave exits, and the call fails with
    PCME_ESTALL instead of hanging; the default bound then works again."""
    import torch
    say("stall path")
    lib = gpu.load()
    w, h, n = 640, 360, 3
    y, _, _ = _clip(gpu, w, h, n)
    mbw, mbh = gpu.mb_dims(w, h)
    dc = gpu.DeviceClip(y, mbw, mbh)
    dc.pad()
    dc.hpel(0, n - 1)
    outs = torch.zeros((n - 1, mbw * mbh * gpu.MB_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    plan = gpu.Plan(_jobs_for(gpu, dc, outs, 16, 1, True, 28))
    old = lib.pcme_set_spin_limit(1)
    try:
        plan.launch()
        with pytest.raises(gpu.PcmeError) as e:
            plan.check()
        assert e.value.code == gpu.PCME_ESTALL
    finally:
        lib.pcme_set_spin_limit(old)
    plan.launch()
    plan.check()
    plan.close()


def test_search_frames_4k_me_range_64(gpu, oracle):
    """3840x2160, me_range 64, QP 27: one P frame against the oracle."""
    say("4k")
    y, _, _ = _clip(gpu, 3840, 2160, 2, nrect=8, rect_lo=128, rect_hi=900, vmax=16.0, pan_max=1.0, zoom=0.002)
    mbw, mbh = gpu.mb_dims(3840, 2160)
    dc = gpu.DeviceClip(y, mbw, mbh)
    got = gpu.search_clip_me(dc, 64, 1, True, 27)
    want = oracle.search_clip_me(y, 64, 1, True, 27)
    assert_same_field(got, want, "4k")
