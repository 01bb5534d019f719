"""Calls from several host threads on several streams at once.

The reference calls its search from up to 64 frame threads concurrently
(enc/mod.rs:207-238) and its output does not depend on the thread count
(tests.rs:212-224).  Here 4 host threads, each with a stream of its own,
call pcme_search_part_batch, pcme_b_analyse_batch and pcme_search_frames
with job tables of their own, several times each, while the others run;
every output is compared with the oracle bit for bit.  A job table shared
between calls would be overwritten by another thread's upload while a
kernel reads it, and the outputs would differ.

Each case prints what it runs before the GPU step."""
from __future__ import annotations

import ctypes
import threading

import numpy as np
import pytest

from _util import assert_same_field
from test_b_oracle import _rand_mbinfo2
from test_gpu_b_analyse import _nb
from test_gpu_search_part import _nb_block, _random_mbs

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _part_jobs(pcme, O, dc, y, seed, n):
    mbw, mbh = dc.mbw, dc.mbh
    rng = np.random.default_rng(100 + seed)
    mbs = _random_mbs(O, mbw, mbh, rng)
    src = O.pad_luma(y[1], mbw, mbh)
    ref = O.Ref(O.pad_luma(y[0], mbw, mbh))
    lam = O.load().o_lambda(22 + 5 * (seed % 4))
    jobs, want = [], []
    for k in range(n):
        mx, my = int(rng.integers(0, mbw)), int(rng.integers(0, mbh))
        part = (k + seed) % 4
        seedmv = (int(rng.integers(-40, 41)), int(rng.integers(-40, 41)))
        vlim = O.load().o_vlim(my, mbh)
        prm = O.oparams(mbw, mbh, 32, 1, 1, vlim)
        src16 = np.ascontiguousarray(src[64 + 16 * my:80 + 16 * my, 64 + 16 * mx:80 + 16 * mx])
        mvs, mvps = (O.omv * 4)(), (O.omv * 4)()
        total = O.load().o_search_part(ctypes.byref(prm), ctypes.cast(mbs, ctypes.c_void_p), mx, my,
                                       ctypes.byref(ref.view), part, src16.ctypes.data, lam, O.omv(*seedmv),
                                       mvs, mvps, None)
        want.append(([(m.x, m.y) for m in mvs], [(m.x, m.y) for m in mvps], int(total)))
        j = pcme.PartJob()
        j.src = dc.plane_ptr(0, 1) + (64 + 16 * my) * dc.stride + 64 + 16 * mx
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
    return jobs, want


def _b_jobs(pcme, O, dc, y, seed, n):
    mbw, mbh = dc.mbw, dc.mbh
    rs = np.random.default_rng(200 + seed)
    mbs = (O.ombinfo2 * (mbw * mbh))()
    for i in range(mbw * mbh):
        _rand_mbinfo2(O, mbs[i], rs)
    src = O.pad_luma(y[1], mbw, mbh)
    refs = [O.Ref(O.pad_luma(y[0], mbw, mbh)), O.Ref(O.pad_luma(y[2], mbw, mbh))]
    lam = O.load().o_lambda(27)
    corner = (0, 3, 12, 15)
    jobs, want = [], []
    for _ in range(n):
        mx, my = int(rs.integers(0, mbw)), int(rs.integers(0, mbh))
        vlim = O.load().o_vlim(my, mbh)
        prm = O.oparams(mbw, mbh, 32, 1, 1, vlim)
        col = O.ombinfo2()
        _rand_mbinfo2(O, col, rs)
        s16 = np.ascontiguousarray(src[64 + 16 * my:80 + 16 * my, 64 + 16 * mx:80 + 16 * mx])
        want.append(O.b_analyse(prm, mbs, col, mx, my, refs[0], refs[1], s16, lam))
        j = pcme.BJob()
        j.src = dc.plane_ptr(0, 1) + (64 + 16 * my) * dc.stride + 64 + 16 * mx
        j.sstride = dc.stride
        for k in range(4):
            j.ref0[k] = dc.plane_ptr(k, 0)
            j.ref1[k] = dc.plane_ptr(k, 2)
        j.mbw, j.mbh, j.mx, j.my = mbw, mbh, mx, my
        j.lam, j.me_range, j.subpel_rounds, j.vlim = lam, 32, 1, vlim
        for l in range(2):
            for i in range(4):
                j.nb[l][i] = _nb(pcme, O, mbs, mbw, mx, my, -1, 0, 4 * i + 3, l)
                j.nb[l][4 + i] = _nb(pcme, O, mbs, mbw, mx, my, 0, -1, 12 + i, l)
            j.nb[l][8] = _nb(pcme, O, mbs, mbw, mx, my, 1, -1, 12, l)
            j.nb[l][9] = _nb(pcme, O, mbs, mbw, mx, my, -1, -1, 15, l)
            for q in range(4):
                j.col_ref[l][q] = col.ref_idx[l][q]
                j.col_mv[l][q] = pcme.Mv(col.mv[l][corner[q]].x, col.mv[l][corner[q]].y)
        j.col_intra = 1 if col.kind in (O.OK_I4x4, O.OK_I8x8, O.OK_I16x16) else 0
        jobs.append(j)
    return jobs, want


def test_four_threads_four_streams(gpu, oracle):
    import torch
    pcme, O = gpu, oracle
    nthreads, reps = 4, 3
    w, h = 320, 240
    mbw, mbh = pcme.mb_dims(w, h)
    say("threads: building jobs")
    clips, parts, bs, frames = [], [], [], []
    for t in range(nthreads):
        y, _, _ = pcme.synth_clip(pcme.ClipSpec(w, h, 3, seed=0x5EED2000 + t, nrect=5, vmax=5.0))
        dc = pcme.DeviceClip(y, mbw, mbh)
        dc.pad()
        dc.hpel()
        clips.append((y, dc))
        parts.append(_part_jobs(pcme, O, dc, y, t, 48 + 16 * t))
        bs.append(_b_jobs(pcme, O, dc, y, t, 24 + 8 * t))
        frames.append(O.search_clip_me(y[:2], 16, 1, True, 24 + 3 * t)[0])
    streams = [torch.cuda.Stream() for _ in range(nthreads)]
    pout = [torch.zeros(len(parts[t][0]) * pcme.PART_OUT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
            for t in range(nthreads)]
    bout = [torch.zeros(len(bs[t][0]) * pcme.B_OUT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
            for t in range(nthreads)]
    fout = [torch.zeros(mbw * mbh * pcme.MB_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
            for _ in range(nthreads)]
    torch.cuda.synchronize()
    errors, results = [], {}
    barrier = threading.Barrier(nthreads)

    def run(t):
        try:
            s = streams[t].cuda_stream
            dc = clips[t][1]
            fjob = pcme.frame_job(dc, 1, 0, fout[t].data_ptr(), 16, 1, True, 24 + 3 * t)
            barrier.wait()
            for r in range(reps):
                pcme.search_part_batch(parts[t][0], pout[t].data_ptr(), s)
                pcme.b_analyse_batch(bs[t][0], bout[t].data_ptr(), s)
                pcme.search_frames([fjob], s, sync=False)
                streams[t].synchronize()
                results[(t, r)] = (pout[t].cpu().numpy().copy(), bout[t].cpu().numpy().copy(),
                                   fout[t].cpu().numpy().copy())
            pcme.search_frames_check(s)
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append((t, repr(e)))

    say("threads: 4 threads x 3 rounds of part/B/frame calls on 4 streams")
    th = [threading.Thread(target=run, args=(t,)) for t in range(nthreads)]
    for x in th:
        x.start()
    for x in th:
        x.join(timeout=300)
    assert not any(x.is_alive() for x in th), "a thread did not finish"
    assert not errors, errors
    for t in range(nthreads):
        for r in range(reps):
            pb, bb, fb = results[(t, r)]
            got = pb.view(pcme.PART_OUT_DTYPE)
            for k, ((mvs, mvps, total), g) in enumerate(zip(parts[t][1], got)):
                assert [tuple(int(v) for v in m) for m in g["mvs"]] == mvs, (t, r, k)
                assert [tuple(int(v) for v in m) for m in g["mvps"]] == mvps, (t, r, k)
                assert int(g["total"]) == total, (t, r, k)
            gb = bb.view(pcme.B_OUT_DTYPE)
            for k, (g, wnt) in enumerate(zip(gb, bs[t][1])):
                assert g.tobytes() == wnt.tobytes(), (t, r, k)
            assert_same_field(fb.view(pcme.MB_DTYPE).reshape(mbh, mbw), frames[t], f"thread {t} round {r}")
