"""pcme_search_frames_noalloc (workspace in a static device array, no
allocation) against pcme_search_frames_sync, bit for bit: two calls in a
row, the second with a smaller input, then an input beyond the array's
limit, which must fail with PCME_EINVAL before any GPU work.  Collected
last (the file name), and passes alone or after the other tests."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from _util import assert_same_field

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _run(gpu, jobs, fn):
    arr = (gpu.FrameJob * len(jobs))(*jobs)
    import torch
    return fn(arr, len(jobs), torch.cuda.current_stream().cuda_stream)


def _field(gpu, out, n, mbh, mbw):
    return out[:n].cpu().numpy().view(gpu.MB_DTYPE).reshape(n, mbh, mbw)



def test_noalloc_matches_search_frames(gpu):
    import torch
    lib = gpu.load()
    for w, h, n, qp in [(640, 368, 6, 27), (320, 240, 3, 32)]:  # the second call is smaller
        say("noalloc vs search_frames_sync", w, h, n, qp)
        y, _, _ = gpu.synth_clip(gpu.ClipSpec(w, h, n, seed=0x5EED3000 + w, nrect=5, vmax=6.0))
        mbw, mbh = gpu.mb_dims(w, h)
        dc = gpu.DeviceClip(y, mbw, mbh)
        dc.prepare()
        nb = mbw * mbh * gpu.MB_DTYPE.itemsize
        a = torch.full((n - 1, nb), 0xA5, dtype=torch.uint8, device="cuda")
        b = torch.full((n - 1, nb), 0x5A, dtype=torch.uint8, device="cuda")
        ja = [gpu.frame_job(dc, t, t - 1, a[t - 1].data_ptr(), 32, 1, True, qp) for t in range(1, n)]
        jb = [gpu.frame_job(dc, t, t - 1, b[t - 1].data_ptr(), 32, 1, True, qp) for t in range(1, n)]
        torch.cuda.synchronize()
        gpu._check(_run(gpu, ja, lib.pcme_search_frames_sync))
        gpu._check(_run(gpu, jb, lib.pcme_search_frames_noalloc))
        torch.cuda.synchronize()
        assert_same_field(_field(gpu, b, n - 1, mbh, mbw), _field(gpu, a, n - 1, mbh, mbw), f"{w}x{h}")


def test_noalloc_rejects_beyond_its_workspace(gpu):
    import torch
    lib = gpu.load()
    say("noalloc over the limit")
    w, h = 1920, 1080
    mbw, mbh = gpu.mb_dims(w, h)
    y = np.zeros((2, h, w), np.uint8)
    dc = gpu.DeviceClip(y, mbw, mbh)
    out = torch.zeros(mbw * mbh * gpu.MB_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    job = gpu.frame_job(dc, 1, 0, out.data_ptr(), 32, 1, True, 27)
    rc = _run(gpu, [job] * 200, lib.pcme_search_frames_noalloc)  # 200 1080p frames: 26 MB of exchange words
    assert rc == gpu.PCME_EINVAL
    assert b"workspace" in lib.pcme_last_error()
