"""The ME-only field of a whole B picture on the GPU (pcme_b_field: one
launch of row walkers advancing as a wavefront, each MB searched over the
MBs committed before it), bit for bit against the oracle's o_b_field
(analyse_b's searches and decision with rd off, raster order), at 320x240
with and without partitions and at 1080p; a one-row and a one-column
picture; and the stall path (pcme_b_field_check).

Each case prints what it runs before the GPU step so a failing run names
its case."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _oracle_cols(oracle, field_p):
    """ombinfo2 of the list-1 picture from its P field (list 0 only)."""
    mbh, mbw = field_p.shape
    cols = (oracle.ombinfo2 * (mbw * mbh))()
    for i, r in enumerate(field_p.reshape(-1)):
        cols[i].kind = oracle.OK_PINTER
        for q in range(4):
            cols[i].ref_idx[0][q] = int(r["ref"])
            cols[i].ref_idx[1][q] = -1
        for b in range(16):
            cols[i].mv[0][b] = oracle.omv(int(r["mv"][b][0]), int(r["mv"][b][1]))
    return cols


def _oracle_field(oracle, y, t, fp, qp, partitions, me_range=32):
    """o_b_field of picture t of y between t-1 and t+1, the co-located MBs
    from P field fp of picture t+1."""
    mbh, mbw = fp.shape
    prm = oracle.oparams(mbw, mbh, me_range, 1, 1 if partitions else 0, 0)
    return oracle.b_field(prm, _oracle_cols(oracle, fp), oracle.Ref(oracle.pad_luma(y[t - 1], mbw, mbh)),
                          oracle.Ref(oracle.pad_luma(y[t + 1], mbw, mbh)), oracle.pad_luma(y[t], mbw, mbh),
                          oracle.load().o_lambda(qp)).reshape(mbh, mbw)


def _assert_same_b(got, want, what):
    mbh, mbw = want.shape
    g = np.ascontiguousarray(got).view(np.uint8).reshape(mbh * mbw, -1)
    wt = np.ascontiguousarray(want).view(np.uint8).reshape(mbh * mbw, -1)
    bad = np.nonzero((g != wt).any(axis=1))[0]
    assert bad.size == 0, f"{what}: {bad.size} of {mbw * mbh} B records differ, first MB {divmod(int(bad[0]), mbw)}"


@pytest.mark.parametrize("w,h,qp,partitions", [(320, 240, 27, True), (320, 240, 32, False), (1920, 1080, 27, True),
                                                (320, 16, 27, True), (16, 240, 22, True)])
def test_b_field_matches_oracle(gpu, oracle, w, h, qp, partitions):
    from pcme import bsearch as B
    say("B field", w, h, qp, partitions)
    y, _, _ = gpu.synth_clip(gpu.ClipSpec(w, h, 3, seed=0x5EED6000 + w + h, nrect=6 if min(w, h) > 16 else 1,
                                          vmax=6.0))
    mbw, mbh = gpu.mb_dims(w, h)
    dc = gpu.DeviceClip(y, mbw, mbh)
    dc.prepare()
    fp = oracle.search_clip_me(y[1:3], 32, 1, True, qp)[0]  # picture 2 coded as P from picture 1
    got = B.b_field(dc, 1, (dc, 0), (dc, 2), B.col_records(fp, mbh, mbw), qp, 32, 1, partitions)
    want = _oracle_field(oracle, y, 1, fp, qp, partitions)
    _assert_same_b(got, want, f"{w}x{h}")
    if min(mbw, mbh) > 1:
        modes = np.bincount(got["mode"].reshape(-1), minlength=4)
        assert modes[0] > 0 and modes[1] > 0 and (modes[2] + modes[3] > 0) == partitions


def test_b_fields_one_launch_matches_oracle(gpu, oracle):
    """pcme_b_fields: five B pictures of three clips (two sizes, QP 22-37,
    with and without partitions, me_range 16 and 32) in one launch, each
    bit for bit its oracle field; then the same jobs again in another
    order, the records unchanged."""
    import torch
    from pcme import bsearch as B
    clips = []
    for k, (w, h, n) in enumerate(((320, 240, 5), (176, 144, 3), (640, 368, 3))):
        y, _, _ = gpu.synth_clip(gpu.ClipSpec(w, h, n, seed=0x5EED6200 + k, nrect=5, vmax=6.0))
        mbw, mbh = gpu.mb_dims(w, h)
        dc = gpu.DeviceClip(y, mbw, mbh)
        dc.prepare()
        clips.append((y, dc))
    # (clip, B picture, qp, partitions, me_range)
    cases = [(0, 1, 27, True, 32), (0, 3, 32, True, 32), (1, 1, 37, False, 32), (2, 1, 22, True, 16),
             (0, 2, 27, True, 24)]
    jobs, outs, want, keep = [], [], [], []
    for ci, t, qp, part, rng in cases:
        y, dc = clips[ci]
        fp = oracle.search_clip_me(y[t:t + 2], rng, 1, True, qp)[0]
        col = B.col_records(fp, dc.mbh, dc.mbw)
        dcol = torch.as_tensor(np.ascontiguousarray(col.reshape(-1)).view(np.uint8)).to("cuda")
        o = torch.full((dc.mbw * dc.mbh * gpu.B_MB_DTYPE.itemsize,), 0x5A, dtype=torch.uint8, device="cuda")
        jobs.append(B.b_field_job(dc, t, (dc, t - 1), (dc, t + 1), dcol.data_ptr(), qp, rng, 1, part))
        outs.append(o)
        keep.append(dcol)
        want.append(_oracle_field(oracle, y, t, fp, qp, part, rng))
    say("B fields, five pictures in one launch")
    B.b_fields_launch(jobs, [o.data_ptr() for o in outs])
    B.b_field_check()
    got = [o.cpu().numpy().view(gpu.B_MB_DTYPE).reshape(w_.shape) for o, w_ in zip(outs, want)]
    for k, (g, w_) in enumerate(zip(got, want)):
        _assert_same_b(g, w_, f"picture {k} {cases[k]}")
    say("B fields, reversed order")
    for o in outs:
        o.fill_(0xA5)
    B.b_fields_launch(jobs[::-1], [o.data_ptr() for o in outs[::-1]])
    B.b_field_check()
    for k, (o, g) in enumerate(zip(outs, got)):
        assert o.cpu().numpy().tobytes() == g.tobytes(), f"picture {k} changed with the order"


def test_b_field_stall_reports_estall(gpu, oracle):
    """With the dependency-wait bound forced to one poll, rows that must wait
    give up, every walker leaves, and pcme_b_field_check reports
    PCME_ESTALL (once: the flag is then clear); with the default bound the
    next field is whole and equal to one computed before."""
    import torch
    from pcme import bsearch as B
    say("B field stall path")
    lib = gpu.load()
    w, h, qp = 320, 240, 27
    y, _, _ = gpu.synth_clip(gpu.ClipSpec(w, h, 3, seed=0x5EED6100, nrect=6, vmax=6.0))
    mbw, mbh = gpu.mb_dims(w, h)
    dc = gpu.DeviceClip(y, mbw, mbh)
    dc.prepare()
    fp = oracle.search_clip_me(y[1:3], 32, 1, True, qp)[0]
    col = B.col_records(fp, mbh, mbw)
    first = B.b_field(dc, 1, (dc, 0), (dc, 2), col, qp)
    dcol = torch.as_tensor(np.ascontiguousarray(col.reshape(-1)).view(np.uint8)).to("cuda")
    dout = torch.zeros(mbw * mbh * gpu.B_MB_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    old = lib.pcme_set_spin_limit(1)
    try:
        B.b_field_launch(dc, 1, (dc, 0), (dc, 2), dcol.data_ptr(), dout.data_ptr(), qp)
        with pytest.raises(gpu.PcmeError) as e:
            B.b_field_check()
        assert e.value.code == gpu.PCME_ESTALL
        B.b_field_check()  # cleared
    finally:
        lib.pcme_set_spin_limit(old)
    again = B.b_field(dc, 1, (dc, 0), (dc, 2), col, qp)
    assert again.tobytes() == first.tobytes()
