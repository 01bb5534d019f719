"""GPU parity of the B macroblock motion search (pcme_b_analyse_batch:
direct_spatial, the direct prediction's cost, list-0/list-1 16x16
search_part, the bi-average cost and direction, b_partition 16x8/8x16)
against the oracle's o_b_analyse, record for record, over random two-list
neighbourhoods and co-located MBs.

Each case prints what it runs before the GPU step so a failing run names
its case."""
from __future__ import annotations

import numpy as np
import pytest

from test_b_oracle import _rand_mbinfo2

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _nb(pcme, O, mbs, mbw, mx, my, dx, dy, r, l):
    b = pcme.NbBlock()
    nx, ny = mx + dx, my + dy
    if nx < 0 or nx >= mbw or ny < 0:
        b.ref = -2
        return b
    m = mbs[ny * mbw + nx]
    if m.kind in (O.OK_I4x4, O.OK_I8x8, O.OK_I16x16):
        b.ref = -1
        return b
    rf = m.ref_idx[l][((r >> 2) >> 1) * 2 + ((r & 3) >> 1)]
    b.ref = rf
    if rf >= 0:
        b.mv = pcme.Mv(m.mv[l][r].x, m.mv[l][r].y)
    return b


@pytest.mark.parametrize("seed", [0, 1])
def test_b_analyse_batch_matches_oracle(gpu, oracle, seed):
    import torch
    pcme, O = gpu, oracle
    w, h, qp = 192, 128, 27
    mbw, mbh = pcme.mb_dims(w, h)
    say("b_analyse batch", seed)
    y, _, _ = pcme.synth_clip(pcme.ClipSpec(w, h, 3, seed=0x5EED1800 + seed, nrect=4, rect_lo=16, rect_hi=64,
                                            vmax=5.0))
    dc = pcme.DeviceClip(y, mbw, mbh)
    dc.pad()
    dc.hpel()
    torch.cuda.synchronize()
    rs = np.random.default_rng(seed + 7)
    mbs = (O.ombinfo2 * (mbw * mbh))()
    for i in range(mbw * mbh):
        _rand_mbinfo2(O, mbs[i], rs)
    src = O.pad_luma(y[1], mbw, mbh)
    refs = [O.Ref(O.pad_luma(y[0], mbw, mbh)), O.Ref(O.pad_luma(y[2], mbw, mbh))]
    lam = O.load().o_lambda(qp)
    corner = (0, 3, 12, 15)
    jobs, want = [], []
    for my in range(mbh):
        vlim = O.load().o_vlim(my, mbh)
        prm = O.oparams(mbw, mbh, 32, 1, 1, vlim)
        for mx in range(mbw):
            col = O.ombinfo2()
            _rand_mbinfo2(O, col, rs)
            if rs.random() < 0.3:
                for l in range(2):
                    for b in range(16):
                        col.mv[l][b] = O.omv(int(rs.integers(-1, 2)), int(rs.integers(-1, 2)))
            s16 = np.ascontiguousarray(src[64 + 16 * my:80 + 16 * my, 64 + 16 * mx:80 + 16 * mx])
            want.append(O.b_analyse(prm, mbs, col, mx, my, refs[0], refs[1], s16, lam))
            j = pcme.BJob()
            j.src = dc.plane_ptr(0, 1) + (pcme.LUMA_PAD + 16 * my) * dc.stride + pcme.LUMA_PAD + 16 * mx
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
    out = torch.zeros(len(jobs) * pcme.B_OUT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    pcme.b_analyse_batch(jobs, out.data_ptr())
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(pcme.B_OUT_DTYPE)
    n_direct = sum(int(r["direct_ok"]) for r in want)
    assert n_direct > 0
    for k, (g, wnt) in enumerate(zip(got, want)):
        for f in pcme.B_OUT_DTYPE.names:
            assert np.array_equal(g[f], wnt[f]), (k, divmod(k, mbw)[::-1], f, g[f], wnt[f])


def test_b_picture_through_the_job_builder(gpu, oracle):
    """A whole 352x288 B picture: pcme.bsearch.b_jobs from random committed
    fields (intra MBs, refs -1/0/1 per list, MVs within +-16 px) and
    co-located MBs, against the oracle's o_b_frame_mt on the same fields."""
    import torch
    from pcme import bsearch as B
    pcme, O = gpu, oracle
    w, h, qp = 352, 288, 32
    mbw, mbh = pcme.mb_dims(w, h)
    say("b picture 352x288")
    y, _, _ = pcme.synth_clip(pcme.ClipSpec(w, h, 3, seed=0x5EED1A00, nrect=5, vmax=6.0))
    dc = pcme.DeviceClip(y, mbw, mbh)
    dc.pad()
    dc.hpel()
    rs = np.random.default_rng(77)
    mv = rs.integers(-64, 65, (2, mbh, mbw, 16, 2)).astype(np.int16)
    ref = rs.choice([-1, 0, 0, 0, 1], (2, mbh, mbw, 4)).astype(np.int8)
    intra = rs.random((mbh, mbw)) < 0.1
    col_mv = rs.integers(-2, 3, (mbh, mbw, 2, 16, 2)).astype(np.int16)
    col_ref = rs.choice([-1, 0, 1], (mbh, mbw, 2, 4)).astype(np.int8)
    col_intra = rs.random((mbh, mbw)) < 0.1
    jobs = B.b_jobs(dc, 1, (dc, 0), (dc, 2), mv, ref, intra, col_mv, col_ref, col_intra, qp, 24, 1)
    n = mbw * mbh
    out = torch.zeros(n * pcme.B_OUT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    B.b_analyse(jobs, out.data_ptr())
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(pcme.B_OUT_DTYPE)
    mbs = np.zeros(n, [("kind", "i1"), ("ref_idx", "i1", (2, 4)), ("pad", "u1"), ("mv", "<i2", (2, 16, 2))])
    mbs["kind"] = np.where(intra.reshape(n), O.OK_I16x16, O.OK_PINTER)
    mbs["ref_idx"] = ref.transpose(1, 2, 0, 3).reshape(n, 2, 4)
    mbs["mv"] = mv.transpose(1, 2, 0, 3, 4).reshape(n, 2, 16, 2)
    cols = np.zeros_like(mbs)
    cols["kind"] = np.where(col_intra.reshape(n), O.OK_I16x16, O.OK_PINTER)
    cols["ref_idx"] = col_ref.reshape(n, 2, 4)
    cols["mv"] = col_mv.reshape(n, 2, 16, 2)
    planes = dc.planes[0, :3].cpu().numpy()
    prm = O.oparams(mbw, mbh, 24, 1, 1, 0)
    _, want = O.b_frame_mt(prm, (O.ombinfo2 * n).from_buffer(mbs), (O.ombinfo2 * n).from_buffer(cols),
                           O.Ref(planes[0]), O.Ref(planes[2]), planes[1], O.load().o_lambda(qp), 4)
    for k in range(n):
        assert got[k].tobytes() == want[k].tobytes(), (k, divmod(k, mbw)[::-1])
