"""How often the GPU's float QP offsets change an MB's QP.

The AQ offsets (lookahead.rs:116-139) and the macroblock-tree offsets
(mbtree.rs:225-231) end in log2, which the GPU evaluates with its own
log2f: they match the reference's C-library log2f within 1e-5, not to the
bit.  The encoder turns their sum into an MB QP as
round(qpf + o).clamp(0, 51) (enc/mod.rs:941-947, f64, round half away from
zero), so an MB's QP changes only when qpf + o lies within |delta o| of a
half-integer.  For a frame QP qpf with a uniformly distributed fraction the
expected number of changed MBs is sum |delta o| over the MBs.

This test measures, on 1080p frames of the benchmark's generator, the
offsets' differences, that expectation per frame, and the changed MB QPs
over a grid of 2,001 frame QPs, and bounds the expectation at one MB per
200 frames (measured: 0.0023 per 1080p frame, i.e. one MB QP moved by one
step in about 440 frames)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def round_away(x):
    f = np.floor(x)
    return np.where(x - f >= 0.5, f + 1.0, f)


def test_offset_differences_rarely_change_an_mb_qp(gpu, oracle):
    import torch
    from pcme import TREE_BLOCK_DTYPE
    from pcme import lookahead as L
    n = 16
    say("qp flips: 1080p,", n, "frames, AQ + tree offsets")
    y, u, v = gpu.synth_clip(gpu.ClipSpec(1920, 1080, n, nrect=12, rect_lo=64, rect_hi=400, vmax=8.0,
                                          pan_max=1.0, seed=0x5EED0000))
    mbw, mbh = gpu.mb_dims(1920, 1080)
    dc = gpu.DeviceClip(y, mbw, mbh, u, v)
    dc.pad()
    aq = torch.stack([L.aq_offsets(dc, t) for t in range(n)])
    lr = L.lowres_frames(dc)
    # an IPPP queue: frame k predicts from frame k-1 (node k-1)
    triples = [(0, None, None)] + [(k, k - 1, None) for k in range(1, n)]
    blocks = L.analyse_batch(lr, triples)
    refs = [(-1, -1)] + [(k - 1, -1) for k in range(1, n)]
    nodes = [L.Node(blocks[k], tuple(None if r < 0 else r for r in refs[k])) for k in range(n)]
    prop = L.propagate(nodes, mbw, mbh)
    tree = torch.stack([L.offsets(blocks[k], prop[k]) for k in range(n)])
    torch.cuda.synchronize()
    g_aq = aq.cpu().numpy().reshape(n, -1)
    g_tree = tree.cpu().numpy().reshape(n, -1)
    host = [b for b in blocks.cpu().numpy().view(TREE_BLOCK_DTYPE)]
    want_prop = oracle.mbtree_propagate(host, refs, mbw, mbh)
    P, C = gpu.LUMA_PAD, gpu.CHROMA_PAD
    pl = dc.planes[0].cpu().numpy()
    cp = dc.cplanes.cpu().numpy()
    w_aq = np.stack([oracle.aq_offsets(pl[t, P:P + 16 * mbh, P:P + 16 * mbw], cp[0, t, C:C + 8 * mbh, C:C + 8 * mbw],
                                       cp[1, t, C:C + 8 * mbh, C:C + 8 * mbw], mbw, mbh, 0.6).reshape(-1)
                     for t in range(n)])
    w_tree = np.stack([oracle.mbtree_offsets(host[k], want_prop[k]).reshape(-1) for k in range(n)])
    g = (g_aq.astype(np.float32) + g_tree.astype(np.float32)).astype(np.float64)
    w = (w_aq.astype(np.float32) + w_tree.astype(np.float32)).astype(np.float64)
    d = np.abs(g - w)
    differ = int(np.count_nonzero(d))
    per_frame = float(d.sum()) / n
    qpfs = np.linspace(18.0, 42.0, 2001) + 0.000123
    flips = 0
    for q in qpfs:
        flips += int(np.count_nonzero(np.clip(round_away(q + g), 0, 51) != np.clip(round_away(q + w), 0, 51)))
    print(f"[qp flips] {g.size} MB offsets: {differ} differ (aq {int(np.count_nonzero(g_aq != w_aq))}, "
          f"tree {int(np.count_nonzero(g_tree != w_tree))}), max |delta| {d.max():.3g}; expected changed MB QPs "
          f"per frame {per_frame:.3g}; changed over {qpfs.size} frame QPs x {g.size} MBs: {flips}", flush=True)
    assert d.max() < 2e-5
    assert per_frame < 5e-3
