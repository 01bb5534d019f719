"""GPU parity of the lowres lookahead (pcme_lowres, pcme_complexity,
pcme_mbtree_analyse) against the CPU oracle, bit-exact: lowres frames,
complexity sums and per-block records (intra, inter, MV), and the
macroblock-tree blocks (intra, inter, list, both MVs) for list 0, list 1,
bi and intra-only jobs.

Each case prints what it runs before the GPU step so a failing run names
its case."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _device_clip(pcme, w, h, n, seed, **kw):
    y, _, _ = pcme.synth_clip(pcme.ClipSpec(w, h, n, seed=seed, **kw))
    mbw, mbh = pcme.mb_dims(w, h)
    dc = pcme.DeviceClip(y, mbw, mbh)
    dc.pad()
    return dc


def _oracle_lowres(oracle, dc):
    planes = dc.planes[0].cpu().numpy()
    return np.stack([oracle.lowres(planes[i], 16 * dc.mbw, 16 * dc.mbh) for i in range(dc.n)])


@pytest.mark.parametrize("w,h", [(1920, 1080), (352, 288), (100, 70), (16, 16)])
def test_lowres_matches_oracle(gpu, oracle, w, h):
    import torch
    from pcme import lookahead as L
    say("lowres", w, h)
    dc = _device_clip(gpu, w, h, 3, 0x5EED1100 + w)
    got = L.lowres_frames(dc)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(got.cpu().numpy(), _oracle_lowres(oracle, dc))


CLIPS = [
    # (w, h, frames, synth kwargs)
    (640, 360, 5, dict(nrect=6, rect_lo=32, rect_hi=160, vmax=6.0)),
    (352, 288, 4, dict(nrect=4, vmax=12.0, zoom=0.01)),
    (48, 32, 3, dict(nrect=2, rect_lo=8, rect_hi=24, vmax=3.0)),
]


@pytest.mark.parametrize("ci", range(len(CLIPS)))
def test_complexity_matches_oracle(gpu, oracle, ci):
    import torch
    from pcme import lookahead as L
    w, h, n, kw = CLIPS[ci]
    say("complexity", w, h, n)
    dc = _device_clip(gpu, w, h, n, 0x5EED1200 + ci, **kw)
    lr = L.lowres_frames(dc)
    pairs = [(0, None)] + [(t, t - 1) for t in range(1, n)] + [(n - 1, 0)]
    sums, blocks = L.complexity_batch(lr, pairs, with_blocks=True)
    torch.cuda.synchronize()
    sums = sums.cpu().numpy()
    lrn = lr.cpu().numpy()
    for k, (c, p) in enumerate(pairs):
        intra, inter, bi, bp, mv = oracle.complexity(lrn[c], None if p is None else lrn[p])
        assert (int(sums[k, 0]), int(sums[k, 1])) == (intra, inter), (k, c, p)
        g = blocks[k].reshape(-1)
        np.testing.assert_array_equal(g["intra"], bi)
        np.testing.assert_array_equal(g["inter"], bp)
        np.testing.assert_array_equal(g["mv"], mv.reshape(-1, 2))


def test_complexity_1080p_pair_and_single_call(gpu, oracle):
    import torch
    from pcme import lookahead as L
    say("complexity 1080p")
    dc = _device_clip(gpu, 1920, 1080, 2, 0x5EED1300, nrect=12, rect_lo=64, rect_hi=400, vmax=8.0)
    lr = L.lowres_frames(dc)
    c = L.complexity(L.LowRes(lr[1]), L.LowRes(lr[0]))
    torch.cuda.synchronize()
    lrn = lr.cpu().numpy()
    intra, inter, *_ = oracle.complexity(lrn[1], lrn[0])
    assert (c.intra, c.inter) == (intra, inter)
    c0 = L.complexity(lr[0])
    assert c0.intra == c0.inter == oracle.complexity(lrn[0], None)[0]


def test_complexity_edge_content(gpu, oracle):
    """Flat frames (every candidate ties), noise (the diamond wanders), a
    frame of one block, and more jobs than one launch holds."""
    import torch
    from pcme import lookahead as L
    say("complexity edges")
    rng = np.random.default_rng(5)
    frames = [np.zeros((32, 48), np.uint8), np.full((32, 48), 255, np.uint8),
              rng.integers(0, 256, (32, 48)).astype(np.uint8), rng.integers(0, 256, (32, 48)).astype(np.uint8)]
    fr = torch.as_tensor(np.stack(frames)).cuda()
    pairs = [(i, j) for i in range(4) for j in range(4)] + [(i, None) for i in range(4)]
    pairs = pairs * 4  # 80 jobs: two launches
    sums, blocks = L.complexity_batch(fr, pairs, with_blocks=True)
    torch.cuda.synchronize()
    sums = sums.cpu().numpy()
    for k, (c, p) in enumerate(pairs):
        intra, inter, bi, bp, mv = oracle.complexity(frames[c], None if p is None else frames[p])
        assert (int(sums[k, 0]), int(sums[k, 1])) == (intra, inter), (c, p)
        np.testing.assert_array_equal(blocks[k].reshape(-1)["mv"], mv.reshape(-1, 2))
    one = torch.as_tensor(rng.integers(0, 256, (2, 8, 8)).astype(np.uint8)).cuda()
    c = L.complexity(one[1], one[0])
    o = oracle.complexity(one[1].cpu().numpy(), one[0].cpu().numpy())
    assert (c.intra, c.inter) == o[:2]


def _check_tree(got, want, tag):
    for f in ("intra", "inter", "list"):
        bad = np.nonzero(got[f] != want[f])[0]
        assert bad.size == 0, (tag, f, bad[:5], got[bad[:5]], want[bad[:5]])
    np.testing.assert_array_equal(got["mv"], want["mv"], err_msg=str(tag))


@pytest.mark.parametrize("ci", range(len(CLIPS)))
def test_mbtree_analyse_matches_oracle(gpu, oracle, ci):
    import torch
    from pcme import TREE_BLOCK_DTYPE
    from pcme import lookahead as L
    w, h, n, kw = CLIPS[ci]
    say("mbtree analyse", w, h, n)
    dc = _device_clip(gpu, w, h, n, 0x5EED1400 + ci, **kw)
    lr = L.lowres_frames(dc)
    triples = [(0, None, None)] + [(t, t - 1, None) for t in range(1, n)] + \
              [(1, 0, n - 1), (n - 2, None, n - 1), (n - 1, n - 1, n - 1)]
    blocks = L.analyse_batch(lr, triples)
    torch.cuda.synchronize()
    got = blocks.cpu().numpy().view(TREE_BLOCK_DTYPE)
    lrn = lr.cpu().numpy()
    for k, (c, a, b) in enumerate(triples):
        want = oracle.tree_analyse(lrn[c], None if a is None else lrn[a], None if b is None else lrn[b])
        _check_tree(got[k], want, (k, c, a, b))


def test_mbtree_analyse_1080p(gpu, oracle):
    import torch
    from pcme import lookahead as L
    say("mbtree analyse 1080p")
    dc = _device_clip(gpu, 1920, 1080, 3, 0x5EED1500, nrect=12, rect_lo=64, rect_hi=400, vmax=8.0)
    lr = L.lowres_frames(dc)
    got = L.analyse(L.LowRes(lr[1]), L.LowRes(lr[0]), L.LowRes(lr[2]))
    torch.cuda.synchronize()
    lrn = lr.cpu().numpy()
    _check_tree(got, oracle.tree_analyse(lrn[1], lrn[0], lrn[2]), "1080p bi")


def test_mbtree_analyse_edge_content(gpu, oracle):
    import torch
    from pcme import TREE_BLOCK_DTYPE
    from pcme import lookahead as L
    say("mbtree edges")
    rng = np.random.default_rng(6)
    frames = [np.zeros((24, 40), np.uint8), np.full((24, 40), 255, np.uint8),
              rng.integers(0, 256, (24, 40)).astype(np.uint8), rng.integers(0, 256, (24, 40)).astype(np.uint8)]
    fr = torch.as_tensor(np.stack(frames)).cuda()
    triples = [(i, j, k) for i in range(4) for j in (None, 0, 2) for k in (None, 1, 3)] * 2  # 72 jobs: two launches
    got = L.analyse_batch(fr, triples).cpu().numpy().view(TREE_BLOCK_DTYPE)
    torch.cuda.synchronize()
    for n_, (c, a, b) in enumerate(triples):
        want = oracle.tree_analyse(frames[c], None if a is None else frames[a], None if b is None else frames[b])
        _check_tree(got[n_], want, (c, a, b))


def test_lookahead_replays_in_a_hip_graph(gpu, oracle):
    """pcme_lowres + pcme_complexity + pcme_mbtree_analyse captured once in a
    graph and replayed three times, with eager work and synchronisation
    between the replays; every replay gives the eager (and oracle) result,
    including the complexity sums the clear kernel resets."""
    import torch
    from pcme import TREE_BLOCK_DTYPE
    from pcme import lookahead as L
    say("lookahead graph replay")
    dc = _device_clip(gpu, 320, 240, 4, 0x5EED1B00, nrect=4, vmax=5.0)
    lr = L.lowres_frames(dc)
    pairs = [(0, None)] + [(t, t - 1) for t in range(1, 4)]
    triples = [(0, None, None)] + [(t, t - 1, None) for t in range(1, 4)]
    sums = torch.zeros((4, 2), dtype=torch.int64, device="cuda")
    blocks = torch.zeros((4, 20 * 15 * TREE_BLOCK_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    cj = L.complexity_jobs(lr, pairs, sums)
    tj = L.tree_jobs(lr, triples, blocks)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
            import pcme
            pcme._check(pcme.load().pcme_lowres(dc.plane_ptr(0, 0), dc.pitch, dc.stride, pcme.LUMA_PAD, 320, 240,
                                                lr.data_ptr(), 160 * 120, 4, s.cuda_stream))
            L.run_complexity(cj, s.cuda_stream)
            L.run_analyse(tj, s.cuda_stream)
    torch.cuda.synchronize()
    lrn = lr.cpu().numpy()
    want_sums = np.array([oracle.complexity(lrn[c], None if p is None else lrn[p])[:2] for c, p in pairs])
    want_blocks = [oracle.tree_analyse(lrn[c], None if a is None else lrn[a], None) for c, a, _ in triples]
    for rep in range(3):
        sums.fill_(-1)
        blocks.fill_(0xAB)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(sums.cpu().numpy().astype(np.uint64), want_sums.astype(np.uint64)), rep
        got = blocks.cpu().numpy().view(TREE_BLOCK_DTYPE)
        for k in range(4):
            _check_tree(got[k], want_blocks[k], (rep, k))


@pytest.mark.parametrize("w,h", [(1920, 1080), (352, 288)])
def test_aq_offsets_match_oracle(gpu, oracle, w, h):
    """pcme_aq_offsets against the oracle's aq_offsets on the coded-area
    planes, within 1e-5 (float32 results; the device's and the C library's
    log2 may differ in the last place)."""
    import torch
    from pcme import lookahead as L
    say("aq offsets", w, h)
    y, u, v = gpu.synth_clip(gpu.ClipSpec(w, h, 2, seed=0x5EED1D00 + w, nrect=6, vmax=4.0))
    mbw, mbh = gpu.mb_dims(w, h)
    dc = gpu.DeviceClip(y, mbw, mbh, u, v)
    dc.pad()
    got = L.aq_offsets(dc, 1)
    torch.cuda.synchronize()
    pl = dc.planes[0, 1].cpu().numpy()
    cp = dc.cplanes[:, 1].cpu().numpy()
    P, C = gpu.LUMA_PAD, gpu.CHROMA_PAD
    want = oracle.aq_offsets(pl[P:P + 16 * mbh, P:P + 16 * mbw], cp[0, C:C + 8 * mbh, C:C + 8 * mbw],
                             cp[1, C:C + 8 * mbh, C:C + 8 * mbw], mbw, mbh, 0.6)
    g = got.cpu().numpy()
    np.testing.assert_allclose(g, want, atol=1e-5, rtol=0)
    assert np.abs(g).max() > 0.1  # a textured clip: offsets are not all zero
