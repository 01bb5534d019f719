"""GPU parity of the macroblock-tree propagation (pcme_mbtree_propagate)
and offsets (pcme_mbtree_offsets) against the CPU oracle: the propagated
costs bit-exact (compared as float32 bit patterns) on random nodes --
large and negative MVs, bi blocks, references that feed nothing, two
lists on one reference, more nodes than one window launch takes -- and on
the encoder's lookahead queue built from real analyse() output at 640x360
and 1080p; the offsets within 1e-5 (the GPU's log2); a HIP-graph replay.

Each case prints what it runs before the GPU step so a failing run names
its case."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _random_nodes(rng, n, bw, bh, mvmax):
    from oracle import oracle as O
    nodes, refs = [], []
    for i in range(n):
        b = np.zeros(bw * bh, O.OTREE_DTYPE)
        b["intra"] = rng.integers(0, 3000, bw * bh)
        b["inter"] = rng.integers(0, 3500, bw * bh)
        b["list"] = rng.integers(0, 4, bw * bh)
        b["mv"] = rng.integers(-mvmax, mvmax + 1, (bw * bh, 2, 2))
        nodes.append(b)
        r0 = int(rng.integers(-1, max(i, 1)))
        r1 = r0 if rng.random() < 0.2 else int(rng.integers(-1, n))
        refs.append((r0, r1))
    return nodes, refs


def _run(nodes, refs, bw, bh, stream=None):
    import torch
    from pcme import lookahead as L
    dev = [torch.as_tensor(b.view(np.uint8).copy()).cuda() for b in nodes]
    ln = [L.Node(d, tuple(None if r < 0 else r for r in rf)) for d, rf in zip(dev, refs)]
    prop = L.propagate(ln, bw, bh, stream=stream)
    torch.cuda.synchronize()
    return prop.cpu().numpy(), dev


def _bits_equal(got, want, what):
    g = np.ascontiguousarray(got, np.float32).view(np.uint32)
    w = np.ascontiguousarray(want, np.float32).view(np.uint32)
    bad = np.argwhere(g != w)
    assert bad.size == 0, (what, len(bad), bad[:5].tolist(), got.ravel()[:0])


@pytest.mark.parametrize("seed,n,bw,bh,mvmax", [
    (0, 5, 7, 5, 20),
    (1, 40, 120, 68, 48),     # a 1080p lowres grid, a 40-frame queue
    (2, 12, 30, 17, 600),     # MVs across the whole picture
    (3, 3, 1, 1, 9),
    (4, 70, 8, 6, 16),        # two window launches
    (5, 4, 120, 68, 300),     # source rectangles beyond the LDS stage: the per-cell path
])
def test_propagate_random_nodes(gpu, oracle, seed, n, bw, bh, mvmax):
    say("propagate random", seed, n, bw, bh, mvmax)
    rng = np.random.default_rng(0x7EE0 + seed)
    nodes, refs = _random_nodes(rng, n, bw, bh, mvmax)
    got, _ = _run(nodes, refs, bw, bh)
    want = oracle.mbtree_propagate(nodes, refs, bw, bh)
    assert want.any()
    _bits_equal(got, want, (seed, n, bw, bh))


def _queue(n, bframes):
    """(frame, ref0 frame, ref1 frame) in the lookahead queue's order -- the
    IDR, then per mini-GOP its P anchor followed by its B frames
    (mod.rs:741-797) -- and the node index of each reference."""
    order, anchor, t = [(0, None, None)], 0, 1
    while t < n:
        p = min(t + bframes, n - 1)
        order.append((p, anchor, None))
        order += [(b, anchor, p) for b in range(t, p)]
        anchor, t = p, p + 1
    node = {f: i for i, (f, _, _) in enumerate(order)}
    return order, [tuple(-1 if r is None else node[r] for r in (a, b)) for _, a, b in order]


@pytest.mark.parametrize("w,h,n,bframes", [(640, 360, 13, 2), (1920, 1080, 9, 3), (352, 288, 6, 0)])
def test_propagate_encoder_queue(gpu, oracle, w, h, n, bframes):
    import torch
    from pcme import TREE_BLOCK_DTYPE
    from pcme import lookahead as L
    say("propagate queue", w, h, n, bframes)
    y, _, _ = gpu.synth_clip(gpu.ClipSpec(w, h, n, seed=0x7EE100 + w, nrect=6, vmax=7.0))
    mbw, mbh = gpu.mb_dims(w, h)
    dc = gpu.DeviceClip(y, mbw, mbh)
    dc.pad()
    lr = L.lowres_frames(dc)
    order, refs = _queue(n, bframes)
    blocks = L.analyse_batch(lr, order)
    nodes = [L.Node(blocks[k], tuple(None if r < 0 else r for r in refs[k])) for k in range(len(order))]
    prop = L.propagate(nodes, mbw, mbh)
    off = L.offsets(blocks[0], prop[0])
    torch.cuda.synchronize()
    host = [b for b in blocks.cpu().numpy().view(TREE_BLOCK_DTYPE)]
    want = oracle.mbtree_propagate(host, refs, mbw, mbh)
    assert want[0].any()
    _bits_equal(prop.cpu().numpy(), want, (w, h))
    np.testing.assert_allclose(off.cpu().numpy(), oracle.mbtree_offsets(host[0], want[0]), rtol=1e-6, atol=1e-5)


def test_offsets_match_oracle(gpu, oracle):
    import torch
    from pcme import lookahead as L
    say("mbtree offsets")
    rng = np.random.default_rng(0x7EE2)
    b = np.zeros(5000, oracle.OTREE_DTYPE)
    b["intra"] = rng.integers(0, 5000, b.size)
    b["intra"][:50] = 0
    prop = (rng.random(b.size) * rng.choice([0, 1, 100, 1e5], b.size)).astype(np.float32)
    got = L.offsets(torch.as_tensor(b.view(np.uint8).copy()).cuda(), torch.as_tensor(prop).cuda())
    torch.cuda.synchronize()
    np.testing.assert_allclose(got.cpu().numpy(), oracle.mbtree_offsets(b, prop), rtol=1e-6, atol=1e-5)


def test_propagate_edge_calls(gpu, oracle):
    import torch
    from pcme import lookahead as L
    say("propagate edges")
    assert L.propagate([], 4, 4).shape == (0, 16)
    rng = np.random.default_rng(0x7EE3)
    nodes, refs = _random_nodes(rng, 1, 6, 4, 10)
    got, _ = _run(nodes, refs, 6, 4)
    assert not got.any()
    nodes, refs = _random_nodes(rng, 6, 10, 7, 30)
    first, dev = _run(nodes, refs, 10, 7)
    out = torch.full((6, 70), 7.0, device="cuda")
    ln = [L.Node(d, tuple(None if r < 0 else r for r in rf)) for d, rf in zip(dev, refs)]
    L.propagate(ln, 10, 7, out=out)
    torch.cuda.synchronize()
    _bits_equal(out.cpu().numpy(), first, "second call into a dirty buffer")
    _bits_equal(first, oracle.mbtree_propagate(nodes, refs, 10, 7), "edge call")


def test_propagate_replays_in_a_hip_graph(gpu, oracle):
    """pcme_mbtree_propagate + pcme_mbtree_offsets captured once and replayed
    three times into buffers dirtied between the replays; every replay gives
    the oracle's result (the propagation clears its output with a kernel)."""
    import torch
    from pcme import lookahead as L
    say("propagate graph replay")
    rng = np.random.default_rng(0x7EE4)
    nodes, refs = _random_nodes(rng, 9, 40, 23, 24)
    _, dev = _run(nodes, refs, 40, 23)       # first call sizes the window buffer outside the capture
    ln = [L.Node(d, tuple(None if r < 0 else r for r in rf)) for d, rf in zip(dev, refs)]
    prop = torch.zeros((9, 40 * 23), device="cuda")
    off = torch.zeros(40 * 23, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    import pcme
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s, capture_error_mode="thread_local"):
            L.propagate(ln, 40, 23, out=prop, stream=s.cuda_stream)
            pcme._check(pcme.load().pcme_mbtree_offsets(dev[0].data_ptr(), prop.data_ptr(), 40 * 23,
                                                        off.data_ptr(), s.cuda_stream))
    torch.cuda.synchronize()
    want = oracle.mbtree_propagate(nodes, refs, 40, 23)
    want_off = oracle.mbtree_offsets(nodes[0], want[0])
    for rep in range(3):
        prop.fill_(3.0)
        off.fill_(5.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _bits_equal(prop.cpu().numpy(), want, ("replay", rep))
        np.testing.assert_allclose(off.cpu().numpy(), want_off, rtol=1e-6, atol=1e-5)


@pytest.mark.parametrize("seed,nq,bw,bh", [(0, 5, 20, 12), (1, 70, 6, 5)])
def test_propagate_independent_queues_in_one_call(gpu, oracle, seed, nq, bw, bh):
    """Several queues in one call share the launches (70 queues: a round of
    more than 64 jobs); each queue's costs equal the oracle's for that
    queue alone, bit for bit."""
    import torch
    from pcme import lookahead as L
    say("propagate queues", seed, nq, bw, bh)
    rng = np.random.default_rng(0x7EE5 + seed)
    queues, host = [], []
    for q in range(nq):
        n = int(rng.integers(1, 9))
        if q % 3 == 0:   # the encoder's P/B queue shape
            order, refs = _queue(n, 2)
            nodes, _ = _random_nodes(rng, len(order), bw, bh, 20)
        else:
            nodes, refs = _random_nodes(rng, n, bw, bh, 20)
        dev = [torch.as_tensor(b.view(np.uint8).copy()).cuda() for b in nodes]
        queues.append([L.Node(d, tuple(None if r < 0 else r for r in rf)) for d, rf in zip(dev, refs)])
        host.append((nodes, refs))
    props = L.propagate_queues(queues, bw, bh)
    torch.cuda.synchronize()
    for q, ((nodes, refs), pr) in enumerate(zip(host, props)):
        _bits_equal(pr.cpu().numpy(), oracle.mbtree_propagate(nodes, refs, bw, bh), ("queue", q))
