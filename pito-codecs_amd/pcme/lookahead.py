"""The encoder's lowres lookahead on the GPU, behind libpcme's C ABI.

Mirrors the reference's interface (pito-codecs src/h264/enc/):

* ``LowRes.new(frame)``       lookahead.rs:10-33, the 2x2-averaged luma
* ``complexity(cur, prev)``   lookahead.rs:52-108, the scenecut input
                              (``Complexity {intra, inter}``)
* ``analyse(cur, l0, l1)``    mbtree.rs:113-159, the per-block costs of the
                              macroblock tree (``Costs.blocks``)
* ``aq_offsets(frame)``       lookahead.rs:116-139, the per-MB AQ offsets
* ``propagate(nodes, bw, bh)`` mbtree.rs:167-223, the propagated cost of
                              every lookahead node
* ``offsets(costs, prop)``    mbtree.rs:225-231, the tree's QP offsets

plus batched forms that take many frames in one launch.  Results are
bit-identical to the reference, the float propagation included (the AQ
and tree offsets, which end in the GPU's log2, within 1e-5).  There is no CPU path: without libpcme or a
gfx950 device every call raises ``PcmeError``.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import (CPLX_BLOCK_DTYPE, TREE_BLOCK_DTYPE, CplxJob, DeviceClip, LUMA_PAD, PcmeError, PCME_EINVAL,
               TreeJob, TreeNode, _check, _stream_ptr, _torch, load)


@dataclass
class Complexity:
    """lookahead.rs:35-39"""
    intra: int
    inter: int


class LowRes:
    """A lowres frame resident in HBM (lookahead.rs:4-8): ``data`` is a
    [h, w] uint8 tensor on the GPU."""

    def __init__(self, data):
        self.data = data
        self.h, self.w = (int(s) for s in data.shape)

    @staticmethod
    def new(clip: DeviceClip, i: int, stream=None) -> "LowRes":
        """LowRes::new of frame i of a prepared clip."""
        return LowRes(lowres_frames(clip, i, 1, stream)[0])


def lowres_frames(clip: DeviceClip, first: int = 0, count: int | None = None, stream=None):
    """LowRes::new of frames [first, first+count) of a clip whose luma planes
    are padded: a [count, 8*mbh, 8*mbw] uint8 tensor."""
    torch = clip.torch
    count = clip.n - first if count is None else count
    w, h = 16 * clip.mbw, 16 * clip.mbh
    out = torch.empty((max(count, 0), h // 2, w // 2), dtype=torch.uint8, device=clip.raw.device)
    if count > 0:
        _check(load().pcme_lowres(clip.plane_ptr(0, first), clip.pitch, clip.stride, LUMA_PAD, w, h,
                                  out.data_ptr(), (h // 2) * (w // 2), count, _stream_ptr(stream)))
    return out


def _ptr(x) -> int | None:
    if x is None:
        return None
    return (x.data if isinstance(x, LowRes) else x).data_ptr()


def _dims(x) -> tuple[int, int]:
    d = x.data if isinstance(x, LowRes) else x
    return int(d.shape[-1]), int(d.shape[-2])


def complexity_jobs(frames, pairs, sums, blocks=None) -> list[CplxJob]:
    """Jobs of pcme_complexity: pairs = [(cur, prev or None)] indices into
    frames [n, h, w]; sums [len(pairs), 2] int64; blocks [len(pairs),
    nblocks * 12] uint8 or None."""
    n, h, w = (int(s) for s in frames.shape)
    jobs = []
    for k, (c, p) in enumerate(pairs):
        if not (0 <= c < n) or (p is not None and not (0 <= p < n)):
            raise PcmeError(PCME_EINVAL, f"complexity job {k}: frame index out of range")
        jobs.append(CplxJob(frames[c].data_ptr(), None if p is None else frames[p].data_ptr(),
                            None if blocks is None else blocks[k].data_ptr(), sums[k].data_ptr(), w, h))
    return jobs


def run_complexity(jobs: list[CplxJob], stream=None) -> None:
    arr = (CplxJob * len(jobs))(*jobs)
    _check(load().pcme_complexity(arr, len(jobs), _stream_ptr(stream)))


def complexity_batch(frames, pairs, with_blocks: bool = False, stream=None):
    """complexity() of every (cur, prev) index pair over frames [n, h, w]
    (device).  Returns (sums [k, 2] int64 tensor, blocks [k, bh, bw]
    CPLX_BLOCK_DTYPE array or None); synchronises when blocks are asked for."""
    torch = _torch()
    _, h, w = (int(s) for s in frames.shape)
    nb = (w // 8) * (h // 8)
    sums = torch.empty((len(pairs), 2), dtype=torch.int64, device=frames.device)
    blocks = torch.empty((len(pairs), nb * CPLX_BLOCK_DTYPE.itemsize), dtype=torch.uint8,
                         device=frames.device) if with_blocks else None
    run_complexity(complexity_jobs(frames, pairs, sums, blocks), stream)
    if blocks is None:
        return sums, None
    return sums, blocks.cpu().numpy().view(CPLX_BLOCK_DTYPE).reshape(len(pairs), h // 8, w // 8)


def complexity(cur, prev=None, stream=None) -> Complexity:
    """lookahead::complexity of one lowres frame against the previous one."""
    torch = _torch()
    w, h = _dims(cur)
    if prev is not None and _dims(prev) != (w, h):
        raise PcmeError(PCME_EINVAL, "complexity: cur and prev differ in size")
    sums = torch.empty(2, dtype=torch.int64, device=(cur.data if isinstance(cur, LowRes) else cur).device)
    run_complexity([CplxJob(_ptr(cur), _ptr(prev), None, sums.data_ptr(), w, h)], stream)
    a, b = sums.cpu().tolist()
    return Complexity(int(a), int(b))


def tree_jobs(frames, triples, blocks) -> list[TreeJob]:
    """Jobs of pcme_mbtree_analyse: triples = [(cur, l0 or None, l1 or None)]
    indices into frames [n, h, w]; blocks [len(triples), nblocks * 20]."""
    n, h, w = (int(s) for s in frames.shape)
    jobs = []
    for k, (c, a, b) in enumerate(triples):
        for i in (c, a, b):
            if i is not None and not (0 <= i < n):
                raise PcmeError(PCME_EINVAL, f"analyse job {k}: frame index out of range")
        jobs.append(TreeJob(frames[c].data_ptr(), None if a is None else frames[a].data_ptr(),
                            None if b is None else frames[b].data_ptr(), blocks[k].data_ptr(), w, h))
    return jobs


def run_analyse(jobs: list[TreeJob], stream=None) -> None:
    arr = (TreeJob * len(jobs))(*jobs)
    _check(load().pcme_mbtree_analyse(arr, len(jobs), _stream_ptr(stream)))


def analyse_batch(frames, triples, stream=None):
    """mbtree::analyse of every (cur, l0, l1) over frames [n, h, w] (device):
    the blocks as a [k, nblocks * 20] uint8 device tensor (view it with
    TREE_BLOCK_DTYPE)."""
    torch = _torch()
    _, h, w = (int(s) for s in frames.shape)
    nb = (w // 8) * (h // 8)
    blocks = torch.empty((len(triples), nb * TREE_BLOCK_DTYPE.itemsize), dtype=torch.uint8, device=frames.device)
    run_analyse(tree_jobs(frames, triples, blocks), stream)
    return blocks


def analyse(cur, l0=None, l1=None, stream=None) -> np.ndarray:
    """mbtree::analyse of one lowres frame: its Costs.blocks as a
    [bh*bw] TREE_BLOCK_DTYPE array."""
    torch = _torch()
    w, h = _dims(cur)
    for r in (l0, l1):
        if r is not None and _dims(r) != (w, h):
            raise PcmeError(PCME_EINVAL, "analyse: reference size differs from cur")
    nb = (w // 8) * (h // 8)
    dev = (cur.data if isinstance(cur, LowRes) else cur).device
    blocks = torch.empty(nb * TREE_BLOCK_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    run_analyse([TreeJob(_ptr(cur), _ptr(l0), _ptr(l1), blocks.data_ptr(), w, h)], stream)
    return blocks.cpu().numpy().view(TREE_BLOCK_DTYPE)


@dataclass
class Node:
    """mbtree.rs:162-165: ``costs`` is a node's analyse() output resident on
    the GPU (a uint8 tensor of bw*bh TREE_BLOCK_DTYPE records, e.g. a row of
    analyse_batch), ``refs`` the node index each list refers to (None: no
    reference; one at or after the node feeds nothing)."""
    costs: object
    refs: tuple = (None, None)


def propagate(nodes: list[Node], bw: int, bh: int, out=None, stream=None):
    """mbtree::propagate: the propagated cost of every node as an [n, bh*bw]
    float32 tensor on the GPU (written into `out` when given)."""
    torch = _torch()
    n = len(nodes)
    nb = bw * bh
    for k, nd in enumerate(nodes):
        if int(nd.costs.numel()) < nb * TREE_BLOCK_DTYPE.itemsize:
            raise PcmeError(PCME_EINVAL, f"propagate: node {k} holds fewer than {nb} blocks")
    if out is None:
        dev = nodes[0].costs.device if n else "cuda"
        out = torch.empty((n, nb), dtype=torch.float32, device=dev)
    arr = (TreeNode * max(n, 1))()
    for k, nd in enumerate(nodes):
        arr[k].blocks = nd.costs.data_ptr()
        for l in (0, 1):
            r = nd.refs[l]
            arr[k].ref[l] = -1 if r is None else int(r)
    _check(load().pcme_mbtree_propagate(arr, n, bw, bh, out.data_ptr() if n else None, _stream_ptr(stream)))
    return out


def tree_nodes(queues: list[list[Node]]):
    """The pcme_tree_node table of several independent queues laid out one
    after the other (refs local to each queue; a reference at or after its
    node, or past the queue's end, feeds nothing): (ctypes array, n)."""
    n = sum(len(q) for q in queues)
    arr = (TreeNode * max(n, 1))()
    base = 0
    for q in queues:
        for k, nd in enumerate(q):
            arr[base + k].blocks = nd.costs.data_ptr()
            for l in (0, 1):
                r = nd.refs[l]
                arr[base + k].ref[l] = -1 if (r is None or r < 0 or r >= k) else r + base
        base += len(q)
    return arr, n


def run_propagate(arr, n: int, bw: int, bh: int, prop, stream=None) -> None:
    """pcme_mbtree_propagate over a tree_nodes table into prop (device,
    [n, bh*bw] float32)."""
    _check(load().pcme_mbtree_propagate(arr, n, bw, bh, prop.data_ptr() if n else None, _stream_ptr(stream)))


def propagate_queues(queues: list[list[Node]], bw: int, bh: int, stream=None):
    """propagate of several independent lookahead queues of one block grid
    in one call: the library runs every queue's next node in the same
    launch.  queues = [[Node, ...], ...] with refs local to their queue (a
    reference at or after its node, or past the queue's end, feeds
    nothing).  Returns one [len(q), bh*bw] float32 tensor per queue (views
    of one buffer)."""
    torch = _torch()
    for qi, q in enumerate(queues):
        for k, nd in enumerate(q):
            if int(nd.costs.numel()) < bw * bh * TREE_BLOCK_DTYPE.itemsize:
                raise PcmeError(PCME_EINVAL, f"propagate_queues: queue {qi} node {k} holds fewer than {bw * bh} blocks")
    arr, n = tree_nodes(queues)
    dev = next((nd.costs.device for q in queues for nd in q), "cuda")
    prop = torch.empty((n, bw * bh), dtype=torch.float32, device=dev)
    run_propagate(arr, n, bw, bh, prop, stream)
    out, base = [], 0
    for q in queues:
        out.append(prop[base:base + len(q)])
        base += len(q)
    return out


def offsets(costs, prop, stream=None):
    """mbtree::offsets of one node: its costs (device TREE_BLOCK_DTYPE
    records) and propagated cost (device float32, one per block) -> a
    float32 tensor of the node's per-block QP offsets."""
    torch = _torch()
    n = int(prop.numel())
    if int(costs.numel()) < n * TREE_BLOCK_DTYPE.itemsize:
        raise PcmeError(PCME_EINVAL, "offsets: fewer blocks than propagated costs")
    prop = prop.contiguous()
    out = torch.empty(n, dtype=torch.float32, device=prop.device)
    _check(load().pcme_mbtree_offsets(costs.data_ptr(), prop.data_ptr(), n, out.data_ptr(), _stream_ptr(stream)))
    return out


AQ_STRENGTH = 0.6  # mod.rs:109


def aq_offsets(clip: DeviceClip, t: int, strength: float = AQ_STRENGTH, stream=None):
    """lookahead::aq_offsets of frame t of a padded clip with chroma: a
    [mbh, mbw] float32 tensor on the GPU."""
    from . import CHROMA_PAD
    if not clip.has_chroma:
        raise PcmeError(PCME_EINVAL, "aq_offsets needs the clip's chroma planes")
    torch = clip.torch
    out = torch.empty((clip.mbh, clip.mbw), dtype=torch.float32, device=clip.raw.device)
    y = clip.plane_ptr(0, t) + LUMA_PAD * clip.stride + LUMA_PAD
    u = clip.cplanes[0, t].data_ptr() + CHROMA_PAD * clip.cstride + CHROMA_PAD
    v = clip.cplanes[1, t].data_ptr() + CHROMA_PAD * clip.cstride + CHROMA_PAD
    _check(load().pcme_aq_offsets(y, clip.stride, u, v, clip.cstride, clip.mbw, clip.mbh, float(strength),
                                  out.data_ptr(), _stream_ptr(stream)))
    return out


__all__ = ["Complexity", "LowRes", "aq_offsets", "AQ_STRENGTH", "lowres_frames", "complexity", "complexity_batch", "complexity_jobs",
           "run_complexity", "analyse", "analyse_batch", "tree_jobs", "run_analyse", "Node", "propagate", "propagate_queues",
           "tree_nodes", "run_propagate", "offsets"]
