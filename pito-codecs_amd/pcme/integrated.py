"""Encoder-integrated use of the ME-only field: per-MB verification against
the state the encoder committed, and exact GPU re-analysis where it differs.

The ME-only field (pcme_search_frames) analyses every MB as if its causal
neighbours had all been committed as P-inter, reference 0, with the field's
own MVs.  An MB's analysis (analyse_p's skip MV, 16x16 search and partition
choice, pito-codecs src/h264/enc/slice.rs:1884-1945) reads its neighbours
only through the ten 4x4 blocks nb_part (slice.rs:392-432) looks at: left MB
blocks 3/7/11/15, top MB 12-15, top-right 12, top-left 15.  Where the
encoder committed something else (early skip, intra, RD, another reference),
``verify`` flags the MBs whose ten blocks differ; only those can differ from
the encoder's own analysis.  ``reanalyse`` runs them again on the committed
state with pcme_search_part_batch: pskip_mv (slice.rs:484-494) on the host,
the 16x16 search_part seeded with it, the 16x8 / 8x16 / 8x8 search_part
seeded with mv16 when partitions are on and cost16 > 1024, and the strict-<
mode choice -- the same records the field holds.
"""
from __future__ import annotations

import numpy as np

from . import LUMA_PAD, MB_DTYPE, PART_OUT_DTYPE, DeviceClip, PartJob, Mv, NbBlock, lam as lam_of, search_part_batch
from .bsearch import _NB, vlim_rows

INTRA_KINDS = (1, 2, 3)  # I4x4, I8x8, I16x16 (mbinfo.rs:18-45 order: None, I4x4, I8x8, I16x16, PSkip, PInter)


def nb_blocks(kind, ref_idx, mv) -> np.ndarray:
    """[mbh, mbw, 10, 3] (ref, mv.x, mv.y) of the ten neighbour blocks per
    MB, list 0, as nb_part reads the committed MbInfo: -2 outside the
    picture, (-1, 0) for an intra MB, else the 8x8's ref and, for ref >= 0,
    the block's MV.  kind [mbh, mbw], ref_idx [mbh, mbw, 4], mv [mbh, mbw,
    16, 2]."""
    kind = np.asarray(kind)
    ref_idx = np.asarray(ref_idx, np.int64)
    mv = np.asarray(mv, np.int64)
    mbh, mbw = kind.shape
    my, mx = np.mgrid[0:mbh, 0:mbw]
    out = np.zeros((mbh, mbw, 10, 3), np.int64)
    intra = np.isin(kind, INTRA_KINDS)
    for i, ((dx, dy), blk) in enumerate(_NB):
        nx, ny = mx + dx, my + dy
        avail = (nx >= 0) & (nx < mbw) & (ny >= 0)
        cx, cy = np.clip(nx, 0, mbw - 1), np.clip(ny, 0, mbh - 1)
        q8 = ((blk >> 2) >> 1) * 2 + ((blk & 3) >> 1)
        r = np.where(intra[cy, cx], -1, ref_idx[cy, cx, q8])
        r = np.where(avail, r, -2)
        out[:, :, i, 0] = r
        for k in (0, 1):
            out[:, :, i, 1 + k] = np.where(r >= 0, mv[cy, cx, blk, k], 0)
    return out


def field_state(field: np.ndarray):
    """(kind, ref_idx, mv) the ME-only field assumes: every MB P-inter,
    reference 0, its committed MVs."""
    mbh, mbw = field.shape
    return (np.full((mbh, mbw), 5), np.zeros((mbh, mbw, 4), np.int64), field["mv"].astype(np.int64))


def verify(field: np.ndarray, kind, ref_idx, mv) -> np.ndarray:
    """[mbh, mbw] bool: True where the field's record is the encoder's
    analysis for the committed state (the MB's ten neighbour blocks read
    the same)."""
    return np.all(nb_blocks(kind, ref_idx, mv) == nb_blocks(*field_state(field)), axis=(2, 3))


def _median(a, b, c):
    return max(min(a, b), min(max(a, b), c))


def _mvp16(nb) -> tuple[int, int]:
    """mvp (slice.rs:434-482) of the 16x16 partition, rf 0, Normal shape."""
    a, b, c = nb[0], nb[4], nb[8] if nb[8][0] >= -1 else nb[9]
    av = [x[0] >= -1 for x in (a, b, c)]
    a, b, c = [tuple(int(v) for v in x) for x in (a, b, c)]
    if not av[1] and not av[2] and av[0]:
        b = c = a
        av = [True, True, True]
    a, b, c = [(x if ok else (-1, 0, 0)) for x, ok in zip((a, b, c), av)]
    hits = [x for x in (a, b, c) if x[0] == 0]
    if len(hits) == 1:
        return hits[0][1], hits[0][2]
    return _median(a[1], b[1], c[1]), _median(a[2], b[2], c[2])


def pskip_mv(nb) -> tuple[int, int]:
    """pskip_mv (slice.rs:484-494) from the MB's ten neighbour blocks."""
    a, b = nb[0], nb[4]
    if a[0] < -1 or b[0] < -1:
        return 0, 0
    if (a[0] == 0 and a[1] == 0 and a[2] == 0) or (b[0] == 0 and b[1] == 0 and b[2] == 0):
        return 0, 0
    return _mvp16(nb)


def _job(clip, t, ref_t, mx, my, part, seed, nb, lam, vlim, me_range, subpel_rounds):
    j = PartJob()
    j.src = clip.plane_ptr(0, t) + (LUMA_PAD + 16 * my) * clip.stride + LUMA_PAD + 16 * mx
    j.sstride = clip.stride
    j.part = part
    j.ref, j.ref_h, j.ref_v, j.ref_c = (clip.plane_ptr(k, ref_t) for k in range(4))
    j.mbw, j.mbh, j.mx, j.my = clip.mbw, clip.mbh, mx, my
    j.seed = Mv(int(seed[0]), int(seed[1]))
    j.lam, j.me_range, j.subpel_rounds, j.vlim = lam, me_range, subpel_rounds, int(vlim)
    blocks = [NbBlock(Mv(int(b[1]), int(b[2])), int(b[0])) for b in nb]
    for i in range(4):
        j.left[i] = blocks[i]
        j.top[i] = blocks[4 + i]
    j.top_right, j.top_left = blocks[8], blocks[9]
    return j


def _run(jobs, torch, device):
    out = torch.zeros(len(jobs) * PART_OUT_DTYPE.itemsize, dtype=torch.uint8, device=device)
    search_part_batch(jobs, out.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy().view(PART_OUT_DTYPE)


_RECT = {0: [(0, 0, 16, 16)], 1: [(0, 0, 16, 8), (0, 8, 16, 8)], 2: [(0, 0, 8, 16), (8, 0, 8, 16)],
         3: [(0, 0, 8, 8), (8, 0, 8, 8), (0, 8, 8, 8), (8, 8, 8, 8)]}


def reanalyse(clip: DeviceClip, t: int, ref_t: int, mbs, kind, ref_idx, mv, qp: int, me_range: int = 32,
              subpel_rounds: int = 1, partitions: bool = True) -> np.ndarray:
    """MB records (MB_DTYPE) of the MBs `mbs` [(mx, my), ...] of frame t
    against frame ref_t, analysed on the committed state (kind, ref_idx, mv)
    of the picture: two rounds of pcme_search_part_batch (16x16, then the
    partition modes), the strict-< choice on the host."""
    torch = clip.torch
    dev = clip.raw.device
    nbs = nb_blocks(kind, ref_idx, mv)
    lam = lam_of(qp)
    vl = vlim_rows(clip.mbh)
    out = np.zeros(len(mbs), MB_DTYPE)
    if not mbs:
        return out
    skips = [pskip_mv(nbs[my, mx]) for mx, my in mbs]
    r16 = _run([_job(clip, t, ref_t, mx, my, 0, s, nbs[my, mx], lam, vl[my], me_range, subpel_rounds)
                for (mx, my), s in zip(mbs, skips)], torch, dev)
    split = [k for k in range(len(mbs)) if partitions and int(r16[k]["total"]) > 1024]
    rp = {}
    if split:
        jobs = [_job(clip, t, ref_t, mbs[k][0], mbs[k][1], part, tuple(r16[k]["mvs"][0]), nbs[mbs[k][1], mbs[k][0]],
                     lam, vl[mbs[k][1]], me_range, subpel_rounds) for k in split for part in (1, 2, 3)]
        res = _run(jobs, torch, dev)
        for i, k in enumerate(split):
            rp[k] = res[3 * i:3 * i + 3]
    pen = {1: (lam * 3) >> 8, 2: (lam * 3) >> 8, 3: (lam * 10) >> 8}
    for k in range(len(mbs)):
        o = out[k]
        o["cost"][:] = 0xFFFFFFFF
        o["cost"][0] = r16[k]["total"]
        best, part, rec = int(r16[k]["total"]) + (lam >> 8), 0, r16[k]
        for m, r in zip((1, 2, 3), rp.get(k, [])):
            o["cost"][m] = r["total"]
            if int(r["total"]) + pen[m] < best:
                best, part, rec = int(r["total"]) + pen[m], m, r
        o["mvs"], o["mvps"] = rec["mvs"], rec["mvps"]
        for pi, (x, y, w, h) in enumerate(_RECT[part]):
            for by in range(y // 4, (y + h) // 4):
                for bx in range(x // 4, (x + w) // 4):
                    o["mv"][by * 4 + bx] = rec["mvs"][pi]
        o["best_cost"] = best
        o["skip_mv"] = skips[k]
        o["mv16"] = r16[k]["mvs"][0]
        o["part"] = part
    return out


__all__ = ["nb_blocks", "field_state", "verify", "pskip_mv", "reanalyse"]
