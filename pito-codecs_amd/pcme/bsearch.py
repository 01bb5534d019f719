"""B-frame motion search on the GPU, behind libpcme's C ABI.

``b_jobs`` builds the pcme_b_job table of a whole B picture from the
encoder's committed motion state (the two lists' MV fields, reference
indices, intra flags) and the co-located MBs of the list-1 picture, the way
``nb_part`` (pito-codecs src/h264/enc/slice.rs:392-432) and
``direct_spatial`` (slice.rs:2030-2084) read them; ``b_analyse`` runs
pcme_b_analyse_batch over it.  One record per MB (``B_OUT_DTYPE``) holds
what analyse_b's search produces: the direct MVs and cost, the list-0 /
list-1 16x16 searches with the direction costs, and b_partition's 16x8 and
8x16 searches.  There is no CPU path: without libpcme or a gfx950 device
every call raises ``PcmeError``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import B_OUT_DTYPE, BJob, DeviceClip, LUMA_PAD, _check, _stream_ptr, lam, load

NB_DTYPE = np.dtype([("mv", "<i2", (2,)), ("ref", "i1"), ("reserved", "u1", (3,))])
B_JOB_DTYPE = np.dtype([
    ("src", "<u8"), ("sstride", "<i4"), ("reserved0", "<i4"), ("ref0", "<u8", (4,)), ("ref1", "<u8", (4,)),
    ("mbw", "<i4"), ("mbh", "<i4"), ("mx", "<i4"), ("my", "<i4"), ("lam", "<u4"), ("me_range", "<i4"),
    ("subpel_rounds", "<i4"), ("reserved1", "<i4"), ("vlim", "<i8"), ("nb", NB_DTYPE, (2, 10)),
    ("col_mv", "<i2", (2, 4, 2)), ("col_ref", "i1", (2, 4)), ("col_intra", "<i4"), ("reserved2", "<i4"),
])
assert B_JOB_DTYPE.itemsize == ctypes.sizeof(BJob)

# (neighbour MB offset, 4x4 block) of the ten neighbour blocks: left[4],
# top[4], top_right, top_left.
_NB = [((-1, 0), 3), ((-1, 0), 7), ((-1, 0), 11), ((-1, 0), 15), ((0, -1), 12), ((0, -1), 13), ((0, -1), 14),
       ((0, -1), 15), ((1, -1), 12), ((-1, -1), 15)]
_CORNER = (0, 3, 12, 15)


def vlim_rows(mbh: int) -> np.ndarray:
    """SliceEnc::vlim of every MB row (frame.rs:152-156, picture.rs:261-265)."""
    need = 16 * (np.arange(mbh) + 1) + 96
    return np.where(need >= 16 * mbh, 16 * mbh + 44, need).astype(np.int64)


def b_jobs(cur: DeviceClip, t: int, ref0: tuple[DeviceClip, int], ref1: tuple[DeviceClip, int], mv, ref_idx,
           intra, col_mv, col_ref, col_intra, qp: int, me_range: int = 32, subpel_rounds: int = 1) -> np.ndarray:
    """The pcme_b_job table of B picture t of `cur`.

    mv [2, mbh, mbw, 16, 2] and ref_idx [2, mbh, mbw, 4] are the committed
    MbInfo.mv / ref_idx of the picture's MBs per list (read for the causal
    neighbours only), intra [mbh, mbw] their intra flags; col_mv [mbh, mbw,
    2, 16, 2], col_ref [mbh, mbw, 2, 4] and col_intra [mbh, mbw] describe the
    list-1 picture's MbInfo at the same positions."""
    mbw, mbh = cur.mbw, cur.mbh
    n = mbw * mbh
    mv = np.asarray(mv, np.int16)
    ref_idx = np.asarray(ref_idx, np.int8)
    intra = np.asarray(intra, bool)
    jobs = np.zeros(n, B_JOB_DTYPE)
    my, mx = np.divmod(np.arange(n), mbw)
    jobs["src"] = cur.plane_ptr(0, t) + (LUMA_PAD + 16 * my) * cur.stride + LUMA_PAD + 16 * mx
    jobs["sstride"] = cur.stride
    for k in range(4):
        jobs["ref0"][:, k] = ref0[0].plane_ptr(k, ref0[1])
        jobs["ref1"][:, k] = ref1[0].plane_ptr(k, ref1[1])
    jobs["mbw"], jobs["mbh"], jobs["mx"], jobs["my"] = mbw, mbh, mx, my
    jobs["lam"] = lam(qp)
    jobs["me_range"], jobs["subpel_rounds"] = me_range, subpel_rounds
    jobs["vlim"] = vlim_rows(mbh)[my]
    for i, ((dx, dy), blk) in enumerate(_NB):
        nx, ny = mx + dx, my + dy
        avail = (nx >= 0) & (nx < mbw) & (ny >= 0)
        cx, cy = np.clip(nx, 0, mbw - 1), np.clip(ny, 0, mbh - 1)
        q8 = ((blk >> 2) >> 1) * 2 + ((blk & 3) >> 1)
        for l in range(2):
            r = np.where(intra[cy, cx], -1, ref_idx[l, cy, cx, q8])
            r = np.where(avail, r, -2)
            m = np.where((r >= 0)[:, None], mv[l, cy, cx, blk], 0)
            jobs["nb"]["ref"][:, l, i] = r
            jobs["nb"]["mv"][:, l, i] = m
    jobs["col_mv"] = np.asarray(col_mv, np.int16).reshape(n, 2, 16, 2)[:, :, list(_CORNER)]
    jobs["col_ref"] = np.asarray(col_ref, np.int8).reshape(n, 2, 4)
    jobs["col_intra"] = np.asarray(col_intra, bool).reshape(n).astype(np.int32)
    return jobs


def b_analyse(jobs: np.ndarray, out_ptr: int, stream=None) -> None:
    """pcme_b_analyse_batch over a B_JOB_DTYPE table; out_ptr (device)
    receives len(jobs) B_OUT_DTYPE records."""
    arr = np.ascontiguousarray(jobs, B_JOB_DTYPE)
    _check(load().pcme_b_analyse_batch(arr.ctypes.data_as(ctypes.POINTER(BJob)), len(arr), out_ptr,
                                       _stream_ptr(stream)))


def b_field_job(cur: DeviceClip, t: int, ref0: tuple[DeviceClip, int], ref1: tuple[DeviceClip, int],
                col_ptr: int, qp: int, me_range: int = 32, subpel_rounds: int = 1, partitions: bool = True):
    """The pcme_b_field_job of B picture t of `cur` between list-0 picture
    ref0 and list-1 picture ref1 ((clip, index) pairs, prepared); col_ptr:
    mbw*mbh B_COL_DTYPE records on the device."""
    from . import BFieldJob
    (c0, t0), (c1, t1) = ref0, ref1
    return BFieldJob(cur.plane_ptr(0, t), (ctypes.c_void_p * 4)(*[c0.plane_ptr(k, t0) for k in range(4)]),
                     (ctypes.c_void_p * 4)(*[c1.plane_ptr(k, t1) for k in range(4)]), col_ptr, cur.mbw, cur.mbh,
                     int(qp), int(me_range), int(subpel_rounds), 1 if partitions else 0)


def b_field_launch(cur: DeviceClip, t: int, ref0: tuple[DeviceClip, int], ref1: tuple[DeviceClip, int],
                   col_ptr: int, out_ptr: int, qp: int, me_range: int = 32, subpel_rounds: int = 1,
                   partitions: bool = True, stream=None) -> None:
    """pcme_b_field on device buffers: col_ptr holds mbw*mbh B_COL_DTYPE
    records, out_ptr receives mbw*mbh B_MB_DTYPE records (stream order, no
    synchronisation; b_field_check reports a stall)."""
    job = b_field_job(cur, t, ref0, ref1, col_ptr, qp, me_range, subpel_rounds, partitions)
    _check(load().pcme_b_field(ctypes.byref(job), out_ptr, _stream_ptr(stream)))


def b_fields_launch(jobs, out_ptrs, stream=None) -> None:
    """pcme_b_fields: the B field jobs (b_field_job) in one launch, picture
    i's records to out_ptrs[i] (stream order, no synchronisation)."""
    from . import BFieldJob
    assert len(jobs) == len(out_ptrs)
    arr = (BFieldJob * len(jobs))(*jobs)
    outs = (ctypes.c_void_p * len(jobs))(*out_ptrs)
    _check(load().pcme_b_fields(arr, len(jobs), outs, _stream_ptr(stream)))


def b_field_check(stream=None) -> None:
    """pcme_b_field_check: waits for the stream; PcmeError(PCME_ESTALL) if a
    B field launch on the device stalled since the previous check."""
    _check(load().pcme_b_field_check(_stream_ptr(stream)))


def b_field(cur: DeviceClip, t: int, ref0: tuple[DeviceClip, int], ref1: tuple[DeviceClip, int], col, qp: int,
            me_range: int = 32, subpel_rounds: int = 1, partitions: bool = True, stream=None):
    """pcme_b_field: the ME-only field of B picture t of `cur` (all of its
    MBs, each searched over the MBs committed before it) between list-0
    picture ref0 and list-1 picture ref1 ((clip, index) pairs, prepared).
    col: [mbh, mbw] B_COL_DTYPE records of the list-1 picture's MBs (host).
    Returns the [mbh, mbw] B_MB_DTYPE records."""
    from . import B_COL_DTYPE, B_MB_DTYPE
    torch = cur.torch
    dev = cur.raw.device
    c = np.ascontiguousarray(np.asarray(col).reshape(-1).view(B_COL_DTYPE))
    assert c.size == cur.mbw * cur.mbh
    dcol = torch.as_tensor(c.view(np.uint8)).to(dev)
    out = torch.empty(cur.mbw * cur.mbh * B_MB_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    b_field_launch(cur, t, ref0, ref1, dcol.data_ptr(), out.data_ptr(), qp, me_range, subpel_rounds, partitions,
                   stream)
    b_field_check(stream)
    torch.cuda.synchronize()
    return out.cpu().numpy().view(B_MB_DTYPE).reshape(cur.mbh, cur.mbw)


def col_records(field_p, mbh: int, mbw: int) -> np.ndarray:
    """B_COL_DTYPE records of a list-1 picture coded as a P picture with the
    ME-only field (MB_DTYPE records): list 0 = its MVs and reference indices,
    list 1 unused, never intra."""
    from . import B_COL_DTYPE
    f = np.asarray(field_p).reshape(mbh, mbw)
    col = np.zeros((mbh, mbw), B_COL_DTYPE)
    col["mv"][:, :, 0] = f["mv"][:, :, list(_CORNER)]
    col["ref"][:, :, 0] = f["ref"][:, :, None]
    col["ref"][:, :, 1] = -1
    return col


__all__ = ["B_JOB_DTYPE", "B_OUT_DTYPE", "NB_DTYPE", "b_jobs", "b_analyse", "b_field", "b_field_launch", "col_records",
           "vlim_rows"]
