"""pcme -- Python binding of libpcme, the gfx950 H.264 P-frame motion search.

libpcme's C ABI (include/pcme.h) is the product boundary; this module is a
thin ctypes layer over it plus device-memory plumbing through PyTorch
(tensors on ``cuda:0`` are only used as allocations and streams).

The search is the reference encoder's ``SliceEnc::search`` /
``search_part`` / partition choice (pito-codecs src/h264/enc/slice.rs:
1590-1723, 1810-1868, 1921-1945) run as an ME-only field over whole P
frames.  There is no CPU fallback: without the built library or a gfx950
device every call raises ``PcmeError``.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "lib", "libpcme.so"))

LUMA_PAD = 64
CHROMA_PAD = 32

PCME_OK = 0
PCME_EINVAL = -1
PCME_ENODEV = -2
PCME_EHIP = -3
PCME_ESTALL = -4
PCME_ENOMEM = -5

P16x16, P16x8, P8x16, P8x8 = 0, 1, 2, 3


class PcmeError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"pcme error {code}: {msg}")
        self.code = code


# --------------------------------------------------------------- C structs

class Mv(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int16), ("y", ctypes.c_int16)]


class Params(ctypes.Structure):
    _fields_ = [("mbw", ctypes.c_int32), ("mbh", ctypes.c_int32),
                ("me_range", ctypes.c_int32), ("subpel_rounds", ctypes.c_int32),
                ("partitions", ctypes.c_int32), ("qp", ctypes.c_int32)]


class RefPlanes(ctypes.Structure):
    _fields_ = [("luma", ctypes.c_void_p), ("h", ctypes.c_void_p), ("v", ctypes.c_void_p), ("c", ctypes.c_void_p)]


MAX_REFS = 4


class PlanesJob(ctypes.Structure):
    """pcme_planes_job: the arguments of one pcme_luma_planes call."""
    _fields_ = [("src", ctypes.c_void_p), ("full", ctypes.c_void_p), ("hp", ctypes.c_void_p),
                ("vp", ctypes.c_void_p), ("cp", ctypes.c_void_p), ("src_pitch", ctypes.c_int64),
                ("pitch", ctypes.c_int64), ("sstride", ctypes.c_int32), ("w", ctypes.c_int32),
                ("h", ctypes.c_int32), ("n", ctypes.c_int32), ("mbw", ctypes.c_int32), ("mbh", ctypes.c_int32)]


class FrameJob(ctypes.Structure):
    _fields_ = [("src", ctypes.c_void_p), ("ref", ctypes.c_void_p),
                ("ref_h", ctypes.c_void_p), ("ref_v", ctypes.c_void_p),
                ("ref_c", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("hint", ctypes.c_void_p), ("prm", Params),
                ("nref", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("older", RefPlanes * (MAX_REFS - 1))]


assert ctypes.sizeof(FrameJob) == 184


class SearchJob(ctypes.Structure):
    _fields_ = [("src", ctypes.c_void_p), ("sstride", ctypes.c_int32),
                ("ref", ctypes.c_void_p), ("ref_h", ctypes.c_void_p),
                ("ref_v", ctypes.c_void_p), ("ref_c", ctypes.c_void_p),
                ("mbw", ctypes.c_int32), ("mbh", ctypes.c_int32),
                ("bx", ctypes.c_int32), ("by", ctypes.c_int32),
                ("w", ctypes.c_int32), ("h", ctypes.c_int32),
                ("mvp", Mv), ("cands", Mv * 2), ("ncands", ctypes.c_int32),
                ("lam", ctypes.c_uint32), ("me_range", ctypes.c_int32),
                ("subpel_rounds", ctypes.c_int32), ("vlim", ctypes.c_int64)]


class NbBlock(ctypes.Structure):
    """pcme_nb_block: ref -2 unavailable, -1 no reference of the list, >= 0 index."""
    _fields_ = [("mv", Mv), ("ref", ctypes.c_int8), ("reserved", ctypes.c_uint8 * 3)]


class PartJob(ctypes.Structure):
    _fields_ = [("src", ctypes.c_void_p), ("sstride", ctypes.c_int32), ("part", ctypes.c_int32),
                ("ref", ctypes.c_void_p), ("ref_h", ctypes.c_void_p),
                ("ref_v", ctypes.c_void_p), ("ref_c", ctypes.c_void_p),
                ("mbw", ctypes.c_int32), ("mbh", ctypes.c_int32), ("mx", ctypes.c_int32), ("my", ctypes.c_int32),
                ("seed", Mv), ("lam", ctypes.c_uint32), ("me_range", ctypes.c_int32),
                ("subpel_rounds", ctypes.c_int32), ("vlim", ctypes.c_int64),
                ("left", NbBlock * 4), ("top", NbBlock * 4), ("top_right", NbBlock), ("top_left", NbBlock)]


assert ctypes.sizeof(NbBlock) == 8 and ctypes.sizeof(PartJob) == 168


class BJob(ctypes.Structure):
    _fields_ = [("src", ctypes.c_void_p), ("sstride", ctypes.c_int32), ("reserved0", ctypes.c_int32),
                ("ref0", ctypes.c_void_p * 4), ("ref1", ctypes.c_void_p * 4),
                ("mbw", ctypes.c_int32), ("mbh", ctypes.c_int32), ("mx", ctypes.c_int32), ("my", ctypes.c_int32),
                ("lam", ctypes.c_uint32), ("me_range", ctypes.c_int32), ("subpel_rounds", ctypes.c_int32),
                ("reserved1", ctypes.c_int32), ("vlim", ctypes.c_int64),
                ("nb", (NbBlock * 10) * 2), ("col_mv", (Mv * 4) * 2), ("col_ref", (ctypes.c_int8 * 4) * 2),
                ("col_intra", ctypes.c_int32), ("reserved2", ctypes.c_int32)]


assert ctypes.sizeof(BJob) == 328


class BFieldJob(ctypes.Structure):
    _fields_ = [("src", ctypes.c_void_p), ("ref0", ctypes.c_void_p * 4), ("ref1", ctypes.c_void_p * 4),
                ("col", ctypes.c_void_p), ("mbw", ctypes.c_int32), ("mbh", ctypes.c_int32), ("qp", ctypes.c_int32),
                ("me_range", ctypes.c_int32), ("subpel_rounds", ctypes.c_int32), ("partitions", ctypes.c_int32)]


assert ctypes.sizeof(BFieldJob) == 104


# pcme_skip_job flags
SKIP_T8X8, SKIP_ADAPTIVE_T8, SKIP_CABAC = 1, 2, 4


class SkipJob(ctypes.Structure):
    """pcme_skip_job: the early P-skip gate of one MB (source blocks and
    reference planes on the device)."""
    _fields_ = [("src_y", ctypes.c_void_p), ("src_u", ctypes.c_void_p), ("src_v", ctypes.c_void_p),
                ("ystride", ctypes.c_int32), ("cstride", ctypes.c_int32),
                ("ref", ctypes.c_void_p * 4), ("ref_u", ctypes.c_void_p), ("ref_v", ctypes.c_void_p),
                ("mbw", ctypes.c_int32), ("mbh", ctypes.c_int32), ("mx", ctypes.c_int32), ("my", ctypes.c_int32),
                ("skip_mv", Mv), ("qp", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("chroma_qp_offset", ctypes.c_int32), ("vlim", ctypes.c_int64)]


assert ctypes.sizeof(SkipJob) == 120

# pcme_skip_out (16 bytes)
SKIP_OUT_DTYPE = np.dtype([("sad", "<u4"), ("thresh", "<u4"), ("mv_ok", "u1"), ("coded", "u1"), ("cbp", "u1"),
                           ("skip", "u1"), ("t8x8", "u1"), ("reserved", "u1", (3,))])
assert SKIP_OUT_DTYPE.itemsize == 16


class Tune(ctypes.Structure):
    """pcme_tune: a plan's kernel configuration (see include/pcme.h)."""
    _fields_ = [("waves", ctypes.c_int32), ("waves_per_group", ctypes.c_int32), ("waves_per_cu", ctypes.c_int32),
                ("reuse_regions", ctypes.c_int32), ("ms", ctypes.c_float)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


# The kernel configurations libpcme builds: (waves_per_group, waves_per_cu, reuse_regions).
CONFIGS = ((1, 12, 1), (2, 12, 1), (4, 12, 1), (1, 8, 1), (2, 8, 1), (1, 12, 0), (4, 16, 0), (4, 16, 1))


class SynthCfg(ctypes.Structure):
    _fields_ = [("w", ctypes.c_int32), ("h", ctypes.c_int32),
                ("frames", ctypes.c_int32), ("nrect", ctypes.c_int32),
                ("rect_lo", ctypes.c_int32), ("rect_hi", ctypes.c_int32),
                ("vmax", ctypes.c_float), ("pan_max", ctypes.c_float),
                ("zoom", ctypes.c_float), ("noise_sigma", ctypes.c_float),
                ("seed", ctypes.c_uint64)]


class CplxJob(ctypes.Structure):
    _fields_ = [("cur", ctypes.c_void_p), ("prev", ctypes.c_void_p), ("blocks", ctypes.c_void_p),
                ("sums", ctypes.c_void_p), ("w", ctypes.c_int32), ("h", ctypes.c_int32)]


class TreeJob(ctypes.Structure):
    _fields_ = [("cur", ctypes.c_void_p), ("ref0", ctypes.c_void_p), ("ref1", ctypes.c_void_p),
                ("blocks", ctypes.c_void_p), ("w", ctypes.c_int32), ("h", ctypes.c_int32)]


class TreeNode(ctypes.Structure):
    """pcme_tree_node: a node's blocks (device) and its lists' reference
    node indices (-1: none)."""
    _fields_ = [("blocks", ctypes.c_void_p), ("ref", ctypes.c_int32 * 2)]


# pcme_part_out (36 bytes): search_part's (mvs, mvps, total)
PART_OUT_DTYPE = np.dtype([("mvs", "<i2", (4, 2)), ("mvps", "<i2", (4, 2)), ("total", "<u4")])
assert PART_OUT_DTYPE.itemsize == 36

# pcme_b_out (136 bytes): one B macroblock's motion search
B_OUT_DTYPE = np.dtype([
    ("direct_mv", "<i2", (2, 4, 2)), ("direct_ref", "i1", (2, 4)), ("direct_ok", "<u4"), ("direct_cost", "<u4"),
    ("mv16", "<i2", (2, 2)), ("mvp16", "<i2", (2, 2)), ("cost16", "<u4", (2,)), ("cost", "<u4", (3,)),
    ("dir", "<u4"), ("part_mv", "<i2", (2, 2, 2, 2)), ("part_cost", "<u4", (2,)), ("part_dir", "u1", (2, 2)),
    ("reserved", "u1", (4,)),
])
assert B_OUT_DTYPE.itemsize == 136
# pcme_b_col (44 bytes) and pcme_b_mb (280 bytes)
B_COL_DTYPE = np.dtype([("mv", "<i2", (2, 4, 2)), ("ref", "i1", (2, 4)), ("intra", "<i4")])
assert B_COL_DTYPE.itemsize == 44
B_MB_DTYPE = np.dtype([("search", B_OUT_DTYPE), ("mv", "<i2", (2, 16, 2)), ("ref", "i1", (2, 4)),
                       ("best_cost", "<u4"), ("mode", "u1"), ("dir", "u1"), ("reserved", "u1", (2,))])
assert B_MB_DTYPE.itemsize == 280

# pcme_cplx_block (12 bytes) and pcme_tree_block (20 bytes) as numpy records
CPLX_BLOCK_DTYPE = np.dtype([("intra", "<u4"), ("inter", "<u4"), ("mv", "<i2", (2,))])
assert CPLX_BLOCK_DTYPE.itemsize == 12
TREE_BLOCK_DTYPE = np.dtype([("intra", "<u4"), ("inter", "<u4"), ("mv", "<i2", (2, 2)), ("list", "u1"),
                             ("reserved", "u1", (3,))])
assert TREE_BLOCK_DTYPE.itemsize == 20

# pcme_mb_out as a numpy record (128 bytes)
MB_DTYPE = np.dtype([
    ("mv", "<i2", (16, 2)), ("mvs", "<i2", (4, 2)), ("mvps", "<i2", (4, 2)),
    ("cost", "<u4", (4,)), ("best_cost", "<u4"), ("skip_mv", "<i2", (2,)),
    ("mv16", "<i2", (2,)), ("part", "u1"), ("ref", "u1"), ("reserved", "u1", (2,)),
])
assert MB_DTYPE.itemsize == 128

# pcme_hint as a numpy record (40 bytes)
HINT_DTYPE = np.dtype([
    ("cand", "<i2", (8, 2)), ("conf_part", "<u2"), ("conf_mv", "<u2"),
    ("part", "u1"), ("ncand", "u1"), ("prune", "u1"), ("reserved", "u1"),
])
assert HINT_DTYPE.itemsize == 40

# Every symbol include/pcme.h and include/pcme_synth.h declare.
EXPORTS = (
    "pcme_version", "pcme_last_error", "pcme_device_count", "pcme_lambda",
    "pcme_qpel_table", "pcme_pad_planes", "pcme_hpel", "pcme_luma_planes", "pcme_luma_planes_batch", "pcme_search_frames",
    "pcme_search_frames_sync", "pcme_search_frames_check",
    "pcme_search_frames_noalloc", "pcme_search_batch", "pcme_pred_frame",
    "pcme_synth_clip", "pcme_plan_create", "pcme_plan_launch", "pcme_plan_check",
    "pcme_plan_set_grid", "pcme_plan_autotune", "pcme_plan_destroy", "pcme_set_spin_limit",
    "pcme_plan_get_config", "pcme_plan_set_config", "pcme_plan_autotune_ex",
    "pcme_net_create", "pcme_net_destroy", "pcme_net_get_params", "pcme_net_set_params",
    "pcme_net_inputs", "pcme_net_forward", "pcme_net_grad", "pcme_net_train_step", "pcme_net_hints",
    "pcme_lowres", "pcme_complexity", "pcme_mbtree_analyse", "pcme_search_part_batch", "pcme_b_analyse_batch",
    "pcme_b_field",
    "pcme_b_field_check",
    "pcme_b_fields",
    "pcme_aq_offsets", "pcme_mbtree_propagate", "pcme_mbtree_offsets", "pcme_skip_gate_batch",
)

_lib = None


def load() -> ctypes.CDLL:
    """Load libpcme.so (built in-tree).  PyTorch is imported first so that
    the library binds to the same HIP runtime as torch's allocations."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        import torch  # noqa: F401  (shares its libamdhip64 with libpcme)
    except Exception:  # pragma: no cover - torch is part of the image
        pass
    if not os.path.exists(LIB_PATH):
        raise PcmeError(PCME_ENODEV, f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    c = ctypes
    vp = c.c_void_p
    sig = {
        "pcme_version": (c.c_int, []),
        "pcme_last_error": (c.c_char_p, []),
        "pcme_device_count": (c.c_int, []),
        "pcme_lambda": (c.c_uint32, [c.c_int32]),
        "pcme_qpel_table": (None, [c.POINTER(c.c_uint32)]),
        "pcme_pad_planes": (c.c_int, [vp, c.c_int64, c.c_int32, c.c_int32, c.c_int32, vp, c.c_int64,
                                      c.c_int32, c.c_int32, c.c_int32, c.c_int32, c.c_int32, vp]),
        "pcme_hpel": (c.c_int, [vp, vp, vp, vp, c.c_int64, c.c_int32, c.c_int32, c.c_int32, vp]),
        "pcme_luma_planes": (c.c_int, [vp, c.c_int64, c.c_int32, c.c_int32, c.c_int32, vp, vp, vp, vp, c.c_int64,
                                       c.c_int32, c.c_int32, c.c_int32, vp]),
        "pcme_luma_planes_batch": (c.c_int, [c.POINTER(PlanesJob), c.c_int32, vp]),
        "pcme_search_frames": (c.c_int, [c.POINTER(FrameJob), c.c_int, vp]),
        "pcme_search_frames_sync": (c.c_int, [c.POINTER(FrameJob), c.c_int, vp]),
        "pcme_search_frames_check": (c.c_int, [vp]),
        "pcme_search_frames_noalloc": (c.c_int, [c.POINTER(FrameJob), c.c_int, vp]),
        "pcme_search_batch": (c.c_int, [c.POINTER(SearchJob), c.c_int, vp, vp, vp]),
        "pcme_search_part_batch": (c.c_int, [c.POINTER(PartJob), c.c_int, vp, vp]),
        "pcme_b_analyse_batch": (c.c_int, [c.POINTER(BJob), c.c_int, vp, vp]),
        "pcme_b_field": (c.c_int, [c.POINTER(BFieldJob), vp, vp]),
        "pcme_b_field_check": (c.c_int, [vp]),
        "pcme_b_fields": (c.c_int, [c.POINTER(BFieldJob), c.c_int, c.POINTER(vp), vp]),
        "pcme_skip_gate_batch": (c.c_int, [c.POINTER(SkipJob), c.c_int, vp, vp]),
        "pcme_pred_frame": (c.c_int, [vp, vp, vp, vp, vp, vp, vp, c.c_int32, c.c_int32, vp, vp, vp, vp]),
        "pcme_synth_clip": (c.c_int, [c.POINTER(SynthCfg), vp, vp, vp, c.c_int32]),
        "pcme_plan_create": (c.c_int, [c.POINTER(FrameJob), c.c_int, c.POINTER(vp)]),
        "pcme_plan_launch": (c.c_int, [vp, vp]),
        "pcme_plan_check": (c.c_int, [vp, vp]),
        "pcme_plan_set_grid": (c.c_int, [vp, c.c_int32]),
        "pcme_plan_autotune": (c.c_int, [vp, vp, c.c_int32, c.POINTER(c.c_int32), c.POINTER(c.c_float)]),
        "pcme_plan_destroy": (None, [vp]),
        "pcme_plan_get_config": (c.c_int, [vp, c.POINTER(Tune)]),
        "pcme_plan_set_config": (c.c_int, [vp, c.POINTER(Tune)]),
        "pcme_plan_autotune_ex": (c.c_int, [vp, vp, c.c_int32, c.POINTER(Tune)]),
        "pcme_set_spin_limit": (c.c_uint32, [c.c_uint32]),
        "pcme_lowres": (c.c_int, [vp, c.c_int64, c.c_int32, c.c_int32, c.c_int32, c.c_int32, vp, c.c_int64,
                                  c.c_int32, vp]),
        "pcme_complexity": (c.c_int, [c.POINTER(CplxJob), c.c_int32, vp]),
        "pcme_aq_offsets": (c.c_int, [vp, c.c_int32, vp, vp, c.c_int32, c.c_int32, c.c_int32, c.c_float, vp, vp]),
        "pcme_mbtree_analyse": (c.c_int, [c.POINTER(TreeJob), c.c_int32, vp]),
        "pcme_mbtree_propagate": (c.c_int, [c.POINTER(TreeNode), c.c_int32, c.c_int32, c.c_int32, vp, vp]),
        "pcme_mbtree_offsets": (c.c_int, [vp, vp, c.c_int32, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != PCME_OK:
        msg = load().pcme_last_error().decode(errors="replace")
        raise PcmeError(rc, msg)


def lam(qp: int) -> int:
    """Lambdas::new().sad[qp] (slice.rs:52-68)."""
    return int(load().pcme_lambda(int(qp)))


def qpel_table() -> list[int]:
    out = (ctypes.c_uint32 * 4)()
    load().pcme_qpel_table(out)
    return list(out)


def device_count() -> int:
    return int(load().pcme_device_count())


# ----------------------------------------------------------- synthetic video

@dataclass
class ClipSpec:
    """A seeded synthetic clip (include/pcme_synth.h)."""
    w: int
    h: int
    frames: int
    nrect: int = 4
    rect_lo: int = 24
    rect_hi: int = 80
    vmax: float = 3.0
    pan_max: float = 1.0
    zoom: float = 0.0
    noise_sigma: float = 2.0
    seed: int = 0x5EED0000


def synth_clip(spec: ClipSpec, threads: int = 0):
    """Generate the clip on the host: returns (Y[f,h,w], U[f,h/2,w/2], V)."""
    lib = load()
    y = np.empty((spec.frames, spec.h, spec.w), np.uint8)
    u = np.empty((spec.frames, spec.h // 2, spec.w // 2), np.uint8)
    v = np.empty_like(u)
    cfg = SynthCfg(spec.w, spec.h, spec.frames, spec.nrect, spec.rect_lo, spec.rect_hi,
                   spec.vmax, spec.pan_max, spec.zoom, spec.noise_sigma, spec.seed)
    rc = lib.pcme_synth_clip(ctypes.byref(cfg), y.ctypes.data, u.ctypes.data, v.ctypes.data, threads)
    _check(rc)
    return y, u, v


# --------------------------------------------------------------- geometry

def mb_dims(w: int, h: int) -> tuple[int, int]:
    return (w + 15) // 16, (h + 15) // 16


def luma_geom(mbw: int, mbh: int) -> tuple[int, int]:
    """(stride, rows) of a padded luma plane (picture.rs:150-155)."""
    return 16 * mbw + 2 * LUMA_PAD, 16 * mbh + 2 * LUMA_PAD


def chroma_geom(mbw: int, mbh: int) -> tuple[int, int]:
    return 8 * mbw + 2 * CHROMA_PAD, 8 * mbh + 2 * CHROMA_PAD


# ---------------------------------------------------------- device plumbing

def _torch():
    import torch
    if not torch.cuda.is_available():
        raise PcmeError(PCME_ENODEV, "no GPU visible to torch")
    return torch


def _stream_ptr(stream) -> int | None:
    if stream is None:
        torch = _torch()
        return torch.cuda.current_stream().cuda_stream
    return int(stream)


class DeviceClip:
    """A clip resident in HBM: padded luma planes, their half-pel planes and
    the padded chroma planes, laid out frame after frame."""

    def __init__(self, y_frames, mbw: int, mbh: int, u_frames=None, v_frames=None):
        torch = _torch()
        self.torch = torch
        self.n, self.h, self.w = (int(s) for s in y_frames.shape)
        self.mbw, self.mbh = mbw, mbh
        self.stride, self.rows = luma_geom(mbw, mbh)
        self.pitch = self.stride * self.rows
        dev = torch.device("cuda", torch.cuda.current_device())
        self.raw = torch.as_tensor(np.ascontiguousarray(y_frames)).to(dev) if not torch.is_tensor(y_frames) \
            else y_frames.contiguous().to(dev)
        self.planes = torch.empty((4, self.n, self.rows, self.stride), dtype=torch.uint8, device=dev)
        self.has_chroma = u_frames is not None
        if self.has_chroma:
            self.cstride, self.crows = chroma_geom(mbw, mbh)
            self.craw = [torch.as_tensor(np.ascontiguousarray(a)).to(dev) if not torch.is_tensor(a)
                         else a.contiguous().to(dev) for a in (u_frames, v_frames)]
            self.cplanes = torch.empty((2, self.n, self.crows, self.cstride), dtype=torch.uint8, device=dev)

    def plane_ptr(self, k: int, i: int) -> int:
        return self.planes[k, i].data_ptr()

    def pad(self, stream=None) -> None:
        """Import + edge-extend every frame (mod.rs:842-873, frame.rs:58-97)."""
        lib, s = load(), _stream_ptr(stream)
        _check(lib.pcme_pad_planes(self.raw.data_ptr(), self.h * self.w, self.w, self.w, self.h,
                                   self.planes[0].data_ptr(), self.pitch, self.stride,
                                   16 * self.mbw, 16 * self.mbh, LUMA_PAD, self.n, s))
        self.pad_chroma(stream)

    def pad_chroma(self, stream=None) -> None:
        """The padded chroma planes of every frame (pcme_pad_planes); nothing
        for a luma-only clip."""
        lib, s = load(), _stream_ptr(stream)
        if self.has_chroma:
            ch, cw = self.h // 2, self.w // 2
            for k in range(2):
                _check(lib.pcme_pad_planes(self.craw[k].data_ptr(), ch * cw, cw, cw, ch,
                                           self.cplanes[k].data_ptr(), self.crows * self.cstride,
                                           self.cstride, 8 * self.mbw, 8 * self.mbh, CHROMA_PAD,
                                           self.n, s))

    def prepare(self, stream=None) -> None:
        """pad() and hpel() of every frame: the luma planes in one fused pass
        (pcme_luma_planes), the chroma planes padded."""
        lib, s = load(), _stream_ptr(stream)
        _check(lib.pcme_luma_planes(self.raw.data_ptr(), self.h * self.w, self.w, self.w, self.h,
                                    self.plane_ptr(0, 0), self.plane_ptr(1, 0), self.plane_ptr(2, 0),
                                    self.plane_ptr(3, 0), self.pitch, self.n, self.mbw, self.mbh, s))
        self.pad_chroma(stream)

    def hpel(self, first: int = 0, count: int | None = None, stream=None) -> None:
        """Half-pel planes of frames [first, first+count) (picture.rs:318-380)."""
        count = self.n - first if count is None else count
        if count <= 0:
            return
        lib, s = load(), _stream_ptr(stream)
        _check(lib.pcme_hpel(self.plane_ptr(0, first), self.plane_ptr(1, first), self.plane_ptr(2, first),
                             self.plane_ptr(3, first), self.pitch, count, self.mbw, self.mbh, s))


def prepare_clips(clips: list[DeviceClip], stream=None) -> None:
    """DeviceClip.prepare of every clip, the luma planes of all of them
    through pcme_luma_planes_batch (one launch per 16 clips); a single clip
    through pcme_luma_planes, whose 2-D grid runs a 1080p clip's 60 frames in
    0.19 ms against 0.20 ms for the batch kernel's 1-D one."""
    if not clips:
        return
    if len(clips) == 1:
        clips[0].prepare(stream)
        return
    jobs = (PlanesJob * len(clips))(*[
        PlanesJob(dc.raw.data_ptr(), dc.plane_ptr(0, 0), dc.plane_ptr(1, 0), dc.plane_ptr(2, 0),
                  dc.plane_ptr(3, 0), dc.h * dc.w, dc.pitch, dc.w, dc.w, dc.h, dc.n, dc.mbw, dc.mbh)
        for dc in clips])
    _check(load().pcme_luma_planes_batch(jobs, len(clips), _stream_ptr(stream)))
    for dc in clips:
        dc.pad_chroma(stream)


def frame_job(clip: DeviceClip, cur: int, ref: int, out_ptr: int, me_range: int, subpel_rounds: int,
              partitions: bool, qp: int, ref_clip: DeviceClip | None = None, hint_ptr: int | None = None,
              older: list[int] | tuple = ()) -> FrameJob:
    """One P frame: picture `cur` of `clip` against picture `ref` (of
    ref_clip, default the same clip); `older` lists further reference
    pictures of the clip (reference indices 1, 2, ...), for the
    multi-reference search."""
    rc = clip if ref_clip is None else ref_clip
    if len(older) > MAX_REFS - 1:
        raise PcmeError(PCME_EINVAL, f"at most {MAX_REFS} references")
    j = FrameJob(clip.plane_ptr(0, cur), rc.plane_ptr(0, ref), rc.plane_ptr(1, ref),
                 rc.plane_ptr(2, ref), rc.plane_ptr(3, ref), out_ptr, hint_ptr,
                 Params(clip.mbw, clip.mbh, me_range, subpel_rounds, 1 if partitions else 0, qp))
    j.nref = 1 + len(older)
    for k, t in enumerate(older):
        j.older[k] = RefPlanes(rc.plane_ptr(0, t), rc.plane_ptr(1, t), rc.plane_ptr(2, t), rc.plane_ptr(3, t))
    return j


def search_frames(jobs: list[FrameJob], stream=None, sync: bool = True) -> None:
    """pcme_search_frames[_sync] over a list of FrameJob."""
    lib = load()
    arr = (FrameJob * len(jobs))(*jobs)
    s = _stream_ptr(stream)
    fn = lib.pcme_search_frames_sync if sync else lib.pcme_search_frames
    _check(fn(arr, len(jobs), s))


def search_frames_check(stream=None) -> None:
    """pcme_search_frames_check: raises PcmeError(PCME_ESTALL) if any
    pcme_search_frames launch of this device stalled since the last check."""
    _check(load().pcme_search_frames_check(_stream_ptr(stream)))


class Plan:
    """A reusable frame-search plan (pcme_plan_*): set up once, launch many
    times; launches are graph-capturable."""

    def __init__(self, jobs: list[FrameJob]):
        lib = load()
        self._jobs = (FrameJob * len(jobs))(*jobs)
        h = ctypes.c_void_p()
        _check(lib.pcme_plan_create(self._jobs, len(jobs), ctypes.byref(h)))
        self.handle = h

    def launch(self, stream=None) -> None:
        _check(load().pcme_plan_launch(self.handle, _stream_ptr(stream)))

    def check(self, stream=None) -> None:
        _check(load().pcme_plan_check(self.handle, _stream_ptr(stream)))

    def set_grid(self, waves: int) -> int:
        rc = load().pcme_plan_set_grid(self.handle, int(waves))
        if rc < 0:
            _check(rc)
        return rc

    def autotune(self, reps: int = 3, stream=None) -> tuple[int, float]:
        w = ctypes.c_int32()
        ms = ctypes.c_float()
        _check(load().pcme_plan_autotune(self.handle, _stream_ptr(stream), int(reps), ctypes.byref(w),
                                         ctypes.byref(ms)))
        return int(w.value), float(ms.value)

    def get_config(self) -> Tune:
        t = Tune()
        _check(load().pcme_plan_get_config(self.handle, ctypes.byref(t)))
        return t

    def set_config(self, waves: int = 0, waves_per_group: int = 1, waves_per_cu: int = 12,
                   reuse_regions: int = 1) -> Tune:
        t = Tune(int(waves), int(waves_per_group), int(waves_per_cu), int(reuse_regions), 0.0)
        _check(load().pcme_plan_set_config(self.handle, ctypes.byref(t)))
        return self.get_config()

    def autotune_ex(self, reps: int = 3, stream=None) -> Tune:
        """Every built kernel configuration over the grid sizes; keeps and
        returns the fastest."""
        t = Tune()
        _check(load().pcme_plan_autotune_ex(self.handle, _stream_ptr(stream), int(reps), ctypes.byref(t)))
        return t

    def close(self) -> None:
        if self.handle:
            load().pcme_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def search_batch(jobs: list[SearchJob], out_mv_ptr: int, out_cost_ptr: int, stream=None) -> None:
    lib = load()
    arr = (SearchJob * len(jobs))(*jobs)
    _check(lib.pcme_search_batch(arr, len(jobs), out_mv_ptr, out_cost_ptr, _stream_ptr(stream)))


def search_part_batch(jobs: list[PartJob], out_ptr: int, stream=None) -> None:
    """pcme_search_part_batch: out_ptr (device) receives len(jobs)
    PART_OUT_DTYPE records."""
    lib = load()
    arr = (PartJob * len(jobs))(*jobs)
    _check(lib.pcme_search_part_batch(arr, len(jobs), out_ptr, _stream_ptr(stream)))


def b_analyse_batch(jobs: list[BJob], out_ptr: int, stream=None) -> None:
    """pcme_b_analyse_batch: out_ptr (device) receives len(jobs)
    B_OUT_DTYPE records."""
    lib = load()
    arr = (BJob * len(jobs))(*jobs)
    _check(lib.pcme_b_analyse_batch(arr, len(jobs), out_ptr, _stream_ptr(stream)))


def skip_gate_batch(jobs: list[SkipJob], out_ptr: int, stream=None) -> None:
    """pcme_skip_gate_batch: out_ptr (device) receives len(jobs)
    SKIP_OUT_DTYPE records."""
    lib = load()
    arr = (SkipJob * len(jobs))(*jobs)
    _check(lib.pcme_skip_gate_batch(arr, len(jobs), out_ptr, _stream_ptr(stream)))


def pred_frame(clip: DeviceClip, ref: int, mb_ptr: int, pred_y_ptr: int, pred_u_ptr: int,
               pred_v_ptr: int, stream=None) -> None:
    if not clip.has_chroma:
        raise PcmeError(PCME_EINVAL, "clip has no chroma planes")
    lib = load()
    _check(lib.pcme_pred_frame(clip.plane_ptr(0, ref), clip.plane_ptr(1, ref), clip.plane_ptr(2, ref),
                               clip.plane_ptr(3, ref), clip.cplanes[0, ref].data_ptr(),
                               clip.cplanes[1, ref].data_ptr(), mb_ptr, clip.mbw, clip.mbh,
                               pred_y_ptr, pred_u_ptr, pred_v_ptr, _stream_ptr(stream)))


def search_clip_mr(clip: DeviceClip, nref: int, me_range: int, subpel_rounds: int, partitions: bool, qp: int,
                   stream=None, prepared: bool = False):
    """Multi-reference ME-only field of every P frame of an IPPP clip: frame
    t against its min(t, nref) previous frames, the newest first.  Returns
    the (n-1, mbh, mbw) MB_DTYPE records."""
    torch = clip.torch
    if not prepared:
        clip.pad(stream)
        clip.hpel(0, clip.n - 1, stream)
    nb = clip.n - 1
    out = torch.empty((max(nb, 1), clip.mbh * clip.mbw * MB_DTYPE.itemsize), dtype=torch.uint8,
                      device=clip.raw.device)
    jobs = [frame_job(clip, t, t - 1, out[t - 1].data_ptr(), me_range, subpel_rounds, partitions, qp,
                      older=[t - 1 - k for k in range(1, min(t, nref))])
            for t in range(1, clip.n)]
    search_frames(jobs, stream, sync=True)
    return out[:nb].cpu().numpy().view(MB_DTYPE).reshape(nb, clip.mbh, clip.mbw)


def search_clip_me(clip: DeviceClip, me_range: int, subpel_rounds: int, partitions: bool, qp: int,
                   stream=None):
    """ME-only field of every P frame of an IPPP clip (frame t searched
    against frame t-1).  Returns the (n-1, mbh, mbw) MB_DTYPE records."""
    torch = clip.torch
    clip.pad(stream)
    clip.hpel(0, clip.n - 1, stream)
    nb = clip.n - 1
    out = torch.empty((max(nb, 1), clip.mbh * clip.mbw * MB_DTYPE.itemsize), dtype=torch.uint8,
                      device=clip.raw.device)
    jobs = [frame_job(clip, t, t - 1, out[t - 1].data_ptr(), me_range, subpel_rounds, partitions, qp)
            for t in range(1, clip.n)]
    search_frames(jobs, stream, sync=True)
    return out[:nb].cpu().numpy().view(MB_DTYPE).reshape(nb, clip.mbh, clip.mbw)
