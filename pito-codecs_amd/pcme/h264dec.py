"""The minimal H.264 decoder (include/pcme_dec.h, libpcme_h264.so) under the
reference's names: ``Decoder`` with ``decode(packet)`` and ``flush()``
returning decoded frames (src/h264/dec/decoder.rs), ``decode_annexb`` and
``frames_digest`` as the reference's test harness uses them
(src/h264/tests.rs:121-170, src/testref.rs:543-578)."""
from __future__ import annotations

import ctypes
import hashlib
import os
from dataclasses import dataclass

import numpy as np

from . import PCME_OK, PcmeError

LIB_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib", "libpcme_h264.so")

# pcme_h264_mb (144 bytes)
MB_DTYPE = np.dtype([("kind", "i1"), ("part", "i1"), ("sub", "i1", (4,)), ("ref", "i1", (4,)), ("qp", "u1"),
                     ("reserved", "u1"), ("mv", "<i2", (16, 2)), ("mvp", "<i2", (16, 2)), ("slice", "<i4")])
assert MB_DTYPE.itemsize == 144
KIND_I4x4, KIND_I16x16, KIND_IPCM, KIND_P, KIND_PSKIP = range(5)

EXPORTS = ("pcme_h264dec_create", "pcme_h264dec_destroy", "pcme_h264dec_decode", "pcme_h264dec_flush",
           "pcme_h264dec_frame", "pcme_h264dec_frame_mbs", "pcme_h264_mc_luma", "pcme_h264dec_last_error")


class _Info(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("poc", ctypes.c_int32),
                ("frame_num", ctypes.c_int32)]


_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PcmeError(-2, f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    c, vp = ctypes, ctypes.c_void_p
    sig = {
        "pcme_h264dec_create": (c.c_int, [c.POINTER(vp)]),
        "pcme_h264dec_destroy": (None, [vp]),
        "pcme_h264dec_decode": (c.c_int, [vp, c.c_char_p, c.c_size_t]),
        "pcme_h264dec_flush": (c.c_int, [vp]),
        "pcme_h264dec_frame": (c.c_int, [vp, c.POINTER(_Info), vp, c.c_int32, vp, vp, c.c_int32]),
        "pcme_h264dec_frame_mbs": (c.c_int, [vp, vp, c.c_int32]),
        "pcme_h264_mc_luma": (c.c_int, [vp, c.c_int32, c.c_int32, c.c_int32, c.c_int32, c.c_int32, c.c_int32,
                                        c.c_int32, c.c_int32, c.c_int32, vp, c.c_int32]),
        "pcme_h264dec_last_error": (c.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc: int) -> None:
    if rc != PCME_OK:
        raise PcmeError(rc, load().pcme_h264dec_last_error().decode(errors="replace"))


@dataclass
class Frame:
    """A decoded I420 picture (cropped) with its macroblocks."""
    width: int
    height: int
    poc: int
    frame_num: int
    y: np.ndarray
    u: np.ndarray
    v: np.ndarray
    mbs: np.ndarray  # MB_DTYPE records in raster order


class Decoder:
    """Decoder (dec/decoder.rs): decode() takes Annex-B bytes of whole NAL
    units (an access unit, or more) and returns the frames they completed;
    flush() returns the rest."""

    def __init__(self):
        self._lib = load()
        self._h = ctypes.c_void_p()
        _check(self._lib.pcme_h264dec_create(ctypes.byref(self._h)))

    def _drain(self) -> list[Frame]:
        out = []
        while True:
            info = _Info()
            r = self._lib.pcme_h264dec_frame(self._h, ctypes.byref(info), None, 0, None, None, 0)
            if r < 0:
                _check(r)
            if r == 0:
                return out
            n = self._lib.pcme_h264dec_frame_mbs(self._h, None, 0)
            mbs = np.zeros(n, MB_DTYPE)
            if n:
                self._lib.pcme_h264dec_frame_mbs(self._h, mbs.ctypes.data, n)
            w, h = info.width, info.height
            y = np.empty((h, w), np.uint8)
            u = np.empty((h // 2, w // 2), np.uint8)
            v = np.empty((h // 2, w // 2), np.uint8)
            r = self._lib.pcme_h264dec_frame(self._h, ctypes.byref(info), y.ctypes.data, w, u.ctypes.data,
                                             v.ctypes.data, w // 2)
            if r != 1:
                _check(r if r < 0 else -1)
            out.append(Frame(w, h, info.poc, info.frame_num, y, u, v, mbs))

    def decode(self, data: bytes) -> list[Frame]:
        data = bytes(data)
        _check(self._lib.pcme_h264dec_decode(self._h, data, len(data)))
        return self._drain()

    def flush(self) -> list[Frame]:
        _check(self._lib.pcme_h264dec_flush(self._h))
        return self._drain()

    def close(self) -> None:
        if self._h:
            self._lib.pcme_h264dec_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def access_units(data: bytes) -> list[bytes]:
    """Split an Annex-B stream into access units the way the reference's
    test harness does (tests.rs:141-170): a unit starts at the parameter
    sets / SEI before a slice whose first_mb_in_slice is 0, or at that
    slice."""
    starts, pending, i = [], None, 0
    while i + 3 < len(data):
        if data[i] == 0 and data[i + 1] == 0 and data[i + 2] == 1:
            at = i - 1 if i > 0 and data[i - 1] == 0 else i
            t = data[i + 3] & 0x1F
            first_mb_zero = i + 4 < len(data) and (data[i + 4] & 0x80) != 0
            if 6 <= t <= 9 and pending is None:
                pending = at
            elif t in (1, 5) and first_mb_zero:
                starts.append(pending if pending is not None else at)
                pending = None
            i += 3
        else:
            i += 1
    starts.append(len(data))
    return [data[a:b] for a, b in zip(starts, starts[1:])]


def decode_annexb(data: bytes, per_access_unit: bool = True) -> list[Frame]:
    """tests.rs:121-136: every frame of a stream, fed access unit by access
    unit (or as one buffer), then flushed."""
    dec = Decoder()
    frames = []
    for au in (access_units(data) if per_access_unit else [data]):
        frames += dec.decode(au)
    frames += dec.flush()
    dec.close()
    return frames


def frames_digest(frames: list[Frame]) -> str:
    """testref.rs Digest::frame: sha256 over every plane's visible rows."""
    h = hashlib.sha256()
    for f in frames:
        for p in (f.y, f.u, f.v):
            h.update(np.ascontiguousarray(p).tobytes())
    return h.hexdigest()


def mc_luma(plane: np.ndarray, x: int, y: int, mv, bw: int, bh: int) -> np.ndarray:
    """8.4.2.2.1: the luma prediction of a bw x bh block at (x, y) of an
    unpadded picture, displaced by the quarter-sample MV."""
    p = np.ascontiguousarray(plane, np.uint8)
    out = np.empty((bh, bw), np.uint8)
    _check(load().pcme_h264_mc_luma(p.ctypes.data, p.shape[1], p.shape[1], p.shape[0], x, y, int(mv[0]),
                                    int(mv[1]), bw, bh, out.ctypes.data, bw))
    return out
