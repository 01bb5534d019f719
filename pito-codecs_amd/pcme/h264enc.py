"""The field coder (include/pcme_enc.h, libpcme_h264.so): codes a motion
field -- the ME-only field of the GPU search, plain or pruned -- into a
Baseline CAVLC H.264 stream whose reconstruction the decoder of pcme_dec.h
reproduces exactly, so a field's rate and distortion are measured on real
bits.  No counterpart in the reference (its encoder decides with RD, CABAC
and intra analysis)."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import MB_DTYPE, PCME_OK, PcmeError
from . import h264dec

EXPORTS = ("pcme_h264enc_create", "pcme_h264enc_destroy", "pcme_h264enc_picture", "pcme_h264enc_recon",
           "pcme_h264enc_set_refs", "pcme_h264enc_stats")


def load() -> ctypes.CDLL:
    lib = h264dec.load()
    if not hasattr(lib, "_enc_bound"):
        c, vp = ctypes, ctypes.c_void_p
        sig = {
            "pcme_h264enc_create": (c.c_int, [c.c_int32, c.c_int32, c.POINTER(vp)]),
            "pcme_h264enc_destroy": (None, [vp]),
            "pcme_h264enc_picture": (c.c_int, [vp, vp, c.c_int32, vp, vp, c.c_int32, vp, c.c_int32, vp, c.c_size_t,
                                               c.POINTER(c.c_size_t)]),
            "pcme_h264enc_recon": (c.c_int, [vp, vp, c.c_int32, vp, vp, c.c_int32]),
            "pcme_h264enc_set_refs": (c.c_int, [vp, c.c_int32]),
            "pcme_h264enc_stats": (c.c_int, [vp, c.POINTER(c.c_int32), c.POINTER(c.c_int32),
                                             c.POINTER(c.c_int32)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        lib._enc_bound = True
    return lib


def _check(rc: int) -> None:
    if rc != PCME_OK:
        raise PcmeError(rc, load().pcme_h264dec_last_error().decode(errors="replace"))


class Encoder:
    """One coder per clip: picture() codes an IDR (field None) or a P
    picture with the given field (mbw*mbh MB_DTYPE records)."""

    def __init__(self, width: int, height: int, nref: int = 1):
        self._lib = load()
        self.width, self.height = width, height
        self.mbw, self.mbh = (width + 15) // 16, (height + 15) // 16
        self._h = ctypes.c_void_p()
        _check(self._lib.pcme_h264enc_create(width, height, ctypes.byref(self._h)))
        if nref != 1:
            _check(self._lib.pcme_h264enc_set_refs(self._h, int(nref)))
        self._buf = np.empty(16 * self.mbw * 16 * self.mbh * 3 + (1 << 16), np.uint8)

    def picture(self, y, u, v, field=None, qp: int = 27) -> bytes:
        y, u, v = (np.ascontiguousarray(a, np.uint8) for a in (y, u, v))
        assert y.shape == (self.height, self.width) and u.shape == v.shape == (self.height // 2, self.width // 2)
        fp = None
        if field is not None:
            f = np.ascontiguousarray(np.asarray(field).reshape(-1))
            assert f.dtype.itemsize == MB_DTYPE.itemsize and f.size == self.mbw * self.mbh
            fp = f.ctypes.data
        n = ctypes.c_size_t()
        _check(self._lib.pcme_h264enc_picture(self._h, y.ctypes.data, self.width, u.ctypes.data, v.ctypes.data,
                                              self.width // 2, fp, int(qp), self._buf.ctypes.data, self._buf.size,
                                              ctypes.byref(n)))
        return self._buf[:n.value].tobytes()

    def recon(self):
        y = np.empty((self.height, self.width), np.uint8)
        u = np.empty((self.height // 2, self.width // 2), np.uint8)
        v = np.empty_like(u)
        _check(self._lib.pcme_h264enc_recon(self._h, y.ctypes.data, self.width, u.ctypes.data, v.ctypes.data,
                                            self.width // 2))
        return y, u, v

    def stats(self) -> dict:
        s, i, p = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _check(self._lib.pcme_h264enc_stats(self._h, ctypes.byref(s), ctypes.byref(i), ctypes.byref(p)))
        return {"skip": s.value, "inter": i.value, "pcm": p.value}

    def close(self) -> None:
        if self._h:
            self._lib.pcme_h264enc_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def psnr(a, b) -> float:
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    mse = float(np.mean(d * d))
    return 99.0 if mse == 0 else 10.0 * np.log10(255.0 * 255.0 / mse)


@dataclass
class Coded:
    """A clip coded with one field per P picture: the stream, the P pictures'
    bits and their luma PSNR, and the reconstructions."""
    stream: bytes
    p_bits: int
    p_psnr: float
    psnr: list
    recon: list


def code_clip(y, u, v, fields, qp: int, nref: int = 1) -> Coded:
    """IDR of frame 0 (I_PCM), then P pictures 1..n-1, picture t coded with
    fields[t-1] (with nref > 1 a multi-reference field: picture t's list 0
    is the reconstructions of t-1 .. t-nref).  The rate counts the P
    pictures only (the IDR is the same lossless picture whatever the field)
    and the distortion is their mean luma PSNR."""
    n, h, w = y.shape
    enc = Encoder(w, h, nref)
    out = [enc.picture(y[0], u[0], v[0], None, qp)]
    rec, ps = [enc.recon()], []
    for t in range(1, n):
        out.append(enc.picture(y[t], u[t], v[t], fields[t - 1], qp))
        r = enc.recon()
        rec.append(r)
        ps.append(psnr(r[0], y[t]))
    enc.close()
    p_bits = 8 * sum(len(b) for b in out[1:])
    return Coded(b"".join(out), p_bits, float(np.mean(ps)) if ps else 99.0, ps, rec)
