"""numpy restatement of the learned search pruner (include/pcme_net.h).

TEST INFRASTRUCTURE ONLY: the tests use it as the checker of the HIP
training and inference kernels -- never the product path.  The pruner has
no counterpart in the reference; this file restates the network that
pcme_net.h defines, in float64:

  * `inputs`   the 24x24x2 input of a macroblock (integer arithmetic, exact);
  * `forward`  conv1-4 + ReLU, global average pool, fc -- with
               `emulate=True` the weights and the conv1-3 activations are
               rounded to bf16 (round to nearest even) as on the GPU, the
               arithmetic stays float64;
  * `loss_and_grad`  mean softmax cross-entropy of both heads and its
               gradient by hand-written backpropagation (rounding treated as
               the identity).

How it is pinned (tests/test_cnn_oracle.py): the convolution against a
brute-force loop, the gradient against central finite differences in
float64, the bf16 rounding against float32 bit patterns, the softmax
gradient against its closed form (rows sum to zero).
"""
from __future__ import annotations

import numpy as np

CH = 32
NCLS = 293
MVC = 289
MV_SIDE = 17
NPARAM = 38021
# (name, shape) in the order of the flat parameter vector
LAYOUT = [("w1", (32, 9, 2)), ("b1", (32,)), ("w2", (32, 9, 32)), ("b2", (32,)),
          ("w3", (32, 9, 32)), ("b3", (32,)), ("w4", (32, 9, 32)), ("b4", (32,)),
          ("wf", (NCLS, 32)), ("bf", (NCLS,))]
STRIDES = (2, 2, 1, 2)
LUMA_PAD = 64


def bf16(x) -> np.ndarray:
    """Round to the nearest bf16 (ties to even), returned as float64."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(flat: np.ndarray) -> dict:
    flat = np.asarray(flat, np.float64)
    assert flat.size == NPARAM
    out, o = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = flat[o:o + n].reshape(shape)
        o += n
    return out


def pack(p: dict) -> np.ndarray:
    return np.concatenate([np.asarray(p[name], np.float64).reshape(-1) for name, _ in LAYOUT])


def inputs(src: np.ndarray, ref: np.ndarray, mbw: int, mbh: int, mbs) -> np.ndarray:
    """[n, 24, 24, 2] network inputs of macroblocks `mbs` (raster indices)
    of padded luma planes src/ref: channel 0 = the 48x48 reference window
    around the MB 2x2-averaged with rounding, channel 1 = the MB averaged
    the same way at rows/cols 8..15; value (v - 128) / 64."""
    out = np.zeros((len(mbs), 24, 24, 2))
    for i, m in enumerate(mbs):
        my, mx = divmod(int(m), mbw)
        X, Y = 16 * mx + LUMA_PAD, 16 * my + LUMA_PAD
        w = ref[Y - 16:Y + 32, X - 16:X + 32].astype(np.int64)
        d = (w[0::2, 0::2] + w[0::2, 1::2] + w[1::2, 0::2] + w[1::2, 1::2] + 2) >> 2
        out[i, :, :, 0] = (d - 128) / 64.0
        s = src[Y:Y + 16, X:X + 16].astype(np.int64)
        e = (s[0::2, 0::2] + s[0::2, 1::2] + s[1::2, 0::2] + s[1::2, 1::2] + 2) >> 2
        out[i, 8:16, 8:16, 1] = (e - 128) / 64.0
    return out


def _taps(x: np.ndarray, stride: int, hout: int):
    """The nine shifted views of a zero-padded [n, h, h, c] map."""
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    for ty in range(3):
        for tx in range(3):
            yield 3 * ty + tx, xp[:, ty:ty + stride * (hout - 1) + 1:stride, tx:tx + stride * (hout - 1) + 1:stride, :]


def conv(x: np.ndarray, w: np.ndarray, b: np.ndarray, stride: int) -> np.ndarray:
    """3x3 convolution, pad 1: y[n, oy, ox, co] = b[co] + sum_{tap, ci}
    w[co, tap, ci] x[n, oy*s + ty - 1, ox*s + tx - 1, ci]."""
    h = x.shape[1]
    hout = (h - 1) // stride + 1
    y = np.broadcast_to(b, (x.shape[0], hout, hout, w.shape[0])).copy()
    for tap, xs in _taps(x, stride, hout):
        y += xs @ w[:, tap, :].T
    return y


def conv_bwd(x: np.ndarray, w: np.ndarray, dy: np.ndarray, stride: int):
    """Gradients of conv w.r.t. its input, weights and bias."""
    h = x.shape[1]
    hout = dy.shape[1]
    dxp = np.zeros((x.shape[0], h + 2, h + 2, x.shape[3]))
    dw = np.zeros_like(w)
    for tap, xs in _taps(x, stride, hout):
        ty, tx = divmod(tap, 3)
        dw[:, tap, :] = np.einsum("nyxo,nyxi->oi", dy, xs)
        dxp[:, ty:ty + stride * (hout - 1) + 1:stride, tx:tx + stride * (hout - 1) + 1:stride, :] += dy @ w[:, tap, :]
    db = dy.sum(axis=(0, 1, 2))
    return dxp[:, 1:-1, 1:-1, :], dw, db


def forward(x: np.ndarray, p: dict, emulate: bool = True) -> dict:
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    acts = {"x": x}
    a = x
    for k in range(4):
        z = conv(a, r(p[f"w{k + 1}"]), p[f"b{k + 1}"], STRIDES[k])
        a = np.maximum(z, 0.0)
        if k < 3:
            a = r(a)
        acts[f"a{k + 1}"] = a
    g = a.reshape(a.shape[0], 9, CH).mean(axis=1)
    acts["g"] = g
    acts["logits"] = g @ r(p["wf"]).T + p["bf"]
    return acts


def _softmax(z: np.ndarray) -> np.ndarray:
    e = np.exp(z - z.max(axis=1, keepdims=True))
    return e / e.sum(axis=1, keepdims=True)


def loss_and_grad(x: np.ndarray, p: dict, mv: np.ndarray, part: np.ndarray, emulate: bool = True):
    """(loss_total, loss_mv, loss_part, flat gradient)."""
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    n = x.shape[0]
    acts = forward(x, p, emulate)
    z = acts["logits"]
    pm, pp = _softmax(z[:, :MVC]), _softmax(z[:, MVC:])
    lmv = -np.log(pm[np.arange(n), mv]).mean()
    lpt = -np.log(pp[np.arange(n), part]).mean()
    dz = np.zeros_like(z)
    dz[:, :MVC] = pm
    dz[np.arange(n), mv] -= 1.0
    dz[:, MVC:] = pp
    dz[np.arange(n), MVC + part] -= 1.0
    dz /= n
    grads = {"wf": dz.T @ acts["g"], "bf": dz.sum(axis=0)}
    dg = dz @ r(p["wf"])
    a4 = acts["a4"]
    da = np.repeat(dg[:, None, :] / 9.0, 9, axis=1).reshape(a4.shape)
    ins = [acts["x"], acts["a1"], acts["a2"], acts["a3"]]
    for k in range(3, -1, -1):
        dzk = da * (acts[f"a{k + 1}"] > 0)
        dx, dw, db = conv_bwd(ins[k], r(p[f"w{k + 1}"]), dzk, STRIDES[k])
        grads[f"w{k + 1}"], grads[f"b{k + 1}"] = dw, db
        da = dx
    return lmv + lpt, lmv, lpt, pack(grads)


def mv_class(mv16: np.ndarray) -> np.ndarray:
    """Motion class of quarter-pel MVs [..., 2]: the nearest point of the
    17x17 grid of even integer-pel offsets within +-16 (clamped)."""
    c = np.clip((np.asarray(mv16, np.int64) + 4) >> 3, -8, 8) + 8
    return c[..., 1] * MV_SIDE + c[..., 0]


def class_mv(c) -> np.ndarray:
    """Quarter-pel MV of a motion class."""
    c = np.asarray(c, np.int64)
    cy, cx = np.divmod(c, MV_SIDE)
    return np.stack([8 * cx - 64, 8 * cy - 64], axis=-1)
