"""A rate-distortion model of P-frame luma coding, to put a BD-rate on a
motion field without the rest of the encoder.

The reference encoder's bitstream is out of reach here (its transform,
entropy coders and RD decisions are outside the ported path), so the pruned
search is judged by what its motion field costs to code under one fixed
model, the same for every field compared:

* prediction: the chosen partitions' motion-compensated luma
  (pcme_pred_frame, the reference's pred_block);
* residual: the H.264 4x4 forward core transform, quantised with the
  standard multiplication factors and the inter rounding offset
  (f = 2^qbits / 6), dequantised and inverse-transformed with the standard
  rescaling factors and (x + 32) >> 6 (H.264 8.5.12, 8.6.1), added to the
  prediction and clipped to 0..255;
* distortion: the sum of squared errors against the source, as PSNR;
* rate: per 4x4 block, 1 bit when every level is zero, else each nonzero
  level in zig-zag order as ue(zero run) + se(level) plus 1 bit of block
  end; per MB ue(mb_type) (0 for P16x16, 1 for P16x8, 2 for P8x16, 3 for
  P8x8) and se(mvd) of both components of every partition's MV against its
  predictor (the record's mvps), as the reference's mv_bits counts them.

BD-rate is Bjontegaard's: log10(rate) fitted as a cubic in PSNR per field
over the QP points, integrated over the common PSNR range; the result is
the average rate difference of the test field over the anchor, in percent.
bd_rate_pchip interpolates piecewise-cubically (monotone) instead of
fitting one cubic, the variant later test conditions use.
"""
from __future__ import annotations

import numpy as np

# Core transform (H.264 8.6.1 forward matrix) and zig-zag scan of a 4x4 block.
CF = np.array([[1, 1, 1, 1], [2, 1, -1, -2], [1, -1, -1, 1], [1, -2, 2, -1]], np.int64)
ZIGZAG = np.array([0, 1, 4, 8, 5, 2, 3, 6, 9, 12, 13, 10, 7, 11, 14, 15])

# Quantisation (MF) and rescaling (V) factors by QP % 6 for the three
# position classes: (0,0)-type (both coordinates even), (1,1)-type (both
# odd), the rest.
_MF = np.array([[13107, 5243, 8066], [11916, 4660, 7490], [10082, 4194, 6554],
                [9362, 3647, 5825], [8192, 3355, 5243], [7282, 2893, 4559]], np.int64)
_V = np.array([[10, 16, 13], [11, 18, 14], [13, 20, 16], [14, 23, 18], [16, 25, 20], [18, 29, 23]], np.int64)


def _pos_class() -> np.ndarray:
    i, j = np.meshgrid(np.arange(4), np.arange(4), indexing="ij")
    return np.where((i % 2 == 0) & (j % 2 == 0), 0, np.where((i % 2 == 1) & (j % 2 == 1), 1, 2))


_CLS = _pos_class()


def mf(qp: int) -> np.ndarray:
    return _MF[qp % 6][_CLS]


def vscale(qp: int) -> np.ndarray:
    return _V[qp % 6][_CLS]


def forward(res: np.ndarray) -> np.ndarray:
    """Core transform of [..., 4, 4] residual blocks: CF . X . CF^T."""
    return np.einsum("ij,...jk,lk->...il", CF, res.astype(np.int64), CF)


def quantise(w: np.ndarray, qp: int) -> np.ndarray:
    """Levels of transform coefficients, inter offset."""
    qbits = 15 + qp // 6
    f = (1 << qbits) // 6
    return np.sign(w) * ((np.abs(w) * mf(qp) + f) >> qbits)


def reconstruct(z: np.ndarray, qp: int) -> np.ndarray:
    """Residual of levels: rescale, inverse core transform, (x + 32) >> 6."""
    d = z * vscale(qp) << (qp // 6)
    # inverse transform rows then columns (8.5.12.2), with the half-weights
    def inv1(x, axis):
        x = np.moveaxis(x, axis, -1)
        e = x[..., 0] + x[..., 2]
        f = x[..., 0] - x[..., 2]
        g = (x[..., 1] >> 1) - x[..., 3]
        h = x[..., 1] + (x[..., 3] >> 1)
        y = np.stack([e + h, f + g, f - g, e - h], axis=-1)
        return np.moveaxis(y, -1, axis)
    r = inv1(inv1(d, -1), -2)
    return (r + 32) >> 6


def ue_bits(v: np.ndarray) -> np.ndarray:
    """Exp-Golomb ue(v) length: 2 floor(log2(v + 1)) + 1."""
    v = np.asarray(v, np.int64)
    return 2 * np.floor(np.log2(v + 1)).astype(np.int64) + 1


def se_bits(v: np.ndarray) -> np.ndarray:
    """Exp-Golomb se(v) length (k = 2v - 1 for v > 0, -2v otherwise)."""
    v = np.asarray(v, np.int64)
    return ue_bits(np.where(v > 0, 2 * v - 1, -2 * v))


def block_bits(z: np.ndarray) -> np.ndarray:
    """Bits of [..., 4, 4] level blocks under the run/level model."""
    s = z.reshape(*z.shape[:-2], 16)[..., ZIGZAG]
    nz = s != 0
    # zero run before each nonzero level in scan order
    idx = np.arange(16)
    last = np.where(nz, idx, -1)
    prev = np.maximum.accumulate(np.concatenate([np.full(s.shape[:-1] + (1,), -1), last[..., :-1]], -1), axis=-1)
    run = idx - prev - 1
    bits = np.where(nz, ue_bits(np.maximum(run, 0)) + se_bits(s), 0).sum(-1)
    return np.where(nz.any(-1), bits + 1, 1)


PART_COUNT = np.array([1, 2, 2, 4])


def mv_bits(field: np.ndarray) -> np.ndarray:
    """Per-MB side bits: ue(mb_type) + se(mvd.x) + se(mvd.y) per partition."""
    f = field.reshape(-1)
    part = f["part"].astype(np.int64)
    bits = ue_bits(part)
    mvs, mvps = f["mvs"].astype(np.int64), f["mvps"].astype(np.int64)
    d = se_bits(mvs[..., 0] - mvps[..., 0]) + se_bits(mvs[..., 1] - mvps[..., 1])
    used = np.arange(4)[None, :] < PART_COUNT[part][:, None]
    return bits + (d * used).sum(-1)


def code_frame(src_mbs: np.ndarray, pred_mbs: np.ndarray, field: np.ndarray, qp: int):
    """(bits, sse) of one frame: src_mbs, pred_mbs [nmb, 16, 16] uint8."""
    s = src_mbs.astype(np.int64)
    p = pred_mbs.astype(np.int64)
    blk = (s - p).reshape(-1, 4, 4, 4, 4).transpose(0, 1, 3, 2, 4)  # [mb, by, bx, 4, 4]
    z = quantise(forward(blk), qp)
    rec = np.clip(p + reconstruct(z, qp).transpose(0, 1, 3, 2, 4).reshape(-1, 16, 16), 0, 255)
    sse = float(((rec - s) ** 2).sum())
    bits = float(block_bits(z).sum() + mv_bits(field).sum())
    return bits, sse


def psnr(sse: float, npix: int) -> float:
    return 10.0 * np.log10(255.0 ** 2 * npix / max(sse, 1e-9))


def bd_rate(anchor: list[tuple[float, float]], test: list[tuple[float, float]]) -> float:
    """Bjontegaard delta rate (%) of `test` over `anchor`, each a list of
    (rate, psnr) points (at least 4)."""
    ra, pa = np.log10([r for r, _ in anchor]), np.array([q for _, q in anchor])
    rt, pt = np.log10([r for r, _ in test]), np.array([q for _, q in test])
    ca, ct = np.polyfit(pa, ra, 3), np.polyfit(pt, rt, 3)
    lo, hi = max(pa.min(), pt.min()), min(pa.max(), pt.max())
    ia, it = np.polyint(ca), np.polyint(ct)
    da = (np.polyval(ia, hi) - np.polyval(ia, lo)) / (hi - lo)
    dt = (np.polyval(it, hi) - np.polyval(it, lo)) / (hi - lo)
    return float((10.0 ** (dt - da) - 1.0) * 100.0)


def bd_rate_pchip(anchor: list[tuple[float, float]], test: list[tuple[float, float]]) -> float:
    """BD-rate (%) with monotone piecewise-cubic interpolation of log10(rate)
    over PSNR, integrated numerically over the common PSNR range."""
    from scipy.interpolate import PchipInterpolator

    def curve(pts):
        r = np.log10([p[0] for p in pts])
        q = np.array([p[1] for p in pts])
        o = np.argsort(q)
        return PchipInterpolator(q[o], r[o]), q.min(), q.max()

    fa, la, ha = curve(anchor)
    ft, lt, ht = curve(test)
    lo, hi = max(la, lt), min(ha, ht)
    xs = np.linspace(lo, hi, 4001)
    da = np.trapezoid(fa(xs), xs) / (hi - lo)
    dt = np.trapezoid(ft(xs), xs) / (hi - lo)
    return float((10.0 ** (dt - da) - 1.0) * 100.0)
