"""Writes the golden vectors of the motion search: deterministic numpy
frames (no libm-dependent generator) and the ME-only fields computed by the
independent Python restatement tests/pysearch.py.  Run from the repo root:
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import pysearch  # noqa: E402

DTYPE = np.dtype([
    ("mv", "<i2", (16, 2)), ("mvs", "<i2", (4, 2)), ("mvps", "<i2", (4, 2)),
    ("cost", "<u4", (4,)), ("best_cost", "<u4"), ("skip_mv", "<i2", (2,)),
    ("mv16", "<i2", (2,)), ("part", "u1"), ("ref", "u1"), ("reserved", "u1", (2,)),
])


def frames(seed, w, h, n, motion):
    """Block-smoothed uniform noise, a flat rectangle moving by `motion`
    px/frame over a background panned by 1 px/frame, integer rounding."""
    rs = np.random.default_rng(seed)
    base = rs.integers(0, 256, (h + 48, w + 48)).astype(np.int64)
    sm = np.zeros_like(base)
    for dy in range(-2, 3):
        for dx in range(-2, 3):
            sm += np.roll(np.roll(base, dy, 0), dx, 1)
    sm //= 25
    out = []
    for t in range(n):
        f = sm[24 - t:24 - t + h, 24 + t:24 + t + w].copy()
        x0, y0 = 8 + motion[0] * t, 6 + motion[1] * t
        f[max(y0, 0):max(y0 + 12, 0), max(x0, 0):max(x0 + 14, 0)] = 200
        f += rs.integers(-2, 3, (h, w))
        out.append(np.clip(f, 0, 255).astype(np.uint8))
    return np.stack(out)


CASES = [
    # name, seed, w, h, frames, motion, me_range, subpel, partitions, qp
    ("g48x32_r16", 1, 48, 32, 3, (3, 1), 16, 1, 1, 28),
    ("g64x48_r8_s0", 2, 64, 48, 3, (-2, 3), 8, 0, 1, 22),
    ("g32x32_r32_s2", 3, 32, 32, 2, (5, -4), 32, 2, 1, 37),
    ("g40x24_odd_np", 4, 40, 24, 2, (1, 2), 16, 1, 0, 27),
    ("g16x16_single", 5, 16, 16, 3, (2, 2), 16, 1, 1, 26),
]


def main():
    for name, seed, w, h, n, motion, rng, sub, parts, qp in CASES:
        f = frames(seed, w, h, n, motion)
        exp = np.stack([pysearch.to_records(pysearch.Frame(f[t], f[t - 1], rng, sub, bool(parts), qp).run(), DTYPE)
                        for t in range(1, n)])
        np.savez_compressed(os.path.join(HERE, name + ".npz"), frames=f,
                            params=np.array([rng, sub, parts, qp], np.int32),
                            expected=exp.view(np.uint8))
        print(name, f.shape, exp.shape)


if __name__ == "__main__":
    main()
