"""Shared helpers of the test suite."""
from __future__ import annotations

import numpy as np

FIELDS = ("mv", "mvs", "mvps", "cost", "best_cost", "skip_mv", "mv16", "part", "ref")


def assert_same_field(got: np.ndarray, want: np.ndarray, what: str = "") -> None:
    """Byte-identical MB records; on mismatch name the first MB and field."""
    assert got.shape == want.shape, (got.shape, want.shape)
    gb = np.ascontiguousarray(got).view(np.uint8).reshape(got.size, -1)
    wb = np.ascontiguousarray(want).view(np.uint8).reshape(want.size, -1)
    if np.array_equal(gb, wb):
        return
    bad = np.nonzero((gb != wb).any(axis=1))[0]
    i = int(bad[0])
    g = got.reshape(-1)[i]
    w = want.reshape(-1)[i]
    idx = np.unravel_index(i, got.shape)
    diffs = [f for f in FIELDS if not np.array_equal(g[f], w[f])]
    detail = "; ".join(f"{f}: got {g[f].tolist()} want {w[f].tolist()}" for f in diffs)
    raise AssertionError(f"{what}: {len(bad)} of {got.size} MB records differ; first at {idx}: {detail}")


def scene_clip(scene: str, w: int, h: int, n: int) -> np.ndarray:
    """The reference encoder tests' synthetic luma scenes
    (src/h264/tests_enc.rs:16-62: hash noise, gradient, moving checker)."""
    M = (1 << 64) - 1

    def hsh(x, y, i, k):
        s = ((x * 0x9E3779B97F4A7C15) & M) ^ ((y * 0xC2B2AE3D27D4EB4F) & M) ^ ((i * 0x165667B19E3779F9) & M) ^ k
        s ^= s >> 29
        s = (s * 0xBF58476D1CE4E5B9) & M
        s ^= s >> 32
        return s & 0xFFFFFFFF

    out = np.zeros((n, h, w), np.uint8)
    xs, ys = np.meshgrid(np.arange(w, dtype=np.uint64), np.arange(h, dtype=np.uint64))
    for i in range(n):
        if scene == "flat":
            v = np.full((h, w), 96)
        elif scene == "gradient":
            v = (np.arange(w)[None, :] * 200 // max(w, 1) + np.arange(h)[:, None] * 50 // max(h, 1) + 2 * i)
        else:
            with np.errstate(over="ignore"):
                s = (xs * np.uint64(0x9E3779B97F4A7C15)) ^ (ys * np.uint64(0xC2B2AE3D27D4EB4F)) ^ \
                    np.uint64((i * 0x165667B19E3779F9) & M) ^ np.uint64(0)
                s ^= s >> np.uint64(29)
                s = s * np.uint64(0xBF58476D1CE4E5B9)
                s ^= s >> np.uint64(32)
                s = (s & np.uint64(0xFFFFFFFF)).astype(np.int64)
            if scene == "noise":
                v = 108 + s % 41
            else:  # motion: checkerboard moving (3, 2) px/frame plus noise
                mx = np.arange(w)[None, :] + 3 * i
                my = np.arange(h)[:, None] + 2 * i
                checker = ((mx // 8 + my // 8) % 2)
                v = 50 + checker * 90 + (np.arange(w)[None, :] + np.arange(h)[:, None]) % 7 + s % 13
        out[i] = np.minimum(v, 255).astype(np.uint8)
    return out
