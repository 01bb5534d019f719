"""Pins of the oracle's AQ offsets (lookahead.rs:116-139 aq_offsets, dsp.rs
block_var): block_var against its closed form, the offsets of planes whose
variances are known (flat: all zero; one textured MB: strength times its
log2 term minus the picture mean), the +-8 clamp, and a float32 numpy
restatement on random planes (within 1e-5: numpy's and the C library's
log2 may differ in the last place)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("w,h", [(16, 16), (8, 8)])
def test_block_var_closed_form(w, h):
    rng = np.random.default_rng(w)
    for trial in range(20):
        a = rng.integers(0, 256, (h, w)).astype(np.uint8) if trial else np.full((h, w), 255, np.uint8)
        s = int(a.astype(np.int64).sum())
        ss = int((a.astype(np.int64) ** 2).sum())
        assert O.block_var(a) == ss - (s * s) // (w * h)
    assert O.block_var(np.full((h, w), 255, np.uint8)) == 0


def test_flat_picture_has_zero_offsets():
    y = np.full((48, 64), 90, np.uint8)
    c = np.full((24, 32), 128, np.uint8)
    assert not O.aq_offsets(y, c, c, 4, 3, 0.6).any()


def test_one_textured_mb():
    mbw, mbh, k = 4, 3, 5
    y = np.full((48, 64), 90, np.uint8)
    c = np.full((24, 32), 128, np.uint8)
    my, mx = divmod(k, mbw)
    blk = (np.indices((16, 16)).sum(0) % 2 * 255).astype(np.uint8)  # checkerboard 0/255
    y[16 * my:16 * my + 16, 16 * mx:16 * mx + 16] = blk
    var = 128 * 255 ** 2 - (128 * 255) ** 2 // 256  # 4,161,600
    assert O.block_var(blk) == var
    e = np.log2(np.float32(var / 256 + 1.0))
    mean = e / (mbw * mbh)
    got = O.aq_offsets(y, c, c, mbw, mbh, 0.6)
    want = np.full((mbh, mbw), -0.6 * mean)
    want[my, mx] = 0.6 * (e - mean)
    np.testing.assert_allclose(got, want, atol=1e-5)
    big = O.aq_offsets(y, c, c, mbw, mbh, 10.0)
    assert big[my, mx] == 8.0  # 10 * (13.99 - 1.17) clamps


def test_random_planes_against_float32_restatement():
    rng = np.random.default_rng(7)
    mbw, mbh = 6, 4
    y = rng.integers(0, 256, (64, 96)).astype(np.uint8)
    y[:, :48] = (y[:, :48] // 16 + 100).astype(np.uint8)  # low-variance half
    u = rng.integers(0, 256, (32, 48)).astype(np.uint8)
    v = rng.integers(100, 140, (32, 48)).astype(np.uint8)

    def var(a):
        a = a.astype(np.int64)
        return int((a ** 2).sum()) - int(a.sum()) ** 2 // a.size

    e = np.zeros(mbw * mbh, np.float32)
    for my in range(mbh):
        for mx in range(mbw):
            vy = np.float32(var(y[16 * my:16 * my + 16, 16 * mx:16 * mx + 16])) / np.float32(256)
            vu = np.float32(var(u[8 * my:8 * my + 8, 8 * mx:8 * mx + 8])) / np.float32(64)
            vv = np.float32(var(v[8 * my:8 * my + 8, 8 * mx:8 * mx + 8])) / np.float32(64)
            e[my * mbw + mx] = np.log2(vy + np.float32(0.5) * (vu + vv) + np.float32(1))
    mean = np.float32(0)
    for x in e:
        mean = np.float32(mean + x)
    mean = mean / np.float32(e.size)
    want = np.clip(np.float32(0.6) * (e - mean), -8, 8).reshape(mbh, mbw)
    np.testing.assert_allclose(O.aq_offsets(y, u, v, mbw, mbh, 0.6), want, atol=1e-5)
