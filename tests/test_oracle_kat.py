"""Pin the oracle's leaf functions against closed forms and the H.264
specification, each restated here from the mathematics (numpy), not from
the oracle's code."""
from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

# Lambdas.sad for QP 0..51, the published closed-form values.
LAMBDA_TABLE = [
    59, 66, 74, 83, 94, 105, 118, 132, 149, 167, 187, 210, 236, 265, 297, 334, 375, 421, 472, 530,
    595, 668, 749, 841, 944, 1060, 1189, 1335, 1499, 1682, 1888, 2119, 2379, 2670, 2997, 3364,
    3776, 4239, 4758, 5341, 5995, 6729, 7553, 8478, 9516, 10681, 11989, 13457, 15105, 16955, 19031, 21362,
]


def P(a):
    return np.ascontiguousarray(a).ctypes.data


def test_lambda_closed_form_and_table(oracle):
    lib = oracle.load()
    for qp in range(52):
        v = max(16, int(math.floor(256 * math.sqrt(0.85 * 2 ** ((qp - 12) / 3)) + 0.5)))
        assert lib.o_lambda(qp) == v == LAMBDA_TABLE[qp], qp


def test_se_bits_is_exp_golomb_length(oracle):
    lib = oracle.load()
    for v in range(-5000, 5001):
        k = 2 * v - 1 if v > 0 else -2 * v  # se(v) -> ue(k) mapping
        assert lib.o_se_bits(v) == 2 * ((k + 1).bit_length() - 1) + 1, v
    mv = oracle.omv(7, -3)
    mvp = oracle.omv(-2, 5)
    assert lib.o_mv_bits(mv, mvp) == lib.o_se_bits(9) + lib.o_se_bits(-8)


def test_median(oracle):
    lib = oracle.load()
    rs = np.random.default_rng(1)
    for _ in range(2000):
        a, b, c = (int(x) for x in rs.integers(-3000, 3000, 3))
        assert lib.o_median(a, b, c) == sorted([a, b, c])[1]


@pytest.mark.parametrize("w,h", [(16, 16), (16, 8), (8, 16), (8, 8), (4, 4)])
def test_sad_closed_form(oracle, w, h):
    lib = oracle.load()
    rs = np.random.default_rng(w * 100 + h)
    for trial in range(50):
        a = rs.integers(0, 256, (h, 32), dtype=np.uint8)
        b = rs.integers(0, 256, (h, 40), dtype=np.uint8)
        if trial == 0:
            a[:] = 0
            b[:] = 255
        want = int(np.abs(a[:, :w].astype(int) - b[:, :w].astype(int)).sum())
        assert lib.o_sad(P(a), 32, P(b), 40, w, h) == want
        if w == 16:
            assert lib.o_sad16(P(a), 32, P(b), 40, h) == want


H4 = np.array([[1, 1, 1, 1], [1, -1, 1, -1], [1, 1, -1, -1], [1, -1, -1, 1]])


def satd_textbook(a, b):
    d = a.astype(int) - b.astype(int)
    tot = 0
    for y in range(0, d.shape[0], 4):
        for x in range(0, d.shape[1], 4):
            t = H4 @ d[y:y + 4, x:x + 4] @ H4.T
            s = int(np.abs(t).sum())
            assert s % 2 == 0
            tot += s // 2
    return tot


@pytest.mark.parametrize("w,h", [(16, 16), (16, 8), (8, 16), (8, 8), (4, 4)])
def test_satd_is_hadamard(oracle, w, h):
    lib = oracle.load()
    rs = np.random.default_rng(w + 7 * h)
    for trial in range(60):
        a = rs.integers(0, 256, (h, w), dtype=np.uint8)
        b = rs.integers(0, 256, (h, w), dtype=np.uint8)
        if trial == 1:
            a[:] = 255
            b[:] = 0
        if trial == 2:
            a[:] = 0
            b[::2, ::3] = 255
        assert lib.o_satd(P(a), w, P(b), w, w, h) == satd_textbook(a, b)


def test_satd_constant_residual_kat(oracle):
    lib = oracle.load()
    for a in (0, 1, 5, 100, 255):
        x = np.full((16, 16), a, np.uint8)
        z = np.zeros((16, 16), np.uint8)
        assert lib.o_satd(P(x), 16, P(z), 16, 16, 16) == 8 * a * 16


# ---------------------------------------------------------------- planes

def pad_reference(frame, mbw, mbh, pad):
    """Import (replicate into the MB-aligned area) then edge-extend."""
    h, w = frame.shape
    unit = 16 if pad == 64 else 8
    al = np.pad(frame, ((0, unit * mbh - h), (0, unit * mbw - w)), mode="edge")
    return np.pad(al, pad, mode="edge")


@pytest.mark.parametrize("w,h", [(48, 32), (50, 30), (17, 3), (320, 240)])
def test_pad_is_edge_replication(oracle, w, h):
    rs = np.random.default_rng(w * h)
    f = rs.integers(0, 256, (h, w), dtype=np.uint8)
    mbw, mbh = (w + 15) // 16, (h + 15) // 16
    np.testing.assert_array_equal(oracle.pad_luma(f, mbw, mbh), pad_reference(f, mbw, mbh, 64))
    c = rs.integers(0, 256, (h // 2 or 1, w // 2 or 1), dtype=np.uint8)
    np.testing.assert_array_equal(oracle.pad_chroma(c, mbw, mbh), pad_reference(c, mbw, mbh, 32))


TAP = np.array([1, -5, 20, 20, -5, 1])


def hpel_spec(G):
    """H.264 8.4.2.2.1: b (horizontal 6-tap), h (vertical), j (vertical
    6-tap over the unrounded horizontal intermediates), samples outside the
    plane taken from the nearest edge sample."""
    G = G.astype(np.int64)
    Gp = np.pad(G, 3, mode="edge")
    R, C = G.shape
    b1 = sum(TAP[i] * Gp[3:3 + R, 1 + i:1 + i + C] for i in range(6))   # x-2 .. x+3
    h1 = sum(TAP[i] * Gp[1 + i:1 + i + R, 3:3 + C] for i in range(6))
    b1p = np.pad(b1, 3, mode="edge")
    j1 = sum(TAP[i] * b1p[1 + i:1 + i + R, 3:3 + C] for i in range(6))
    clip = lambda v: np.clip(v, 0, 255).astype(np.uint8)  # noqa: E731
    return clip((b1 + 16) >> 5), clip((h1 + 16) >> 5), clip((j1 + 512) >> 10)


@pytest.mark.parametrize("shape", [(40, 36), (7, 9), (160, 200)])
def test_hpel_matches_spec(oracle, shape):
    rs = np.random.default_rng(shape[0])
    G = rs.integers(0, 256, shape, dtype=np.uint8)
    G[0, :] = 255
    G[:, -1] = 0
    got = oracle.hpel(G)
    want = hpel_spec(G)
    for k in range(3):
        np.testing.assert_array_equal(got[k], want[k], err_msg="bhj"[k])


def test_hpel_impulse_kat(oracle):
    """A single 255 on zeros: b at the two positions left of the impulse
    column pair gives the tap weights (clipped); j the outer product."""
    G = np.zeros((32, 32), np.uint8)
    G[16, 16] = 255
    hb, hh, hj = oracle.hpel(G)
    row = hb[16, 12:20].astype(int)
    want = [0] * 8
    for x in range(12, 20):
        tap_idx = 16 - (x - 2)  # sample 16 is tap (16 - (x-2)) of position x
        if 0 <= tap_idx < 6:
            want[x - 12] = int(np.clip((TAP[tap_idx] * 255 + 16) >> 5, 0, 255))
    assert row.tolist() == want
    assert hh[12:20, 16].tolist() == want


# The two samples averaged at each quarter position, H.264 8.4.2.2.1, as
# (plane, dx, dy) with planes G=0, b=1, h=2, j=3.
QSPEC = {
    (0, 0): ((0, 0, 0), (0, 0, 0)), (1, 0): ((0, 0, 0), (1, 0, 0)), (2, 0): ((1, 0, 0), (1, 0, 0)),
    (3, 0): ((1, 0, 0), (0, 1, 0)), (0, 1): ((0, 0, 0), (2, 0, 0)), (1, 1): ((1, 0, 0), (2, 0, 0)),
    (2, 1): ((1, 0, 0), (3, 0, 0)), (3, 1): ((1, 0, 0), (2, 1, 0)), (0, 2): ((2, 0, 0), (2, 0, 0)),
    (1, 2): ((2, 0, 0), (3, 0, 0)), (2, 2): ((3, 0, 0), (3, 0, 0)), (3, 2): ((3, 0, 0), (2, 1, 0)),
    (0, 3): ((2, 0, 0), (0, 0, 1)), (1, 3): ((2, 0, 0), (1, 0, 1)), (2, 3): ((3, 0, 0), (1, 0, 1)),
    (3, 3): ((2, 1, 0), (1, 0, 1)),
}


def test_mc_luma_matches_spec(oracle):
    lib = oracle.load()
    rs = np.random.default_rng(5)
    f = rs.integers(0, 256, (64, 80), dtype=np.uint8)
    mbw, mbh = 5, 4
    plane = oracle.pad_luma(f, mbw, mbh)
    ref = oracle.Ref(plane)
    planes = [plane.astype(int), *[p.astype(int) for p in hpel_spec(plane)]]
    stride = plane.shape[1]
    for trial in range(300):
        w, h = [(16, 16), (16, 8), (8, 16), (8, 8)][trial % 4]
        x = int(rs.integers(0, 80 - w + 1))
        y = int(rs.integers(0, 64 - h + 1))
        mvx = int(rs.integers(-160, 161))
        mvy = int(rs.integers(-160, 161))
        out = np.zeros((16, 16), np.uint8)
        lib.o_mc_luma(ctypes.byref(ref.view), x, y, oracle.omv(mvx, mvy), w, h, P(out), 16)
        (pa, dxa, dya), (pb, dxb, dyb) = QSPEC[(mvx & 3, mvy & 3)]
        X = x + (mvx >> 2) + 64
        Y = y + (mvy >> 2) + 64
        A = planes[pa][Y + dya:Y + dya + h, X + dxa:X + dxa + w]
        B = planes[pb][Y + dyb:Y + dyb + h, X + dxb:X + dxb + w]
        np.testing.assert_array_equal(out[:h, :w], ((A + B + 1) >> 1).astype(np.uint8), err_msg=str(trial))
    assert stride == 80 + 128


def test_mc_chroma_matches_spec(oracle):
    """H.264 8.4.2.2.2: ((8-xF)(8-yF)A + xF(8-yF)B + (8-xF)yF C + xF yF D + 32) >> 6."""
    lib = oracle.load()
    rs = np.random.default_rng(9)
    u = rs.integers(0, 256, (32, 40), dtype=np.uint8)
    v = rs.integers(0, 256, (32, 40), dtype=np.uint8)
    pu, pv = oracle.pad_chroma(u, 5, 4), oracle.pad_chroma(v, 5, 4)
    for trial in range(200):
        w, h = [(8, 8), (8, 4), (4, 8), (4, 4)][trial % 4]
        x = int(rs.integers(0, 40 - w + 1))
        y = int(rs.integers(0, 32 - h + 1))
        mvx = int(rs.integers(-150, 151))
        mvy = int(rs.integers(-150, 151))
        ou = np.zeros((8, 8), np.uint8)
        ov = np.zeros((8, 8), np.uint8)
        lib.o_mc_chroma(P(pu), P(pv), pu.shape[1], 32, x, y, oracle.omv(mvx, mvy), w, h, P(ou), P(ov), 8)
        fx, fy = mvx & 7, mvy & 7
        X, Y = x + (mvx >> 3) + 32, y + (mvy >> 3) + 32
        for pl, o in ((pu, ou), (pv, ov)):
            q = pl.astype(int)
            A = q[Y:Y + h, X:X + w]
            B = q[Y:Y + h, X + 1:X + 1 + w]
            C = q[Y + 1:Y + 1 + h, X:X + w]
            D = q[Y + 1:Y + 1 + h, X + 1:X + 1 + w]
            want = ((8 - fx) * (8 - fy) * A + fx * (8 - fy) * B + (8 - fx) * fy * C + fx * fy * D + 32) >> 6
            np.testing.assert_array_equal(o[:h, :w], want.astype(np.uint8))


def test_vlim_rule(oracle):
    """need = 16*(my+1) + 96 rows of the reference; the whole picture once
    need reaches the coded height, then the limit is h + 64 - 20."""
    lib = oracle.load()
    assert [lib.o_vlim(my, 68) for my in (0, 1, 60, 61, 62, 67)] == [112, 128, 1072, 16 * 68 + 44, 1132, 1132]
    assert lib.o_vlim(0, 15) == 112 and lib.o_vlim(8, 15) == 240 + 44
    assert lib.o_vlim(0, 1) == 60


def test_mv_ok_boundaries(oracle):
    lib = oracle.load()
    p = oracle.oparams(5, 4, 16, 1, 1, 100)  # pw = 80
    ok = lambda x, y, w, h, mx, my: lib.o_mv_ok(ctypes.byref(p), x, y, w, h, oracle.omv(mx, my))  # noqa: E731
    assert ok(0, 0, 16, 16, -44 * 4, 0) and not ok(0, 0, 16, 16, -45 * 4, 0)
    assert ok(0, 0, 16, 16, -44 * 4 - 3 + 3, 0) and not ok(0, 0, 16, 16, -44 * 4 - 1, 0)
    assert ok(64, 0, 16, 16, 44 * 4 + 3, 0) and not ok(64, 0, 16, 16, 45 * 4, 0)
    assert ok(0, 0, 16, 16, 0, -44 * 4) and not ok(0, 0, 16, 16, 0, -45 * 4)
    assert ok(0, 16, 16, 16, 0, (99 - 32) * 4) and not ok(0, 16, 16, 16, 0, (100 - 32) * 4)
    p2 = oracle.oparams(200, 200, 16, 1, 1, 10000)
    ok2 = lambda mx, my: lib.o_mv_ok(ctypes.byref(p2), 1000, 1000, 16, 16, oracle.omv(mx, my))  # noqa: E731
    assert ok2(1999, -1999) and not ok2(2000, 0) and not ok2(0, -2000)
