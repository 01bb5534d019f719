"""Pins of the P-frame coding model (pcme.rdmodel) that puts a BD-rate on a
motion field: the transform against its definition, the quantiser tables
against the transform's norms, reconstruction error against the step size,
the bit counts against hand-coded Exp-Golomb sequences, BD-rate against
curves with a known rate ratio."""
from __future__ import annotations

import numpy as np
import pytest

from pcme import rdmodel as R


def test_forward_transform_matches_its_definition():
    rng = np.random.default_rng(1)
    x = rng.integers(-255, 256, (5, 4, 4))
    got = R.forward(x)
    for n in range(5):
        for i in range(4):
            for l in range(4):
                want = sum(R.CF[i, j] * x[n, j, k] * R.CF[l, k] for j in range(4) for k in range(4))
                assert got[n, i, l] == want


def test_quantiser_tables_match_the_transform_norms():
    """MF * V = 2^17 times the squared-norm ratio of each position class
    (1, 0.64, 0.8 for (even, even), (odd, odd), mixed), for every QP % 6,
    to within the tables' own rounding (0.014%); a one-digit typo in an
    entry moves its product by more than 0.2%."""
    for q in range(6):
        prod = R.mf(q) * R.vscale(q)
        want = (2 ** 17) * np.where(R._CLS == 0, 1.0, np.where(R._CLS == 1, 0.64, 0.8))
        assert np.all(np.abs(prod / want - 1.0) < 5e-4), (q, prod)


@pytest.mark.parametrize("qp", [0, 12, 22, 27, 32, 37, 45])
def test_reconstruction_error_is_bounded_by_the_step(qp):
    rng = np.random.default_rng(qp)
    x = rng.integers(-60, 61, (4000, 4, 4))
    r = R.reconstruct(R.quantise(R.forward(x), qp), qp)
    qstep = 0.625 * 2 ** (qp / 6)
    err = np.abs(r - x)
    assert err.max() <= max(1, 2.5 * qstep), (qp, err.max(), qstep)
    if qp == 0:
        assert err.mean() < 0.2
    elif qp <= 37:
        # a wrong quantiser or rescale scale moves the mean error off the
        # step size's ~0.35 (uniform input, inter dead zone)
        assert 0.25 < err.mean() / qstep < 0.45, (qp, err.mean() / qstep)


def test_dc_only_level_reconstructs_flat():
    for qp in (10, 28, 40):
        z = np.zeros((4, 4), np.int64)
        z[0, 0] = 3
        d = 3 * R.vscale(qp)[0, 0] << (qp // 6)
        assert np.all(R.reconstruct(z, qp) == (d + 32) >> 6)


def test_exp_golomb_lengths():
    assert list(R.ue_bits([0, 1, 2, 3, 6, 7, 14, 15])) == [1, 3, 3, 5, 5, 7, 7, 9]
    assert list(R.se_bits([0, 1, -1, 2, -2, 3, -3, 4])) == [1, 3, 3, 5, 5, 5, 5, 7]


def test_block_bits_hand_coded():
    z = np.zeros((3, 4, 4), np.int64)
    z[1, 0, 0] = 1            # DC level 1: ue(0) + se(1) + end = 1 + 3 + 1
    z[2, 2, 0] = -2           # zig-zag position 3 (after 3 zeros): ue(3) + se(-2) + end = 5 + 5 + 1
    z[2, 0, 1] = 1            # position 1 too: runs become 1 then 1: ue(1)+se(1) + ue(1)+se(-2) + 1
    b = R.block_bits(z)
    assert b[0] == 1 and b[1] == 5
    assert b[2] == (3 + 3) + (3 + 5) + 1


def test_mv_bits_of_a_two_partition_mb():
    from pcme import MB_DTYPE
    f = np.zeros(1, MB_DTYPE)
    f["part"] = 1
    f["mvs"][0, 0] = (4, -2)
    f["mvps"][0, 0] = (0, 0)   # mvd (4, -2): se(4) + se(-2) = 7 + 5
    f["mvs"][0, 1] = (8, 8)
    f["mvps"][0, 1] = (8, 9)   # mvd (0, -1): 1 + 3
    f["mvs"][0, 2] = (100, 100)  # unused partitions are not counted
    assert R.mv_bits(f)[0] == 3 + (7 + 5) + (1 + 3)


def test_bd_rate_of_known_rate_ratios():
    base = [(1000.0, 30.0), (1800.0, 33.0), (3500.0, 36.0), (7000.0, 39.5)]
    assert abs(R.bd_rate(base, base)) < 1e-9
    assert abs(R.bd_rate(base, [(1.1 * r, q) for r, q in base]) - 10.0) < 1e-6
    assert abs(R.bd_rate(base, [(0.8 * r, q) for r, q in base]) + 20.0) < 1e-6
    assert abs(R.bd_rate_pchip(base, base)) < 1e-9
    assert abs(R.bd_rate_pchip(base, [(1.1 * r, q) for r, q in base]) - 10.0) < 1e-6


def test_coding_a_perfect_prediction_costs_only_side_bits():
    from pcme import MB_DTYPE
    rng = np.random.default_rng(4)
    src = rng.integers(0, 256, (6, 16, 16)).astype(np.uint8)
    f = np.zeros(6, MB_DTYPE)
    bits, sse = R.code_frame(src, src.copy(), f, 27)
    assert sse == 0.0
    assert bits == 6 * 16 * 1 + 6 * (1 + 1 + 1)  # 16 empty blocks + ue(0) + se(0) + se(0) per MB
