"""The oracle's early P-skip gate (oracle/skip_oracle.c) pinned by the
H.264 definitions its pieces restate: the forward core transforms as
matrix products, the quantiser tables through the identity that makes
quantisation and dequantisation inverse, the dequantise+inverse-transform
reconstruction within the quantiser's step, the decimation score on
hand-counted runs, the chroma QP table, and the gate's decisions on flat
pictures whose levels are worked out by hand."""
from __future__ import annotations

import numpy as np
import pytest

# H.264 forward 4x4 core transform (8.5.12 is its inverse): Y = Cf X Cf^T
CF4 = np.array([[1, 1, 1, 1], [2, 1, -1, -2], [1, -1, -1, 1], [1, -2, 2, -1]])
# the 8x8 forward transform whose inverse is the 8x8 transform of 8.5.13
# (every row of the inverse's basis, scaled by 8)
CF8 = np.array([
    [8, 8, 8, 8, 8, 8, 8, 8],
    [12, 10, 6, 3, -3, -6, -10, -12],
    [8, 4, -4, -8, -8, -4, 4, 8],
    [10, -3, -12, -6, 6, 12, 3, -10],
    [8, -8, -8, 8, 8, -8, -8, 8],
    [6, -12, 3, 10, -10, -3, 12, -6],
    [4, -8, 8, -4, -4, 8, -8, 4],
    [3, -6, 10, -12, 12, -10, 6, -3],
])


def test_fdct4_is_the_core_transform(oracle):
    rs = np.random.default_rng(1)
    for _ in range(200):
        x = rs.integers(-255, 256, (4, 4))
        np.testing.assert_array_equal(oracle.fdct4(x), CF4 @ x @ CF4.T)


def test_cf8_rows_are_orthogonal():
    """The 8x8 basis: mutually orthogonal rows (the premise of the
    round-trip tests below)."""
    g = CF8 @ CF8.T
    assert np.count_nonzero(g - np.diag(np.diag(g))) == 0


def test_quant_tables_invert_the_dequantiser(oracle):
    """4x4: mf * v * w = 2^21 within 0.05%, where w = 16, 25, 20 are the
    squared norms (x 64) of positions (even, even), (odd, odd) and mixed;
    8x8: mf8 * v8 * |CF8 row x|^2 * |CF8 row y|^2 = 2^36 within 0.02%
    (squared row norms 512, 578, 320)."""
    q = oracle.quant_tables()
    for m in range(6):
        for i in range(16):
            x, y = i & 3, i >> 2
            w = 16 if (x % 2 == 0 and y % 2 == 0) else (25 if (x % 2 and y % 2) else 20)
            prod = int(q["mf4"][m, i]) * int(q["v4"][m, i]) * w
            assert abs(prod - 2 ** 21) / 2 ** 21 < 5e-4, (m, i, prod)
        norms = (CF8 * CF8).sum(1)  # 512, 578, 320, ...
        for i in range(64):
            x, y = i & 7, i >> 3
            prod = float(q["mf8"][m, i]) * float(q["v8"][m, i]) * norms[x] * norms[y]
            assert abs(prod - 2 ** 36) / 2 ** 36 < 2e-4, (m, i)
    # the standard's first MF and V rows (8.5.9 / the usual quantiser tables)
    assert list(q["mf4"][0, [0, 5, 1]]) == [13107, 5243, 8066]
    assert list(q["v4"][0, [0, 5, 1]]) == [10, 16, 13]


@pytest.mark.parametrize("qp", [12, 22, 27, 32, 37])
def test_reconstruction_within_the_step(oracle, qp):
    """Without decimation the 4x4 and 8x8 paths reconstruct the source
    within the quantiser's step (about 0.625 * 2^(qp/6) per pixel, a few
    steps at worst): a wrong table entry, position map or zigzag breaks it."""
    rs = np.random.default_rng(qp)
    step = 0.625 * 2 ** (qp / 6)
    for _ in range(20):
        pred = rs.integers(30, 226, (16, 16)).astype(np.uint8)
        src = np.clip(pred.astype(int) + rs.integers(-40, 41, (16, 16)), 0, 255).astype(np.uint8)
        _, _, r4 = oracle.code_luma4(src, pred, qp, decimate=False)
        _, _, r8 = oracle.code_luma8(src, pred, qp, decimate=False)
        for r in (r4, r8):
            err = np.abs(r.astype(int) - src.astype(int))
            assert err.mean() < 0.6 * step + 0.5 and err.max() <= 4 * step + 2, (qp, err.mean(), err.max())


def test_decimate_score_counts_runs(oracle):
    # slice.rs:149-165: a level of magnitude > 1 scores 64; each +-1 scores
    # T[run of zeros before it] with T = 3, 2, 2, 1, 1, 1, 0, ...
    assert oracle.decimate_score([1] + [0] * 15) == 3
    assert oracle.decimate_score([0, -1] + [0] * 14) == 2
    assert oracle.decimate_score([0] * 5 + [1] + [0] * 10) == 1
    assert oracle.decimate_score([0] * 6 + [1] + [0] * 9) == 0
    assert oracle.decimate_score([1, 1, 0, 0, 1] + [0] * 11) == 3 + 3 + 2
    assert oracle.decimate_score([0, 0, 2] + [0] * 13) == 64
    assert oracle.decimate_score([0] * 64) == 0


def test_chroma_qp_table(oracle):
    # H.264 Table 8-15 (QPc as a function of qPI)
    tail = [29, 30, 31, 32, 32, 33, 34, 34, 35, 35, 36, 36, 37, 37, 37, 38, 38, 38, 39, 39, 39, 39]
    want = list(range(30)) + tail
    assert [oracle.load().o_chroma_qp(q) for q in range(52)] == want
    assert oracle.load().o_chroma_qp(-3) == 0 and oracle.load().o_chroma_qp(60) == 39


def test_flat_residual_levels_by_hand(oracle):
    """A constant luma residual d gives each 4x4 block one DC coefficient
    16d (the other 15 are 0).  At QP 27: qbits 19, MF 9362, f = 2^19*43/256
    = 88,064, so level = (16|d|*9362 + 88064) >> 19: d = 2 -> 387,648 >> 19
    = 0 (cbp 0), d = 3 -> 537,440 >> 19 = 1.  A level-1 DC per block scores
    3 per block, 12 per 8x8, 48 in all: nothing is decimated, cbp 0xF.
    8x8 at QP 27: the DC is 64*8*... = 8*8*64 d (CF8 row 0 has entries 8),
    mf8[3][0] = round(2^36 / (512*512*28)) = 9362, qbits 26, f = 2^26*43/256
    = 11,272,192: d = 3 -> (12,288*9362 + 11272192) >> 26 = 1 per 8x8
    (score 3 < 4: decimated, cbp 0); d = 6 -> 241,352,704 >> 26 = 3 (score 64,
    kept)."""
    pred = np.full((16, 16), 100, np.uint8)
    for d, want4 in ((2, 0), (3, 0xF), (-3, 0xF)):
        cbp, lv, _ = oracle.code_luma4((pred.astype(int) + d).astype(np.uint8), pred, 27)
        assert cbp == want4, d
        assert (lv[:, 1:] == 0).all() and (lv[:, 0] == (0 if want4 == 0 else np.sign(d))).all()
    assert oracle.quant_tables()["mf8"][3, 0] == 9362
    for d, want8, lvl in ((3, 0, 0), (6, 0xF, 3)):
        cbp, lv, _ = oracle.code_luma8((pred.astype(int) + d).astype(np.uint8), pred, 27)
        assert cbp == want8, d
        assert (lv[:, 0] == lvl).all() and (lv[:, 1:] == 0).all()


def test_gate_on_flat_pictures(oracle):
    """Flat reference 100, flat source 100 + d, chroma equal: the skip MV
    (0, 0) predicts 100 everywhere, sad16 = 256|d| against the QP 27 bound
    2*16*v4[3][0]*2^4 = 7,168, and the 4x4 path codes nothing for |d| <= 2
    (see above): d = 2 is an early PSkip, d = 3 has cbp 0x0F and is not;
    with d = 28, sad = 7,168 is not below the bound and code_inter never
    runs (cbp reported as 0xFF)."""
    mbw = mbh = 2
    ref_y = np.full((16 * mbh + 128, 16 * mbw + 128), 100, np.uint8)
    ref_c = np.full((8 * mbh + 64, 8 * mbw + 64), 128, np.uint8)
    ref = oracle.Ref(ref_y)
    prm = oracle.oparams(mbw, mbh, 16, 1, 1, oracle.load().o_vlim(0, mbh))
    su = np.full((8, 8), 128, np.uint8)
    for d, skip, coded, cbp in ((2, 1, 1, 0), (3, 0, 1, 0x0F), (28, 0, 0, 0xFF), (0, 1, 1, 0)):
        sy = np.full((16, 16), 100 + d, np.uint8)
        r = oracle.skip_gate(prm, ref, ref_c, ref_c, 0, 0, (0, 0), sy, su, su, 27, 0)
        assert r["mv_ok"] == 1 and r["thresh"] == 7168 and r["sad"] == 256 * d
        assert (r["skip"], r["coded"], r["cbp"]) == (skip, coded, cbp), d
    # an illegal skip MV (x0 = -45) is not tested at all
    r = oracle.skip_gate(prm, ref, ref_c, ref_c, 0, 0, (-180, 0), np.full((16, 16), 100, np.uint8), su, su, 27, 0)
    assert (r["mv_ok"], r["coded"], r["skip"]) == (0, 0, 0)


def test_chroma_dc_alone(oracle):
    """A constant chroma residual d: every 4x4 block's AC is zero, the DC
    coefficients 16d form a 2x2 Hadamard of [64d, 0, 0, 0]; at QP 27 (QPc
    27): qbits 20, MF 9362, f = 2^20*43/256 = 176,128, level = (64|d|*9362
    + 176128) >> 20: d = 1 -> 0 (cbp 0), d = 2 -> 1 (chroma DC only: cbp 1)."""
    z = np.full(64, 128, np.uint8)
    assert oracle.code_chroma(z + 1, z, z, z, 27) == 0
    assert oracle.code_chroma(z + 2, z, z, z, 27) == 1
    assert oracle.code_chroma(z, z - 2, z, z, 27) == 1


def test_adaptive_transform_choice_by_hand(oracle):
    """code_inter's adaptive choice j8 <= j4, j = ssd(src, rec) + lam * bits,
    on residuals whose levels, reconstructions and rates follow by hand
    (lam = 0.85 * 2^((qp-12)/3); a level v costs 3 + 2 log2|v|, a zero 0.4,
    a coded 4x4 block 1 more; the 8x8 transform of a flat block is 4096 d at
    DC, of the +-1 step s = (1,1,1,1,-1,-1,-1,-1) the odd rows 62, -22, 14, -10):
    * QP 12: the top-left 8x8 holds +6 in its first and last 4x4 block, 0 in
      the other two.  4x4: DC level (96*13107 + 22016) >> 17 = 9, dequantised
      9*40 = 360 -> (360+32) >> 6 = 6: exact, j4 = 0.85 * (2*(1 + 9.34 + 15*0.4)
      + 2*(1 + 16*0.4)) = 40.4.  8x8: the block is 3*(1 + s x s), so DC
      2048*6 -> level 9, and the odd-odd coefficients 3*o[k]*o[l] with MF
      round(2^36 / (578^2*18)) = 11428, f = 2^24*43/256: (1,1) 11532 -> 8,
      (1,3) 4092 -> 2 (twice), (3,3), (1,5), (1,7) -> 1 (five), (3,5) 924 -> 0;
      bits 9.34 + 9 + 2*5 + 5*3 + 55*0.4 = 65.34, j8 >= 0.85 * 65.34 = 55.5 >
      j4: 4x4 wins, cbp 0x1 without the 8x8 flag; 8x8 alone keeps it.
    * QP 27, d = 6 everywhere: 4x4 level 1 (score 12, kept), rec pred + 4
      (14 << 4 = 224 -> 4), ssd 1024, j4 = 1024 + 27.2 * 160 = 5376.  8x8
      level 3 (score 64, kept), dequantised (3*448 + 2) >> 2 = 336 -> 5, ssd
      256, bits 4 * (3 + 2 log2 3 + 63 * 0.4), j8 = 3669.  8x8 wins: cbp 0xF
      with the 8x8 flag.
    * QP 27, d = 3: 4x4 level 1 per block (kept, cbp 0xF) but 8x8 level 1
      (score 3 < 4, decimated): j8 = ssd 2304 + 0 < j4 = 256 + 4352, so the
      adaptive choice codes nothing: an early PSkip that 4x4 alone misses.
    * QP 27, d = 2: both transforms code nothing, j8 = j4 (the tie goes to
      8x8, whose flag is only reported with a coded luma block): cbp 0."""
    T8, AD = oracle.SKIP_T8X8, oracle.SKIP_ADAPTIVE_T8
    pred = np.full((16, 16), 100, np.uint8)
    c = np.full((8, 8), 128, np.uint8)

    def run(res, qp, flags):
        return oracle.code_inter((pred.astype(int) + res).astype(np.uint8), c, c, pred, c, c, qp, flags)

    lam12, lam27 = oracle.load().o_lambda_ssd(12), oracle.load().o_lambda_ssd(27)
    assert abs(lam12 - 0.85) < 1e-12 and abs(lam27 - 27.2) < 1e-12
    assert oracle.quant_tables()["mf8"][0, 9] == 11428
    diag = np.zeros((16, 16), int)
    diag[0:4, 0:4] = diag[4:8, 4:8] = 6
    j4 = lam12 * (2 * (1 + 3 + 2 * np.log2(9) + 15 * 0.4) + 2 * (1 + 16 * 0.4))
    j8_min = lam12 * (3 + 2 * np.log2(9) + 9 + 2 * 5 + 5 * 3 + 55 * 0.4)
    assert j4 == pytest.approx(40.36, abs=0.01) and j8_min == pytest.approx(55.54, abs=0.01)
    assert run(diag, 12, T8 | AD) == (0x01, 0)
    assert run(diag, 12, T8) == (0x01, 1) and run(diag, 12, 0) == (0x01, 0)
    _, lv8, _ = oracle.code_luma8((pred.astype(int) + diag).astype(np.uint8), pred, 12)
    assert sorted(np.abs(lv8[0][lv8[0] != 0]).tolist()) == [1, 1, 1, 1, 1, 2, 2, 8, 9]
    j8 = 256 + lam27 * 4 * (3 + 2 * np.log2(3) + 63 * 0.4)
    assert abs(j8 - 3669.0) < 1.0 and 1024 + lam27 * 160 == pytest.approx(5376)
    assert run(6, 27, T8 | AD) == (0x0F, 1)
    assert run(3, 27, T8 | AD) == (0, 0) and run(3, 27, 0) == (0x0F, 0)
    assert run(2, 27, T8 | AD) == (0, 0)
