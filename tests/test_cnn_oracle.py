"""Pins of the pruner's numpy oracle (oracle/cnn_oracle.py) and of the
oracle's pruned search, on the CPU: rounding against torch's bf16,
convolution against a brute-force loop, the hand-written backward pass
against central finite differences, the input builder against a per-pixel
loop, and the pruned search against the full search where they must agree."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import cnn_oracle as C


def _params(seed: int, scale: float = 1.0) -> dict:
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in C.LAYOUT:
        if name.startswith("w"):
            fan = shape[1] * shape[2] if len(shape) == 3 else shape[1]
            p[name] = rng.normal(0, scale * np.sqrt(2.0 / fan), shape)
        else:
            p[name] = rng.normal(0, 0.05, shape)
    return p


def test_bf16_rounding_matches_torch_and_ties_to_even():
    import torch
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 10, 5000), rng.normal(0, 1e-3, 1000), [0.0, -0.0, 1.0, -2.5]])
    want = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(C.bf16(x), want)
    # ties: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 -> even (1.0);
    # 1 + 3*2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6 -> even (1 + 2^-6)
    assert C.bf16(1 + 2.0 ** -8) == 1.0
    assert C.bf16(1 + 3 * 2.0 ** -8) == 1 + 2.0 ** -6
    # every network input value is exact in bf16
    v = (np.arange(256) - 128) / 64.0
    assert np.array_equal(C.bf16(v), v)


@pytest.mark.parametrize("stride", [1, 2])
def test_conv_matches_bruteforce(stride):
    rng = np.random.default_rng(2)
    x = rng.normal(size=(2, 7, 7, 3))
    w = rng.normal(size=(4, 9, 3))
    b = rng.normal(size=4)
    y = C.conv(x, w, b, stride)
    ho = (7 - 1) // stride + 1
    assert y.shape == (2, ho, ho, 4)
    ref = np.zeros_like(y)
    for n in range(2):
        for oy in range(ho):
            for ox in range(ho):
                for co in range(4):
                    s = b[co]
                    for ty in range(3):
                        for tx in range(3):
                            iy, ix = oy * stride + ty - 1, ox * stride + tx - 1
                            if 0 <= iy < 7 and 0 <= ix < 7:
                                s += np.dot(w[co, 3 * ty + tx], x[n, iy, ix])
                    ref[n, oy, ox, co] = s
    assert np.allclose(y, ref, atol=1e-12)


def test_layer_shapes_and_param_count():
    p = _params(3)
    assert C.pack(p).size == C.NPARAM == 38021
    x = np.random.default_rng(3).normal(size=(2, 24, 24, 2))
    a = C.forward(x, p, emulate=False)
    assert a["a1"].shape == (2, 12, 12, 32)
    assert a["a2"].shape == (2, 6, 6, 32)
    assert a["a3"].shape == (2, 6, 6, 32)
    assert a["a4"].shape == (2, 3, 3, 32)
    assert a["logits"].shape == (2, 293)


def test_gradient_matches_finite_differences():
    """Central differences in float64 (no rounding) at 80 sampled parameters
    spread over every tensor, plus the softmax identity of the bias rows."""
    rng = np.random.default_rng(4)
    p = _params(5)
    x = rng.uniform(-2, 2, size=(3, 24, 24, 2))
    mv = rng.integers(0, 289, 3)
    part = rng.integers(0, 4, 3)
    flat = C.pack(p)
    loss, _, _, g = C.loss_and_grad(x, p, mv, part, emulate=False)

    def f(v):
        return C.loss_and_grad(x, C.unpack(v), mv, part, emulate=False)[0]

    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in C.LAYOUT])
    picks = np.concatenate([rng.integers(offs[i], offs[i + 1], 8) for i in range(len(C.LAYOUT))])
    eps = 1e-6
    bad = []
    for i in picks:
        e = np.zeros_like(flat)
        e[i] = eps
        fd = (f(flat + e) - f(flat - e)) / (2 * eps)
        if abs(fd - g[i]) > 1e-6 + 1e-4 * abs(g[i]):
            bad.append((int(i), fd, g[i]))
    assert not bad, bad[:5]
    gb = C.unpack(g)["bf"]
    assert abs(gb[:289].sum()) < 1e-12 and abs(gb[289:].sum()) < 1e-12
    assert np.isfinite(loss)


def test_emulated_forward_rounds_activations():
    p = _params(6)
    x = np.random.default_rng(6).uniform(-2, 2, size=(2, 24, 24, 2))
    a = C.forward(x, p, emulate=True)
    for k in (1, 2, 3):
        assert np.array_equal(C.bf16(a[f"a{k}"]), a[f"a{k}"])
    b = C.forward(x, p, emulate=False)
    assert np.allclose(a["logits"], b["logits"], rtol=0.05, atol=0.05)


def test_inputs_match_per_pixel_loop():
    rng = np.random.default_rng(7)
    mbw, mbh = 5, 4
    stride, rows = 16 * mbw + 128, 16 * mbh + 128
    src = rng.integers(0, 256, (rows, stride), dtype=np.uint8)
    ref = rng.integers(0, 256, (rows, stride), dtype=np.uint8)
    mbs = [0, 7, 19]
    got = C.inputs(src, ref, mbw, mbh, mbs)
    for i, m in enumerate(mbs):
        my, mx = divmod(m, mbw)
        X, Y = 16 * mx + 64, 16 * my + 64
        for oy in range(24):
            for ox in range(24):
                r = [int(ref[Y - 16 + 2 * oy + a, X - 16 + 2 * ox + b]) for a in (0, 1) for b in (0, 1)]
                assert got[i, oy, ox, 0] == ((sum(r) + 2) // 4 - 128) / 64
                if 8 <= oy < 16 and 8 <= ox < 16:
                    s = [int(src[Y + 2 * (oy - 8) + a, X + 2 * (ox - 8) + b]) for a in (0, 1) for b in (0, 1)]
                    assert got[i, oy, ox, 1] == ((sum(s) + 2) // 4 - 128) / 64
                else:
                    assert got[i, oy, ox, 1] == 0.0


def test_mv_class_is_nearest_grid_point():
    import pcme.pruner as P
    m = np.stack(np.meshgrid(np.arange(-80, 81), np.arange(-80, 81)), -1).reshape(-1, 2)
    c = C.mv_class(m)
    assert np.array_equal(c, P.mv_class(m))
    back = C.class_mv(c)
    inside = np.all(np.abs(m) <= 64, axis=1)
    assert np.all(np.abs(back[inside] - m[inside]) <= 4)
    assert np.all(np.abs(back) <= 64)
    assert set(np.unique(c)) == set(range(289))


# ------------------------------------------------------- pruned search oracle

def _frame_pair(seed, w=96, h=64):
    import pcme
    spec = pcme.ClipSpec(w, h, 2, nrect=3, rect_lo=16, rect_hi=40, vmax=4.0, seed=seed)
    y, _, _ = pcme.synth_clip(spec)
    return y


def test_pruned_search_without_prune_flags_is_the_full_search(oracle):
    y = _frame_pair(0x5EED0100)
    mbw, mbh = 6, 4
    src, ref = oracle.pad_luma(y[1], mbw, mbh), oracle.Ref(oracle.pad_luma(y[0], mbw, mbh))
    full = oracle.search_frame(src, ref, mbw, mbh, 16, 1, True, 28)
    h = np.zeros(mbw * mbh, oracle.OHINT_DTYPE)
    h["cand"] = 40  # arbitrary, ignored without the prune flag
    h["ncand"] = 8
    got = oracle.search_frame(src, ref, mbw, mbh, 16, 1, True, 28, hints=h)
    assert np.array_equal(got, full)


def test_pruned_search_invariants(oracle):
    """With the prune flag set: only P16x16 and the hinted modes are searched
    (the others read as not searched), the partitions only when cost16 >
    1024, and the chosen cost is the chosen mode's total plus its penalty."""
    y = _frame_pair(0x5EED0101)
    mbw, mbh = 6, 4
    src, ref = oracle.pad_luma(y[1], mbw, mbh), oracle.Ref(oracle.pad_luma(y[0], mbw, mbh))
    rng = np.random.default_rng(9)
    h = np.zeros(mbw * mbh, oracle.OHINT_DTYPE)
    h["prune"] = 1
    h["ncand"] = 8
    h["part"] = rng.integers(0, 4, mbw * mbh)
    h["cand"] = (rng.integers(-8, 9, (mbw * mbh, 8, 2)) * 4).astype(np.int16)
    got = oracle.search_frame(src, ref, mbw, mbh, 16, 1, True, 28, hints=h).reshape(-1)
    for i in range(mbw * mbh):
        hp = int(h["part"][i])
        searched = {0} | ({1, 2} if hp in (1, 2) else ({3} if hp == 3 else set()))
        for m in range(4):
            if m not in searched:
                assert got["cost"][i][m] == 0xFFFFFFFF
        if got["cost"][i][0] <= 1024:
            assert list(got["cost"][i][1:]) == [0xFFFFFFFF] * 3
        assert got["part"][i] in searched
    # the chosen cost is the searched mode's total plus its penalty
    lam = oracle.load().o_lambda(28)
    pen = [lam >> 8, (lam * 3) >> 8, (lam * 3) >> 8, (lam * 10) >> 8]
    for i in range(mbw * mbh):
        pt = int(got["part"][i])
        assert int(got["best_cost"][i]) == int(got["cost"][i][pt]) + pen[pt]


# ------------------------------------------------- what the network buys

def _gain_fields(oracle, w, h, nf, seed, qp=27):
    """Plain, CNN-pruned and network-free pruned fields of a held-out clip
    (the oracle's exact searches; hints from the oracle network on the
    committed weights), with the oracle's work counters."""
    import pcme
    import pcme.pruner as P
    spec = pcme.ClipSpec(w, h, nf, nrect=8, rect_lo=24, rect_hi=120, vmax=6.0, pan_max=1.0, seed=seed)
    y, _, _ = pcme.synth_clip(spec)
    mbw, mbh = pcme.mb_dims(w, h)
    pads = [oracle.pad_luma(f, mbw, mbh) for f in y]
    prm = C.unpack(P.load_weights(P.WEIGHTS, qp).astype(np.float64))
    out = {k: [] for k in ("plain", "cnn", "none")}
    ctr = {k: oracle.ocounters() for k in out}
    for t in range(1, nf):
        z = C.forward(C.inputs(pads[t], pads[t - 1], mbw, mbh, range(mbw * mbh)), prm, emulate=True)["logits"]
        zm, zp = z[:, :C.MVC], z[:, C.MVC:]
        top = np.argsort(-zm, axis=1, kind="stable")[:, :8]
        pp = np.exp(zp - zp.max(1, keepdims=True))
        pp /= pp.sum(1, keepdims=True)
        hc = np.zeros(mbw * mbh, oracle.OHINT_DTYPE)
        hc["cand"] = C.class_mv(top).astype(np.int16)
        hc["part"] = np.argmax(zp, 1)
        hc["ncand"] = 5
        pm = np.exp(zm - zm.max(1, keepdims=True))
        pm /= pm.sum(1, keepdims=True)
        mass = np.take_along_axis(pm, top[:, :5], 1).sum(1)
        hc["prune"] = (pp.max(1) >= 0.6) & (mass >= 0.6)
        h0 = np.zeros(mbw * mbh, oracle.OHINT_DTYPE)  # prune every MB, no candidates, P16x16 only
        h0["prune"] = 1
        ref = oracle.Ref(pads[t - 1])
        for k, hh in (("plain", None), ("cnn", hc), ("none", h0)):
            out[k].append(oracle.search_frame(pads[t], ref, mbw, mbh, 32, 1, True, qp, counters=ctr[k], hints=hh))
    work = {k: c.sad_px + 4 * c.satd_px for k, c in ctr.items()}
    return {k: np.stack(v) for k, v in out.items()}, work


def test_network_gain_over_the_search_without_it(oracle):
    """On a held-out 640x368 clip the network's hints (5 candidates; fallback
    where either head's confidence is below 0.6) keep the J-cost within 1% of the
    plain search at well under half its work, and lose several times less
    than the same pruned search given no hints at all (the method without the
    network: predictor and search_part candidates, P16x16 only)."""
    import pcme.pruner as P
    f, work = _gain_fields(oracle, 640, 368, 4, 0x5EED9101)
    j_cnn = P.judge(f["plain"], f["cnn"])
    j_none = P.judge(f["plain"], f["none"])
    assert work["cnn"] < 0.5 * work["plain"], work
    assert j_cnn["jcost_increase"] < 0.01, j_cnn
    assert j_cnn["jcost_increase"] * 4 < j_none["jcost_increase"], (j_cnn, j_none)
    assert j_cnn["agreement"] > j_none["agreement"] and j_cnn["part_agreement"] > j_none["part_agreement"]
