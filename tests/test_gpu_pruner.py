"""GPU tests of the learned search pruner (include/pcme_net.h): its HIP
training and inference kernels against the numpy oracle, its hints against
its own forward pass, and the pruned search against the C oracle's
restatement of it, bit for bit.

Each case prints what it runs before the GPU step so a failing run names
its case."""
from __future__ import annotations

import numpy as np
import pytest

from _util import assert_same_field
from oracle import cnn_oracle as C

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _dclip(pcme, w, h, n, seed, **kw):
    spec = pcme.ClipSpec(w, h, n, seed=seed, **kw)
    y, _, _ = pcme.synth_clip(spec)
    mbw, mbh = pcme.mb_dims(w, h)
    dc = pcme.DeviceClip(y, mbw, mbh)
    dc.pad()
    dc.hpel(0, n - 1)
    return y, dc


def _bf16_bits(v: np.ndarray) -> np.ndarray:
    return (np.asarray(v, np.float32).view(np.uint32) >> 16).astype(np.uint16)


def _inputs_all(P, dc, t):
    import torch
    nmb = dc.mbw * dc.mbh
    x = torch.empty((nmb, 1152), dtype=torch.int16, device="cuda")
    mbs = torch.arange(nmb, dtype=torch.int32, device="cuda")
    P.inputs(dc, t, t - 1, mbs, x)
    torch.cuda.synchronize()
    return x


def test_inputs_kernel_bit_exact(gpu):
    import pcme.pruner as P
    say("net inputs 100x70")
    _, dc = _dclip(gpu, 100, 70, 2, 0x5EED0200, nrect=3, rect_lo=16, rect_hi=40)
    x = _inputs_all(P, dc, 1).cpu().numpy().view(np.uint16).reshape(-1, 24, 24, 2)
    planes = dc.planes[0].cpu().numpy()
    want = C.inputs(planes[1], planes[0], dc.mbw, dc.mbh, range(dc.mbw * dc.mbh))
    np.testing.assert_array_equal(x, _bf16_bits(want))


def _net(P, seed=11):
    return P.Net(seed=seed, max_batch=256)


def test_forward_matches_oracle(gpu):
    import torch
    import pcme.pruner as P
    say("net forward 64 MBs")
    _, dc = _dclip(gpu, 128, 64, 2, 0x5EED0201, nrect=3, rect_lo=16, rect_hi=40)
    net = _net(P)
    x = _inputs_all(P, dc, 1)[:32 * ((dc.mbw * dc.mbh) // 32)]
    n = x.shape[0]
    logits = torch.empty((n, 293), dtype=torch.float32, device="cuda")
    net.forward(x, n, logits)
    torch.cuda.synchronize()
    got = logits.cpu().numpy().astype(np.float64)
    xs = x.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    xin = xs.view(np.float32).astype(np.float64).reshape(n, 24, 24, 2)
    want = C.forward(xin, C.unpack(net.get_params()), emulate=True)["logits"]
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    # bf16 activations: fp32 vs fp64 accumulation may flip a rounding
    assert err < 1e-2, err
    assert np.mean(np.argmax(got[:, :289], 1) == np.argmax(want[:, :289], 1)) > 0.9


def test_gradient_matches_oracle(gpu):
    """The HIP backward pass (bf16 MFMA products, fp32 accumulation) against
    the oracle's float64 backpropagation of the same bf16-rounded forward."""
    import torch
    import pcme.pruner as P
    say("net gradient 64 MBs")
    _, dc = _dclip(gpu, 128, 64, 2, 0x5EED0202, nrect=3, rect_lo=16, rect_hi=40)
    net = _net(P, seed=12)
    x = _inputs_all(P, dc, 1)
    n = 32 * (x.shape[0] // 32)
    rng = np.random.default_rng(3)
    mv = torch.as_tensor(rng.integers(0, 289, x.shape[0]).astype(np.int32), device="cuda")
    part = torch.as_tensor(rng.integers(0, 4, x.shape[0]).astype(np.int32), device="cuda")
    idx = torch.as_tensor(rng.permutation(x.shape[0])[:n].astype(np.int32), device="cuda")
    loss = torch.zeros(3, dtype=torch.float32, device="cuda")
    grad = torch.zeros(P.NPARAM, dtype=torch.float32, device="cuda")
    net.grad(x, mv, part, idx, n, loss, grad)
    torch.cuda.synchronize()
    g = grad.cpu().numpy().astype(np.float64)
    ii = idx.cpu().numpy()
    xs = x.cpu().numpy().view(np.uint16)[ii].astype(np.uint32) << 16
    xin = xs.view(np.float32).astype(np.float64).reshape(n, 24, 24, 2)
    lt, lm, lp, want = C.loss_and_grad(xin, C.unpack(net.get_params()), mv.cpu().numpy()[ii],
                                       part.cpu().numpy()[ii], emulate=True)
    lv = loss.cpu().numpy()
    assert abs(lv[0] - lt) < 1e-3 * abs(lt) and abs(lv[1] - lm) < 1e-3 * abs(lm)
    gu, wu = C.unpack(g), C.unpack(want)
    for name, _ in C.LAYOUT:
        a, b = gu[name].ravel(), wu[name].ravel()
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        assert rel < 3e-2 and cos > 0.999, (name, rel, cos)


def test_training_reduces_loss_and_is_deterministic(gpu):
    import torch
    import pcme.pruner as P
    say("net training 320x240 x 5 frames, 2 nets")
    _, dc = _dclip(gpu, 320, 240, 6, 0x5EED0203, nrect=4, rect_lo=24, rect_hi=80)
    data, _ = P.label_clip(dc, 16, 28)
    hs, ps = [], []
    for _ in range(2):
        net = P.Net(seed=21, max_batch=256)
        hs.append(P.train(net, data, epochs=8, batch=256, lr=3e-3, seed=5))
        ps.append(net.get_params())
        net.close()
    first, last = hs[0][0][1], hs[0][-1][1]
    assert last < 0.7 * first, hs[0]
    np.testing.assert_array_equal(ps[0], ps[1])


def _topk(z, k):
    order = np.lexsort((np.arange(z.size), -z))  # logit desc, class asc
    return order[:k]


def test_hints_match_forward_logits(gpu):
    """The fused inference kernel's logits equal pcme_net_forward's bit for
    bit, so its top-k classes and partition argmax are the same."""
    import torch
    import pcme
    import pcme.pruner as P
    say("net hints vs forward 192x96")
    _, dc = _dclip(gpu, 192, 96, 2, 0x5EED0204, nrect=3, rect_lo=16, rect_hi=48)
    net = _net(P, seed=13)
    nmb = dc.mbw * dc.mbh  # 72
    x = _inputs_all(P, dc, 1)
    n = 32 * ((nmb + 31) // 32)
    xp = torch.cat([x, x[: n - nmb]]) if n > nmb else x
    logits = torch.empty((n, 293), dtype=torch.float32, device="cuda")
    net.forward(xp, n, logits)
    hints = torch.zeros(nmb * pcme.HINT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    out = torch.empty(nmb * pcme.MB_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    job = pcme.frame_job(dc, 1, 0, out.data_ptr(), 16, 1, True, 28, hint_ptr=hints.data_ptr())
    net.hints([job], ncand=8, thr_part=0.5, thr_mv=0.3)
    torch.cuda.synchronize()
    z = logits.cpu().numpy()[:nmb].astype(np.float64)
    h = hints.cpu().numpy().view(pcme.HINT_DTYPE)
    for i in range(nmb):
        top = _topk(z[i, :289], 8)
        np.testing.assert_array_equal(h["cand"][i], C.class_mv(top), err_msg=f"mb {i}")
        pz = z[i, 289:]
        assert h["part"][i] == int(np.argmax(pz))
        pp = np.exp(pz - pz.max())
        cp = pp.max() / pp.sum()
        e = np.exp(z[i, :289] - z[i, :289].max())
        cm = e[top].sum() / e.sum()
        assert abs(h["conf_part"][i] / 65535 - cp) < 2e-4
        assert abs(h["conf_mv"][i] / 65535 - cm) < 2e-4
        if abs(cp - 0.5) > 1e-4 and abs(cm - 0.3) > 1e-4:
            assert h["prune"][i] == int(cp >= 0.5 and cm >= 0.3)
        assert h["ncand"][i] == 8


@pytest.mark.parametrize("ncand,thr", [(8, 0.0), (3, 0.0), (0, 0.0), (8, 0.45), (5, 0.3)])
def test_pruned_search_matches_oracle(gpu, oracle, ncand, thr):
    """GPU pruned field == the oracle's pruned search on the same hints."""
    import pcme.pruner as P
    say("pruned search 320x240 x 3 frames", ncand, thr)
    y, dc = _dclip(gpu, 320, 240, 4, 0x5EED0205, nrect=4, rect_lo=24, rect_hi=80)
    net = _net(P, seed=14)
    field, hints = P.search_clip(dc, 16, 28, net=net, ncand=ncand, thr_part=thr, thr_mv=thr, prepared=True)
    mbw, mbh = dc.mbw, dc.mbh
    for t in range(1, 4):
        src = oracle.pad_luma(y[t], mbw, mbh)
        ref = oracle.Ref(oracle.pad_luma(y[t - 1], mbw, mbh))
        want = oracle.search_frame(src, ref, mbw, mbh, 16, 1, True, 28, hints=hints[t - 1].reshape(-1))
        assert_same_field(field[t - 1], want, f"frame {t}")
    if thr == 0.0:
        assert np.all(hints["prune"] == 1)


def test_pruned_search_without_confidence_is_the_full_search(gpu):
    import pcme.pruner as P
    say("pruned search, thresholds above 1")
    _, dc = _dclip(gpu, 320, 240, 3, 0x5EED0206, nrect=4, rect_lo=24, rect_hi=80)
    net = _net(P, seed=15)
    plain, _ = P.search_clip(dc, 16, 28, prepared=True)
    field, hints = P.search_clip(dc, 16, 28, net=net, thr_part=1.5, thr_mv=1.5, prepared=True)
    assert np.all(hints["prune"] == 0)
    np.testing.assert_array_equal(field, plain)


@pytest.mark.parametrize("thr_part,thr_mv", [(0.7, 0.0), (0.6, 0.6)])
def test_network_gain_over_the_search_without_it(gpu, thr_part, thr_mv):
    """The product path on a held-out 640x368 clip with the committed
    weights, under the default rule (partition head at 0.7) and under both
    heads at 0.6: the network's hints keep the J-cost within 1% of the plain
    search and lose several times less than the same pruned search given no
    hints (every MB pruned, no candidates, P16x16 only)."""
    import pcme.pruner as P
    say("pruner gain 640x368 x 3 P frames, committed weights", thr_part, thr_mv)
    _, dc = _dclip(gpu, 640, 368, 4, 0x5EED9101, nrect=8, rect_lo=24, rect_hi=120, vmax=6.0, pan_max=1.0)
    net = P.Net(seed=1, max_batch=64, weights=P.WEIGHTS, qp=27)
    plain, _ = P.search_clip(dc, 32, 27, prepared=True)
    cnn, hints = P.search_clip(dc, 32, 27, net=net, ncand=5, thr_part=thr_part, thr_mv=thr_mv, prepared=True)
    none = np.zeros(hints.shape, gpu.HINT_DTYPE)
    none["prune"] = 1
    bare, _ = P.search_clip(dc, 32, 27, prepared=True, hints_in=none)
    j_cnn, j_none = P.judge(plain, cnn), P.judge(plain, bare)
    assert np.mean(hints["prune"]) > 0.5
    assert j_cnn["jcost_increase"] < 0.01, j_cnn
    assert j_cnn["jcost_increase"] * 4 < j_none["jcost_increase"], (j_cnn, j_none)


def test_bd_rate_model_runs_on_product_fields(gpu):
    """pcme.eval_pruner end to end on a small held-out clip: four QPs, three
    fields each, finite (rate, PSNR) points rising with rate, and the plain
    field's BD-rate against itself zero."""
    from pcme import eval_pruner as E, rdmodel as R
    say("bd-rate model 320x240 x 2 P frames, 4 QPs")
    rep = E.evaluate(2, 0x5EED9300, log=lambda s: None, w=320, h=240)
    pts = rep["points"]
    for k in ("plain", "pruned", "no_network"):
        r = [p[0] for p in pts[k]]
        q = [p[1] for p in pts[k]]
        assert all(np.isfinite(r)) and all(np.isfinite(q))
        assert r == sorted(r, reverse=True) and q == sorted(q, reverse=True), (k, pts[k])
    assert abs(R.bd_rate([tuple(p) for p in pts["plain"]], [tuple(p) for p in pts["plain"]])) < 1e-9
    assert np.isfinite(rep["bd_rate_pruned_pct"]) and np.isfinite(rep["bd_rate_no_network_pct"])


def test_hinted_plan_configurations(gpu, oracle):
    """A plan with pruner hints runs the default configuration and (4, 16, 1)
    (the hinted kernel at 16 waves per CU), at its resident grid and a small
    one, bit for bit the oracle's pruned field on the same hints (pruned and
    fallback MBs mixed); autotune_ex keeps one of the two."""
    import torch
    import pcme.pruner as P
    say("hinted plan configurations 320x240 x 3 frames")
    y, dc = _dclip(gpu, 320, 240, 4, 0x5EED0207, nrect=4, rect_lo=24, rect_hi=80)
    net = P.Net(seed=1, max_batch=256, weights=P.WEIGHTS, qp=27)
    nb, nmb = dc.n - 1, dc.mbw * dc.mbh
    out = torch.zeros((nb, nmb * gpu.MB_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    hbuf = torch.zeros((nb, nmb * gpu.HINT_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    jobs = [gpu.frame_job(dc, t, t - 1, out[t - 1].data_ptr(), 16, 1, True, 27, hint_ptr=hbuf[t - 1].data_ptr())
            for t in range(1, dc.n)]
    net.hints(jobs, 5, P.thr_part_for(27), 0.0)
    net.close()
    hints = hbuf.cpu().numpy().view(gpu.HINT_DTYPE).reshape(nb, -1)
    assert 0.0 < float(np.mean(hints["prune"])) < 1.0
    want = []
    for t in range(1, dc.n):
        src = oracle.pad_luma(y[t], dc.mbw, dc.mbh)
        ref = oracle.Ref(oracle.pad_luma(y[t - 1], dc.mbw, dc.mbh))
        want.append(oracle.search_frame(src, ref, dc.mbw, dc.mbh, 16, 1, True, 27, hints=hints[t - 1]))
    plan = gpu.Plan(jobs)
    for wpg, wcu in ((1, 12), (4, 16)):
        for waves in (0, 6):
            say("configuration", wpg, wcu, "waves", waves)
            t = plan.set_config(waves, wpg, wcu, 1)
            assert (t.waves_per_group, t.waves_per_cu) == (wpg, wcu)
            out.zero_()
            plan.launch()
            plan.check()
            got = out.cpu().numpy().view(gpu.MB_DTYPE).reshape(nb, dc.mbh, dc.mbw)
            for k in range(nb):
                assert_same_field(got[k], want[k], f"config {wpg}/{wcu} waves {waves} frame {k + 1}")
    with pytest.raises(gpu.PcmeError):
        plan.set_config(0, 2, 12, 1)
    best = plan.autotune_ex(reps=1)
    assert (best.waves_per_group, best.waves_per_cu) in ((1, 12), (4, 16))
