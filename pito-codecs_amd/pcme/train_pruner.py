"""Train the search pruner on the GPU and write pruner_weights.npz.

    python -m pcme.train_pruner [--clips 8] [--frames 31] [--epochs 3]

Training data: seeded synthetic 1080p clips (12 rectangles of 64-400 px at
up to 8 px/frame, pan up to 1 px/frame -- the generator of the 1080p search
workload, with seeds of their own), every P-frame macroblock labelled by
the plain GPU search (mv16's motion class and the chosen partition) at each
QP of {22, 27, 32, 37}; one network per QP, since the mode decision moves
with lambda.  The networks start from a seeded He initialisation and are
trained by libpcme's kernels (pcme_net_train_step).  Held-out clips (other
seeds) then measure what the pruned search keeps of the plain one.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import ClipSpec, DeviceClip, mb_dims, synth_clip
from . import pruner as P

QPS = (22, 27, 32, 37)
TRAIN_SEED = 0x5EED7000
EVAL_SEED = 0x5EED9000


def spec_1080p(seed: int, frames: int) -> ClipSpec:
    return ClipSpec(1920, 1080, frames, nrect=12, rect_lo=64, rect_hi=400, vmax=8.0, pan_max=1.0, seed=seed)


def log(*a):
    print(*a, flush=True)


def build_sets(clips: int, frames: int, me_range: int):
    """Inputs (shared) and per-QP labels of `clips` training clips."""
    import torch
    xs = []
    labels = {qp: ([], []) for qp in QPS}
    for i in range(clips):
        t0 = time.time()
        y, _, _ = synth_clip(spec_1080p(TRAIN_SEED + i, frames))
        mbw, mbh = mb_dims(1920, 1080)
        clip = DeviceClip(y, mbw, mbh)
        for k, qp in enumerate(QPS):
            ts, _ = P.label_clip(clip, me_range, qp)
            if k == 0:
                xs.append(ts.x)
            labels[qp][0].append(ts.mv)
            labels[qp][1].append(ts.part)
        del clip
        torch.cuda.synchronize()
        log(f"clip {i}: {frames - 1} P-frames labelled in {time.time() - t0:.1f} s")
    x = torch.cat(xs)
    return {qp: P.TrainSet(x, torch.cat(labels[qp][0]), torch.cat(labels[qp][1])) for qp in QPS}


def evaluate(params_by_qp: dict, frames: int, me_range: int, ncand: int, thrs, seed: int = EVAL_SEED) -> dict:
    """Plain vs pruned on a held-out clip, per QP and confidence threshold."""
    out = {}
    y, _, _ = synth_clip(spec_1080p(seed, frames))
    mbw, mbh = mb_dims(1920, 1080)
    for qp in QPS:
        net = P.Net(seed=1, max_batch=64, weights=params_by_qp[qp])
        clip = DeviceClip(y, mbw, mbh)
        plain, _ = P.search_clip(clip, me_range, qp)
        res = {}
        for thr in thrs:
            pruned, hints = P.search_clip(clip, me_range, qp, net=net, ncand=ncand, thr_part=thr, thr_mv=thr,
                                          prepared=True)
            j = P.judge(plain, pruned)
            j.update(P.hint_stats(hints))
            res[str(thr)] = j
        net.close()
        del clip
        out[qp] = res
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--clips", type=int, default=8)
    ap.add_argument("--frames", type=int, default=31)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--me-range", type=int, default=32)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--eval-frames", type=int, default=16)
    ap.add_argument("--ncand", type=int, default=8)
    ap.add_argument("--thr", type=float, nargs="+", default=[0.5, 0.6, 0.7, 0.8, 0.9])
    ap.add_argument("--out", default=P.WEIGHTS)
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        log("training needs the GPU")
        return 2
    t0 = time.time()
    sets = build_sets(a.clips, a.frames, a.me_range)
    log(f"training set: {sets[QPS[0]].n} MBs per QP ({time.time() - t0:.1f} s)")
    params, hist = {}, {}
    for qp in QPS:
        t1 = time.time()
        net = P.Net(seed=a.seed, max_batch=a.batch)
        hist[qp] = P.train(net, sets[qp], epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed + qp,
                           log=lambda s, q=qp: log(f"qp {q}: {s}"))
        torch.cuda.synchronize()
        params[qp] = net.get_params()
        net.close()
        log(f"qp {qp}: trained in {time.time() - t1:.1f} s, final loss {hist[qp][-1][1]:.4f}")
    del sets
    torch.cuda.empty_cache()
    meta = {"qps": np.array(QPS), "train_mbs_per_qp": a.clips * (a.frames - 1) * 8160, "epochs": a.epochs,
            "batch": a.batch, "lr": a.lr, "seed": a.seed, "train_seed": TRAIN_SEED}
    P.save_weights_qps(params, a.out, **meta)
    log(f"wrote {a.out}")
    ev = evaluate(params, a.eval_frames, a.me_range, a.ncand, a.thr)
    rep = {"train": {str(q): [list(h) for h in hist[q]] for q in QPS}, "eval": {str(q): ev[q] for q in QPS},
           "ncand": a.ncand, "thr": a.thr, "eval_frames": a.eval_frames}
    for qp in QPS:
        log(f"qp {qp}: {json.dumps(ev[qp])}")
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    log(f"total {time.time() - t0:.1f} s")
    return 0


if __name__ == "__main__":
    sys.exit(main())
