"""Host side of the learned search pruner (include/pcme_net.h).

The network, its training and its inference all run in libpcme's HIP
kernels; this module holds the ctypes binding, the training loop (batching,
learning-rate schedule, labels taken from the plain GPU search) and the
judge that compares a pruned field with the plain one.  No counterpart in
the reference.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import (MB_DTYPE, HINT_DTYPE, ClipSpec, DeviceClip, FrameJob, PcmeError, PCME_EINVAL, _check,
               _stream_ptr, _torch, frame_job, load, mb_dims, synth_clip)

NPARAM = 38021
NCLS = 293
MVC = 289
MV_SIDE = 17
INPUT = 1152
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pruner_weights.npz")

# Partition-head confidence below which an MB falls back to the full search,
# per QP (one network per QP).  At the high QPs a pruned MB's lost mode or
# motion costs more of the picture's rate, so those networks prune only more
# confident MBs.
THR_PART = {22: 0.7, 27: 0.7, 32: 0.8, 37: 0.9}


def parse_thr(text: str | None):
    """A fallback rule from the command line: None or "" (THR_PART), one
    threshold for every QP ("0.7"), or per-QP thresholds
    ("22:0.7,27:0.7,32:0.8,37:0.9")."""
    if text is None or text == "":
        return None
    if ":" not in text:
        return float(text)
    return {int(k): float(v) for k, v in (kv.split(":") for kv in text.split(","))}


def thr_part_for(qp: int, thr=None) -> float:
    """The fallback threshold of QP `qp` under rule `thr` (see parse_thr):
    THR_PART's entry when `thr` is None, 0.7 for a QP the table lacks."""
    if thr is None:
        thr = THR_PART
    if isinstance(thr, dict):
        return float(thr.get(qp, 0.7))
    return float(thr)

_sig_done = False


def _lib():
    global _sig_done
    lib = load()
    if not _sig_done:
        c = ctypes
        vp = c.c_void_p
        sig = {
            "pcme_net_create": (c.c_int, [c.c_uint64, c.c_int32, c.POINTER(vp)]),
            "pcme_net_destroy": (None, [vp]),
            "pcme_net_get_params": (c.c_int, [vp, vp]),
            "pcme_net_set_params": (c.c_int, [vp, vp]),
            "pcme_net_inputs": (c.c_int, [vp, vp, c.c_int32, c.c_int32, vp, c.c_int32, vp, vp]),
            "pcme_net_forward": (c.c_int, [vp, vp, c.c_int32, vp, vp]),
            "pcme_net_grad": (c.c_int, [vp, vp, vp, vp, vp, c.c_int32, vp, vp, vp]),
            "pcme_net_train_step": (c.c_int, [vp, vp, vp, vp, vp, c.c_int32, c.c_float, vp, vp]),
            "pcme_net_hints": (c.c_int, [vp, c.POINTER(FrameJob), c.c_int32, c.c_int32, c.c_float, c.c_float,
                                         vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _sig_done = True
    return lib


def mv_class(mv16: np.ndarray) -> np.ndarray:
    """Motion class of quarter-pel MVs [..., 2] (nearest even integer-pel
    offset within +-16, clamped): 17 * (cy + 8) + (cx + 8)."""
    c = np.clip((np.asarray(mv16, np.int64) + 4) >> 3, -8, 8) + 8
    return (c[..., 1] * MV_SIDE + c[..., 0]).astype(np.int32)


class Net:
    """A pruner network on the current device (pcme_net_*)."""

    def __init__(self, seed: int = 1, max_batch: int = 1024, weights: str | np.ndarray | None = None,
                 qp: int = 27):
        lib = _lib()
        h = ctypes.c_void_p()
        _check(lib.pcme_net_create(int(seed), int(max_batch), ctypes.byref(h)))
        self.handle = h
        self.max_batch = max_batch
        if weights is not None:
            self.set_params(load_weights(weights, qp) if isinstance(weights, str) else weights)

    def close(self) -> None:
        if getattr(self, "handle", None):
            _lib().pcme_net_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        out = np.empty(NPARAM, np.float32)
        _check(_lib().pcme_net_get_params(self.handle, out.ctypes.data))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        if p.size != NPARAM:
            raise PcmeError(PCME_EINVAL, f"expected {NPARAM} parameters, got {p.size}")
        _check(_lib().pcme_net_set_params(self.handle, p.ctypes.data))

    def forward(self, x, n: int, logits, stream=None) -> None:
        _check(_lib().pcme_net_forward(self.handle, x.data_ptr(), int(n), logits.data_ptr(), _stream_ptr(stream)))

    def grad(self, x, mv, part, idx, n: int, loss, grad=None, stream=None) -> None:
        _check(_lib().pcme_net_grad(self.handle, x.data_ptr(), mv.data_ptr(), part.data_ptr(), idx.data_ptr(),
                                    int(n), loss.data_ptr(), grad.data_ptr() if grad is not None else None,
                                    _stream_ptr(stream)))

    def train_step(self, x, mv, part, idx, n: int, lr: float, loss, stream=None) -> None:
        _check(_lib().pcme_net_train_step(self.handle, x.data_ptr(), mv.data_ptr(), part.data_ptr(),
                                          idx.data_ptr(), int(n), float(lr), loss.data_ptr(), _stream_ptr(stream)))

    def hints(self, jobs: list[FrameJob], ncand: int = 8, thr_part: float = 0.7, thr_mv: float = 0.0,
              stream=None) -> None:
        arr = (FrameJob * len(jobs))(*jobs)
        _check(_lib().pcme_net_hints(self.handle, arr, len(jobs), int(ncand), float(thr_part), float(thr_mv),
                                     _stream_ptr(stream)))


def inputs(clip: DeviceClip, cur: int, ref: int, mbs, out, stream=None) -> None:
    """pcme_net_inputs for macroblocks `mbs` (device int32 tensor) of frame
    `cur` against reference frame `ref` into `out` (device int16/bf16)."""
    _check(_lib().pcme_net_inputs(clip.plane_ptr(0, cur), clip.plane_ptr(0, ref), clip.mbw, clip.mbh,
                                  mbs.data_ptr(), int(mbs.numel()), out.data_ptr(), _stream_ptr(stream)))


def load_weights(path: str = WEIGHTS, qp: int = 27) -> np.ndarray:
    """Parameters of the network trained for the QP nearest to `qp`."""
    with np.load(path, allow_pickle=False) as z:
        if "params" in z.files:
            return np.asarray(z["params"], np.float32)
        qps = [int(k[len("params_qp"):]) for k in z.files if k.startswith("params_qp")]
        best = min(qps, key=lambda q: (abs(q - qp), q))
        return np.asarray(z[f"params_qp{best}"], np.float32)


def save_weights(p: np.ndarray, path: str = WEIGHTS, **meta) -> None:
    np.savez_compressed(path, params=np.asarray(p, np.float32),
                        **{k: np.asarray(v) for k, v in meta.items()})


def save_weights_qps(params: dict, path: str = WEIGHTS, **meta) -> None:
    """One parameter vector per QP (keys params_qp<QP>)."""
    np.savez_compressed(path, **{f"params_qp{int(q)}": np.asarray(v, np.float32) for q, v in params.items()},
                        **{k: np.asarray(v) for k, v in meta.items()})


# ------------------------------------------------------------- training

@dataclass
class TrainSet:
    """Network inputs and plain-search labels, resident on the device."""
    x: object       # torch int16 [N, 1152]
    mv: object      # torch int32 [N]
    part: object    # torch int32 [N]

    @property
    def n(self) -> int:
        return int(self.mv.numel())


def label_clip(clip: DeviceClip, me_range: int, qp: int, stream=None):
    """Plain GPU search of every P frame of a device clip, then the network
    inputs of every MB and its labels (mv16's class, the chosen partition)."""
    from . import search_frames
    torch = clip.torch
    dev = clip.raw.device
    clip.pad(stream)
    clip.hpel(0, clip.n - 1, stream)
    nb = clip.n - 1
    nmb = clip.mbw * clip.mbh
    out = torch.empty((nb, nmb * MB_DTYPE.itemsize), dtype=torch.uint8, device=dev)
    jobs = [frame_job(clip, t, t - 1, out[t - 1].data_ptr(), me_range, 1, True, qp) for t in range(1, clip.n)]
    search_frames(jobs, stream, sync=True)
    field = out.cpu().numpy().view(MB_DTYPE).reshape(nb, nmb)
    x = torch.empty((nb * nmb, INPUT), dtype=torch.int16, device=dev)
    mbs = torch.arange(nmb, dtype=torch.int32, device=dev)
    for t in range(1, clip.n):
        inputs(clip, t, t - 1, mbs, x[(t - 1) * nmb:t * nmb], stream)
    mv = torch.as_tensor(mv_class(field["mv16"].reshape(-1, 2)), device=dev)
    part = torch.as_tensor(field["part"].reshape(-1).astype(np.int32), device=dev)
    return TrainSet(x, mv, part), field


def make_train_set(specs: list[tuple[ClipSpec, int, int]], stream=None) -> TrainSet:
    """Label clips [(spec, qp, me_range)] with the plain GPU search."""
    torch = _torch()
    xs, mvs, parts = [], [], []
    for spec, qp, rng in specs:
        y, _, _ = synth_clip(spec)
        mbw, mbh = mb_dims(spec.w, spec.h)
        clip = DeviceClip(y, mbw, mbh)
        ts, _ = label_clip(clip, rng, qp, stream)
        xs.append(ts.x)
        mvs.append(ts.mv)
        parts.append(ts.part)
        del clip
    return TrainSet(torch.cat(xs), torch.cat(mvs), torch.cat(parts))


def train(net: Net, data: TrainSet, epochs: int = 3, batch: int = 1024, lr: float = 2e-3, seed: int = 7,
          log=None, stream=None) -> list[tuple[int, float, float, float]]:
    """Adam over shuffled minibatches with a cosine learning-rate decay;
    returns the (step, loss, loss_mv, loss_part) history (every 100 steps)."""
    torch = _torch()
    dev = data.x.device
    g = torch.Generator(device="cpu").manual_seed(seed)
    steps_per_epoch = data.n // batch
    total = max(1, epochs * steps_per_epoch)
    loss = torch.zeros(3, dtype=torch.float32, device=dev)
    hist = []
    step = 0
    for ep in range(epochs):
        perm = torch.randperm(data.n, generator=g).to(torch.int32).to(dev)
        for b in range(steps_per_epoch):
            idx = perm[b * batch:(b + 1) * batch]
            rate = lr * 0.5 * (1.0 + np.cos(np.pi * step / total))
            net.train_step(data.x, data.mv, data.part, idx, batch, max(rate, lr * 1e-3), loss, stream)
            if step % 100 == 0 or step == total - 1:
                v = loss.cpu().numpy()
                hist.append((step, float(v[0]), float(v[1]), float(v[2])))
                if log:
                    log(f"step {step}/{total} loss {v[0]:.4f} (mv {v[1]:.4f} part {v[2]:.4f})")
            step += 1
    return hist


# ------------------------------------------------------- pruned search

def search_clip(clip: DeviceClip, me_range: int, qp: int, net: Net | None = None, ncand: int = 8,
                thr_part: float = 0.7, thr_mv: float = 0.0, stream=None, prepared: bool = False,
                hints_in: np.ndarray | None = None, nref: int = 1):
    """ME-only field of every P frame of a device clip: the full search
    (net None), the pruned search with the network's hints, or with given
    hint records `hints_in` ([P frames, mbh, mbw] of HINT_DTYPE); frame t
    searches its min(t, nref) previous frames.  Returns (field, hints) as
    numpy records (hints None for the full search)."""
    from . import search_frames
    torch = clip.torch
    dev = clip.raw.device
    if not prepared:
        clip.pad(stream)
        clip.hpel(0, clip.n - 1, stream)
    nb = clip.n - 1
    nmb = clip.mbw * clip.mbh
    out = torch.empty((nb, nmb * MB_DTYPE.itemsize), dtype=torch.uint8, device=dev)
    hints = None
    hp = [None] * nb
    if hints_in is not None:
        raw = np.ascontiguousarray(np.asarray(hints_in, HINT_DTYPE).reshape(nb, nmb)).view(np.uint8)
        hints = torch.as_tensor(raw.reshape(nb, -1)).to(dev)
        hp = [hints[t].data_ptr() for t in range(nb)]
    elif net is not None:
        hints = torch.zeros((nb, nmb * HINT_DTYPE.itemsize), dtype=torch.uint8, device=dev)
        hp = [hints[t].data_ptr() for t in range(nb)]
    jobs = [frame_job(clip, t, t - 1, out[t - 1].data_ptr(), me_range, 1, True, qp, hint_ptr=hp[t - 1],
                      older=[t - 1 - k for k in range(1, min(t, nref))])
            for t in range(1, clip.n)]
    if net is not None and hints_in is None:
        net.hints(jobs, ncand, thr_part, thr_mv, stream)
    search_frames(jobs, stream, sync=True)
    field = out.cpu().numpy().view(MB_DTYPE).reshape(nb, clip.mbh, clip.mbw)
    h = hints.cpu().numpy().view(HINT_DTYPE).reshape(nb, clip.mbh, clip.mbw) if hints is not None else None
    return field, h


# ---------------------------------------------------------------- judge

def judge(plain: np.ndarray, pruned: np.ndarray) -> dict:
    """The plain search judges the pruned field: fraction of MBs with the
    same decision (partition and every MV), the J-cost increase
    (sum of best_cost, pruned over plain, minus 1) and its per-frame maximum."""
    p = plain.reshape(plain.shape[0], -1) if plain.ndim > 1 else plain[None]
    q = pruned.reshape(pruned.shape[0], -1) if pruned.ndim > 1 else pruned[None]
    same = (p["part"] == q["part"]) & np.all(p["mv"] == q["mv"], axis=(-1, -2))
    jp = p["best_cost"].astype(np.float64).sum(axis=1)
    jq = q["best_cost"].astype(np.float64).sum(axis=1)
    per_frame = jq / jp - 1.0
    return {
        "agreement": float(same.mean()),
        "part_agreement": float((p["part"] == q["part"]).mean()),
        "jcost_increase": float(jq.sum() / jp.sum() - 1.0),
        "jcost_increase_max_frame": float(per_frame.max()),
        "mbs": int(same.size),
    }


def hint_stats(hints: np.ndarray) -> dict:
    h = hints.reshape(-1).view(HINT_DTYPE)
    return {
        "pruned_fraction": float((h["prune"] != 0).mean()),
        "part_hist": np.bincount(h["part"], minlength=4).tolist(),
        "conf_part_mean": float(h["conf_part"].mean() / 65535.0),
        "conf_mv_mean": float(h["conf_mv"].mean() / 65535.0),
    }
