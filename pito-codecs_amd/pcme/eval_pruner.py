"""BD-rate of the learned search pruner on a held-out clip (1080p by
default) at QP 22, 27, 32 and 37.

    python -m pcme.eval_pruner [--frames 16] [--seed 0x5EED9200] [--coder h264|model] [--size WxH] [--nref N]
                               [--out report.json]

For each QP the same frames are searched three ways: the plain search, the
pruned search with that QP's network (5 candidates; full search where
the partition head's confidence is below 0.7) and the same pruned search
with empty hints (the method without the network).  Every field is then
coded, and the (rate, PSNR) points of each method give its BD-rate against
the plain search:

* coder "h264" (the default): the field coder (pcme.h264enc) writes a
  Baseline CAVLC H.264 stream -- an I_PCM IDR, then every P picture with the
  field's partitions and MVs against the previous reconstruction, 4x4
  residual, P_Skip -- and the rate is the P pictures' bits, the distortion
  their mean luma PSNR; every stream is decoded by the minimal decoder
  (pcme.h264dec) and must reproduce the coder's reconstruction bit for bit;
* coder "model": the P-frame coding model of pcme.rdmodel (luma only,
  bits estimated from run/level and MV-difference Exp-Golomb lengths).
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import DeviceClip, HINT_DTYPE, MB_DTYPE, ClipSpec, mb_dims, pred_frame, synth_clip, LUMA_PAD
from . import pruner as P
from . import rdmodel as R

QPS = (22, 27, 32, 37)


# The benchmark workloads' generators by picture size: (rects, smallest and
# largest rect side, largest speed in px/frame, zoom per frame).
GENERATORS = {(352, 288): (8, 24, 120, 6.0, 0.0), (640, 360): (8, 24, 180, 6.0, 0.0),
              (1280, 720): (10, 64, 300, 8.0, 0.0), (1920, 1080): (12, 64, 400, 8.0, 0.0),
              (3840, 2160): (8, 128, 900, 16.0, 0.002)}


def clip_spec(seed: int, frames: int, w: int = 1920, h: int = 1080) -> ClipSpec:
    """A held-out clip of the workload generator of that size (the 1080p one
    for other sizes)."""
    n, lo, hi, v, z = GENERATORS.get((w, h), GENERATORS[(1920, 1080)])
    return ClipSpec(w, h, frames, nrect=n, rect_lo=lo, rect_hi=hi, vmax=v, pan_max=1.0, zoom=z, seed=seed)


def source_mbs(dc: DeviceClip, t: int) -> np.ndarray:
    """[mbw*mbh, 16, 16] luma of frame t (coded area of its padded plane)."""
    p = dc.planes[0, t].cpu().numpy()
    a = p[LUMA_PAD:LUMA_PAD + 16 * dc.mbh, LUMA_PAD:LUMA_PAD + 16 * dc.mbw]
    return a.reshape(dc.mbh, 16, dc.mbw, 16).transpose(0, 2, 1, 3).reshape(-1, 16, 16)


def rd_point(dc: DeviceClip, field: np.ndarray, qp: int) -> tuple[float, float]:
    """(bits, PSNR) of coding every P frame of `field` under the model."""
    torch = dc.torch
    nmb = dc.mbw * dc.mbh
    bits = sse = 0.0
    py = torch.empty(nmb * 256, dtype=torch.uint8, device=dc.raw.device)
    pc = torch.empty(nmb * 64, dtype=torch.uint8, device=dc.raw.device)
    for t in range(1, dc.n):
        rec = np.ascontiguousarray(field[t - 1].reshape(-1).view(MB_DTYPE))
        mb = torch.as_tensor(rec.view(np.uint8)).to(dc.raw.device)
        pred_frame(dc, t - 1, mb.data_ptr(), py.data_ptr(), pc.data_ptr(), pc.data_ptr())
        pred = py.cpu().numpy().reshape(nmb, 16, 16)
        b, s = R.code_frame(source_mbs(dc, t), pred, rec, qp)
        bits += b
        sse += s
    npix = (dc.n - 1) * nmb * 256
    return bits, R.psnr(sse, npix)


def rd_point_h264(y, u, v, field: np.ndarray, qp: int, nref: int = 1) -> tuple[float, float]:
    """(bits, PSNR) of the P pictures of the clip coded into a real stream
    with `field` (nref reference indices); the stream must decode to the
    coder's reconstruction."""
    from . import h264dec, h264enc
    c = h264enc.code_clip(y, u, v, field, qp, nref=nref)
    frames = h264dec.decode_annexb(c.stream, per_access_unit=False)
    if len(frames) != len(c.recon) or not all(
            np.array_equal(f.y, r[0]) and np.array_equal(f.u, r[1]) and np.array_equal(f.v, r[2])
            for f, r in zip(frames, c.recon)):
        raise AssertionError("the coded stream does not decode to the coder's reconstruction")
    return float(c.p_bits), c.p_psnr


def evaluate(frames: int, seed: int, ncand: int = 5, thr_part=None, thr_mv: float = 0.0,
             log=print, w: int = 1920, h: int = 1080, weights: str | None = None, coder: str = "h264",
             nref: int = 1) -> dict:
    y, u, v = synth_clip(clip_spec(seed, frames + 1, w, h))
    mbw, mbh = mb_dims(w, h)
    dc = DeviceClip(y, mbw, mbh, u, v)
    dc.prepare()
    rng = 64 if (w, h) == (3840, 2160) else 32
    if coder != "h264" and nref != 1:
        raise ValueError("the coding model codes one reference")
    pts = {"plain": [], "pruned": [], "no_network": []}
    if nref > 1:
        pts["plain_1ref"] = []
    judge = {}
    for qp in QPS:
        net = P.Net(seed=1, max_batch=64, weights=weights or P.WEIGHTS, qp=qp)
        plain, _ = P.search_clip(dc, rng, qp, prepared=True, nref=nref)
        pruned, hints = P.search_clip(dc, rng, qp, net=net, ncand=ncand, thr_part=P.thr_part_for(qp, thr_part),
                                      thr_mv=thr_mv,
                                      prepared=True, nref=nref)
        none = np.zeros(hints.shape, HINT_DTYPE)
        none["prune"] = 1
        bare, _ = P.search_clip(dc, rng, qp, prepared=True, hints_in=none, nref=nref)
        net.close()
        fields = [("plain", plain, nref), ("pruned", pruned, nref), ("no_network", bare, nref)]
        if nref > 1:
            fields.append(("plain_1ref", P.search_clip(dc, rng, qp, prepared=True)[0], 1))
        for k, f, nr in fields:
            pts[k].append(rd_point_h264(y, u, v, f, qp, nr) if coder == "h264" else rd_point(dc, f, qp))
        judge[qp] = {"pruned": P.judge(plain, pruned), "no_network": P.judge(plain, bare),
                     "pruned_fraction": float(np.mean(hints["prune"] != 0))}
        log(f"qp {qp}: " + ", ".join(f"{k} {pts[k][-1][0] / 1e3:.0f} kbit {pts[k][-1][1]:.3f} dB" for k in pts))
    rep = {
        "frames": frames, "seed": seed, "size": [w, h], "qps": list(QPS), "ncand": ncand,
        "thr_part": {str(q): P.thr_part_for(q, thr_part) for q in QPS}, "thr_mv": thr_mv,
        "nref": nref, "me_range": rng, "coder": coder, "decoded_bit_exact": coder == "h264",
        "points": {k: [[round(r, 1), round(q, 4)] for r, q in v] for k, v in pts.items()},
        "bd_rate_pruned_pct": round(R.bd_rate(pts["plain"], pts["pruned"]), 4),
        "bd_rate_no_network_pct": round(R.bd_rate(pts["plain"], pts["no_network"]), 4),
        "bd_rate_pruned_pchip_pct": round(R.bd_rate_pchip(pts["plain"], pts["pruned"]), 4),
        "bd_rate_no_network_pchip_pct": round(R.bd_rate_pchip(pts["plain"], pts["no_network"]), 4),
        "judge": {str(q): j for q, j in judge.items()},
    }
    if nref > 1:  # the references' own gain: the plain field over nref references against one
        rep["bd_rate_plain_vs_1ref_pct"] = round(R.bd_rate(pts["plain_1ref"], pts["plain"]), 4)
        rep["bd_rate_plain_vs_1ref_pchip_pct"] = round(R.bd_rate_pchip(pts["plain_1ref"], pts["plain"]), 4)
    return rep


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=16, help="P frames per QP")
    ap.add_argument("--seed", type=lambda s: int(s, 0), default=0x5EED9200)
    ap.add_argument("--ncand", type=int, default=5)
    ap.add_argument("--thr", default=None,
                    help="partition-head confidence threshold: one for every QP (0.7), per QP "
                         "(22:0.7,27:0.7,32:0.8,37:0.9), or the per-QP default (omitted)")
    ap.add_argument("--thr-mv", type=float, default=0.0, help="motion-head candidate-mass threshold")
    ap.add_argument("--weights", default=None, help="pruner weights (.npz; default: the committed ones)")
    ap.add_argument("--coder", choices=("h264", "model"), default="h264")
    ap.add_argument("--size", default="1920x1080")
    ap.add_argument("--nref", type=int, default=1, help="references of list 0 (the 4K workload: 4)")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        print("the evaluation needs the GPU", file=sys.stderr)
        return 2
    w, h = (int(t) for t in a.size.split("x"))
    rep = evaluate(a.frames, a.seed, a.ncand, P.parse_thr(a.thr), a.thr_mv, log=lambda s: print(s, file=sys.stderr, flush=True),
                   w=w, h=h, weights=a.weights, coder=a.coder, nref=a.nref)
    print(json.dumps(rep))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(rep, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
