"""Per-macroblock timeline of the frame search on the 1080p workload.

Loads the timeline build of the library (``make -C pito-codecs_amd
timeline``: the frame walkers stamp, for every MB of one job, the moment its
turn begins, the moment its top-right neighbour's exchange words are seen,
the end of its analysis and the publication of its own words) and runs the
plain search of the seeded 1920x1080 clip twice: one P frame alone (68 row
walkers, nothing else on the GPU) and the whole 59-frame step on the given
persistent grid, recording one frame.  Optionally the pruned search (the
committed network's hints at the default rule) instead of the plain one.

Reports, per run: the frame's end, its critical path per MB (end / 254), the
analysis time per MB (mean, by row band, and for pruned and full MBs), the
wait for the row above, and the hand-off latency (from a word's publication,
or from the reader's arrival when it came later, to the reader seeing it).

    python -m pcme.timeline [--grid 2688] [--config 1,12,0] [--frame 58]
                            [--pruned] [--json out.json]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys

import numpy as np

import pcme

TIMELINE_LIB = os.path.normpath(os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "lib",
                                             "libpcme_timeline.so"))
CLOCK_HZ = 100e6  # s_memrealtime


def _load():
    if not os.path.exists(TIMELINE_LIB):
        raise SystemExit(f"{TIMELINE_LIB} is not built: run make -C pito-codecs_amd timeline")
    pcme.LIB_PATH = TIMELINE_LIB
    lib = pcme.load()
    lib.pcme_timeline_set.argtypes = [ctypes.c_int]
    lib.pcme_timeline_read.argtypes = [ctypes.c_void_p, ctypes.c_int]
    return lib


def summarise(stamps: np.ndarray, mbw: int, mbh: int, pruned: np.ndarray | None = None) -> dict:
    """Statistics of one frame's stamps ([mbh*mbw, 4] u64 clock ticks)."""
    t = stamps.astype(np.int64).reshape(mbh, mbw, 4)
    t0 = t[t > 0].min()
    us = (t - t0) / CLOCK_HZ * 1e6
    start, seen, done, pub = us[..., 0], us[..., 1], us[..., 2], us[..., 3]
    ana = done - seen
    # MB (x, y) waits for (x + 1, y - 1): hand-off from the later of the
    # publication and the reader's arrival to the reader seeing the words
    arrive, publ, got = start[1:, :-1], pub[:-1, 1:], seen[1:, :-1]
    blocked = publ > arrive
    handoff = got - np.maximum(arrive, publ)
    path = mbw + 2 * (mbh - 1)
    out = {
        "frame_end_us": round(float(pub.max()), 1),
        "critical_path_mbs": path,
        "per_path_mb_us": round(float(pub.max()) / path, 2),
        "analysis_us_mean": round(float(ana.mean()), 2),
        "analysis_us_by_row_band": [round(float(ana[i:i + 8].mean()), 1) for i in range(0, mbh, 8)],
        "blocked_fraction": round(float(blocked.mean()), 3),
        "wait_for_row_above_us_per_row": round(float(np.clip(publ - arrive, 0, None).sum(1).mean()), 1),
        "handoff_us_median": round(float(np.median(handoff)), 2),
        "handoff_us_p90": round(float(np.quantile(handoff, 0.9)), 2),
        "publish_after_analysis_us_median": round(float(np.median(pub - done)), 2),
        "row_lag_us_mean": round(float(np.diff(seen[:, 0]).mean()), 1),
    }
    if pruned is not None:
        pr = pruned.reshape(mbh, mbw)
        out["pruned_fraction"] = round(float(pr.mean()), 3)
        out["analysis_us_pruned"] = round(float(ana[pr].mean()), 2) if pr.any() else None
        out["analysis_us_full"] = round(float(ana[~pr].mean()), 2) if (~pr).any() else None
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=2688, help="persistent grid of the whole step (waves)")
    ap.add_argument("--config", default="1,12,0", help="waves_per_group,waves_per_cu,reuse_regions")
    ap.add_argument("--frame", type=int, default=58, help="the step's frame job to record (0-58)")
    ap.add_argument("--pruned", action="store_true", help="the pruned search at the default rule")
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    lib = _load()
    import torch
    if not torch.cuda.is_available():
        print("the timeline needs the GPU", file=sys.stderr)
        return 2
    torch.cuda.set_device(0)
    qp, rng = 27, 32
    spec = pcme.ClipSpec(1920, 1080, 60, nrect=12, rect_lo=64, rect_hi=400, vmax=8.0, pan_max=1.0, seed=0x5EED0000)
    y, _, _ = pcme.synth_clip(spec, threads=16)
    mbw, mbh = pcme.mb_dims(spec.w, spec.h)
    dc = pcme.DeviceClip(y, mbw, mbh)
    nb, nmb = dc.n - 1, mbw * mbh
    out = torch.empty((nb, nmb * pcme.MB_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    hints = torch.zeros((nb, nmb * pcme.HINT_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    sp = torch.cuda.current_stream().cuda_stream
    dc.prepare(sp)
    net = None
    if a.pruned:
        from . import pruner as P
        net = P.Net(seed=1, max_batch=32, weights=P.WEIGHTS, qp=qp)
    wpg, wcu, reuse = (int(v) for v in a.config.split(","))
    report = {"workload": "1920x1080 QP 27 me_range 32, seeded clip of the 1080p bench",
              "search": "pruned (default rule)" if a.pruned else "plain", "runs": {}}
    for name, frames, grid, rec in (("one frame alone", [a.frame + 1], mbh, 0),
                                    ("whole step", list(range(1, dc.n)), a.grid, a.frame)):
        jobs = [pcme.frame_job(dc, t, t - 1, out[t - 1].data_ptr(), rng, 1, True, qp,
                               hint_ptr=hints[t - 1].data_ptr() if a.pruned else None) for t in frames]
        if net is not None:
            net.hints(jobs, 5, P.thr_part_for(qp), 0.0, sp)
        plan = pcme.Plan(jobs)
        if a.pruned:
            plan.set_grid(grid)
        else:
            plan.set_config(grid, wpg, wcu, reuse)
        for _ in range(3):
            plan.launch(sp)
        torch.cuda.synchronize()
        pcme._check(lib.pcme_timeline_set(rec))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        plan.launch(sp)
        e1.record()
        torch.cuda.synchronize()
        plan.check(sp)
        st = np.zeros(nmb * 4, dtype=np.uint64)
        pcme._check(lib.pcme_timeline_read(st.ctypes.data, nmb))
        if (st.reshape(-1, 4) == 0).any():
            raise SystemExit(f"{name}: some MB of the recorded frame carries no stamp")
        pr = None
        if a.pruned:
            h = hints[frames[rec] - 1].cpu().numpy().view(pcme.HINT_DTYPE)
            pr = h["prune"] != 0
        r = summarise(st, mbw, mbh, pr)
        r["launch_ms"] = round(e0.elapsed_time(e1), 3)
        r["frames"] = len(frames)
        r["grid"] = grid
        report["runs"][name] = r
        print(f"{name}: " + json.dumps(r), flush=True)
    if net is not None:
        net.close()
    if a.json:
        with open(a.json, "w") as f:
            json.dump(report, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
