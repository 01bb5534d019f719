"""Benchmark: H.264 P-frame motion search (ME-only field) on one MI355X.

Workload (one step): a seeded synthetic 1920x1080 (coded 1920x1088) clip of
60 frames IPPP, QP 27, me_range 32, subpel_rounds 1, all four partition
shapes; 12 rectangles of 64-400 px moving at up to 8 px/frame over a panned
blurred-noise background.  The raw I420 luma frames are resident in HBM when
the timed region starts; a step imports/pads the 60 frames, builds the
half-pel planes of the 59 reference frames and searches every macroblock of
the 59 P frames (481,440 MBs) -- the whole hot path of the reference's
search (src/h264/enc/slice.rs:1590-1945, picture.rs:318-449).

Prints one JSON line (see the field list at the end of main()).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "pito-codecs_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "1080p H.264 ME Mblocks/s: plain GPU vs CNN-pruned at <=1% BD-rate; GB/s"
HBM_PEAK_GBS = 8000.0
# Integer VALU peak: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz = 39.3 T lane-op/s
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12

W, H, FRAMES, QP, RANGE, SUBPEL = 1920, 1080, 60, 27, 32, 1


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-frames", type=int, default=16,
                    help="P frames of the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a kernel of known "
                         "traffic, for scaling the FETCH_SIZE/WRITE_SIZE counters)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the boxes it targets have one MI355X); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import pcme

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    if args.copy_calib_mib:
        a = torch.empty(args.copy_calib_mib << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    spec = pcme.ClipSpec(W, H, FRAMES, nrect=12, rect_lo=64, rect_hi=400, vmax=8.0, pan_max=1.0,
                         zoom=0.0, noise_sigma=2.0, seed=0x5EED0000)
    t0 = time.time()
    y, _, _ = pcme.synth_clip(spec, threads=min(16, os.cpu_count() or 1))
    print(f"[bench] generated {FRAMES} frames in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    mbw, mbh = pcme.mb_dims(W, H)
    clip = pcme.DeviceClip(y, mbw, mbh)
    nP = FRAMES - 1
    mbs_per_step = nP * mbw * mbh
    out = torch.empty((nP, mbw * mbh * pcme.MB_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    jobs = [pcme.frame_job(clip, t, t - 1, out[t - 1].data_ptr(), RANGE, SUBPEL, True, QP) for t in range(1, FRAMES)]
    stream = torch.cuda.current_stream()

    ev = []  # (start, end) events around each search launch

    def step(record: bool):
        clip.pad(stream.cuda_stream)
        clip.hpel(0, nP, stream.cuda_stream)
        if record:
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
        pcme.search_frames(jobs, stream.cuda_stream, sync=False)
        if record:
            b.record(stream)
            ev.append((a, b))

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = 1e3 * dt / args.steps
    value = mbs_per_step * args.steps / dt / 1e6
    search_ms = sum(a.elapsed_time(b) for a, b in ev) / len(ev)
    print(f"[bench] {ms_step:.2f} ms/step, search kernel {search_ms:.2f} ms", file=sys.stderr, flush=True)

    # Stall check + results of the last step.
    pcme.search_frames(jobs, stream.cuda_stream, sync=True)
    gpu_res = out.cpu().numpy().view(pcme.MB_DTYPE).reshape(nP, mbh, mbw)

    # Algorithmic bytes of one search launch (per MB: 256 B source luma,
    # 4 x 256 B reference planes read once, 128 B record, 32 B exchange).
    bytes_per_mb = 256 + 4 * 256 + 128 + 32
    traffic_bytes = bytes_per_mb * mbs_per_step
    achieved_gbs = traffic_bytes / (search_ms * 1e-3) / 1e9
    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "Mblocks/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic (seeded generator, include/pcme_synth.h)",
        "config": {"workload": "plain GPU ME-only search, 1920x1080 60 frames IPPP, 1 ref, QP 27, "
                               "me_range 32 + qpel, partitions 16x16/16x8/8x16/8x8, 12 rects 64-400 px, "
                               "v in [-8,8] px/frame",
                   "mbs_per_step": mbs_per_step, "p_frames": nP, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "search_frames_kernel", "achieved": round(achieved_gbs, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved_gbs / HBM_PEAK_GBS, 5),
                     "traffic": None, "algorithmic_bytes_per_mb": bytes_per_mb,
                     "kernel_ms": round(search_ms, 3)},
    }

    if not args.no_cpu:
        from oracle import oracle as O
        k = min(args.cpu_frames, nP)
        threads = max(1, min(16, os.cpu_count() or 1))
        src = [O.pad_luma(y[t], mbw, mbh) for t in range(1, k + 1)]
        ref = [O.pad_luma(y[t - 1], mbw, mbh) for t in range(1, k + 1)]
        res, secs, ctr = O.search_jobs_mt(src, ref, mbw, mbh, RANGE, SUBPEL, True, QP, threads)
        cpu_val = k * mbw * mbh / secs / 1e6
        parity = all(np.array_equal(np.ascontiguousarray(gpu_res[i]).view(np.uint8),
                                    np.ascontiguousarray(res[i]).view(np.uint8)) for i in range(k))
        # Lane-op count of the same work (the GPU replays the oracle's path):
        # SAD 0.25/px, MC+SATD 3.75/px, half-pel build 9/px of reference.
        per_mb_ops = (0.25 * ctr.sad_px + 3.75 * ctr.satd_px) / max(ctr.mbs, 1)
        valu_tops = per_mb_ops * mbs_per_step / (search_ms * 1e-3) / 1e12
        line["valu_roofline"] = {"bound": "valu", "achieved": round(valu_tops, 3), "peak": round(VALU_PEAK_TOPS, 2),
                                 "unit": "T lane-op/s", "frac": round(valu_tops / VALU_PEAK_TOPS, 4),
                                 "lane_ops_per_mb": round(per_mb_ops, 1)}
        line["cpu_baseline"] = {"value": round(cpu_val, 5), "unit": "Mblocks/s", "cores": threads, "kind": "port",
                                "sample": f"{k} P frames of the same 1080p clip (frames 1..{k}), half-pel build + "
                                          f"search, one frame per thread, C oracle -O3 -march=x86-64-v3",
                                "seconds": round(secs, 3)}
        line["parity_sample"] = bool(parity)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
