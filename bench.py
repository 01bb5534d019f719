"""Benchmark: H.264 P-frame motion search (ME-only field) on one MI355X.

Default workload (one step): a seeded synthetic 1920x1080 (coded 1920x1088)
clip of 60 frames IPPP, QP 27, me_range 32, subpel_rounds 1, all four
partition shapes; 12 rectangles of 64-400 px moving at up to 8 px/frame over
a panned blurred-noise background.  The raw luma frames are resident in HBM
when the timed region starts; a step builds the four padded planes of every
frame (import, edge extension and half-pel planes in one pass) and searches
every macroblock of the P frames (59 x 8,160 = 481,440 MBs) -- the hot path of the reference's
search (src/h264/enc/slice.rs:1590-1945, picture.rs:318-449).

Other workloads (--workload): "batch" = 64 clips of uneven size (16 each
of 352x288, 640x360, 1280x720, 1920x1080), 30 frames each, QP per clip from
{22,27,32,37}, me_range 32, all in one persistent launch; "4k" = 3840x2160,
120 frames, me_range 64, QP 27, 8 rects of 128-900 px at up to 16 px/frame,
0.2%/frame zoom.

The line also carries the CNN-pruned search of the same frames ("pruned")
and the lowres lookahead of the same frames ("lookahead": LowRes::new,
scenecut complexity and macroblock-tree costs, lookahead.rs / mbtree.rs),
each with its own timing.

Prints one JSON line (fields at the end of main()).
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
# Algorithmic HBM bytes of the search per MB (one reference): 256 B source
# luma, 4 x 256 B reference planes read once, 128 B record, 32 B exchange
# words; each further reference adds its four planes.
BYTES_PER_MB = 256 + 4 * 256 + 128 + 32
# Pruner network multiply-adds per MB (include/pcme_net.h): conv1 144 px x
# 32 x 18, conv2 36 x 32 x 288, conv3 36 x 32 x 288, conv4 9 x 32 x 288,
# fc 293 x 32.
NET_MACS_PER_MB = 144 * 32 * 18 + 36 * 32 * 288 * 2 + 9 * 32 * 288 + 293 * 32
MFMA_BF16_PEAK_TFLOPS = 2500.0
# B pictures of the batched B field leg (pcme_b_fields, one launch)
B_BATCH = 16


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("1080p", "batch", "4k"), default="1080p")
    ap.add_argument("--cpu-frames", type=int, default=59, help="P frames of the CPU-baseline sample")
    ap.add_argument("--cpu-frames-1t", type=int, default=4, help="P frames of the one-thread CPU sample")
    ap.add_argument("--grid", type=int, default=0,
                    help="persistent grid in waves (skips the autotuner; 0: autotune)")
    ap.add_argument("--config", default="",
                    help="kernel configuration waves_per_group,waves_per_cu,reuse_regions (with --grid: skips "
                         "the autotuner)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only the timed steps (no graph, critical-path, pruned, lookahead, B or CPU legs): "
                         "the command the profiles are taken with")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-pruned", action="store_true", help="skip the CNN-pruned leg")
    ap.add_argument("--no-lookahead", action="store_true", help="skip the lowres-lookahead leg")
    ap.add_argument("--no-b", action="store_true", help="skip the B-picture search leg")
    ap.add_argument("--nref", type=int, default=0,
                    help="references of list 0 per P frame (0: the workload's own, 4 for 4k, else 1)")
    ap.add_argument("--ncand", type=int, default=5, help="pruner candidates per MB")
    ap.add_argument("--weights", default=None, help="pruner weights (.npz; default: the committed ones)")
    ap.add_argument("--thr", default=None,
                    help="pruner confidence threshold (partition head): one for every QP (0.7), per QP "
                         "(22:0.7,27:0.7,32:0.8,37:0.9), or the per-QP default rule (omitted)")
    ap.add_argument("--overlap", type=int, default=1,
                    help="pruned step with the network and the search overlapped: the frames in this many "
                         "chunks, each chunk's search on a stream of its own once its hints are done (1: the "
                         "network, then one search launch)")
    ap.add_argument("--overlap-grid", type=float, default=1.0,
                    help="with --overlap: each chunk's persistent grid is this fraction of the tuned grid "
                         "divided by the chunk count, so the chunks' walkers are resident together")
    ap.add_argument("--thr-mv", type=float, default=0.0,
                    help="threshold on the motion head's candidate mass (0: the partition head decides)")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a kernel of known "
                         "traffic, for scaling the FETCH_SIZE/WRITE_SIZE counters)")
    return ap.parse_args()


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "search_frames_traffic.md")


def measured_traffic(workload, tune, nref):
    """HBM-side bytes per launch of the search kernel on the workload with
    nref references at the plan's kernel configuration (tune; None: the
    default 1,12,1), from the FETCH_SIZE / WRITE_SIZE runs summarised in
    profiles/ (scaled by a 1 GiB copy measured in the same runs); None for a
    combination not measured there or when the summary is absent."""
    if not os.path.exists(TRAFFIC_SUMMARY):
        return None
    wpg, wcu, reuse = ((tune["waves_per_group"], tune["waves_per_cu"], tune["reuse_regions"])
                       if tune else (1, 12, 1))
    key = f"`search` workload={workload} nref={nref} config={wpg},{wcu},{reuse}: bytes_per_launch="
    with open(TRAFFIC_SUMMARY) as f:
        for ln in f:
            if ln.startswith(key):
                return int(ln[len(key):].split()[0])
    return None


BDRATE_SUMMARY = os.path.join(ROOT, "profiles", "pruner_bdrate_bitstreams.md")
WORKLOAD_SIZES = {"1080p": ["1920x1080"], "4k": ["3840x2160"],
                  "batch": ["352x288", "640x360", "1280x720", "1920x1080"]}


def measured_bd_rate(ncand, thr_part, thr_mv, workload, nref=1):
    """The pruned search's BD-rate against the plain search on real
    bitstreams (the field coder, QP 22/27/32/37, held-out clips of each
    picture size of the workload, every stream decoded bit-exact), from the
    evaluation summarised in profiles/ for the default confidence rule (the
    per-QP thresholds, thr_part None); None for other rules or without the
    summary."""
    if not os.path.exists(BDRATE_SUMMARY) or (ncand, thr_part, thr_mv) != (5, None, 0.0):
        return None
    per = {}
    with open(BDRATE_SUMMARY) as f:
        for ln in f:
            if not ln.startswith("`") or "`: " not in ln:
                continue
            key, rest = ln[1:].split("`: ", 1)
            size = key.split()[0]
            kv = dict(x.split("=") for x in rest.split())
            if int(kv.get("nref", "1")) != nref:
                continue
            per.setdefault(size, []).append((float(kv["bd_rate_cubic_pct"]), float(kv["bd_rate_pchip_pct"])))
    sizes = {z: per[z] for z in WORKLOAD_SIZES[workload] if z in per}
    if not sizes:
        return None
    return {"per_size": {z: {"cubic_pct": [c for c, _ in v], "pchip_pct": [q for _, q in v]} for z, v in sizes.items()},
            "max_pct": max(max(c, q) for v in sizes.values() for c, q in v),
            "coder": "field coder (pcme/h264enc.py), Baseline CAVLC bitstreams decoded bit-exact, QP 22/27/32/37",
            "references": nref,
            "source": "profiles/pruner_bdrate_bitstreams.md"}


def older_refs(t, nref):
    """Pictures of references 1..nref-1 of P frame t (the frames before
    t - 1, nearest first; fewer at the clip's start)."""
    return [t - 1 - k for k in range(1, min(t, nref))]


def clip_specs(pcme, workload):
    """[(ClipSpec, qp, me_range)] of the workload."""
    if workload == "1080p":
        return [(pcme.ClipSpec(1920, 1080, 60, nrect=12, rect_lo=64, rect_hi=400, vmax=8.0, pan_max=1.0,
                               seed=0x5EED0000), 27, 32)]
    if workload == "4k":
        return [(pcme.ClipSpec(3840, 2160, 120, nrect=8, rect_lo=128, rect_hi=900, vmax=16.0, pan_max=1.0,
                               zoom=0.002, seed=0x5EED0000), 27, 64)]
    out = []
    sizes = [(352, 288, 8, 120, 6.0), (640, 360, 8, 180, 6.0), (1280, 720, 10, 300, 8.0), (1920, 1080, 12, 400, 8.0)]
    qps = (22, 27, 32, 37)
    for i in range(64):
        w, h, nrect, hi, v = sizes[i // 16]
        out.append((pcme.ClipSpec(w, h, 30, nrect=nrect, rect_lo=24 if w < 1000 else 64, rect_hi=hi, vmax=v,
                                  pan_max=1.0, seed=0x5EED0000 + i), qps[(i * 7 + 3) % 4], 32))
    return out


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

    specs = clip_specs(pcme, args.workload)
    if args.nref == 0:
        args.nref = 4 if args.workload == "4k" else 1
    if not 1 <= args.nref <= pcme.MAX_REFS:
        sys.exit(f"bench.py: --nref must be 1..{pcme.MAX_REFS}")
    threads = cpu_share()
    t0 = time.time()
    clips, outs, jobs, host = [], [], [], []
    for spec, qp, rng in specs:
        y, _, _ = pcme.synth_clip(spec, threads=threads)
        mbw, mbh = pcme.mb_dims(spec.w, spec.h)
        dc = pcme.DeviceClip(y, mbw, mbh)
        o = torch.empty((spec.frames - 1, mbw * mbh * pcme.MB_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
        jobs += [pcme.frame_job(dc, t, t - 1, o[t - 1].data_ptr(), rng, 1, True, qp, older=older_refs(t, args.nref))
                 for t in range(1, spec.frames)]
        clips.append(dc)
        outs.append(o)
        host.append((y, mbw, mbh, qp, rng))
    print(f"[bench] generated {sum(s.frames for s, _, _ in specs)} frames in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    mbs_per_step = sum((dc.n - 1) * dc.mbw * dc.mbh for dc in clips)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream

    def prepare():
        pcme.prepare_clips(clips, sp)

    plan = pcme.Plan(jobs)
    prepare()
    if args.kernels_only:
        args.no_graph = args.no_pruned = args.no_lookahead = args.no_b = args.no_cpu = True
    tune = fix_or_tune(plan, args, sp)

    ev = []  # (start, end) events around each search launch, on the launch stream

    def step(record: bool):
        prepare()
        if record:
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
        plan.launch(sp)
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
    plan.check(sp)
    ms_step = 1e3 * dt / args.steps
    value = mbs_per_step * args.steps / dt / 1e6
    search_ms = sum(a.elapsed_time(b) for a, b in ev) / len(ev)
    print(f"[bench] {ms_step:.2f} ms/step, search kernel {search_ms:.2f} ms", file=sys.stderr, flush=True)

    graph = None
    if not args.no_graph:
        # The same step captured once as a hipGraph and replayed.
        g = torch.cuda.CUDAGraph()
        s2 = torch.cuda.Stream()
        s2.wait_stream(stream)
        with torch.cuda.stream(s2):
            with torch.cuda.graph(g, stream=s2, capture_error_mode="thread_local"):
                pcme.prepare_clips(clips, s2.cuda_stream)
                plan.launch(s2.cuda_stream)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            g.replay()
        torch.cuda.synchronize()
        gdt = time.perf_counter() - t1
        plan.check(sp)
        graph = {"ms_per_step": round(1e3 * gdt / args.steps, 3),
                 "mblocks_per_s": round(mbs_per_step * args.steps / gdt / 1e6, 4)}
        print(f"[bench] graph replay {graph['ms_per_step']:.2f} ms/step", file=sys.stderr, flush=True)

    # The wavefront's dependency bound: one frame alone (one walker per MB
    # row, nothing else on the GPU) runs its critical path of
    # mbw + 2*(mbh-1) MB analyses back to back; every frame of the step has
    # the same path, so that time is the step's lower bound with unlimited
    # parallel hardware.
    crit = None
    if args.workload in ("1080p", "4k") and not args.kernels_only:
        p1 = pcme.Plan(jobs[:1])
        if tune:  # the step's kernel configuration, one walker per row
            p1.set_config(0, tune["waves_per_group"], tune["waves_per_cu"], tune["reuse_regions"])
        p1.launch(sp)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(5):
            p1.launch(sp)
        e1.record(stream)
        torch.cuda.synchronize()
        p1.check(sp)
        one_ms = e0.elapsed_time(e1) / 5
        dc0 = clips[0]
        path = dc0.mbw + 2 * (dc0.mbh - 1)
        crit = {"single_frame_ms": round(one_ms, 3), "path_mbs": path,
                "mb_latency_us": round(1e3 * one_ms / path, 2), "frames_per_step": len(jobs),
                "search_ms": round(search_ms, 3), "frac": round(one_ms / search_ms, 4)}
        print(f"[bench] one frame alone {one_ms:.2f} ms ({1e3 * one_ms / path:.1f} us per MB on the critical path)",
              file=sys.stderr, flush=True)
        p1.close()
    bytes_per_mb = BYTES_PER_MB + 4 * 256 * (args.nref - 1)
    achieved_gbs = bytes_per_mb * mbs_per_step / (search_ms * 1e-3) / 1e9
    kname = search_kernel_name(tune, args.nref > 1, False)
    refs = f"{args.nref} ref" + ("s" if args.nref > 1 else "")
    wl = {"1080p": f"plain GPU ME-only search, 1920x1080 60 frames IPPP, {refs}, QP 27, me_range 32 + qpel, "
                   "partitions 16x16/16x8/8x16/8x8, 12 rects 64-400 px, v in [-8,8] px/frame",
          "batch": "plain GPU ME-only search, 64 clips (16 each of 352x288, 640x360, 1280x720, 1920x1080), "
                   f"30 frames each IPPP, {refs}, QP per clip from {{22,27,32,37}}, me_range 32, one persistent "
                   "launch",
          "4k": f"plain GPU ME-only search, 3840x2160 120 frames IPPP, {refs}, QP 27, me_range 64 + qpel, "
                "8 rects 128-900 px, v in [-16,16] px/frame, 0.2%/frame zoom"}[args.workload]
    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "Mblocks/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8",
        "data": "synthetic (seeded generator, include/pcme_synth.h)",
        "config": {"workload": wl, "mbs_per_step": mbs_per_step, "p_frames": len(jobs),
                   "parallelism": "single GPU"},
        "roofline": None,  # set below from the oracle's operation counts (cpu_baseline leg)
        "hbm_roofline": {"bound": "hbm", "kernel": kname, "achieved": round(achieved_gbs, 2),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved_gbs / HBM_PEAK_GBS, 5),
                         "traffic": measured_traffic(args.workload, tune, args.nref),
                         "traffic_source": "profiles/search_frames_traffic.md (FETCH_SIZE/WRITE_SIZE runs)"
                         if measured_traffic(args.workload, tune, args.nref) is not None else None,
                         "algorithmic_bytes_per_mb": bytes_per_mb, "kernel_ms": round(search_ms, 3)},
        "autotune": tune,
        "graph": graph,
        "critical_path": crit,
    }

    if not args.no_pruned:
        line["pruned"] = pruned_leg(args, pcme, torch, clips, outs, specs, stream, mbs_per_step, ms_step)

    if not args.no_lookahead:
        line["lookahead"] = lookahead_leg(args, pcme, torch, clips, stream, threads)
    if not args.no_b and args.workload == "1080p":
        line["b_search"] = b_leg(args, pcme, torch, clips[0], outs[0], specs[0][1], specs[0][2], stream, threads)

    if not args.no_cpu:
        cpu, per_mb_ops, parity = cpu_leg(args, pcme, host[-1], outs[-1])
        valu_tops = per_mb_ops * mbs_per_step / (search_ms * 1e-3) / 1e12
        line["roofline"] = {"bound": "valu", "kernel": kname, "achieved": round(valu_tops, 3),
                            "peak": round(VALU_PEAK_TOPS, 2), "unit": "T lane-op/s",
                            "frac": round(valu_tops / VALU_PEAK_TOPS, 4),
                            "traffic": measured_traffic(args.workload, tune, args.nref),
                            "lane_ops_per_mb": round(per_mb_ops, 1), "mbs_per_launch": mbs_per_step,
                            "kernel_ms": round(search_ms, 3),
                            "peak_basis": "256 CU x 64 lanes x 2.4 GHz, one 32-bit lane-op per lane per clock"}
        line["cpu_baseline"] = cpu
        line["parity_sample"] = bool(parity)
    print(json.dumps(line), flush=True)


def fix_or_tune(plan, args, stream) -> dict | None:
    """The plan's kernel configuration: the one --config / --grid name, or
    the autotuner's choice over every built configuration and grid size."""
    if args.grid > 0 or args.config:
        wpg, wcu, reuse = (int(v) for v in (args.config or "1,12,1").split(","))
        t = plan.set_config(args.grid, wpg, wcu, reuse)
        return dict(t.as_dict(), fixed=True, ms=None)
    if args.no_autotune:
        return None
    t = plan.autotune_ex(reps=2, stream=stream)
    print(f"[bench] autotune: {t.waves} waves, {t.waves_per_group} per group, {t.waves_per_cu} per CU, "
          f"reuse {t.reuse_regions}: {t.ms:.2f} ms", file=sys.stderr, flush=True)
    return dict(t.as_dict(), ms=round(t.ms, 3))


def search_kernel_name(tune: dict | None, mr: bool, hinted: bool) -> str:
    """The frame-search kernel a plan launches at configuration ``tune``
    (None: the plan's default, one walker per group at 12 per CU with
    regions reused): the plain search has a kernel per configuration, the
    searches with references or hints one at 12 and one at 16 walkers per
    CU."""
    wpg, wcu, reuse = ((tune["waves_per_group"], tune["waves_per_cu"], tune["reuse_regions"])
                       if tune else (1, 12, 1))
    if mr or hinted:
        base = "search_frames_" + "_".join(n for n, on in (("mr", mr), ("hinted", hinted)) if on)
        return base + ("16_kernel" if wcu == 16 else "_kernel")
    if (wpg, wcu, reuse) == (1, 12, 1):
        return "search_frames_kernel"
    return f"search_frames_var_kernel<{wpg}, {wcu}, {'true' if reuse else 'false'}>"


def cpu_share() -> int:
    """Host threads for the CPU baselines: every core this process may run
    on, or the machine's CPU share for this job where the environment states
    one (OMP_NUM_THREADS), whichever is smaller."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(args, pcme, host, out):
    """The CPU baseline: the C oracle (a restatement of the reference's
    search, -O3 -march=x86-64-v3) on the last clip's P frames, one frame per
    thread like the reference's frame threads (enc/mod.rs:196-243), with
    every core this process may use, and on one thread; each the median of
    3 timed runs without operation counters.  One more (untimed) run counts
    the SAD and MC+SATD pixels, whose lane-op equivalents give the
    roofline's algorithmic work, and gives the frames the GPU's output is
    compared with."""
    import numpy as np
    from oracle import oracle as O
    y, mbw, mbh, qp, rng = host
    nP = y.shape[0] - 1
    k = min(args.cpu_frames, nP)
    pads = [O.pad_luma(y[t], mbw, mbh) for t in range(0, k + 1)]
    src = pads[1:k + 1]
    ref = pads[0:k]
    older = [[pads[u] for u in older_refs(t, args.nref)] for t in range(1, k + 1)]
    nproc = len(os.sched_getaffinity(0))
    threads = max(1, min(cpu_share(), k))
    res, _, ctr = O.search_jobs_mt(src, ref, mbw, mbh, rng, 1, True, qp, threads, count=True, older=older)
    gpu_res = out.cpu().numpy().view(pcme.MB_DTYPE).reshape(nP, mbh, mbw)
    parity = all(np.array_equal(np.ascontiguousarray(gpu_res[i]).view(np.uint8),
                                np.ascontiguousarray(res[i]).view(np.uint8)) for i in range(k))
    # Lane-op count of the same work (the GPU replays the oracle's path):
    # SAD 0.25/px, MC+SATD 3.75/px.
    per_mb_ops = (0.25 * ctr.sad_px + 3.75 * ctr.satd_px) / max(ctr.mbs, 1)
    runs = sorted(O.search_jobs_mt(src, ref, mbw, mbh, rng, 1, True, qp, threads, count=False, older=older)[1]
                  for _ in range(3))
    secs = runs[1]
    k1 = min(args.cpu_frames_1t, k)
    runs1 = sorted(O.search_jobs_mt(src[:k1], ref[:k1], mbw, mbh, rng, 1, True, qp, 1, count=False,
                                    older=older[:k1])[1] for _ in range(3))
    cpu_val = k * mbw * mbh / secs / 1e6
    one_val = k1 * mbw * mbh / runs1[1] / 1e6
    print(f"[bench] cpu baseline: {cpu_val:.3f} Mblocks/s on {threads} threads ({nproc} in the affinity mask), "
          f"{one_val:.4f} on one", file=sys.stderr, flush=True)
    cpu = {"value": round(cpu_val, 5), "unit": "Mblocks/s", "cores": threads, "kind": "port",
           "sample": f"{k} P frames of the last clip of the workload ({y.shape[2]}x{y.shape[1]}), half-pel build + "
                     f"search, one frame per thread, C oracle -O3 -march=x86-64-v3, median of 3 runs without "
                     f"counters",
           "seconds": round(secs, 3), "runs_s": [round(r, 3) for r in runs],
           "nproc": nproc, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "cpu_model": cpu_model(),
           "one_thread": {"value": round(one_val, 5), "unit": "Mblocks/s", "frames": k1,
                          "seconds": round(runs1[1], 3), "runs_s": [round(r, 3) for r in runs1]}}
    return cpu, per_mb_ops, parity


def lookahead_leg(args, pcme, torch, clips, stream, threads):
    """The lowres lookahead of the same frames (lookahead.rs, mbtree.rs): a
    step = LowRes::new of every frame + complexity of every frame against
    its predecessor (the first alone) + the macroblock-tree costs of every
    frame with its predecessor as list 0 + one tree_offsets (mod.rs:819-840)
    with the clip's frames as the queue: propagate over all of them, then
    the first frame's offsets.  The C oracle times the same work on the last
    clip's frames on the host."""
    import numpy as np
    from pcme import lookahead as L
    sp = stream.cuda_stream
    lrs, cjobs, tjobs, sums, tblocks, tnodes, props, toffs = [], [], [], [], [], [], [], []
    blocks_per_step = 0
    for dc in clips:
        lr = L.lowres_frames(dc, stream=sp)
        n = dc.n
        s = torch.zeros((n, 2), dtype=torch.int64, device="cuda")
        nb = dc.mbw * dc.mbh
        tb = torch.empty((n, nb * pcme.TREE_BLOCK_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
        cjobs += L.complexity_jobs(lr, [(0, None)] + [(t, t - 1) for t in range(1, n)], s)
        tjobs += L.tree_jobs(lr, [(0, None, None)] + [(t, t - 1, None) for t in range(1, n)], tb)
        lrs.append(lr)
        sums.append(s)
        tblocks.append(tb)
        tnodes.append([L.Node(tb[t], (t - 1 if t else None, None)) for t in range(n)])
        props.append(None)  # set by each step
        toffs.append(torch.empty(nb, dtype=torch.float32, device="cuda"))
        blocks_per_step += n * nb
    groups = {}
    for k, dc in enumerate(clips):
        groups.setdefault((dc.mbw, dc.mbh), []).append(k)
    # node tables and output buffers built once, like the job tables above
    tables = {}
    for (mbw, mbh), ks in groups.items():
        arr, n = L.tree_nodes([tnodes[k] for k in ks])
        buf = torch.empty((n, mbw * mbh), dtype=torch.float32, device="cuda")
        tables[(mbw, mbh)] = (arr, n, buf)
        base = 0
        for k in ks:
            props[k] = buf[base:base + clips[k].n]
            base += clips[k].n
    ev = []

    def lowres_all():
        for dc, lr in zip(clips, lrs):
            w, h = 16 * dc.mbw, 16 * dc.mbh
            pcme._check(pcme.load().pcme_lowres(dc.plane_ptr(0, 0), dc.pitch, dc.stride, pcme.LUMA_PAD, w, h,
                                                lr.data_ptr(), (w // 2) * (h // 2), dc.n, sp))

    def step(record):
        if record:
            e = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
            e[0].record(stream)
        lowres_all()
        if record:
            e[1].record(stream)
        L.run_complexity(cjobs, sp)
        if record:
            e[2].record(stream)
        L.run_analyse(tjobs, sp)
        if record:
            e[3].record(stream)
        # clips of one size go in one call (independent queues share launches)
        for (mbw, mbh), (arr, n, buf) in tables.items():
            L.run_propagate(arr, n, mbw, mbh, buf, sp)
        for k in range(len(clips)):
            nb = clips[k].mbw * clips[k].mbh
            pcme._check(pcme.load().pcme_mbtree_offsets(tblocks[k][0].data_ptr(), props[k][0].data_ptr(), nb,
                                                        toffs[k].data_ptr(), sp))
        if record:
            e[4].record(stream)
            ev.append(e)

    for dc in clips:
        dc.prepare(sp)
    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = [sum(e[i].elapsed_time(e[i + 1]) for e in ev) / len(ev) for i in range(4)]
    ms_step = 1e3 * dt / args.steps
    in_bytes = sum(dc.n * 256 * dc.mbw * dc.mbh for dc in clips)  # coded luma read once
    lowres_gbs = 1.25 * in_bytes / (ms[0] * 1e-3) / 1e9            # + 1/4 written
    out = {
        "workload": "LowRes::new + complexity (vs previous frame) + mbtree analyse (list 0 = previous frame) "
                    "of every frame + mbtree propagate over the frames as one queue + offsets of the first",
        "frames": sum(dc.n for dc in clips), "blocks_per_step": blocks_per_step,
        "ms_per_step": round(ms_step, 3), "mblocks_per_s": round(blocks_per_step / (ms_step * 1e-3) / 1e6, 4),
        "lowres_ms": round(ms[0], 4), "complexity_ms": round(ms[1], 4), "analyse_ms": round(ms[2], 4),
        "propagate_ms": round(ms[3], 4),
        "lowres_roofline": {"bound": "hbm", "kernel": "lowres_kernel", "achieved": round(lowres_gbs, 1),
                            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(lowres_gbs / HBM_PEAK_GBS, 4),
                            "algorithmic_bytes_per_block": 320},
    }
    print(f"[bench] lookahead: {ms_step:.3f} ms/step (lowres {ms[0]:.3f}, complexity {ms[1]:.3f}, "
          f"analyse {ms[2]:.3f}, propagate {ms[3]:.3f} ms)", file=sys.stderr, flush=True)
    if not args.no_cpu:
        from oracle import oracle as O
        dc, lr = clips[-1], lrs[-1]
        k = min(args.cpu_frames + 1, dc.n)
        frames = lr[:k].cpu().numpy()
        # repeated until about a second of wall time (~16 s of CPU work)
        secs, reps = 0.0, 0
        while secs < 1.0 and reps < 50:
            s1, csums, cblocks = O.lookahead_mt(frames, threads)
            secs += s1
            reps += 1
        nb = dc.mbw * dc.mbh
        cpu_val = reps * (k - 1) * nb / secs / 1e6
        gs = sums[-1][:k].cpu().numpy().astype(np.uint64)
        gb = tblocks[-1][:k].cpu().numpy().view(pcme.TREE_BLOCK_DTYPE).reshape(k, nb)
        parity = bool(np.array_equal(gs[1:], csums[1:]) and
                      np.array_equal(gb[1:].view(np.uint8), np.ascontiguousarray(cblocks[1:]).view(np.uint8)))
        out["cpu_baseline"] = {"value": round(cpu_val, 5), "unit": "Mblocks/s", "cores": threads, "kind": "port",
                               "sample": f"complexity + mbtree analyse of {k - 1} lowres frames "
                                         f"({8 * dc.mbw}x{8 * dc.mbh}) of the last clip, one frame per thread, "
                                         f"C oracle -O3 -march=x86-64-v3, {reps} repetitions",
                               "seconds": round(secs, 3)}
        # tree_offsets of the last clip: propagate is single-threaded in the
        # reference (mbtree.rs:167-223); repeated until about half a second
        gb_all = [b for b in tblocks[-1].cpu().numpy().view(pcme.TREE_BLOCK_DTYPE).reshape(dc.n, nb)]
        refs = [(t - 1 if t else -1, -1) for t in range(dc.n)]
        secs, reps = 0.0, 0
        while secs < 0.5 and reps < 200:
            t1 = time.perf_counter()
            cprop = O.mbtree_propagate(gb_all, refs, dc.mbw, dc.mbh)
            coff = O.mbtree_offsets(gb_all[0], cprop[0])
            secs += time.perf_counter() - t1
            reps += 1
        gp = props[-1].cpu().numpy()
        parity = parity and bool(np.array_equal(gp.view(np.uint32), cprop.view(np.uint32)) and
                                 np.allclose(toffs[-1].cpu().numpy(), coff, rtol=1e-6, atol=1e-5))
        out["propagate_cpu"] = {"ms": round(1e3 * secs / reps, 4), "cores": 1, "kind": "port",
                                "sample": f"propagate over the last clip's {dc.n} nodes + offsets of the first, "
                                          f"C oracle, {reps} repetitions"}
        out["parity_sample"] = parity
    return out


def b_leg(args, pcme, torch, dc, out, qp, me_range, stream, threads):
    """The B-picture search of frame 1 of the clip between frames 0 (list 0)
    and 2 (list 1), all 8,160 MBs in one pcme_b_analyse_batch call (job
    upload included).  The causal neighbours' motion is the plain P field of
    frame 1 for list 0 and its mirror for list 1 (every MB inter, reference
    0); the co-located MBs are the P field of frame 2.  The C oracle runs the
    same MBs on the host (o_b_frame_mt) and the records are compared."""
    import numpy as np
    from pcme import bsearch as B
    from oracle import oracle as O
    mbw, mbh = dc.mbw, dc.mbh
    f = out.cpu().numpy().view(pcme.MB_DTYPE).reshape(-1, mbh, mbw)
    mv0 = f[0]["mv"].astype(np.int16)
    mv = np.stack([mv0, -mv0])
    ref = np.zeros((2, mbh, mbw, 4), np.int8)
    intra = np.zeros((mbh, mbw), bool)
    col_mv = np.zeros((mbh, mbw, 2, 16, 2), np.int16)
    col_mv[:, :, 0] = f[1]["mv"]
    col_ref = np.zeros((mbh, mbw, 2, 4), np.int8)
    col_ref[:, :, 1] = -1
    jobs = B.b_jobs(dc, 1, (dc, 0), (dc, 2), mv, ref, intra, col_mv, col_ref, intra, qp, me_range, 1)
    nmb = mbw * mbh
    res = torch.empty(nmb * pcme.B_OUT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    sp = stream.cuda_stream
    for _ in range(max(1, args.warmup)):
        B.b_analyse(jobs, res.data_ptr(), sp)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = max(3, args.steps)
    a.record(stream)
    for _ in range(reps):
        B.b_analyse(jobs, res.data_ptr(), sp)
    b.record(stream)
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    got = res.cpu().numpy().view(pcme.B_OUT_DTYPE)
    o = {"workload": f"B picture {16 * mbw}x{16 * mbh} (frame 1 between 0 and 2), {nmb} MBs: direct_spatial + "
                     "direct cost, list-0/list-1 16x16 search_part, bi cost, b_partition 16x8/8x16",
         "mbs": nmb, "ms_per_picture": round(ms, 3), "mblocks_per_s": round(nmb / (ms * 1e-3) / 1e6, 4),
         "search_part_calls_per_mb": 6, "searches_per_mb": 10}
    print(f"[bench] B picture: {ms:.3f} ms ({nmb} MBs)", file=sys.stderr, flush=True)
    if not args.no_cpu:
        n = nmb
        mbs = np.zeros(n, [("kind", "i1"), ("ref_idx", "i1", (2, 4)), ("pad", "u1"), ("mv", "<i2", (2, 16, 2))])
        mbs["kind"] = O.OK_PINTER
        mbs["ref_idx"] = ref.transpose(1, 2, 0, 3).reshape(n, 2, 4)
        mbs["mv"] = mv.transpose(1, 2, 0, 3, 4).reshape(n, 2, 16, 2)
        cols = np.zeros_like(mbs)
        cols["kind"] = O.OK_PINTER
        cols["ref_idx"] = col_ref.reshape(n, 2, 4)
        cols["mv"] = col_mv.reshape(n, 2, 16, 2)
        assert mbs.dtype.itemsize == ctypes_sizeof_ombinfo2(O)
        planes = dc.planes[0, :3].cpu().numpy()
        r0, r1 = O.Ref(planes[0]), O.Ref(planes[2])
        prm = O.oparams(mbw, mbh, me_range, 1, 1, 0)
        secs, reps = 0.0, 0
        while secs < 1.0 and reps < 50:
            s1, want = O.b_frame_mt(prm, (O.ombinfo2 * n).from_buffer(mbs), (O.ombinfo2 * n).from_buffer(cols), r0,
                                    r1, planes[1], pcme.lam(qp), threads)
            secs += s1
            reps += 1
        o["cpu_baseline"] = {"value": round(reps * n / secs / 1e6, 5), "unit": "Mblocks/s", "cores": threads,
                             "kind": "port",
                             "sample": f"the same {n} MBs, MB rows over threads, C oracle -O3 -march=x86-64-v3, "
                                       f"{reps} repetitions",
                             "seconds": round(secs, 3)}
        o["parity_sample"] = bool(got.tobytes() == want.tobytes())
    o["field"] = b_field_leg(args, pcme, torch, dc, f, qp, me_range, stream)
    return o


def b_field_leg(args, pcme, torch, dc, fields, qp, me_range, stream):
    """The ME-only field of the same B picture as one dependent whole
    (pcme_b_field: each MB searched over the MBs committed before it, the
    rows walked as a wavefront in one launch), the co-located MBs the P field of
    frame 2; the C oracle's o_b_field (raster order, one thread: the
    picture is one dependent chain) times the same picture and the records
    are compared.  Then B_BATCH pictures of the clip (1, 3, 5, ..., each
    between its neighbours, the co-located MBs the P field of the next
    frame) in one pcme_b_fields launch, the walkers taking their rows side
    by side; picture 1's records are compared with the single field's."""
    import numpy as np
    from pcme import bsearch as B
    from oracle import oracle as O
    mbw, mbh = dc.mbw, dc.mbh
    field2 = fields[1]
    col = B.col_records(field2, mbh, mbw)
    n = mbw * mbh
    dcol = torch.as_tensor(np.ascontiguousarray(col.reshape(-1)).view(np.uint8)).to("cuda")
    dout = torch.empty(n * pcme.B_MB_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    sp = stream.cuda_stream

    def run():
        B.b_field_launch(dc, 1, (dc, 0), (dc, 2), dcol.data_ptr(), dout.data_ptr(), qp, me_range, 1, True, sp)

    for _ in range(max(1, args.warmup)):
        run()
    reps = max(3, args.steps)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        run()
    b.record(stream)
    torch.cuda.synchronize()
    B.b_field_check(sp)  # no timed launch stalled
    ms = a.elapsed_time(b) / reps
    got = dout.cpu().numpy().view(pcme.B_MB_DTYPE).reshape(mbh, mbw)
    diag = mbw + 2 * (mbh - 1)
    o = {"workload": f"ME-only field of the B picture, {n} MBs, one launch of row walkers (critical path {diag} MBs), "
                     "co-located MBs from the P field of frame 2",
         "ms_per_picture": round(ms, 3), "mblocks_per_s": round(n / (ms * 1e-3) / 1e6, 4),
         "modes": [int(x) for x in np.bincount(got["mode"].reshape(-1), minlength=4)]}
    print(f"[bench] B field: {ms:.3f} ms per picture", file=sys.stderr, flush=True)
    nb = min(B_BATCH, (dc.n - 1) // 2)
    ts = [1 + 2 * i for i in range(nb)]
    dcols = [torch.as_tensor(np.ascontiguousarray(B.col_records(fields[t], mbh, mbw).reshape(-1)).view(np.uint8))
             .to("cuda") for t in ts]
    douts = torch.empty((nb, n * pcme.B_MB_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
    bjobs = [B.b_field_job(dc, t, (dc, t - 1), (dc, t + 1), dcol.data_ptr(), qp, me_range, 1, True)
             for t, dcol in zip(ts, dcols)]
    optrs = [douts[i].data_ptr() for i in range(nb)]
    for _ in range(max(1, args.warmup)):
        B.b_fields_launch(bjobs, optrs, sp)
    a.record(stream)
    for _ in range(reps):
        B.b_fields_launch(bjobs, optrs, sp)
    b.record(stream)
    torch.cuda.synchronize()
    B.b_field_check(sp)
    bms = a.elapsed_time(b) / reps
    same = bool(douts[0].cpu().numpy().tobytes() == np.ascontiguousarray(got).tobytes())
    o["batch"] = {"workload": f"{nb} B pictures of the clip (frames {ts[0]}, {ts[1]}, ..., {ts[-1]}) in one "
                              "pcme_b_fields launch", "pictures": nb, "ms_per_launch": round(bms, 3),
                  "ms_per_picture": round(bms / nb, 3), "mblocks_per_s": round(nb * n / (bms * 1e-3) / 1e6, 4),
                  "speedup_vs_one_per_launch": round(ms * nb / bms, 2),
                  "picture_1_same_as_single_field": same}
    print(f"[bench] B fields: {nb} pictures in {bms:.3f} ms ({bms / nb:.3f} ms per picture)", file=sys.stderr,
          flush=True)
    if not args.no_cpu:
        cols = (O.ombinfo2 * n)()
        for i, r in enumerate(np.asarray(field2).reshape(-1)):
            cols[i].kind = O.OK_PINTER
            for q in range(4):
                cols[i].ref_idx[0][q] = int(r["ref"])
                cols[i].ref_idx[1][q] = -1
            for k in range(16):
                cols[i].mv[0][k] = O.omv(int(r["mv"][k][0]), int(r["mv"][k][1]))
        planes = dc.planes[0, :3].cpu().numpy()
        r0, r1 = O.Ref(planes[0]), O.Ref(planes[2])
        prm = O.oparams(mbw, mbh, me_range, 1, 1, 0)
        t0 = time.perf_counter()
        want = O.b_field(prm, cols, r0, r1, planes[1], pcme.lam(qp))
        secs = time.perf_counter() - t0
        o["cpu_baseline"] = {"value": round(n / secs / 1e6, 5), "unit": "Mblocks/s", "cores": 1, "kind": "port",
                             "sample": "the same B picture, C oracle o_b_field -O3 -march=x86-64-v3, one run",
                             "seconds": round(secs, 3)}
        o["parity_sample"] = bool(np.ascontiguousarray(got).tobytes() == want.tobytes())
    return o


def ctypes_sizeof_ombinfo2(O):
    import ctypes
    return ctypes.sizeof(O.ombinfo2)


def pruned_leg(args, pcme, torch, clips, outs, specs, stream, mbs_per_step, plain_ms_step):
    """The same frames through the CNN-pruned search: a step = pad + half-pel
    planes + the fused network inference (hints) + the pruned search; the
    plain search of the step above judges the result."""
    import numpy as np
    import pcme.pruner as P
    wpath = args.weights or P.WEIGHTS
    if not os.path.exists(wpath):
        return {"skipped": "no trained weights (pcme/pruner_weights.npz)"}
    sp = stream.cuda_stream
    thr_mv = args.thr_mv
    thr_rule = P.parse_thr(args.thr)
    nets, jobs, pouts, hints, groups, per_clip = {}, [], [], [], {}, []
    for dc, (spec, qp, rng) in zip(clips, specs):
        if qp not in nets:
            nets[qp] = P.Net(seed=1, max_batch=32, weights=wpath, qp=qp)
        nb, nmb = dc.n - 1, dc.mbw * dc.mbh
        o = torch.empty((nb, nmb * pcme.MB_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
        h = torch.zeros((nb, nmb * pcme.HINT_DTYPE.itemsize), dtype=torch.uint8, device="cuda")
        js = [pcme.frame_job(dc, t, t - 1, o[t - 1].data_ptr(), rng, 1, True, qp, hint_ptr=h[t - 1].data_ptr(),
                             older=older_refs(t, args.nref))
              for t in range(1, dc.n)]
        groups.setdefault(qp, []).extend(js)
        per_clip.append((qp, js))
        jobs += js
        pouts.append(o)
        hints.append(h)
    plan = pcme.Plan(jobs)

    def prepare():
        pcme.prepare_clips(clips, sp)

    def infer():
        for qp, js in groups.items():
            nets[qp].hints(js, args.ncand, P.thr_part_for(qp, thr_rule), thr_mv, sp)

    prepare()
    infer()
    if args.kernels_only:
        args.no_graph = args.no_pruned = args.no_lookahead = args.no_b = args.no_cpu = True
    tune = fix_or_tune(plan, args, sp)
    ev_h, ev_s = [], []

    def step(record):
        prepare()
        if record:
            a, b, c = (torch.cuda.Event(enable_timing=True) for _ in range(3))
            a.record(stream)
        infer()
        if record:
            b.record(stream)
        plan.launch(sp)
        if record:
            c.record(stream)
            ev_h.append((a, b))
            ev_s.append((b, c))

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    plan.check(sp)
    hints_ms = sum(a.elapsed_time(b) for a, b in ev_h) / len(ev_h)
    search_ms = sum(a.elapsed_time(b) for a, b in ev_s) / len(ev_s)
    ms_step = 1e3 * dt / args.steps
    serial_ms = ms_step
    overlap = None
    if args.overlap > 1:
        serial_out = [o.clone() for o in pouts]
        overlap = overlapped_pruned(args, pcme, torch, P, wpath, per_clip, thr_rule, stream, prepare, tune)
        torch.cuda.synchronize()
        overlap["records_identical"] = all(bool(torch.equal(a, b)) for a, b in zip(serial_out, pouts))
        if not overlap["records_identical"]:
            raise SystemExit("[bench] the overlapped pruned step's records differ from the serial step's")
        ms_step = overlap["ms_per_step"]
    plain = np.concatenate([o.cpu().numpy().view(pcme.MB_DTYPE).reshape(-1) for o in outs])
    pruned = np.concatenate([o.cpu().numpy().view(pcme.MB_DTYPE).reshape(-1) for o in pouts])
    hs = np.concatenate([h.cpu().numpy().view(pcme.HINT_DTYPE).reshape(-1) for h in hints])
    j = P.judge(plain[None], pruned[None])
    tflops = 2.0 * NET_MACS_PER_MB * mbs_per_step / (hints_ms * 1e-3) / 1e12
    for n in nets.values():
        n.close()
    print(f"[bench] pruned: {ms_step:.2f} ms/step (serial {serial_ms:.2f}: hints {hints_ms:.2f} ms, "
          f"search {search_ms:.2f} ms), "
          f"agreement {j['agreement']:.4f}, J-cost +{100 * j['jcost_increase']:.3f}%", file=sys.stderr, flush=True)
    return {
        "mblocks_per_s": round(mbs_per_step / (ms_step * 1e-3) / 1e6, 4),
        "ms_per_step": round(ms_step, 3),
        "speedup_vs_plain": round(plain_ms_step / ms_step, 3),
        "hints_ms": round(hints_ms, 3), "search_ms": round(search_ms, 3),
        # the network, then the search, one launch each
        "serial_ms_per_step": round(serial_ms, 3),
        "serial_speedup_vs_plain": round(plain_ms_step / serial_ms, 3),
        "overlap": overlap,
        "decision_agreement": round(j["agreement"], 5),
        "part_agreement": round(j["part_agreement"], 5),
        # hinted mode of the pruned MBs: 16x16 only / 16x8+8x16 (two modes) / 8x8
        "pruned_modes": [round(float(x), 4) for x in
                         np.array([1, 0, 0, 0, 0, 1, 1, 0, 0, 0, 0, 1]).reshape(3, 4) @
                         np.bincount(hs["part"][hs["prune"] != 0], minlength=4) /
                         max(int((hs["prune"] != 0).sum()), 1)],
        "jcost_increase": round(j["jcost_increase"], 6),
        "pruned_fraction": round(float((hs["prune"] != 0).mean()), 5),
        "ncand": args.ncand,
        "confidence_threshold": {str(q): P.thr_part_for(q, thr_rule) for q in sorted(nets)},
        "mv_mass_threshold": thr_mv,
        "autotune": tune,
        "search_kernel": search_kernel_name(tune, args.nref > 1, True),
        "net_roofline": {"bound": "mfma", "kernel": "hints_kernel", "achieved": round(tflops, 3),
                         "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 5), "macs_per_mb": NET_MACS_PER_MB},
        "bd_rate": measured_bd_rate(args.ncand, thr_rule, thr_mv, args.workload, args.nref),
    }


def overlapped_pruned(args, pcme, torch, P, wpath, per_clip, thr_rule, stream, prepare, tune):
    """The pruned step with the network and the search overlapped.  Every
    clip's P frames are cut into ``args.overlap`` contiguous ranges; chunk i
    holds range i of every clip and has its own network job table and frame
    plan.  The networks of all chunks run in order on the main stream; chunk
    i's search launches on a stream of its own as soon as chunk i's hints are
    done, so it runs beside the network of the later chunks and beside the
    other chunks' searches.  Every wait is on work enqueued before it, and
    the searches' walkers wait only on tickets of their own launch, so no
    launch can wait on another.  The records equal the serial step's: a
    frame's search reads only its own planes and hints."""
    k = args.overlap
    chunks = [{} for _ in range(k)]
    for qp, js in per_clip:
        n = len(js)
        for i in range(k):
            part = js[i * n // k:(i + 1) * n // k]
            if part:
                chunks[i].setdefault(qp, []).extend(part)
    chunks = [c for c in chunks if c]
    nets = {(qp, i): P.Net(seed=1, max_batch=32, weights=wpath, qp=qp)
            for i, c in enumerate(chunks) for qp in c}
    plans = [pcme.Plan([j for js in c.values() for j in js]) for c in chunks]
    for pl in plans:
        if tune:
            waves = max(64, int(tune["waves"] * args.overlap_grid / len(plans)))
            pl.set_config(waves, tune["waves_per_group"], tune["waves_per_cu"], tune["reuse_regions"])
    streams = [torch.cuda.Stream() for _ in chunks]
    sp = stream.cuda_stream

    def step():
        prepare()
        done = []
        for i, c in enumerate(chunks):
            for qp, js in c.items():
                nets[(qp, i)].hints(js, args.ncand, P.thr_part_for(qp, thr_rule), args.thr_mv, sp)
            e = torch.cuda.Event()
            e.record(stream)
            done.append(e)
        for pl, st, e in zip(plans, streams, done):
            st.wait_event(e)
            pl.launch(st.cuda_stream)
        for st in streams:
            stream.wait_stream(st)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    waves_of = [pl.get_config().waves for pl in plans]
    for pl, st in zip(plans, streams):
        pl.check(st.cuda_stream)
        pl.close()
    for n in nets.values():
        n.close()
    ms = 1e3 * dt / args.steps
    print(f"[bench] pruned, network and search overlapped in {len(chunks)} chunks: {ms:.2f} ms/step",
          file=sys.stderr, flush=True)
    return {"chunks": len(chunks), "ms_per_step": round(ms, 3),
            "waves_per_chunk": [pl_waves for pl_waves in waves_of]}


if __name__ == "__main__":
    main()
