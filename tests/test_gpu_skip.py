"""GPU parity of the early P-skip gate (pcme_skip_gate_batch) with the
oracle's restatement (oracle/skip_oracle.c, slice.rs:1884-1920 and the
code_inter cbp it calls): every field of every MB's outcome bit-exact, over
QPs 0-51, the 4x4 / 8x8 / adaptive transform choices, CAVLC and CABAC
level limits, chroma QP offsets, and skip MVs from zero through sub-pel
offsets to illegal ones past the padding.

Each case prints what it runs before the GPU step."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

T8, AD, CABAC = 1, 2, 4
FLAGS = (0, T8, T8 | AD, CABAC, T8 | AD | CABAC)
QPS = (0, 6, 12, 18, 22, 27, 32, 37, 44, 51)
FIELDS = ("sad", "thresh", "mv_ok", "coded", "cbp", "skip", "t8x8")


def say(*a):
    print("[case]", *a, flush=True)


class Scene:
    """A two-frame clip resident on the device (padded luma + half-pel
    planes + padded chroma) and the oracle's host copies of the same
    reference and source."""

    def __init__(self, gpu, oracle, w, h, noise, seed):
        import torch
        spec = gpu.ClipSpec(w, h, 2, nrect=6, rect_lo=16, rect_hi=80, vmax=3.0, pan_max=1.0,
                            noise_sigma=noise, seed=seed)
        y, u, v = gpu.synth_clip(spec)
        self.mbw, self.mbh = gpu.mb_dims(w, h)
        self.dc = gpu.DeviceClip(y, self.mbw, self.mbh, u, v)
        self.dc.prepare()
        torch.cuda.synchronize()
        self.ref = oracle.Ref(oracle.pad_luma(y[0], self.mbw, self.mbh))
        self.ref_u = oracle.pad_chroma(u[0], self.mbw, self.mbh)
        self.ref_v = oracle.pad_chroma(v[0], self.mbw, self.mbh)
        self.src_y = oracle.pad_luma(y[1], self.mbw, self.mbh)
        self.src_u = oracle.pad_chroma(u[1], self.mbw, self.mbh)
        self.src_v = oracle.pad_chroma(v[1], self.mbw, self.mbh)
        # the device's padded planes are the oracle's
        assert np.array_equal(self.dc.planes[0, 0].cpu().numpy(), self.ref.planes[0])
        assert np.array_equal(self.dc.cplanes[0, 0].cpu().numpy(), self.ref_u)
        assert np.array_equal(self.dc.cplanes[1, 1].cpu().numpy(), self.src_v)

    def mvs(self, rs, per_mb):
        """Skip MVs per MB: (0, 0), small sub-pel ones, and wide ones of
        which some leave the legal box."""
        out = []
        for my in range(self.mbh):
            for mx in range(self.mbw):
                c = [(0, 0)]
                for k in range(per_mb - 1):
                    span = 24 if k % 2 == 0 else 260
                    c.append(tuple(int(t) for t in rs.integers(-span, span + 1, 2)))
                out.append((mx, my, c))
        return out

    def jobs(self, gpu, oracle, cases, qp, flags, offset):
        dc, js, want = self.dc, [], []
        ys, cs = dc.stride, dc.cstride
        y1 = dc.planes[0, 1].data_ptr()
        u1, v1 = dc.cplanes[0, 1].data_ptr(), dc.cplanes[1, 1].data_ptr()
        for mx, my, mvl in cases:
            vlim = oracle.load().o_vlim(my, self.mbh)
            prm = oracle.oparams(self.mbw, self.mbh, 16, 1, 1, vlim)
            sy = self.src_y[64 + 16 * my:80 + 16 * my, 64 + 16 * mx:80 + 16 * mx]
            su = self.src_u[32 + 8 * my:40 + 8 * my, 32 + 8 * mx:40 + 8 * mx]
            sv = self.src_v[32 + 8 * my:40 + 8 * my, 32 + 8 * mx:40 + 8 * mx]
            for mv in mvl:
                j = gpu.SkipJob()
                j.src_y = y1 + (64 + 16 * my) * ys + 64 + 16 * mx
                j.src_u = u1 + (32 + 8 * my) * cs + 32 + 8 * mx
                j.src_v = v1 + (32 + 8 * my) * cs + 32 + 8 * mx
                j.ystride, j.cstride = ys, cs
                j.ref[:] = [dc.plane_ptr(k, 0) for k in range(4)]
                j.ref_u, j.ref_v = dc.cplanes[0, 0].data_ptr(), dc.cplanes[1, 0].data_ptr()
                j.mbw, j.mbh, j.mx, j.my = self.mbw, self.mbh, mx, my
                j.skip_mv = gpu.Mv(*mv)
                j.qp, j.flags, j.chroma_qp_offset, j.vlim = qp, flags, offset, vlim
                js.append(j)
                want.append(oracle.skip_gate(prm, self.ref, self.ref_u, self.ref_v, mx, my, mv, sy, su, sv, qp,
                                             flags, offset))
        return js, want


def run(gpu, jobs):
    import torch
    out = torch.zeros(len(jobs) * gpu.SKIP_OUT_DTYPE.itemsize, dtype=torch.uint8, device="cuda")
    gpu.skip_gate_batch(jobs, out.data_ptr())
    torch.cuda.synchronize()
    return out.cpu().numpy().view(gpu.SKIP_OUT_DTYPE)


def compare(got, want, label):
    for f in FIELDS:
        g = got[f]
        w = np.array([r[f] for r in want])
        bad = np.nonzero(g != w)[0]
        assert bad.size == 0, f"{label}: {f} differs at job {bad[0]}: gpu {g[bad[0]]} oracle {w[bad[0]]} " \
                              f"({bad.size} jobs)"


@pytest.fixture(scope="module")
def scenes(gpu, oracle):
    return [Scene(gpu, oracle, 320, 192, 0.5, 0x5EED0101), Scene(gpu, oracle, 176, 144, 2.0, 0x5EED0102)]


@pytest.mark.parametrize("flags", FLAGS)
def test_skip_gate_matches_oracle(gpu, oracle, scenes, flags):
    seen = {"skip": 0, "coded_cbp": 0, "not_coded": 0, "illegal": 0, "t8x8": 0, "t4x4": 0}
    for si, sc in enumerate(scenes):
        rs = np.random.default_rng(1000 * flags + si)
        for qp in QPS:
            offset = (0, -12, 12, 3)[qp % 4]
            cases = sc.mvs(rs, 3)
            jobs, want = sc.jobs(gpu, oracle, cases, qp, flags, offset)
            say(f"skip gate scene {si} qp {qp} flags {flags} offset {offset}: {len(jobs)} MBs")
            got = run(gpu, jobs)
            compare(got, want, f"scene {si} qp {qp} flags {flags}")
            seen["skip"] += int(got["skip"].sum())
            seen["illegal"] += int((got["mv_ok"] == 0).sum())
            seen["not_coded"] += int(((got["mv_ok"] == 1) & (got["coded"] == 0)).sum())
            cb = (got["coded"] == 1) & (got["cbp"] != 0)
            seen["coded_cbp"] += int(cb.sum())
            lum = cb & ((got["cbp"] & 15) != 0)
            seen["t8x8"] += int((lum & (got["t8x8"] == 1)).sum())
            seen["t4x4"] += int((lum & (got["t8x8"] == 0)).sum())
    # every outcome of the gate occurs (and the transform the flags allow)
    for k in ("skip", "coded_cbp", "not_coded", "illegal"):
        assert seen[k] > 0, (k, seen)
    if flags & T8:
        assert seen["t8x8"] > 0, seen
    if flags & AD or not flags & T8:
        assert seen["t4x4"] > 0, seen
    if not flags & T8:
        assert seen["t8x8"] == 0, seen


def test_skip_gate_two_streams_and_repeat(gpu, oracle, scenes):
    """Two batches in flight on two streams, then the first again: the
    per-call workspaces keep each call's job table."""
    import torch
    sc = scenes[0]
    rs = np.random.default_rng(7)
    ja, wa = sc.jobs(gpu, oracle, sc.mvs(rs, 2), 27, T8 | AD, 0)
    jb, wb = sc.jobs(gpu, oracle, sc.mvs(rs, 2), 37, CABAC, 2)
    say("skip gate, two streams,", len(ja), "+", len(jb), "MBs")
    n = gpu.SKIP_OUT_DTYPE.itemsize
    oa = torch.zeros(len(ja) * n, dtype=torch.uint8, device="cuda")
    ob = torch.zeros(len(jb) * n, dtype=torch.uint8, device="cuda")
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    gpu.skip_gate_batch(ja, oa.data_ptr(), stream=s1.cuda_stream)
    gpu.skip_gate_batch(jb, ob.data_ptr(), stream=s2.cuda_stream)
    torch.cuda.synchronize()
    compare(oa.cpu().numpy().view(gpu.SKIP_OUT_DTYPE), wa, "stream 1")
    compare(ob.cpu().numpy().view(gpu.SKIP_OUT_DTYPE), wb, "stream 2")
    compare(run(gpu, ja), wa, "repeat")
