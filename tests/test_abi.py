"""The C-ABI library: loads, exports every declared symbol, host-side
helpers and argument checks (no GPU needed)."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for hdr in ("pcme.h", "pcme_synth.h", "pcme_net.h"):
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(pcme_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(pcme_lib):
    lib = pcme_lib.load()
    names = declared_symbols()
    assert names, "no declarations parsed"
    assert names == set(pcme_lib.EXPORTS), names ^ set(pcme_lib.EXPORTS)
    for n in sorted(names):
        assert hasattr(lib, n), n


def test_version_and_lambda(pcme_lib, oracle):
    assert pcme_lib.load().pcme_version() == 1
    lib = oracle.load()
    for qp in range(52):
        assert pcme_lib.lam(qp) == lib.o_lambda(qp)


def test_qpel_table_matches_spec():
    """The packed table the kernels use equals H.264 8.4.2.2.1's choice of
    the two samples averaged for each quarter position (planes: 0 = G full,
    1 = b (H), 2 = h (V), 3 = j (C); offsets +1 in x or y mean the sample to
    the right / below)."""
    import pcme
    words = pcme.qpel_table()
    got = [(w >> (8 * i)) & 0xFF for w in words for i in range(4)]

    def enc(a, b):
        return a[0] | a[1] << 2 | a[2] << 3 | b[0] << 4 | b[1] << 6 | b[2] << 7

    G, H, V, C = 0, 1, 2, 3
    spec = {
        (0, 0): ((G, 0, 0), (G, 0, 0)),  # G
        (1, 0): ((G, 0, 0), (H, 0, 0)),  # a = (G + b + 1) >> 1
        (2, 0): ((H, 0, 0), (H, 0, 0)),  # b
        (3, 0): ((H, 0, 0), (G, 1, 0)),  # c = (H + b + 1) >> 1, H = G right
        (0, 1): ((G, 0, 0), (V, 0, 0)),  # d = (G + h + 1) >> 1
        (1, 1): ((H, 0, 0), (V, 0, 0)),  # e = (b + h + 1) >> 1
        (2, 1): ((H, 0, 0), (C, 0, 0)),  # f = (b + j + 1) >> 1
        (3, 1): ((H, 0, 0), (V, 1, 0)),  # g = (b + m + 1) >> 1, m = h right
        (0, 2): ((V, 0, 0), (V, 0, 0)),  # h
        (1, 2): ((V, 0, 0), (C, 0, 0)),  # i = (h + j + 1) >> 1
        (2, 2): ((C, 0, 0), (C, 0, 0)),  # j
        (3, 2): ((C, 0, 0), (V, 1, 0)),  # k = (j + m + 1) >> 1
        (0, 3): ((V, 0, 0), (G, 0, 1)),  # n = (M + h + 1) >> 1, M = G below
        (1, 3): ((V, 0, 0), (H, 0, 1)),  # p = (h + s + 1) >> 1, s = b below
        (2, 3): ((C, 0, 0), (H, 0, 1)),  # q = (j + s + 1) >> 1
        (3, 3): ((V, 1, 0), (H, 0, 1)),  # r = (m + s + 1) >> 1
    }
    for (fx, fy), (a, b) in spec.items():
        q = fy * 4 + fx
        assert got[q] == enc(a, b), f"q={q} ({fx},{fy})"


def test_synth_is_deterministic_and_in_range(pcme_lib):
    spec = pcme_lib.ClipSpec(64, 48, 3, nrect=2, rect_lo=8, rect_hi=20, vmax=3.0)
    a = pcme_lib.synth_clip(spec, threads=1)
    b = pcme_lib.synth_clip(spec, threads=3)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert a[0].shape == (3, 48, 64) and a[1].shape == (3, 24, 32)
    other = pcme_lib.synth_clip(pcme_lib.ClipSpec(64, 48, 3, nrect=2, rect_lo=8, rect_hi=20, seed=1))
    assert not np.array_equal(other[0], a[0])


def test_synth_rejects_bad_config(pcme_lib):
    with pytest.raises(pcme_lib.PcmeError):
        pcme_lib.synth_clip(pcme_lib.ClipSpec(63, 48, 1))


def test_compute_entry_points_fail_loudly_without_device(pcme_lib):
    """No silent fallback: without a gfx950 device the search refuses."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    lib = pcme_lib.load()
    job = pcme_lib.FrameJob()
    job.prm = pcme_lib.Params(1, 1, 16, 1, 1, 28)
    for f in ("src", "ref", "ref_h", "ref_v", "ref_c", "out"):
        setattr(job, f, 64)
    arr = (pcme_lib.FrameJob * 1)(job)
    rc = lib.pcme_search_frames(arr, 1, None)
    assert rc == pcme_lib.PCME_ENODEV, rc
    assert b"device" in lib.pcme_last_error()


def test_search_frames_rejects_bad_params(pcme_lib):
    lib = pcme_lib.load()
    job = pcme_lib.FrameJob()
    job.prm = pcme_lib.Params(1, 1, 0, 1, 1, 28)  # me_range 0
    for f in ("src", "ref", "ref_h", "ref_v", "ref_c", "out"):
        setattr(job, f, 64)
    arr = (pcme_lib.FrameJob * 1)(job)
    rc = lib.pcme_search_frames(arr, 1, None)
    assert rc in (pcme_lib.PCME_EINVAL, pcme_lib.PCME_ENODEV)


def test_pruner_without_device_fails_loudly(pcme_lib):
    """pcme_net_create reports PCME_ENODEV where no gfx950 device is visible."""
    if pcme_lib.device_count() > 0:
        pytest.skip("a device is visible")
    import pcme.pruner as P
    with pytest.raises(pcme_lib.PcmeError) as e:
        P.Net(seed=1, max_batch=64)
    assert e.value.code == pcme_lib.PCME_ENODEV


def test_lookahead_entry_points_check_arguments(pcme_lib):
    """Sizes that are not multiples of 8, misaligned frames and missing
    outputs are refused before any device work; well-formed calls without
    a gfx950 device report PCME_ENODEV."""
    lib = pcme_lib.load()
    C, T = pcme_lib.CplxJob, pcme_lib.TreeJob
    bad = [C(64, None, None, 64, 12, 8), C(66, None, None, 64, 8, 8), C(64, None, None, 0, 8, 8),
           C(64, None, None, 68, 8, 8), C(64, 64, 66, 64, 8, 8)]
    for j in bad:
        assert lib.pcme_complexity((C * 1)(j), 1, None) == pcme_lib.PCME_EINVAL
        assert b"pcme_complexity" in lib.pcme_last_error()
    assert lib.pcme_complexity(None, 1, None) == pcme_lib.PCME_EINVAL
    assert lib.pcme_complexity(None, 0, None) == pcme_lib.PCME_OK
    for j in [T(64, None, None, 0, 8, 8), T(64, 65, None, 64, 8, 8), T(64, None, None, 64, 8, 20)]:
        assert lib.pcme_mbtree_analyse((T * 1)(j), 1, None) == pcme_lib.PCME_EINVAL
    assert lib.pcme_lowres(64, 0, 64, 0, 24, 16, 64, 0, 1, None) == pcme_lib.PCME_EINVAL  # w % 16
    assert lib.pcme_lowres(66, 0, 64, 0, 16, 16, 64, 0, 1, None) == pcme_lib.PCME_EINVAL  # misaligned
    assert lib.pcme_lowres(64, 100, 64, 0, 16, 16, 64, 64, 2, None) == pcme_lib.PCME_EINVAL  # pitch < frame
    if pcme_lib.device_count() == 0:
        assert lib.pcme_complexity((C * 1)(C(64, None, None, 64, 8, 8)), 1, None) == pcme_lib.PCME_ENODEV
        assert lib.pcme_mbtree_analyse((T * 1)(T(64, None, None, 64, 8, 8)), 1, None) == pcme_lib.PCME_ENODEV
        assert lib.pcme_lowres(64, 0, 64, 0, 16, 16, 64, 0, 1, None) == pcme_lib.PCME_ENODEV


def test_search_part_batch_checks_arguments(pcme_lib):
    lib = pcme_lib.load()

    def job(**kw):
        j = pcme_lib.PartJob()
        j.src, j.ref, j.ref_h, j.ref_v, j.ref_c = 64, 64, 64, 64, 64
        j.sstride, j.mbw, j.mbh, j.me_range, j.subpel_rounds, j.vlim = 144, 4, 4, 16, 1, 4 * 16 + 44
        for k, v in kw.items():
            setattr(j, k, v)
        return (pcme_lib.PartJob * 1)(j)

    out = 64
    for bad in (job(part=4), job(src=None), job(vlim=16), job(vlim=4 * 16 + 45), job(mx=4), job(sstride=146),
                job(seed=pcme_lib.Mv(4096, 0))):
        assert lib.pcme_search_part_batch(bad, 1, out, None) == pcme_lib.PCME_EINVAL
    j = job()
    j[0].left[2].ref = -3
    assert lib.pcme_search_part_batch(j, 1, out, None) == pcme_lib.PCME_EINVAL
    assert lib.pcme_search_part_batch(job(), 1, 66, None) == pcme_lib.PCME_EINVAL  # misaligned out
    if pcme_lib.device_count() == 0:
        assert lib.pcme_search_part_batch(job(), 1, out, None) == pcme_lib.PCME_ENODEV


def test_skip_gate_batch_checks_arguments(pcme_lib):
    lib = pcme_lib.load()

    def job(**kw):
        j = pcme_lib.SkipJob()
        j.src_y, j.src_u, j.src_v, j.ref_u, j.ref_v = 64, 64, 64, 64, 64
        j.ref[:] = [64, 64, 64, 64]
        j.ystride, j.cstride, j.mbw, j.mbh, j.qp, j.vlim = 16, 8, 4, 4, 27, 4 * 16 + 44
        for k, v in kw.items():
            setattr(j, k, v)
        return (pcme_lib.SkipJob * 1)(j)

    out = 64
    for bad in (job(src_u=None), job(qp=52), job(qp=-1), job(flags=8), job(chroma_qp_offset=13), job(mx=4),
                job(my=-1), job(vlim=16), job(vlim=4 * 16 + 45), job(ystride=15), job(cstride=7)):
        assert lib.pcme_skip_gate_batch(bad, 1, out, None) == pcme_lib.PCME_EINVAL
    j = job()
    j[0].ref[2] = None
    assert lib.pcme_skip_gate_batch(j, 1, out, None) == pcme_lib.PCME_EINVAL
    assert lib.pcme_skip_gate_batch(job(), 1, 66, None) == pcme_lib.PCME_EINVAL  # misaligned out
    assert lib.pcme_skip_gate_batch(job(), 0, out, None) == pcme_lib.PCME_OK
    if pcme_lib.device_count() == 0:
        assert lib.pcme_skip_gate_batch(job(), 1, out, None) == pcme_lib.PCME_ENODEV


def test_b_analyse_batch_checks_arguments(pcme_lib):
    lib = pcme_lib.load()

    def job(**kw):
        j = pcme_lib.BJob()
        j.src, j.sstride, j.mbw, j.mbh, j.me_range, j.subpel_rounds, j.vlim = 64, 144, 4, 4, 16, 1, 4 * 16 + 44
        for k in range(4):
            j.ref0[k] = 64
            j.ref1[k] = 128
        for k, v in kw.items():
            setattr(j, k, v)
        return (pcme_lib.BJob * 1)(j)

    for bad in (job(src=None), job(sstride=18), job(my=4), job(vlim=10), job(me_range=-1)):
        assert lib.pcme_b_analyse_batch(bad, 1, 64, None) == pcme_lib.PCME_EINVAL
    j = job()
    j[0].ref1[2] = None
    assert lib.pcme_b_analyse_batch(j, 1, 64, None) == pcme_lib.PCME_EINVAL
    j = job()
    j[0].col_ref[1][3] = -2
    assert lib.pcme_b_analyse_batch(j, 1, 64, None) == pcme_lib.PCME_EINVAL
    j = job()
    j[0].nb[1][9].ref = -5
    assert lib.pcme_b_analyse_batch(j, 1, 64, None) == pcme_lib.PCME_EINVAL
    if pcme_lib.device_count() == 0:
        assert lib.pcme_b_analyse_batch(job(), 1, 64, None) == pcme_lib.PCME_ENODEV


def test_b_fields_check_arguments(pcme_lib):
    """pcme_b_field / pcme_b_fields refuse bad pictures and outputs before
    touching a device; no pictures is a no-op."""
    import ctypes
    lib = pcme_lib.load()

    def job(**kw):
        j = pcme_lib.BFieldJob()
        j.src, j.col, j.mbw, j.mbh, j.qp, j.me_range, j.subpel_rounds, j.partitions = 64, 128, 4, 3, 27, 16, 1, 1
        for k in range(4):
            j.ref0[k] = 256
            j.ref1[k] = 512
        for k, v in kw.items():
            setattr(j, k, v)
        return j

    def fields(*jobs, outs=None):
        arr = (pcme_lib.BFieldJob * max(1, len(jobs)))(*jobs)
        o = (ctypes.c_void_p * max(1, len(jobs)))(*(outs if outs is not None else [1024] * len(jobs)))
        return lib.pcme_b_fields(arr, len(jobs), o, None)

    assert lib.pcme_b_fields(None, 0, None, None) == pcme_lib.PCME_OK
    assert lib.pcme_b_fields(None, -1, None, None) == pcme_lib.PCME_EINVAL
    assert lib.pcme_b_fields(None, 1, None, None) == pcme_lib.PCME_EINVAL
    for bad in (job(src=None), job(col=None), job(mbw=0), job(mbh=5000), job(qp=52), job(me_range=-1),
                job(subpel_rounds=9), job(src=66)):
        assert fields(job(), bad) == pcme_lib.PCME_EINVAL
        assert lib.pcme_b_field(ctypes.byref(bad), 1024, None) == pcme_lib.PCME_EINVAL
    j = job()
    j.ref1[3] = None
    assert fields(j) == pcme_lib.PCME_EINVAL
    assert fields(job(), job(), outs=[1024, None]) == pcme_lib.PCME_EINVAL
    assert fields(job(), outs=[1028]) == pcme_lib.PCME_EINVAL  # records are 8-byte aligned
    assert fields(*[job()] * 1025) == pcme_lib.PCME_EINVAL  # at most 1024 pictures
    assert fields(job(mbw=4096, mbh=4096)) == pcme_lib.PCME_EINVAL  # at most 2^22 MBs
    if pcme_lib.device_count() == 0:
        assert fields(job(), job(mbw=7)) == pcme_lib.PCME_ENODEV


def test_aq_offsets_checks_arguments(pcme_lib):
    lib = pcme_lib.load()
    if pcme_lib.device_count() == 0:
        assert lib.pcme_aq_offsets(64, 128, 64, 64, 64, 4, 4, 0.6, 64, None) == pcme_lib.PCME_ENODEV
    assert lib.pcme_aq_offsets(66, 128, 64, 64, 64, 4, 4, 0.6, 64, None) == pcme_lib.PCME_EINVAL  # misaligned y
    assert lib.pcme_aq_offsets(64, 32, 64, 64, 64, 4, 4, 0.6, 64, None) == pcme_lib.PCME_EINVAL  # stride < 16*mbw
    assert lib.pcme_aq_offsets(64, 128, 64, None, 64, 4, 4, 0.6, 64, None) == pcme_lib.PCME_EINVAL


def test_mbtree_propagate_and_offsets_check_arguments(pcme_lib):
    lib = pcme_lib.load()
    TN = pcme_lib.TreeNode
    assert ctypes.sizeof(TN) == 16
    nodes = (TN * 2)()
    nodes[0].blocks, nodes[1].blocks = 64, 128
    nodes[0].ref[:] = [-1, -1]
    nodes[1].ref[:] = [0, -1]
    if pcme_lib.device_count() == 0:
        assert lib.pcme_mbtree_propagate(nodes, 2, 4, 4, 256, None) == pcme_lib.PCME_ENODEV
        assert lib.pcme_mbtree_offsets(64, 256, 16, 512, None) == pcme_lib.PCME_ENODEV
    assert lib.pcme_mbtree_propagate(nodes, 0, 4, 4, None, None) == pcme_lib.PCME_OK
    assert lib.pcme_mbtree_propagate(None, 2, 4, 4, 256, None) == pcme_lib.PCME_EINVAL
    assert lib.pcme_mbtree_propagate(nodes, -1, 4, 4, 256, None) == pcme_lib.PCME_EINVAL
    assert lib.pcme_mbtree_propagate(nodes, 2, 0, 4, 256, None) == pcme_lib.PCME_EINVAL    # empty grid
    assert lib.pcme_mbtree_propagate(nodes, 2, 4, 4, None, None) == pcme_lib.PCME_EINVAL   # no prop
    assert lib.pcme_mbtree_propagate(nodes, 2, 4, 4, 258, None) == pcme_lib.PCME_EINVAL    # misaligned prop
    nodes[1].ref[1] = -2
    assert lib.pcme_mbtree_propagate(nodes, 2, 4, 4, 256, None) == pcme_lib.PCME_EINVAL    # ref below -1
    nodes[1].ref[1] = -1
    nodes[1].blocks = None
    assert lib.pcme_mbtree_propagate(nodes, 2, 4, 4, 256, None) == pcme_lib.PCME_EINVAL    # null blocks
    assert b"node 1" in lib.pcme_last_error()
    assert lib.pcme_mbtree_offsets(64, 256, 0, 512, None) == pcme_lib.PCME_OK
    assert lib.pcme_mbtree_offsets(None, 256, 16, 512, None) == pcme_lib.PCME_EINVAL
    assert lib.pcme_mbtree_offsets(64, 256, -1, 512, None) == pcme_lib.PCME_EINVAL
    assert lib.pcme_mbtree_offsets(64, 257, 16, 512, None) == pcme_lib.PCME_EINVAL


def test_propagate_queues_rejects_short_nodes(pcme_lib):
    """A node whose blocks tensor is shorter than the grid is refused on the
    host, before any device call."""
    import torch
    from pcme import lookahead as L
    short = torch.zeros(19 * 20, dtype=torch.uint8)
    with pytest.raises(pcme_lib.PcmeError):
        L.propagate_queues([[L.Node(short)]], 5, 4)


def test_tree_nodes_lays_out_independent_queues(pcme_lib):
    """tree_nodes concatenates queues and shifts their refs; a ref at or
    after its node, past its queue, or None becomes -1."""
    import torch
    from pcme import lookahead as L
    b = torch.zeros(8 * 20, dtype=torch.uint8)
    q0 = [L.Node(b), L.Node(b, (0, None)), L.Node(b, (1, 0))]
    q1 = [L.Node(b, (None, 1)), L.Node(b, (0, 5)), L.Node(b, (1, 0))]
    arr, n = L.tree_nodes([q0, q1])
    assert n == 6
    refs = [tuple(arr[k].ref) for k in range(n)]
    assert refs == [(-1, -1), (0, -1), (1, 0), (-1, -1), (3, -1), (4, 3)]
    assert all(arr[k].blocks == b.data_ptr() for k in range(n))
