"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the
benchmark's CPU-baseline leg, as the checker -- never by the product path.
The functions restate pito-codecs src/h264/enc/{dsp,picture,slice}.rs; see
the header of pcme_oracle.c for the line map and for what pins each one.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
LUMA_PAD = 64
CHROMA_PAD = 32

OK_NONE, OK_I4x4, OK_I8x8, OK_I16x16, OK_PSKIP, OK_PINTER = range(6)


class omv(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int16), ("y", ctypes.c_int16)]


class ombinfo(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int8), ("ref_idx", ctypes.c_int8 * 4), ("mv", omv * 16)]


class ombinfo2(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int8), ("ref_idx", (ctypes.c_int8 * 4) * 2), ("mv", (omv * 16) * 2)]


class oref(ctypes.Structure):
    _fields_ = [("pl", ctypes.c_void_p * 4), ("stride", ctypes.c_int), ("pad", ctypes.c_int)]


class oparams(ctypes.Structure):
    _fields_ = [("mbw", ctypes.c_int), ("mbh", ctypes.c_int), ("me_range", ctypes.c_int),
                ("subpel_rounds", ctypes.c_int), ("partitions", ctypes.c_int), ("vlim", ctypes.c_long)]


class ocounters(ctypes.Structure):
    _fields_ = [("sad_px", ctypes.c_uint64), ("satd_px", ctypes.c_uint64),
                ("sad_evals", ctypes.c_uint64), ("satd_evals", ctypes.c_uint64),
                ("searches", ctypes.c_uint64), ("mbs", ctypes.c_uint64)]


class ojob(ctypes.Structure):
    _fields_ = [("src", ctypes.c_void_p), ("ref", ctypes.c_void_p), ("out", ctypes.c_void_p),
                ("mbw", ctypes.c_int), ("mbh", ctypes.c_int), ("me_range", ctypes.c_int),
                ("subpel_rounds", ctypes.c_int), ("partitions", ctypes.c_int), ("qp", ctypes.c_int),
                ("nref", ctypes.c_int), ("older", ctypes.c_void_p * 3)]


# omb_out has the layout of pcme_mb_out
OMB_DTYPE = np.dtype([
    ("mv", "<i2", (16, 2)), ("mvs", "<i2", (4, 2)), ("mvps", "<i2", (4, 2)),
    ("cost", "<u4", (4,)), ("best_cost", "<u4"), ("skip_mv", "<i2", (2,)),
    ("mv16", "<i2", (2,)), ("part", "u1"), ("ref", "u1"), ("reserved", "u1", (2,)),
])

# otree_block has the layout of pcme_tree_block (mbtree.rs Block)
OTREE_DTYPE = np.dtype([("intra", "<u4"), ("inter", "<u4"), ("mv", "<i2", (2, 2)), ("list", "u1"),
                        ("reserved", "u1", (3,))])

# ob_out has the layout of pcme_b_out
OB_DTYPE = np.dtype([
    ("direct_mv", "<i2", (2, 4, 2)), ("direct_ref", "i1", (2, 4)), ("direct_ok", "<u4"), ("direct_cost", "<u4"),
    ("mv16", "<i2", (2, 2)), ("mvp16", "<i2", (2, 2)), ("cost16", "<u4", (2,)), ("cost", "<u4", (3,)),
    ("dir", "<u4"), ("part_mv", "<i2", (2, 2, 2, 2)), ("part_cost", "<u4", (2,)), ("part_dir", "u1", (2, 2)),
    ("reserved", "u1", (4,)),
])
assert OB_DTYPE.itemsize == 136
# ob_mb has the layout of pcme_b_mb
OBMB_DTYPE = np.dtype([("search", OB_DTYPE), ("mv", "<i2", (2, 16, 2)), ("ref", "i1", (2, 4)),
                       ("best_cost", "<u4"), ("mode", "u1"), ("dir", "u1"), ("reserved", "u1", (2,))])
assert OBMB_DTYPE.itemsize == 280

# ohint has the layout of pcme_hint
OHINT_DTYPE = np.dtype([
    ("cand", "<i2", (8, 2)), ("conf_part", "<u2"), ("conf_mv", "<u2"),
    ("part", "u1"), ("ncand", "u1"), ("prune", "u1"), ("reserved", "u1"),
])

# oskip_out has the layout of pcme_skip_out
OSKIP_DTYPE = np.dtype([("sad", "<u4"), ("thresh", "<u4"), ("mv_ok", "u1"), ("coded", "u1"), ("cbp", "u1"),
                        ("skip", "u1"), ("t8x8", "u1"), ("reserved", "u1", (3,))])
SKIP_T8X8, SKIP_ADAPTIVE_T8, SKIP_CABAC = 1, 2, 4

_lib = None


def build() -> None:
    subprocess.run(["make", "-B", "-C", HERE], check=True, stdout=subprocess.DEVNULL)


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    c = ctypes
    P = c.c_void_p
    lib.o_sad.restype = c.c_uint32
    lib.o_sad.argtypes = [P, c.c_int, P, c.c_int, c.c_int, c.c_int]
    lib.o_sad16.restype = c.c_uint32
    lib.o_sad16.argtypes = [P, c.c_int, P, c.c_int, c.c_int]
    lib.o_satd.restype = c.c_uint32
    lib.o_satd.argtypes = [P, c.c_int, P, c.c_int, c.c_int, c.c_int]
    lib.o_lambda.restype = c.c_uint32
    lib.o_lambda.argtypes = [c.c_int]
    lib.o_se_bits.restype = c.c_uint32
    lib.o_se_bits.argtypes = [c.c_int32]
    lib.o_mv_bits.restype = c.c_uint32
    lib.o_mv_bits.argtypes = [omv, omv]
    lib.o_median.restype = c.c_int16
    lib.o_median.argtypes = [c.c_int16, c.c_int16, c.c_int16]
    lib.o_import_plane.restype = None
    lib.o_import_plane.argtypes = [P, c.c_int, c.c_int, c.c_int, P, c.c_int, c.c_int, c.c_int, c.c_int]
    lib.o_extend_plane.restype = None
    lib.o_extend_plane.argtypes = [P, c.c_int, c.c_int, c.c_int, c.c_int]
    lib.o_hpel.restype = None
    lib.o_hpel.argtypes = [P, c.c_int, c.c_int, P, P, P]
    lib.o_mc_luma.restype = None
    lib.o_mc_luma.argtypes = [c.POINTER(oref), c.c_int, c.c_int, omv, c.c_int, c.c_int, P, c.c_int]
    lib.o_mc_chroma.restype = None
    lib.o_mc_chroma.argtypes = [P, P, c.c_int, c.c_int, c.c_int, c.c_int, omv, c.c_int, c.c_int, P, P, c.c_int]
    lib.o_vlim.restype = c.c_long
    lib.o_vlim.argtypes = [c.c_int, c.c_int]
    lib.o_mv_ok.restype = c.c_int
    lib.o_mv_ok.argtypes = [c.POINTER(oparams), c.c_int, c.c_int, c.c_int, c.c_int, omv]
    lib.o_nb_part.restype = c.c_int
    lib.o_nb_part.argtypes = [c.POINTER(oparams), P, c.POINTER(ombinfo), c.c_uint16, c.c_int, c.c_int,
                              c.c_int, c.c_int, c.POINTER(c.c_int8), c.POINTER(omv)]
    lib.o_mvp.restype = omv
    lib.o_mvp.argtypes = [c.POINTER(oparams), P, c.POINTER(ombinfo), c.c_uint16, c.c_int, c.c_int,
                          c.c_int, c.c_int, c.c_int, c.c_int8, c.c_int]
    lib.o_pskip_mv.restype = omv
    lib.o_pskip_mv.argtypes = [c.POINTER(oparams), P, c.c_int, c.c_int]
    lib.o_search.restype = omv
    lib.o_search.argtypes = [c.POINTER(oparams), c.POINTER(oref), c.c_int, c.c_int, c.c_int, c.c_int, P,
                             c.c_int, omv, c.POINTER(omv), c.c_int, c.c_uint32, c.POINTER(c.c_uint32),
                             c.POINTER(ocounters)]
    lib.o_search_part.restype = c.c_uint32
    lib.o_search_part.argtypes = [c.POINTER(oparams), P, c.c_int, c.c_int, c.POINTER(oref), c.c_int, P,
                                  c.c_uint32, omv, c.POINTER(omv), c.POINTER(omv), c.POINTER(ocounters)]
    lib.o_analyse_mb.restype = None
    lib.o_analyse_mb.argtypes = [c.POINTER(oparams), P, c.c_int, c.c_int, c.POINTER(oref), P, c.c_int,
                                 c.c_int, c.c_uint32, P, c.POINTER(ocounters)]
    lib.o_search_ex.restype = omv
    lib.o_search_ex.argtypes = [c.POINTER(oparams), c.POINTER(oref), c.c_int, c.c_int, c.c_int, c.c_int, P,
                                c.c_int, omv, c.POINTER(omv), c.c_int, c.POINTER(omv), c.c_int, c.c_uint32,
                                c.POINTER(c.c_uint32), c.POINTER(ocounters)]
    lib.o_search_frame_ex.restype = None
    lib.o_search_frame_ex.argtypes = [c.c_int, c.c_int, c.c_int, c.c_int, c.c_int, c.c_uint32, P, c.c_int,
                                      c.c_int, c.POINTER(oref), P, P, c.POINTER(ocounters)]
    lib.o_search_frame_mr.restype = None
    lib.o_search_frame_mr.argtypes = [c.c_int, c.c_int, c.c_int, c.c_int, c.c_int, c.c_uint32, P, c.c_int,
                                      c.c_int, c.POINTER(oref), c.c_int, P, P, c.POINTER(ocounters)]
    lib.o_ref_bits.restype = c.c_uint32
    lib.o_ref_bits.argtypes = [c.c_int, c.c_int]
    lib.o_search_frame.restype = None
    lib.o_search_frame.argtypes = [c.c_int, c.c_int, c.c_int, c.c_int, c.c_int, c.c_uint32, P, c.c_int,
                                   c.c_int, c.POINTER(oref), P, c.POINTER(ocounters)]
    lib.o_search_jobs_mt.restype = c.c_double
    lib.o_search_jobs_mt.argtypes = [c.POINTER(ojob), c.c_int, c.c_int, c.POINTER(ocounters)]
    lib.o_lowres.restype = None
    lib.o_lowres.argtypes = [P, c.c_int, c.c_int, c.c_int, c.c_int, P]
    lib.o_lowres_intra.restype = c.c_uint32
    lib.o_lowres_intra.argtypes = [P, c.c_int, c.c_int, c.c_int]
    lib.o_complexity.restype = None
    lib.o_complexity.argtypes = [P, P, c.c_int, c.c_int, c.POINTER(c.c_uint64), c.POINTER(c.c_uint64), P, P, P]
    lib.o_tree_search.restype = omv
    lib.o_tree_search.argtypes = [P, P, c.c_int, c.c_int, c.c_int, c.c_int, omv]
    lib.o_tree_analyse.restype = None
    lib.o_tree_analyse.argtypes = [P, P, P, c.c_int, c.c_int, P]
    lib.o_direct_spatial.restype = None
    lib.o_direct_spatial.argtypes = [c.POINTER(oparams), P, P, c.c_int, c.c_int, P, P]
    lib.o_b_analyse.restype = None
    lib.o_b_analyse.argtypes = [c.POINTER(oparams), P, P, c.c_int, c.c_int, c.POINTER(oref), c.POINTER(oref), P,
                                c.c_uint32, P]
    lib.o_b_field.restype = None
    lib.o_b_field.argtypes = [c.POINTER(oparams), P, c.POINTER(oref), c.POINTER(oref), P, c.c_int, c.c_int,
                              c.c_uint32, P]
    lib.o_b_frame_mt.restype = c.c_double
    lib.o_b_frame_mt.argtypes = [c.POINTER(oparams), P, P, c.POINTER(oref), c.POINTER(oref), P, c.c_int, c.c_int,
                                 c.c_uint32, c.c_int, P]
    lib.o_mbtree_propagate.restype = None
    lib.o_mbtree_propagate.argtypes = [c.POINTER(c.c_void_p), P, c.c_int, c.c_int, c.c_int, P]
    lib.o_mbtree_offsets.restype = None
    lib.o_mbtree_offsets.argtypes = [P, P, c.c_int, P]
    lib.o_block_var.restype = c.c_uint32
    lib.o_block_var.argtypes = [P, c.c_int, c.c_int, c.c_int]
    lib.o_aq_offsets.restype = None
    lib.o_aq_offsets.argtypes = [P, c.c_int, P, P, c.c_int, c.c_int, c.c_int, c.c_float, P]
    lib.o_chroma_qp.restype = c.c_int
    lib.o_chroma_qp.argtypes = [c.c_int]
    lib.o_lambda_ssd.restype = c.c_double
    lib.o_lambda_ssd.argtypes = [c.c_int]
    lib.o_quant_tables.restype = None
    lib.o_quant_tables.argtypes = [P]
    lib.o_fdct4.restype = None
    lib.o_fdct4.argtypes = [P, P]
    lib.o_decimate_score.restype = c.c_uint32
    lib.o_decimate_score.argtypes = [P, c.c_int]
    lib.o_code_luma4.restype = c.c_int
    lib.o_code_luma4.argtypes = [P, P, c.c_int, c.c_int, c.c_int, P, P]
    lib.o_code_luma8.restype = c.c_int
    lib.o_code_luma8.argtypes = [P, P, c.c_int, c.c_int, P, P]
    lib.o_code_chroma.restype = c.c_int
    lib.o_code_chroma.argtypes = [P, P, P, P, c.c_int, c.c_int, c.c_int, c.c_int]
    lib.o_code_inter.restype = c.c_int
    lib.o_code_inter.argtypes = [P, P, P, P, P, P, c.c_int, c.c_int, c.c_int, c.POINTER(c.c_int)]
    lib.o_skip_gate.restype = None
    lib.o_skip_gate.argtypes = [c.POINTER(oparams), c.POINTER(oref), P, P, c.c_int, c.c_int, c.c_int, omv, P, P, P,
                                c.c_int, c.c_int, c.c_int, P]
    lib.o_lookahead_mt.restype = c.c_double
    lib.o_lookahead_mt.argtypes = [P, c.c_int, c.c_int, c.c_int, c.c_int, P, P]
    _lib = lib
    return lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


# ------------------------------------------------------------ convenience

def pad_luma(frame: np.ndarray, mbw: int, mbh: int) -> np.ndarray:
    """Import (mod.rs:842-873) + extend (frame.rs:58-97) one luma plane."""
    lib = load()
    h, w = frame.shape
    stride, rows = 16 * mbw + 2 * LUMA_PAD, 16 * mbh + 2 * LUMA_PAD
    out = np.zeros((rows, stride), np.uint8)
    f = np.ascontiguousarray(frame)
    lib.o_import_plane(_p(f), w, w, h, _p(out), stride, LUMA_PAD, 16 * mbw, 16 * mbh)
    lib.o_extend_plane(_p(out), stride, LUMA_PAD, 16 * mbw, 16 * mbh)
    return out


def pad_chroma(frame: np.ndarray, mbw: int, mbh: int) -> np.ndarray:
    lib = load()
    h, w = frame.shape
    stride, rows = 8 * mbw + 2 * CHROMA_PAD, 8 * mbh + 2 * CHROMA_PAD
    out = np.zeros((rows, stride), np.uint8)
    f = np.ascontiguousarray(frame)
    lib.o_import_plane(_p(f), w, w, h, _p(out), stride, CHROMA_PAD, 8 * mbw, 8 * mbh)
    lib.o_extend_plane(_p(out), stride, CHROMA_PAD, 8 * mbw, 8 * mbh)
    return out


def hpel(plane: np.ndarray):
    """(H, V, C) half-pel planes of a padded plane (picture.rs:318-380)."""
    lib = load()
    rows, stride = plane.shape
    p = np.ascontiguousarray(plane)
    h = np.empty_like(p)
    v = np.empty_like(p)
    c = np.empty_like(p)
    lib.o_hpel(_p(p), stride, rows, _p(h), _p(v), _p(c))
    return h, v, c


class Ref:
    """A reference picture: padded plane + its half-pel planes, kept alive."""

    def __init__(self, padded: np.ndarray):
        self.planes = [np.ascontiguousarray(padded), *hpel(padded)]
        self.view = oref()
        for i, a in enumerate(self.planes):
            self.view.pl[i] = _p(a)
        self.view.stride = padded.shape[1]
        self.view.pad = LUMA_PAD


def search_frame(src_padded: np.ndarray, ref: Ref, mbw: int, mbh: int, me_range: int,
                 subpel_rounds: int, partitions: bool, qp: int, counters: ocounters | None = None,
                 hints: np.ndarray | None = None):
    """ME-only field of one frame; with `hints` (mbw*mbh OHINT_DTYPE records)
    the pruned search of the learned pruner."""
    lib = load()
    out = np.zeros(mbw * mbh, OMB_DTYPE)
    s = np.ascontiguousarray(src_padded)
    h = None
    if hints is not None:
        h = np.ascontiguousarray(hints.reshape(-1).view(OHINT_DTYPE))
        assert h.size == mbw * mbh
    lib.o_search_frame_ex(mbw, mbh, me_range, subpel_rounds, 1 if partitions else 0, lib.o_lambda(qp),
                          _p(s), s.shape[1], LUMA_PAD, ctypes.byref(ref.view), _p(h) if h is not None else None,
                          _p(out), ctypes.byref(counters) if counters is not None else None)
    return out.reshape(mbh, mbw)


def search_frame_mr(src_padded: np.ndarray, refs: list, mbw: int, mbh: int, me_range: int, subpel_rounds: int,
                    partitions: bool, qp: int, hints: np.ndarray | None = None,
                    counters: ocounters | None = None):
    """ME-only field of one frame over several references (refs[0] the
    newest, each a Ref): o_search_frame_mr."""
    lib = load()
    out = np.zeros(mbw * mbh, OMB_DTYPE)
    s = np.ascontiguousarray(src_padded)
    views = (oref * len(refs))(*[r.view for r in refs])
    h = None
    if hints is not None:
        h = np.ascontiguousarray(hints.reshape(-1).view(OHINT_DTYPE))
        assert h.size == mbw * mbh
    lib.o_search_frame_mr(mbw, mbh, me_range, subpel_rounds, 1 if partitions else 0, lib.o_lambda(qp), _p(s),
                          s.shape[1], LUMA_PAD, views, len(refs), _p(h) if h is not None else None, _p(out),
                          ctypes.byref(counters) if counters is not None else None)
    return out.reshape(mbh, mbw)


def search_clip_mr(y_frames: np.ndarray, nref: int, me_range: int, subpel_rounds: int, partitions: bool, qp: int,
                   hints: np.ndarray | None = None):
    """ME-only field of every P frame of an IPPP clip over the nref previous
    frames (fewer at the start: frame t has min(t, nref) references)."""
    n, h, w = y_frames.shape
    mbw, mbh = (w + 15) // 16, (h + 15) // 16
    padded = [pad_luma(y_frames[i], mbw, mbh) for i in range(n)]
    refs = [None] * n
    res = []
    for t in range(1, n):
        for k in range(max(0, t - nref), t):
            if refs[k] is None:
                refs[k] = Ref(padded[k])
        rl = [refs[t - 1 - k] for k in range(min(t, nref))]
        res.append(search_frame_mr(padded[t], rl, mbw, mbh, me_range, subpel_rounds, partitions, qp,
                                   None if hints is None else hints[t - 1]))
    return np.stack(res) if res else np.zeros((0, mbh, mbw), OMB_DTYPE)


def search_clip_me(y_frames: np.ndarray, me_range: int, subpel_rounds: int, partitions: bool, qp: int,
                   counters: ocounters | None = None):
    """ME-only field of every P frame of an IPPP clip, frame t vs t-1."""
    n, h, w = y_frames.shape
    mbw, mbh = (w + 15) // 16, (h + 15) // 16
    padded = [pad_luma(y_frames[i], mbw, mbh) for i in range(n)]
    res = []
    for t in range(1, n):
        ref = Ref(padded[t - 1])
        res.append(search_frame(padded[t], ref, mbw, mbh, me_range, subpel_rounds, partitions, qp, counters))
    return np.stack(res) if res else np.zeros((0, mbh, mbw), OMB_DTYPE)


def search_jobs_mt(srcs, refs, mbw, mbh, me_range, subpel_rounds, partitions, qp, threads, count=True,
                   older=None):
    """Frame-threaded CPU run (half-pel build + search per job); returns
    (results, wall_seconds, counters); with count False the run keeps no
    counters (a timed run) and the counters returned are zero.  older[i]:
    padded luma of job i's further references (indices 1, 2, ...)."""
    lib = load()
    n = len(srcs)
    outs = [np.zeros(mbw * mbh, OMB_DTYPE) for _ in range(n)]
    keep = [np.ascontiguousarray(a) for a in list(srcs) + list(refs)]
    olds = [[np.ascontiguousarray(a) for a in (older[i] if older else [])] for i in range(n)]
    jobs = (ojob * n)()
    for i in range(n):
        jobs[i] = ojob(_p(keep[i]), _p(keep[n + i]), _p(outs[i]), mbw, mbh, me_range, subpel_rounds,
                       1 if partitions else 0, qp, 1 + len(olds[i]))
        for k, a in enumerate(olds[i]):
            jobs[i].older[k] = _p(a)
    ctr = ocounters()
    secs = lib.o_search_jobs_mt(jobs, n, threads, ctypes.byref(ctr) if count else None)
    return [o.reshape(mbh, mbw) for o in outs], secs, ctr


# ------------------------------------------------ lookahead (lookahead.rs, mbtree.rs)

def lowres(plane_padded: np.ndarray, pw: int, ph: int, pad: int = LUMA_PAD) -> np.ndarray:
    """LowRes::new of a padded luma plane whose coded area is pw x ph."""
    p = np.ascontiguousarray(plane_padded, np.uint8)
    out = np.empty((ph // 2, pw // 2), np.uint8)
    load().o_lowres(_p(p), p.shape[1], pad, pw, ph, _p(out))
    return out


def lowres_intra(cur: np.ndarray, x0: int, y0: int) -> int:
    """mbtree::intra_cost of the 8x8 block at (x0, y0) of a lowres frame."""
    c = np.ascontiguousarray(cur, np.uint8)
    return int(load().o_lowres_intra(_p(c), c.shape[1], x0, y0))


def complexity(cur: np.ndarray, prev: np.ndarray | None):
    """lookahead::complexity -> (intra, inter, per-block intra, per-block
    inter, per-block MV [bh, bw, 2])."""
    c = np.ascontiguousarray(cur, np.uint8)
    h, w = c.shape
    nb = (w // 8) * (h // 8)
    bi = np.zeros(nb, np.uint32)
    bp = np.zeros(nb, np.uint32)
    mv = np.zeros((nb, 2), np.int16)
    a, b = ctypes.c_uint64(), ctypes.c_uint64()
    pp = None
    if prev is not None:
        pp = np.ascontiguousarray(prev, np.uint8)
        assert pp.shape == c.shape
    load().o_complexity(_p(c), _p(pp) if pp is not None else None, w, h, ctypes.byref(a), ctypes.byref(b),
                        _p(bi), _p(bp), _p(mv))
    return a.value, b.value, bi, bp, mv.reshape(h // 8, w // 8, 2)


def tree_search(cur: np.ndarray, ref: np.ndarray, x0: int, y0: int, seed=(0, 0)) -> tuple[int, int]:
    c = np.ascontiguousarray(cur, np.uint8)
    r = np.ascontiguousarray(ref, np.uint8)
    m = load().o_tree_search(_p(c), _p(r), c.shape[1], c.shape[0], x0, y0, omv(*seed))
    return m.x, m.y


def tree_analyse(cur: np.ndarray, l0: np.ndarray | None, l1: np.ndarray | None) -> np.ndarray:
    """mbtree::analyse -> [bh*bw] OTREE_DTYPE records."""
    c = np.ascontiguousarray(cur, np.uint8)
    h, w = c.shape
    out = np.zeros((w // 8) * (h // 8), OTREE_DTYPE)
    r0 = np.ascontiguousarray(l0, np.uint8) if l0 is not None else None
    r1 = np.ascontiguousarray(l1, np.uint8) if l1 is not None else None
    load().o_tree_analyse(_p(c), _p(r0) if r0 is not None else None, _p(r1) if r1 is not None else None, w, h,
                          _p(out))
    return out


def lookahead_mt(lr: np.ndarray, threads: int):
    """complexity and analyse (list 0 = previous frame) of frames 1..n-1 of
    [n, h, w] lowres frames on `threads` threads -> (seconds, sums [n, 2],
    blocks [n, bh*bw])."""
    a = np.ascontiguousarray(lr, np.uint8)
    n, h, w = a.shape
    sums = np.zeros((n, 2), np.uint64)
    blocks = np.zeros((n, (w // 8) * (h // 8)), OTREE_DTYPE)
    sec = load().o_lookahead_mt(_p(a), w, h, n, threads, _p(sums), _p(blocks))
    return sec, sums, blocks


# ------------------------------------------------------- B macroblocks (slice.rs)

def direct_spatial(prm: oparams, mbs, col: ombinfo2, mx: int, my: int):
    """direct_spatial -> (refs int8[2, 4], mvs int16[2, 4, 2])."""
    refs = np.zeros((2, 4), np.int8)
    mvs = np.zeros((2, 4, 2), np.int16)
    load().o_direct_spatial(ctypes.byref(prm), ctypes.cast(mbs, ctypes.c_void_p), ctypes.addressof(col), mx, my,
                            _p(refs), _p(mvs))
    return refs, mvs


def b_analyse(prm: oparams, mbs, col: ombinfo2, mx: int, my: int, r0: Ref, r1: Ref, src16: np.ndarray,
              lam: int) -> np.ndarray:
    """The motion search of analyse_b for one MB -> one OB_DTYPE record."""
    out = np.zeros(1, OB_DTYPE)
    s = np.ascontiguousarray(src16, np.uint8)
    load().o_b_analyse(ctypes.byref(prm), ctypes.cast(mbs, ctypes.c_void_p), ctypes.addressof(col), mx, my,
                       ctypes.byref(r0.view), ctypes.byref(r1.view), _p(s), lam, _p(out))
    return out[0]


def b_frame_mt(prm: oparams, mbs, cols, r0: Ref, r1: Ref, src_padded: np.ndarray, lam: int, threads: int):
    """o_b_analyse of every MB of a frame on `threads` threads -> (seconds,
    [mbh*mbw] OB_DTYPE records)."""
    out = np.zeros(prm.mbw * prm.mbh, OB_DTYPE)
    s = np.ascontiguousarray(src_padded, np.uint8)
    sec = load().o_b_frame_mt(ctypes.byref(prm), ctypes.cast(mbs, ctypes.c_void_p), ctypes.cast(cols, ctypes.c_void_p),
                              ctypes.byref(r0.view), ctypes.byref(r1.view), _p(s), s.shape[1], LUMA_PAD, lam,
                              threads, _p(out))
    return sec, out


def b_field(prm: oparams, cols, r0: Ref, r1: Ref, src_padded: np.ndarray, lam: int) -> np.ndarray:
    """o_b_field: the ME-only field of one B picture -> [mbh*mbw] OBMB_DTYPE
    records (cols: mbw*mbh ombinfo2 of the list-1 picture)."""
    out = np.zeros(prm.mbw * prm.mbh, OBMB_DTYPE)
    s = np.ascontiguousarray(src_padded, np.uint8)
    load().o_b_field(ctypes.byref(prm), ctypes.cast(cols, ctypes.c_void_p), ctypes.byref(r0.view),
                     ctypes.byref(r1.view), _p(s), s.shape[1], LUMA_PAD, lam, _p(out))
    return out


# -------------------------------------------------------- AQ (lookahead.rs)

def block_var(a: np.ndarray) -> int:
    b = np.ascontiguousarray(a, np.uint8)
    return int(load().o_block_var(_p(b), b.shape[1], b.shape[1], b.shape[0]))


def aq_offsets(y: np.ndarray, u: np.ndarray, v: np.ndarray, mbw: int, mbh: int, strength: float) -> np.ndarray:
    """aq_offsets of a picture given as its coded-area planes (y 16*mbh x
    16*mbw, u and v 8*mbh x 8*mbw) -> [mbh, mbw] float32."""
    yy = np.ascontiguousarray(y, np.uint8)
    uu = np.ascontiguousarray(u, np.uint8)
    vv = np.ascontiguousarray(v, np.uint8)
    out = np.zeros(mbw * mbh, np.float32)
    load().o_aq_offsets(_p(yy), yy.shape[1], _p(uu), _p(vv), uu.shape[1], mbw, mbh, strength, _p(out))
    return out.reshape(mbh, mbw)


# ------------------------------------------------ macroblock tree (mbtree.rs)

def mbtree_propagate(blocks: list[np.ndarray], refs: list[tuple[int, int]], bw: int, bh: int) -> np.ndarray:
    """propagate over nodes whose Costs.blocks are OTREE_DTYPE arrays
    (bh*bw each) and refs [(ref0, ref1)] (node index or -1) -> [n, bh*bw]
    float32."""
    n = len(blocks)
    arrs = [np.ascontiguousarray(b, OTREE_DTYPE) for b in blocks]
    ptrs = (ctypes.c_void_p * n)(*[_p(a) for a in arrs])
    r = np.array(refs, np.int32).reshape(n, 2)
    prop = np.zeros((n, bw * bh), np.float32)
    load().o_mbtree_propagate(ptrs, _p(r), n, bw, bh, _p(prop))
    return prop


def mbtree_offsets(blocks: np.ndarray, prop: np.ndarray) -> np.ndarray:
    b = np.ascontiguousarray(blocks, OTREE_DTYPE)
    p = np.ascontiguousarray(prop, np.float32)
    out = np.zeros(b.size, np.float32)
    load().o_mbtree_offsets(_p(b), _p(p), b.size, _p(out))
    return out


# ------------------------------------------- early P-skip gate (skip_oracle.c)

def quant_tables():
    """QuantTables::new -> dict of int32 arrays mf4/v4 [6,16], mf8/v8 [6,64]."""
    buf = np.zeros(6 * 16 * 2 + 6 * 64 * 2, np.int32)
    load().o_quant_tables(_p(buf))
    mf4, v4 = buf[:96].reshape(6, 16), buf[96:192].reshape(6, 16)
    mf8, v8 = buf[192:576].reshape(6, 64), buf[576:].reshape(6, 64)
    return {"mf4": mf4, "v4": v4, "mf8": mf8, "v8": v8}


def fdct4(d) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(d, np.int32).reshape(16))
    o = np.zeros(16, np.int32)
    load().o_fdct4(_p(a), _p(o))
    return o.reshape(4, 4)


def decimate_score(levels) -> int:
    a = np.ascontiguousarray(np.asarray(levels, np.int16))
    return int(load().o_decimate_score(_p(a), a.size))


def code_luma4(src, pred, qp, decimate=True, cabac=True):
    s = np.ascontiguousarray(src, np.uint8).reshape(256)
    p = np.ascontiguousarray(pred, np.uint8).reshape(256)
    lv = np.zeros((16, 16), np.int16)
    rec = np.zeros(256, np.uint8)
    cbp = load().o_code_luma4(_p(s), _p(p), qp, int(decimate), int(cabac), _p(lv), _p(rec))
    return cbp, lv, rec.reshape(16, 16)


def code_luma8(src, pred, qp, decimate=True):
    s = np.ascontiguousarray(src, np.uint8).reshape(256)
    p = np.ascontiguousarray(pred, np.uint8).reshape(256)
    lv = np.zeros((4, 64), np.int16)
    rec = np.zeros(256, np.uint8)
    cbp = load().o_code_luma8(_p(s), _p(p), qp, int(decimate), _p(lv), _p(rec))
    return cbp, lv, rec.reshape(16, 16)


def code_chroma(su, sv, pu, pv, qp, offset=0, intra=False, cabac=True) -> int:
    a = [np.ascontiguousarray(x, np.uint8).reshape(64) for x in (su, sv, pu, pv)]
    return int(load().o_code_chroma(*[_p(x) for x in a], qp, offset, int(intra), int(cabac)))


def code_inter(sy, su, sv, py, pu, pv, qp, flags, offset=0):
    """code_inter's (cbp, t8x8)."""
    a = [np.ascontiguousarray(x, np.uint8).reshape(-1) for x in (sy, su, sv, py, pu, pv)]
    t8 = ctypes.c_int()
    cbp = load().o_code_inter(*[_p(x) for x in a], qp, flags, offset, ctypes.byref(t8))
    return int(cbp), int(t8.value)


def skip_gate(prm: oparams, ref: Ref, ref_u: np.ndarray, ref_v: np.ndarray, mx: int, my: int, skip_mv,
              src_y, src_u, src_v, qp: int, flags: int, offset: int = 0) -> np.ndarray:
    """The early P-skip gate of one MB -> one OSKIP_DTYPE record.  ref_u /
    ref_v are the padded chroma planes of the reference."""
    out = np.zeros(1, OSKIP_DTYPE)
    u = np.ascontiguousarray(ref_u, np.uint8)
    v = np.ascontiguousarray(ref_v, np.uint8)
    sy = np.ascontiguousarray(src_y, np.uint8)
    su = np.ascontiguousarray(src_u, np.uint8)
    sv = np.ascontiguousarray(src_v, np.uint8)
    load().o_skip_gate(ctypes.byref(prm), ctypes.byref(ref.view), _p(u), _p(v), u.shape[1], mx, my, omv(*skip_mv),
                       _p(sy), _p(su), _p(sv), qp, flags, offset, _p(out))
    return out[0]
