"""Pin the oracle's predictor and search against an independent
restatement (tests/pysearch.py), committed golden vectors and invariants."""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

import pysearch
from _util import assert_same_field

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ------------------------------------------------------------------- MVP

def spec_mvp(nbs, cur, done, x, y, w, rf, shape, part, mx, my, mbw):
    """H.264 8.4.1.3 in the decoder's structure (dec/slicedec.rs:581-681):
    neighbours A, B, C (else D); intra or unavailable-ref neighbours carry
    refIdx -1; directional 16x8/8x16 rules; single-match rule; median."""
    def nb(bx, by):
        if 0 <= bx < 4 and 0 <= by < 4:
            r = by * 4 + bx
            if r not in done:
                return (False, -1, (0, 0))
            return (True, cur["ref"][(by >> 1) * 2 + (bx >> 1)], cur["mv"][r])
        if bx < 0 and by < 0:
            key, b = (mx - 1, my - 1), 15
        elif bx < 0 and by <= 3:
            key, b = (mx - 1, my), by * 4 + 3
        elif bx > 3 and by < 0:
            key, b = (mx + 1, my - 1), 12
        elif 0 <= bx <= 3 and by < 0:
            key, b = (mx, my - 1), 12 + bx
        else:
            return (False, -1, (0, 0))
        if key[0] < 0 or key[1] < 0 or key[0] >= mbw:
            return (False, -1, (0, 0))
        m = nbs[key]
        if m["intra"]:
            return (True, -1, (0, 0))
        ref = m["ref"][((b >> 2) >> 1) * 2 + ((b & 3) >> 1)]
        return (True, ref, m["mv"][b] if ref >= 0 else (0, 0))

    a, b, c = nb(x - 1, y), nb(x, y - 1), nb(x + w, y - 1)
    if not c[0]:
        c = nb(x - 1, y - 1)
    if shape == 1 and part == 0 and b[1] == rf:
        return b[2]
    if shape == 1 and part == 1 and a[1] == rf:
        return a[2]
    if shape == 2 and part == 0 and a[1] == rf:
        return a[2]
    if shape == 2 and part == 1 and c[1] == rf:
        return c[2]
    if not b[0] and not c[0] and a[0]:
        b = c = a
    m = [a[1] == rf, b[1] == rf, c[1] == rf]
    if sum(m) == 1:
        return (a, b, c)[m.index(True)][2]
    med = lambda p, q, r: sorted([p, q, r])[1]  # noqa: E731
    return (med(a[2][0], b[2][0], c[2][0]), med(a[2][1], b[2][1], c[2][1]))


def test_mvp_matches_spec_random_neighbourhoods(oracle):
    lib = oracle.load()
    rs = np.random.default_rng(11)
    mbw, mbh = 4, 3
    shapes = {(1, 0): 1, (1, 1): 2, (2, 0): 3, (2, 1): 4}
    n_checked = 0
    for trial in range(400):
        mbs = (oracle.ombinfo * (mbw * mbh))()
        nbs = {}
        for i in range(mbw * mbh):
            kind = int(rs.choice([oracle.OK_PINTER, oracle.OK_PINTER, oracle.OK_PSKIP, oracle.OK_I16x16,
                                  oracle.OK_I4x4]))
            mbs[i].kind = kind
            refs = [int(rs.choice([-1, 0, 0, 0, 1])) for _ in range(4)]
            mvs = [(int(rs.integers(-64, 65)), int(rs.integers(-64, 65))) for _ in range(16)]
            if rs.random() < 0.2:
                mvs = [(0, 0)] * 16
            for k in range(4):
                mbs[i].ref_idx[k] = refs[k]
            for k in range(16):
                mbs[i].mv[k] = oracle.omv(*mvs[k])
            nbs[(i % mbw, i // mbw)] = dict(intra=kind in (oracle.OK_I4x4, oracle.OK_I16x16), ref=refs, mv=mvs)
        mx, my = int(rs.integers(0, mbw)), int(rs.integers(0, mbh))
        p = oracle.oparams(mbw, mbh, 16, 1, 1, 10000)
        # every partition of every mode, with the mask of the partitions before it
        for mode, parts in pysearch.PARTS.items():
            cur = oracle.ombinfo()
            cref = [-1] * 4
            cmv = [(0, 0)] * 16
            for k in range(4):
                cur.ref_idx[k] = -1
            done = set()
            for q, (x, y, w, h) in enumerate(parts):
                mask = 0
                for r in done:
                    mask |= 1 << r
                shape = shapes.get((mode, q), 0)
                rf = 0
                got = lib.o_mvp(ctypes.byref(p), ctypes.cast(mbs, ctypes.c_void_p), ctypes.byref(cur), mask,
                                mx, my, x // 4, y // 4, w // 4, rf, shape)
                want = spec_mvp(nbs, dict(ref=cref, mv=cmv), done, x // 4, y // 4, w // 4, rf,
                                {0: 0, 1: 1, 2: 1, 3: 2, 4: 2}[shape], q, mx, my, mbw)
                assert (got.x, got.y) == tuple(want), (trial, mode, q, mx, my)
                n_checked += 1
                newmv = (int(rs.integers(-64, 65)), int(rs.integers(-64, 65)))
                for by in range(y // 4, (y + h) // 4):
                    for bx in range(x // 4, (x + w) // 4):
                        r = by * 4 + bx
                        done.add(r)
                        cmv[r] = newmv
                        cur.mv[r] = oracle.omv(*newmv)
                for by8 in range(y // 8, (y + h + 7) // 8):
                    for bx8 in range(x // 8, (x + w + 7) // 8):
                        cref[by8 * 2 + bx8] = 0
                        cur.ref_idx[by8 * 2 + bx8] = 0
        # pskip_mv (8.4.1.1): A or B unavailable -> 0; A or B with ref 0 and
        # zero MV -> 0; else the 16x16 prediction.
        got = lib.o_pskip_mv(ctypes.byref(p), ctypes.cast(mbs, ctypes.c_void_p), mx, my)
        empty = dict(ref=[-1] * 4, mv=[(0, 0)] * 16)
        if mx == 0 or my == 0:
            want = (0, 0)
        else:
            a = nbs[(mx - 1, my)]
            b = nbs[(mx, my - 1)]
            a_ref = -1 if a["intra"] else a["ref"][1]  # block 3 sits in 8x8 #1
            b_ref = -1 if b["intra"] else b["ref"][2]  # block 12 sits in 8x8 #2
            a_mv = a["mv"][3] if a_ref >= 0 else (0, 0)
            b_mv = b["mv"][12] if b_ref >= 0 else (0, 0)
            if (a_ref == 0 and a_mv == (0, 0)) or (b_ref == 0 and b_mv == (0, 0)):
                want = (0, 0)
            else:
                want = spec_mvp(nbs, empty, set(), 0, 0, 4, 0, 0, 0, mx, my, mbw)
        assert (got.x, got.y) == tuple(want), ("pskip", trial, mx, my)
    assert n_checked == 400 * 9


# ---------------------------------------------------------------- search

def _frames(seed, w, h, n, shift=(3, -2)):
    """Deterministic frames from numpy alone: smooth random texture shifted
    by an integer step per frame plus mild noise."""
    rs = np.random.default_rng(seed)
    base = rs.integers(0, 256, (h + 64, w + 64)).astype(float)
    k = np.ones(5) / 5
    for ax in (0, 1):
        base = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), ax, base)
    out = []
    for t in range(n):
        oy, ox = 32 + shift[1] * t, 32 + shift[0] * t
        f = base[oy:oy + h, ox:ox + w] + rs.normal(0, 2, (h, w))
        out.append(np.clip(np.round(f), 0, 255).astype(np.uint8))
    return np.stack(out)


CASES = [
    # (seed, w, h, frames, me_range, subpel, partitions, qp)
    (1, 48, 32, 3, 16, 1, True, 28),
    (2, 64, 48, 2, 8, 0, True, 22),
    (3, 40, 40, 2, 4, 2, False, 37),
    (4, 32, 64, 2, 24, 1, True, 32),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"s{c[0]}_{c[1]}x{c[2]}")
def test_oracle_matches_independent_restatement(oracle, case):
    seed, w, h, n, rng, sub, parts, qp = case
    y = _frames(seed, w, h, n)
    got = oracle.search_clip_me(y, rng, sub, parts, qp)
    for t in range(1, n):
        rows = pysearch.Frame(y[t], y[t - 1], rng, sub, parts, qp).run()
        want = pysearch.to_records(rows, oracle.OMB_DTYPE)
        assert_same_field(got[t - 1], want, f"frame {t}")


def test_golden_vectors(oracle):
    """Committed fixtures (tests/golden/make_golden.py): inputs + expected
    ME-only fields from the independent restatement."""
    files = sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))
    assert files, "no golden vectors"
    for f in files:
        z = np.load(os.path.join(GOLDEN, f))
        frames = z["frames"]
        prm = z["params"]
        got = oracle.search_clip_me(frames, int(prm[0]), int(prm[1]), bool(prm[2]), int(prm[3]))
        want = z["expected"].view(oracle.OMB_DTYPE).reshape(got.shape)
        assert_same_field(got, want, f)


def test_search_invariants(oracle):
    """cost == SATD(mc(best)) + (lam*bits)>>8 exactly; best is legal; the
    integer phase stays within centre0 +- (range + 2)."""
    lib = oracle.load()
    y = _frames(9, 64, 48, 2, shift=(5, 2))
    mbw, mbh = 4, 3
    ref = oracle.Ref(oracle.pad_luma(y[0], mbw, mbh))
    src = oracle.pad_luma(y[1], mbw, mbh)
    rs = np.random.default_rng(3)
    for trial in range(200):
        w, h = [(16, 16), (16, 8), (8, 16), (8, 8)][trial % 4]
        bx = int(rs.integers(0, 64 - w + 1)) // 4 * 4
        by = int(rs.integers(0, 48 - h + 1)) // 4 * 4
        p = oracle.oparams(mbw, mbh, int(rs.choice([2, 8, 16])), int(rs.integers(0, 3)), 1,
                           int(lib.o_vlim(by // 16, mbh)))
        mvp = oracle.omv(int(rs.integers(-40, 41)), int(rs.integers(-40, 41)))
        cands = (oracle.omv * 2)(oracle.omv(int(rs.integers(-40, 41)), int(rs.integers(-40, 41))), oracle.omv(0, 0))
        lam = int(lib.o_lambda(int(rs.integers(10, 45))))
        blk = np.ascontiguousarray(src[64 + by:64 + by + h, 64 + bx:64 + bx + w])
        cost = ctypes.c_uint32()
        m = lib.o_search(ctypes.byref(p), ctypes.byref(ref.view), bx, by, w, h, blk.ctypes.data, w, mvp, cands,
                         2, lam, ctypes.byref(cost), None)
        assert lib.o_mv_ok(ctypes.byref(p), bx, by, w, h, m)
        pred = np.zeros((16, 16), np.uint8)
        lib.o_mc_luma(ctypes.byref(ref.view), bx, by, m, w, h, pred.ctypes.data, 16)
        d = blk.astype(int) - pred[:h, :w].astype(int)
        want = pysearch.satd(d) + ((lam * lib.o_mv_bits(m, mvp)) >> 8)
        assert cost.value == want


def test_search_recovers_integer_shift(oracle):
    """A pure integer translation of a textured picture with no noise is
    found exactly (cost-surface minimum at the true shift, zero residual)."""
    rs = np.random.default_rng(4)
    base = rs.integers(0, 256, (200, 200)).astype(float)
    k = np.ones(7) / 7
    for ax in (0, 1):
        base = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), ax, base)
    base = np.round(base).astype(np.uint8)
    for dx, dy in [(3, -2), (-5, 1), (0, 4), (7, 6)]:
        f0 = base[50:50 + 64, 50:50 + 80]
        f1 = base[50 + dy:50 + dy + 64, 50 + dx:50 + dx + 80]
        res = oracle.search_clip_me(np.stack([f0, f1]), 16, 1, True, 20)[0]
        inner = res[1:-1, 1:-1]
        assert (inner["mv16"][..., 0] == 4 * dx).all() and (inner["mv16"][..., 1] == 4 * dy).all(), (dx, dy)
