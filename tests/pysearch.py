"""A second, independently written restatement of the reference's P-frame
motion search, in plain Python + numpy, for tiny frames.

It follows pito-codecs src/h264/enc/slice.rs (search 1590-1723, search_part
1810-1868, analyse_p 1884-1945, mvp/nb_part 392-482 restated in the
decoder's structure dec/slicedec.rs:581-681, pskip_mv 484-494, mv_ok
496-507) and computes its pixel primitives from the H.264 specification
(8.4.2.2.1 half/quarter samples) and the textbook Hadamard transform rather
than from the oracle's loops.  Test infrastructure only.
"""
from __future__ import annotations

import numpy as np

PAD = 64
TAP = np.array([1, -5, 20, 20, -5, 1])
H4 = np.array([[1, 1, 1, 1], [1, -1, 1, -1], [1, 1, -1, -1], [1, -1, -1, 1]])

# quarter position (fx, fy) -> two (plane, dx, dy) samples; planes G b h j
QSPEC = {
    (0, 0): ((0, 0, 0), (0, 0, 0)), (1, 0): ((0, 0, 0), (1, 0, 0)), (2, 0): ((1, 0, 0), (1, 0, 0)),
    (3, 0): ((1, 0, 0), (0, 1, 0)), (0, 1): ((0, 0, 0), (2, 0, 0)), (1, 1): ((1, 0, 0), (2, 0, 0)),
    (2, 1): ((1, 0, 0), (3, 0, 0)), (3, 1): ((1, 0, 0), (2, 1, 0)), (0, 2): ((2, 0, 0), (2, 0, 0)),
    (1, 2): ((2, 0, 0), (3, 0, 0)), (2, 2): ((3, 0, 0), (3, 0, 0)), (3, 2): ((3, 0, 0), (2, 1, 0)),
    (0, 3): ((2, 0, 0), (0, 0, 1)), (1, 3): ((2, 0, 0), (1, 0, 1)), (2, 3): ((3, 0, 0), (1, 0, 1)),
    (3, 3): ((2, 1, 0), (1, 0, 1)),
}

PARTS = {  # mode -> list of (x, y, w, h) in pixels
    0: [(0, 0, 16, 16)],
    1: [(0, 0, 16, 8), (0, 8, 16, 8)],
    2: [(0, 0, 8, 16), (8, 0, 8, 16)],
    3: [(0, 0, 8, 8), (8, 0, 8, 8), (0, 8, 8, 8), (8, 8, 8, 8)],
}
HEX = [(-2, 0), (2, 0), (-1, -2), (1, -2), (-1, 2), (1, 2)]
SQ = [(-1, 0), (1, 0), (0, -1), (0, 1), (-1, -1), (1, -1), (-1, 1), (1, 1)]


def lam_of(qp):
    import math
    return max(16, int(math.floor(256 * math.sqrt(0.85 * 2 ** ((qp - 12) / 3)) + 0.5)))


def ue_len(k):
    return 2 * ((k + 1).bit_length() - 1) + 1


def se_len(v):
    return ue_len(2 * v - 1 if v > 0 else -2 * v)


def pad_plane(frame, mbw, mbh):
    h, w = frame.shape
    al = np.pad(frame, ((0, 16 * mbh - h), (0, 16 * mbw - w)), mode="edge")
    return np.pad(al, PAD, mode="edge").astype(np.int64)


def half_planes(G):
    Gp = np.pad(G, 3, mode="edge")
    R, C = G.shape
    b1 = sum(TAP[i] * Gp[3:3 + R, 1 + i:1 + i + C] for i in range(6))
    h1 = sum(TAP[i] * Gp[1 + i:1 + i + R, 3:3 + C] for i in range(6))
    b1p = np.pad(b1, 3, mode="edge")
    j1 = sum(TAP[i] * b1p[1 + i:1 + i + R, 3:3 + C] for i in range(6))
    return [np.clip((b1 + 16) >> 5, 0, 255), np.clip((h1 + 16) >> 5, 0, 255), np.clip((j1 + 512) >> 10, 0, 255)]


def satd(d):
    tot = 0
    for y in range(0, d.shape[0], 4):
        for x in range(0, d.shape[1], 4):
            tot += int(np.abs(H4 @ d[y:y + 4, x:x + 4] @ H4.T).sum()) // 2
    return tot


class Frame:
    """ME-only field analysis of one P frame (raster order)."""

    def __init__(self, src_frame, ref_frame, me_range, subpel_rounds, partitions, qp, older_refs=()):
        """older_refs: further references of list 0 after ref_frame (reference
        indices 1, 2, ...), for the multi-reference extension."""
        h, w = src_frame.shape
        self.mbw, self.mbh = (w + 15) // 16, (h + 15) // 16
        self.src = pad_plane(src_frame, self.mbw, self.mbh)
        G = pad_plane(ref_frame, self.mbw, self.mbh)
        self.planes = [G] + half_planes(G)
        self.ref_planes = [self.planes]
        for f in older_refs:
            Gr = pad_plane(f, self.mbw, self.mbh)
            self.ref_planes.append([Gr] + half_planes(Gr))
        self.rf = 0      # reference index being searched
        self.reff = {}   # committed reference index per MB
        self.range = me_range
        self.subpel = subpel_rounds
        self.partitions = partitions
        self.lam = lam_of(qp)
        # committed MV field of the picture, per MB: 16 (x, y) MVs or None
        self.mvf = {}

    # -- legality (slice.rs:496-507) and view limit (frame.rs:152-156)
    def vlim(self, my):
        need = 16 * (my + 1) + 96
        return 16 * self.mbh + 44 if need >= 16 * self.mbh else need

    def ok(self, x, y, w, h, mv, vlim):
        x0, y0 = x + (mv[0] >> 2), y + (mv[1] >> 2)
        return (abs(mv[0]) < 2000 and abs(mv[1]) < 2000 and x0 >= -44 and y0 >= -44
                and x0 + w <= 16 * self.mbw + 44 and y0 + h < vlim)

    def block(self, plane, x, y, w, h):
        return self.planes[plane][y + PAD:y + PAD + h, x + PAD:x + PAD + w]

    def pred(self, x, y, w, h, mv):
        (pa, dxa, dya), (pb, dxb, dyb) = QSPEC[(mv[0] & 3, mv[1] & 3)]
        X, Y = x + (mv[0] >> 2), y + (mv[1] >> 2)
        return (self.block(pa, X + dxa, Y + dya, w, h) + self.block(pb, X + dxb, Y + dyb, w, h) + 1) >> 1

    # -- neighbours in the decoder's structure (dec/slicedec.rs:581-616)
    def nb_motion(self, mx, my, cur, done, x, y):
        """(avail, ref, mv) of the 4x4 block at (x, y) relative to the MB."""
        if x < 0:
            if y < 0:
                mb, bx, by = (mx - 1, my - 1), 3, 3
            elif y > 3:
                return (False, -1, (0, 0))
            else:
                mb, bx, by = (mx - 1, my), 3, y
        elif x > 3:
            if y < 0:
                mb, bx, by = (mx + 1, my - 1), 0, 3
            else:
                return (False, -1, (0, 0))
        elif y < 0:
            mb, bx, by = (mx, my - 1), x, 3
        else:
            r = y * 4 + x
            if r not in done:
                return (False, -1, (0, 0))
            return (True, self.rf, cur[r])
        nx, ny = mb
        if nx < 0 or ny < 0 or nx >= self.mbw:
            return (False, -1, (0, 0))
        return (True, self.reff.get((nx, ny), 0), self.mvf[(nx, ny)][by * 4 + bx])

    def mvp(self, mx, my, cur, done, x, y, w, shape, part):
        a = self.nb_motion(mx, my, cur, done, x - 1, y)
        b = self.nb_motion(mx, my, cur, done, x, y - 1)
        c = self.nb_motion(mx, my, cur, done, x + w, y - 1)
        if not c[0]:
            c = self.nb_motion(mx, my, cur, done, x - 1, y - 1)
        if shape == 1:  # 16x8
            if part == 0 and b[1] == self.rf:
                return b[2]
            if part == 1 and a[1] == self.rf:
                return a[2]
        elif shape == 2:  # 8x16
            if part == 0 and a[1] == self.rf:
                return a[2]
            if part == 1 and c[1] == self.rf:
                return c[2]
        if not b[0] and not c[0] and a[0]:
            b = c = a
        m = [a[1] == self.rf, b[1] == self.rf, c[1] == self.rf]
        if m == [True, False, False]:
            return a[2]
        if m == [False, True, False]:
            return b[2]
        if m == [False, False, True]:
            return c[2]
        med = lambda p, q, r: sorted([p, q, r])[1]  # noqa: E731
        return (med(a[2][0], b[2][0], c[2][0]), med(a[2][1], b[2][1], c[2][1]))

    def pskip(self, mx, my):
        a = self.nb_motion(mx, my, {}, set(), -1, 0)
        b = self.nb_motion(mx, my, {}, set(), 0, -1)
        if not a[0] or not b[0]:
            return (0, 0)
        if a[1] == 0 and a[2] == (0, 0):
            return (0, 0)
        if b[1] == 0 and b[2] == (0, 0):
            return (0, 0)
        return self.mvp(mx, my, {}, set(), 0, 0, 4, 0, 0)

    # -- the greedy search (slice.rs:1590-1723)
    def search(self, bx, by, w, h, mvp, cands, vlim):
        src = self.src[by + PAD:by + PAD + h, bx + PAD:bx + PAD + w]
        lam = self.lam

        def bits(m):
            return (lam * (se_len(m[0] - mvp[0]) + se_len(m[1] - mvp[1]))) >> 8

        def sad(m):
            return int(np.abs(src - self.block(0, bx + (m[0] >> 2), by + (m[1] >> 2), w, h)).sum()) + bits(m)

        def st(m):
            return satd(src - self.pred(bx, by, w, h, m)) + bits(m)

        rnd = lambda m: ((m[0] + 2) & ~3, (m[1] + 2) & ~3)  # noqa: E731
        best = rnd(mvp)
        if not self.ok(bx, by, w, h, best, vlim):
            best = (0, 0)
        cost = sad(best)
        for c in cands:
            c = rnd(c)
            if c != best and self.ok(bx, by, w, h, c, vlim):
                k = sad(c)
                if k < cost:
                    cost, best = k, c
        c0 = best
        for _ in range(self.range):
            centre, moved = best, False
            for dx, dy in HEX:
                m = (centre[0] + 4 * dx, centre[1] + 4 * dy)
                if abs(m[0] - c0[0]) >> 2 > self.range or abs(m[1] - c0[1]) >> 2 > self.range:
                    continue
                if not self.ok(bx, by, w, h, m, vlim):
                    continue
                k = sad(m)
                if k < cost:
                    cost, best, moved = k, m, True
            if not moved:
                break
        for _ in range(2):
            centre = best
            for dx, dy in SQ:
                m = (centre[0] + 4 * dx, centre[1] + 4 * dy)
                if self.ok(bx, by, w, h, m, vlim):
                    k = sad(m)
                    if k < cost:
                        cost, best = k, m
            if best == centre:
                break
        cost = st(best)
        for step, rounds in ((2, 1), (1, 1 + self.subpel)):
            for _ in range(rounds):
                centre = best
                for dx, dy in SQ:
                    m = (centre[0] + step * dx, centre[1] + step * dy)
                    if self.ok(bx, by, w, h, m, vlim):
                        k = st(m)
                        if k < cost:
                            cost, best = k, m
                if best == centre:
                    break
        return best, cost

    def search_part(self, mx, my, mode, seed, vlim):
        cur, done = {}, set()
        mvs, mvps, total = [], [], 0
        for p, (x, y, w, h) in enumerate(PARTS[mode]):
            shape = 1 if mode == 1 else (2 if mode == 2 else 0)
            mvp = self.mvp(mx, my, cur, done, x // 4, y // 4, w // 4, shape, p)
            first = mvs[0] if mvs else (0, 0)
            mv, c = self.search(16 * mx + x, 16 * my + y, w, h, mvp, [seed, first], vlim)
            mvs.append(mv)
            mvps.append(mvp)
            total += c
            for by in range(y // 4, (y + h) // 4):
                for bx in range(x // 4, (x + w) // 4):
                    cur[by * 4 + bx] = mv
                    done.add(by * 4 + bx)
        return mvs, mvps, total

    def use_ref(self, r):
        self.rf = r
        self.planes = self.ref_planes[r]

    def analyse(self, mx, my):
        self.use_ref(0)
        skip = self.pskip(mx, my)
        # every reference, each paying the rate of its index (te(v)); the
        # least total, the lower index on ties; older references whole
        nref = len(self.ref_planes)

        def ref_cost(r):
            bits = 0 if nref == 1 else (1 if nref == 2 else ue_len(r))
            return (self.lam * bits) >> 8

        found = None
        for r in range(nref):
            self.use_ref(r)
            vl = self.vlim(my) if r == 0 else self.vlim(self.mbh - 1)
            m, p, c = self.search_part(mx, my, 0, skip, vl)
            c += ref_cost(r)
            if found is None or c < found[3]:
                found = (r, m, p, c)
        rs, mv16, mvp16, c16 = found
        self.use_ref(rs)
        vlim = self.vlim(my) if rs == 0 else self.vlim(self.mbh - 1)
        best = (c16 + (self.lam >> 8), 0, mv16, mvp16)
        costs = [c16, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF]
        if self.partitions and c16 > 1024:
            for mode, extra in ((1, 3), (2, 3), (3, 10)):
                m, p, c = self.search_part(mx, my, mode, mv16[0], vlim)
                c += len(PARTS[mode]) * ref_cost(rs)
                costs[mode] = c
                cc = c + ((self.lam * extra) >> 8)
                if cc < best[0]:
                    best = (cc, mode, m, p)
        cost, mode, mvs, mvps = best
        field = [None] * 16
        for p, (x, y, w, h) in enumerate(PARTS[mode]):
            for by in range(y // 4, (y + h) // 4):
                for bx in range(x // 4, (x + w) // 4):
                    field[by * 4 + bx] = mvs[p]
        self.mvf[(mx, my)] = field
        self.reff[(mx, my)] = rs
        self.use_ref(0)
        return dict(mv=field, mvs=mvs, mvps=mvps, cost=costs, best_cost=cost, skip_mv=skip, mv16=mv16[0],
                    part=mode, ref=rs)

    def run(self):
        return [[self.analyse(mx, my) for mx in range(self.mbw)] for my in range(self.mbh)]


def to_records(rows, dtype):
    """Convert analyse() dicts to pcme_mb_out records."""
    mbh, mbw = len(rows), len(rows[0])
    out = np.zeros((mbh, mbw), dtype)
    for my in range(mbh):
        for mx in range(mbw):
            r = rows[my][mx]
            o = out[my, mx]
            o["mv"] = np.array(r["mv"])
            mvs = list(r["mvs"]) + [(0, 0)] * (4 - len(r["mvs"]))
            mvps = list(r["mvps"]) + [(0, 0)] * (4 - len(r["mvps"]))
            o["mvs"] = np.array(mvs)
            o["mvps"] = np.array(mvps)
            o["cost"] = np.array(r["cost"], np.uint32)
            o["best_cost"] = r["best_cost"]
            o["skip_mv"] = np.array(r["skip_mv"])
            o["mv16"] = np.array(r["mv16"])
            o["part"] = r["part"]
            o["ref"] = r.get("ref", 0)
            out[my, mx] = o
    return out
