"""Host logic of the B-frame search (pcme.bsearch.b_jobs): the neighbour
blocks of every MB read from the committed fields as nb_part reads them
(slice.rs:409-431: outside the picture -> unavailable, intra -> ref -1 with
MV 0, else the 8x8's ref and, for ref >= 0, the block's MV), the co-located
corner MVs and the per-row vlim; no GPU needed."""
from __future__ import annotations

import numpy as np

from pcme import bsearch as B


class _Clip:
    def __init__(self, mbw, mbh):
        self.mbw, self.mbh, self.stride = mbw, mbh, 16 * mbw + 128

    def plane_ptr(self, k, t):
        return (1 << 20) + (k << 16) + t


def test_b_jobs_neighbour_blocks_and_colocated_mbs():
    mbw, mbh = 5, 4
    rs = np.random.default_rng(0)
    mv = rs.integers(-50, 50, (2, mbh, mbw, 16, 2))
    ref = rs.choice([-1, 0, 1], (2, mbh, mbw, 4))
    intra = rs.random((mbh, mbw)) < 0.2
    colmv = rs.integers(-3, 3, (mbh, mbw, 2, 16, 2))
    colref = rs.choice([-1, 0, 1], (mbh, mbw, 2, 4))
    colintra = rs.random((mbh, mbw)) < 0.2
    c = _Clip(mbw, mbh)
    jobs = B.b_jobs(c, 1, (c, 0), (c, 2), mv, ref, intra, colmv, colref, colintra, 27)

    def nb(mx, my, dx, dy, blk, l):
        nx, ny = mx + dx, my + dy
        if nx < 0 or nx >= mbw or ny < 0:
            return -2, (0, 0)
        if intra[ny, nx]:
            return -1, (0, 0)
        r = ref[l, ny, nx, ((blk >> 2) >> 1) * 2 + ((blk & 3) >> 1)]
        return r, (tuple(mv[l, ny, nx, blk]) if r >= 0 else (0, 0))

    for i in range(mbw * mbh):
        my, mx = divmod(i, mbw)
        assert jobs["mx"][i] == mx and jobs["my"][i] == my
        assert jobs["src"][i] == c.plane_ptr(0, 1) + (64 + 16 * my) * c.stride + 64 + 16 * mx
        assert list(jobs["ref1"][i]) == [c.plane_ptr(k, 2) for k in range(4)]
        for l in range(2):
            for k, ((dx, dy), blk) in enumerate(B._NB):
                r, m = nb(mx, my, dx, dy, blk, l)
                assert jobs["nb"]["ref"][i, l, k] == r and tuple(jobs["nb"]["mv"][i, l, k]) == m, (i, l, k)
        assert np.array_equal(jobs["col_mv"][i], colmv[my, mx][:, [0, 3, 12, 15]])
        assert np.array_equal(jobs["col_ref"][i], colref[my, mx])
        assert jobs["col_intra"][i] == int(colintra[my, mx])
    assert list(B.vlim_rows(20)) == [16 * (y + 1) + 96 if 16 * (y + 1) + 96 < 320 else 364 for y in range(20)]
