"""Pins of the oracle's macroblock-tree propagation and offsets
(mbtree.rs:162-231): closed forms of single edges (zero MV, whole-cell
shifts, the four-cell area split of fractional MVs with floor division,
bi blocks split in half), a three-node chain, the references that feed
nothing, mass conservation, and a float64 scatter over random nodes."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O


def blocks(bw, bh, intra=0, inter=0, lst=0, mv0=(0, 0), mv1=(0, 0)):
    b = np.zeros(bw * bh, O.OTREE_DTYPE)
    b["intra"], b["inter"], b["list"] = intra, inter, lst
    b["mv"][:, 0] = mv0
    b["mv"][:, 1] = mv1
    return b


def one(bw, bh, k, **kw):
    """A node whose only feeding block is k."""
    b = blocks(bw, bh)
    b[k] = blocks(1, 1, **kw)[0]
    return b


def test_zero_mv_edge_moves_intra_minus_inter():
    bw, bh = 5, 4
    b1 = blocks(bw, bh, intra=100, inter=60, lst=1)
    prop = O.mbtree_propagate([blocks(bw, bh), b1], [(-1, -1), (0, -1)], bw, bh)
    assert (prop[0] == 40.0).all() and (prop[1] == 0.0).all()


@pytest.mark.parametrize("mv,cell", [((8, -8), (3, 1)), ((-16, 8), (0, 3)), ((24, 0), None)])
def test_whole_cell_shift(mv, cell):
    bw, bh = 4, 4
    k = 2 * bw + 2
    b1 = one(bw, bh, k, intra=100, inter=60, lst=1, mv0=mv)
    prop = O.mbtree_propagate([blocks(bw, bh), b1], [(-1, -1), (0, -1)], bw, bh)
    want = np.zeros(bw * bh, np.float32)
    if cell is not None:
        want[cell[1] * bw + cell[0]] = 40.0
    np.testing.assert_array_equal(prop[0], want)


@pytest.mark.parametrize("mv,cells", [
    ((2, 6), {(2, 2): 12, (3, 2): 4, (2, 3): 36, (3, 3): 12}),    # fx 2, fy 6
    ((-3, -1), {(1, 1): 3, (2, 1): 5, (1, 2): 21, (2, 2): 35}),   # floor: cell -1, fx 5, fy 7
    ((13, 0), {(3, 2): 24, (4, 2): 40}),                          # fx 5, fy 0: no lower row
])
def test_fractional_area_split(mv, cells):
    # amount 64 -> 1 per unit of area
    bw, bh = 5, 5
    k = 2 * bw + 2
    b1 = one(bw, bh, k, intra=128, inter=64, lst=2, mv1=mv)
    prop = O.mbtree_propagate([blocks(bw, bh), b1], [(-1, -1), (-1, 0)], bw, bh)
    want = np.zeros(bw * bh, np.float32)
    for (x, y), a in cells.items():
        want[y * bw + x] = a
    np.testing.assert_array_equal(prop[0], want)
    assert prop[0].sum() == 64.0


def test_edge_of_picture_drops_the_outside_share():
    bw, bh = 3, 3
    b1 = one(bw, bh, 2, intra=128, inter=64, lst=1, mv0=(4, -4))  # top-right block, half out each way
    prop = O.mbtree_propagate([blocks(bw, bh), b1], [(-1, -1), (0, -1)], bw, bh)
    assert prop[0][2] == 16.0 and prop[0].sum() == 16.0


def test_bi_block_splits_in_half():
    bw, bh = 3, 2
    b2 = blocks(bw, bh, intra=100, inter=36, lst=3, mv0=(0, 0), mv1=(8, 0))
    prop = O.mbtree_propagate([blocks(bw, bh), blocks(bw, bh), b2], [(-1, -1), (0, -1), (0, 1)], bw, bh)
    np.testing.assert_array_equal(prop[0], np.full(bw * bh, 32.0, np.float32))
    want1 = np.array([0, 32, 32, 0, 32, 32], np.float32)
    np.testing.assert_array_equal(prop[1], want1)


def test_same_reference_on_both_lists_adds_both_halves():
    bw, bh = 2, 2
    b1 = blocks(bw, bh, intra=100, inter=36, lst=3)
    prop = O.mbtree_propagate([blocks(bw, bh), b1], [(-1, -1), (0, 0)], bw, bh)
    np.testing.assert_array_equal(prop[0], np.full(4, 64.0, np.float32))


def test_chain_accumulates_into_the_amount():
    bw, bh = 2, 1
    b2 = blocks(bw, bh, intra=100, inter=60, lst=1)
    b1 = blocks(bw, bh, intra=80, inter=20, lst=1)
    prop = O.mbtree_propagate([blocks(bw, bh), b1, b2], [(-1, -1), (0, -1), (1, -1)], bw, bh)
    np.testing.assert_array_equal(prop[2], [0, 0])
    np.testing.assert_array_equal(prop[1], [40, 40])
    np.testing.assert_array_equal(prop[0], [90, 90])  # (80 + 40) * 60 / 80


def test_blocks_and_references_that_feed_nothing():
    bw, bh = 2, 2
    z = blocks(bw, bh)
    cases = [
        (blocks(bw, bh, intra=100, inter=60, lst=0), (0, -1)),   # intra block
        (blocks(bw, bh, intra=0, inter=0, lst=1), (0, -1)),      # zero intra
        (blocks(bw, bh, intra=50, inter=60, lst=1), (0, -1)),    # inter above intra
        (blocks(bw, bh, intra=100, inter=60, lst=1), (-1, 0)),   # list 0 without a reference
        (blocks(bw, bh, intra=100, inter=60, lst=2), (0, 2)),    # list 1 refers forward
        (blocks(bw, bh, intra=100, inter=60, lst=1), (1, -1)),   # refers to itself
    ]
    for b1, refs in cases:
        prop = O.mbtree_propagate([z, b1, z], [(-1, -1), refs, (-1, -1)], bw, bh)
        assert not prop.any(), refs


def random_nodes(rng, n, bw, bh, mvmax):
    out, refs = [], []
    for i in range(n):
        b = np.zeros(bw * bh, O.OTREE_DTYPE)
        b["intra"] = rng.integers(0, 3000, bw * bh)
        b["inter"] = rng.integers(0, 3500, bw * bh)
        b["list"] = rng.integers(0, 4, bw * bh)
        b["mv"] = rng.integers(-mvmax, mvmax + 1, (bw * bh, 2, 2))
        out.append(b)
        refs.append((int(rng.integers(-1, i)) if i else -1, int(rng.integers(-1, n))))
    return out, refs


def test_mass_is_conserved_inside_the_picture():
    """One edge whose MVs keep every landing cell inside: the reference cell
    sums total sum(intra - min(inter, intra)) over the feeding blocks."""
    rng = np.random.default_rng(5)
    bw, bh = 9, 7
    b = blocks(bw, bh)
    b["intra"] = rng.integers(1, 4000, bw * bh)
    b["inter"] = rng.integers(0, 5000, bw * bh)
    b["list"] = rng.integers(0, 2, bw * bh)
    b["mv"][:, 0] = rng.integers(0, 8, (bw * bh, 2))   # into cells (x, y)..(x+1, y+1)
    bx, by = np.arange(bw * bh) % bw, np.arange(bw * bh) // bw
    edge = (bx == bw - 1) | (by == bh - 1)
    b["mv"][edge, 0] = 0
    prop = O.mbtree_propagate([blocks(bw, bh), b], [(-1, -1), (0, -1)], bw, bh)
    feed = (b["list"] == 1)
    intra = b["intra"].astype(np.float64)
    want = (intra - np.minimum(b["inter"].astype(np.float64), intra))[feed].sum()
    assert abs(prop[0].astype(np.float64).sum() - want) <= 1e-5 * want


def scatter64(nodes, refs, bw, bh):
    """The propagation as a float64 scatter: per node (last first) the
    amounts of all blocks at once, np.add.at into the landing cells."""
    n = len(nodes)
    prop = np.zeros((n, bh, bw))
    by, bx = np.divmod(np.arange(bw * bh), bw)
    for i in range(n - 1, -1, -1):
        b = nodes[i]
        intra = b["intra"].astype(np.float64)
        inter = np.minimum(b["inter"].astype(np.float64), intra)
        ok = (b["list"] > 0) & (b["intra"] > 0)
        amount = np.where(ok, (intra + prop[i].ravel()) * (intra - inter) / np.where(ok, intra, 1), 0)
        for l in (0, 1):
            r = refs[i][l]
            if r < 0 or r >= i:
                continue
            use = ok & (amount > 0) & ((b["list"] == 3) | (b["list"] == 1 + l))
            w = np.where(b["list"] == 3, 0.5, 1.0)
            x = 8 * bx + b["mv"][:, l, 0].astype(np.int64)
            y = 8 * by + b["mv"][:, l, 1].astype(np.int64)
            cx, cy, fx, fy = x // 8, y // 8, x % 8, y % 8
            for dx, dy in ((0, 0), (1, 0), (0, 1), (1, 1)):
                area = np.where(dx, fx, 8 - fx) * np.where(dy, fy, 8 - fy)
                tx, ty = cx + dx, cy + dy
                m = use & (area > 0) & (tx >= 0) & (ty >= 0) & (tx < bw) & (ty < bh)
                np.add.at(prop[r], (ty[m], tx[m]), (amount * w / 64 * area)[m])
    return prop.reshape(n, -1)


@pytest.mark.parametrize("seed,n,bw,bh,mvmax", [(0, 4, 6, 5, 12), (1, 7, 11, 9, 40), (2, 3, 1, 1, 9)])
def test_float64_scatter(seed, n, bw, bh, mvmax):
    rng = np.random.default_rng(seed)
    nodes, refs = random_nodes(rng, n, bw, bh, mvmax)
    got = O.mbtree_propagate(nodes, refs, bw, bh).astype(np.float64)
    want = scatter64(nodes, refs, bw, bh)
    np.testing.assert_allclose(got, want, rtol=2e-5, atol=1e-3)
    assert got.any()


def test_offsets_closed_forms():
    b = blocks(4, 1)
    b["intra"] = [100, 100, 0, 0]
    prop = np.array([0, 100, 1, 3], np.float32)
    np.testing.assert_array_equal(O.mbtree_offsets(b, prop), [0.0, -2.0, -2.0, -4.0])
