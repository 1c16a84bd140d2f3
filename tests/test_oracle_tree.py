"""Pins the oracle's sparse-tree build (sparse.rs) and School-B serialisation
(school_b.rs) with invariants checked by independent code."""
import numpy as np
import oracle as O
import pytest
from helpers import (school_b_from_levels, school_b_occupied, single_voxel_school_b,
                     voxels_to_words, words_to_set)

FIXTURES = [
    ("checkerboard", 32, ()),
    ("octant", 32, (0b0110_1001,)),
    ("octant", 128, (0b1000_0001,)),
    ("single", 32, (5, 17, 30)),
    ("solid", 8, ()),
    ("wire", 128, (16, 1)),
    ("dust", 128, (4096,)),
]


@pytest.mark.parametrize("kind,n,params", FIXTURES)
def test_school_b_holds_exactly_the_field(kind, n, params):
    words = O.fixture_words(kind, n, *params)
    t = O.Tree.build(words, n)
    nodes, leaves, bounds, codes = t.export()
    assert school_b_occupied(n, nodes, leaves) == words_to_set(words, n)
    # School-B node count == School-A node count (school_b.rs:495-500)
    assert t.node_count == sum(t.level_counts.values())
    assert len(bounds) == len(leaves) == t.leaf_count


@pytest.mark.parametrize("kind,n,params", FIXTURES[:4])
def test_point_query_matches_field(kind, n, params):
    words = O.fixture_words(kind, n, *params)
    t = O.Tree.build(words, n)
    occ = words_to_set(words, n)
    for z in range(n):
        for y in range(n):
            for x in range(n):
                assert t.is_occupied(x, y, z) == ((x, y, z) in occ)


def test_school_a_levels_are_the_or_reduce_of_sorted_codes():
    words = O.fixture_words("dust", 128, 512)
    t = O.Tree.build(words, 128)
    _, leaves, _, codes = t.export()
    assert np.all(np.diff(codes.astype(np.int64)) > 0)  # Morton-sorted, unique
    for c, leaf in zip(codes, leaves):
        bx, by, bz = O.morton_decode(int(c))
        assert leaf.any()
        assert all(O.words_get(words, 128, bx * 8 + x, by * 8 + y, bz * 8 + z) ==
                   bool((int(leaf[O.encode_brick(x, y, z) >> 6]) >> (O.encode_brick(x, y, z) & 63)) & 1)
                   for x in range(8) for y in range(8) for z in range(8))
    child = codes.astype(np.uint64)
    for L in range(2, 4):  # k = 2 at n = 128: levels 2..3
        lvl = t.level(L)
        parents, first = np.unique(child >> np.uint64(6), return_index=True)
        assert len(lvl) == len(parents)
        for i, p in enumerate(parents):
            sel = child[(child >> np.uint64(6)) == p]
            mask = 0
            for c in sel:
                mask |= 1 << int(c & np.uint64(63))
            assert (int(lvl[i, 0]) | int(lvl[i, 1]) << 32) == mask
            assert int(lvl[i, 2]) == int(first[i])
        child = parents


def test_bounds_are_leaf_bounds():
    words = O.fixture_words("wire", 128, 16, 1)
    t = O.Tree.build(words, 128)
    _, leaves, bounds, _ = t.export()
    for leaf, b in zip(leaves, bounds):
        assert O.leaf_bounds(leaf) == int(b)


def test_single_brick_k0_and_empty():
    t = O.Tree.build(O.fixture_words("single", 8, 5, 2, 2), 8)
    assert (t.leaf_count, t.node_count) == (1, 0)
    e = O.Tree.build(O.fixture_words("empty", 32), 32)
    assert (e.leaf_count, e.node_count) == (0, 0)


@pytest.mark.parametrize("n,v", [(32, (5, 17, 30)), (128, (100, 3, 77)), (512, (300, 301, 302))])
def test_single_voxel_school_b_matches_hand_built(n, v):
    """A single-voxel tree's School-B arrays equal the hand-built root-to-leaf chain."""
    t = O.Tree.build(O.fixture_words("single", n, *v), n)
    nodes, leaves, bounds, _ = t.export()
    hn, hl, hb = single_voxel_school_b(n, v)
    assert np.array_equal(nodes, hn) and np.array_equal(leaves, hl) and np.array_equal(bounds, hb)


def test_row_read_build_equals_probe_build():
    """The byte-per-row brick read used by the 2048^3 checks builds the probe build's tree
    (a row taken from the wrong byte or bit would move voxels)."""
    for kind, n, params in [("dust", 128, (4096,)), ("octant", 32, (0b0110_1001,)),
                            ("wire", 128, (16, 1)), ("checkerboard", 8, ()), ("empty", 32, ()),
                            ("single", 512, (509, 3, 260))]:
        w = O.fixture_words(kind, n, *params)
        a = O.Tree.build(w, n).export()
        b = O.Tree.build(w, n, rows=True).export()
        for x, y in zip(a, b):
            assert np.array_equal(x, y), (kind, n)


def _levels(t, k):
    return {L: t.level(L) for L in range(2, k + 2)}


@pytest.mark.parametrize("kind,n,params,k", [("dust", 512, (20000,), 3), ("wire", 512, (64, 2), 3),
                                              ("octant", 512, (0b0110_1001,), 3)])
def test_school_b_positions_closed_form_k3(kind, n, params, k):
    """Every School-B entry equals its closed-form position (helpers.school_b_from_levels)."""
    t = O.Tree.build(O.fixture_words(kind, n, *params), n)
    nodes, _, _, _ = t.export()
    assert np.array_equal(nodes, school_b_from_levels(k, _levels(t, k)))


def test_school_b_positions_closed_form_k4():
    """At k = 4 (Resolution(2048)) parent-first DFS and breadth-first order differ: the
    level-3 block of the root's first child comes before the level-4 siblings' subtrees. 30,000
    random voxels in a few clusters, so subtrees differ in size."""
    n, k = 2048, 4
    rng = np.random.default_rng(0x5EED)
    centres = rng.integers(0, n, (6, 3))
    vox = np.concatenate([np.clip(c + rng.normal(0, 90, (5000, 3)), 0, n - 1) for c in centres])
    t = O.Tree.build(voxels_to_words(vox.astype(np.int64), n), n, rows=True)
    nodes, leaves, _, _ = t.export()
    levels = _levels(t, k)
    assert all(len(levels[L]) > 1 for L in (2, 3, 4))
    assert np.array_equal(nodes, school_b_from_levels(k, levels))
    bfs = np.concatenate([levels[L] for L in range(k + 1, 1, -1)])  # level order, for contrast
    assert not np.array_equal(nodes[:, :2], bfs[:, :2])


def test_school_b_k4_hand_built_dfs():
    """Three voxels at Resolution(2048): root children A (cell 0) and B (cell 1); A has level-3
    children A1, A2; B has B1; each has one level-2 node and one leaf. By school_b.rs:68-98
    (root at 0; a node's children contiguously at the current end, then each child's subtree)
    the entries are root, A, B, A1, A2, L2(A1), L2(A2), B1, L2(B1), with child_base 1, 3, 7, 5,
    6, leaf 0, leaf 1, 8, leaf 2 -- breadth-first order would put B1 before the level-2 nodes."""
    n = 2048
    vox = np.array([[0, 0, 0], [128, 0, 0], [512, 0, 0]])
    t = O.Tree.build(voxels_to_words(vox, n), n, rows=True)
    nodes, leaves, bounds, codes = t.export()
    want = np.array([[3, 0, 1], [3, 0, 3], [1, 0, 7], [1, 0, 5], [1, 0, 6], [1, 0, 0], [1, 0, 1],
                     [1, 0, 8], [1, 0, 2]], np.uint32)
    assert np.array_equal(nodes, want)
    assert [O.morton_decode(int(c)) for c in codes] == [(0, 0, 0), (16, 0, 0), (64, 0, 0)]
    assert np.all(leaves[:, 0] == 1) and not leaves[:, 1:].any()
