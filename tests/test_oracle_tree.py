"""Pins the oracle's sparse-tree build (sparse.rs) and School-B serialisation
(school_b.rs) with invariants checked by independent code."""
import numpy as np
import oracle as O
import pytest
from helpers import school_b_occupied, single_voxel_school_b, words_to_set

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
