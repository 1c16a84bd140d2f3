"""GPU parity of incremental edits (SparseTree::set_voxel, sparse.rs:384-436): after every
batch the device tree equals the oracle's fresh build of the same edited occupancy (the
reference's own gate, sparse.rs:856-901), traversal over it equals the mirror, and
materials follow set_voxel's rules."""
import numpy as np
import oracle as O
import pytest
from helpers import words_to_set
from vxl import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _apply(words, n, edits):
    """Sequential set_voxel on x-major occupancy words (in-bounds edits only)."""
    w = words.copy()
    for x, y, z, on in edits:
        if x >= n or y >= n or z >= n:
            continue
        lin = int(x) + n * (int(y) + n * int(z))
        if on:
            w[lin >> 5] |= np.uint32(1 << (lin & 31))
        else:
            w[lin >> 5] &= np.uint32(~(1 << (lin & 31)) & 0xFFFFFFFF)
    return w


def _same(gt, ot):
    gn, gl, gb, gc = gt.download()
    on, ol, ob, oc = ot.export()
    assert np.array_equal(gc, oc), "codes"
    assert np.array_equal(gl, ol), "leaves"
    assert np.array_equal(gb, ob), "bounds"
    assert np.array_equal(gn, on), "nodes"


def test_random_toggles_match_fresh_build(ctx):
    for n, base in ((32, O.fixture_words("octant", 32, 0b0110_1001)),
                    (128, O.fixture_words("dust", 128, 3000)), (8, O.fixture_words("single", 8, 5, 2, 2))):
        _say(f"toggles n={n}")
        rng = np.random.default_rng(n)
        words = base.copy()
        gt = ctx.tree_from_words(words, n)
        kinds = set()
        # in place for sure: set the x-neighbour (same brick) of an occupied voxel
        x, y, z = sorted(words_to_set(words, n))[0]
        lin = (x ^ 1) + n * (y + n * z)
        if not (words[lin >> 5] >> (lin & 31)) & 1:
            words = _apply(words, n, [[x ^ 1, y, z, 1]])
            assert ctx.set_voxels(gt, [[x ^ 1, y, z, 1]]) == (1, False)
            kinds.add(False)
            _same(gt, O.Tree.build(words, n))
        for batch in (1, 1, 7, 60, 500, 3000):
            xyz = rng.integers(0, n, (batch, 3))
            occ = set(words_to_set(words, n)) if n <= 32 else None
            ed = np.zeros((batch, 4), np.uint32)
            ed[:, :3] = xyz
            # toggles against the current state, plus a few repeats of the same voxel
            for i, (x, y, z) in enumerate(xyz):
                lin = int(x) + n * (int(y) + n * int(z))
                ed[i, 3] = 0 if (words[lin >> 5] >> (lin & 31)) & 1 else 1
            if batch > 10:
                ed[-3:] = ed[:3]
                ed[-3:, 3] ^= 1  # the last edit of a repeated voxel wins
            before = words
            words = _apply(words, n, ed)
            changed, topo = ctx.set_voxels(gt, ed)
            kinds.add(topo)
            assert changed == int(np.unpackbits((before ^ words).view(np.uint8)).sum())
            ot = O.Tree.build(words, n)
            _same(gt, ot)
            rays = W.box_rays(3000, n, seed=batch)
            hits, _ = ctx.traverse(gt, rays)
            assert np.array_equal(hits, ot.mirror(rays))
        if n >= 32:  # one brick (n = 8) only changes topology when it empties
            assert kinds == {False, True}, f"n={n}: want in-place and rebuild batches, got {kinds}"


def test_edits_edge_cases(ctx):
    _say("edits: out of bounds, empty batch, clear everything, refill an empty tree")
    n = 32
    words = O.fixture_words("dust", n, 40)
    gt = ctx.tree_from_words(words, n)
    assert ctx.set_voxels(gt, np.zeros((0, 4), np.uint32)) == (0, False)
    assert ctx.set_voxels(gt, [[32, 0, 0, 1], [0, 99, 0, 1]]) == (0, False)
    occ = sorted(words_to_set(words, n))
    changed, topo = ctx.set_voxels(gt, [[x, y, z, 0] for x, y, z in occ])
    assert changed == len(occ) and topo and gt.leaf_count == 0
    hits, _ = ctx.traverse(gt, W.box_rays(500, n))
    assert not hits[:, 3].any()
    changed, topo = ctx.set_voxels(gt, [[1, 2, 3, 1], [30, 30, 30, 1]])
    assert changed == 2 and topo and gt.leaf_count == 2
    _same(gt, O.Tree.build(_apply(np.zeros_like(words), n, [[1, 2, 3, 1], [30, 30, 30, 1]]), n))


def test_edits_keep_materials(ctx):
    _say("edits on a tree with materials")
    rng = np.random.default_rng(5)
    n = 32
    vox = np.concatenate([rng.integers(0, 20, (400, 3)), rng.integers(1, 2 ** 16, (400, 1))], 1)
    gt, _ = ctx.tree_from_compact(vox.astype(np.int32), n)
    before = gt.materials().copy()
    _, _, _, codes0 = gt.download()
    # in place: clear one voxel of a leaf, set a new voxel in the same leaf
    x, y, z = (int(v) for v in vox[0, :3])
    leaf0 = int(np.searchsorted(codes0, O.morton_encode(x >> 3, y >> 3, z >> 3)))
    m = O.encode_brick(x & 7, y & 7, z & 7)
    free = [(a, b, c) for a in range(x & ~7, (x & ~7) + 8) for b in range(y & ~7, (y & ~7) + 8)
            for c in range(z & ~7, (z & ~7) + 8)
            if not gt.download()[1][leaf0][O.encode_brick(a & 7, b & 7, c & 7) >> 6]
            >> np.uint64(O.encode_brick(a & 7, b & 7, c & 7) & 63) & np.uint64(1)]
    fa = free[0]
    changed, topo = ctx.set_voxels(gt, [[x, y, z, 0], [fa[0], fa[1], fa[2], 1]])
    assert changed == 2 and not topo
    after = gt.materials()
    mf = O.encode_brick(fa[0] & 7, fa[1] & 7, fa[2] & 7)
    assert after[leaf0][mf] == 0                      # a newly set voxel reads 0
    assert after[leaf0][m] == before[leaf0][m]        # a cleared slot keeps its value
    mask = np.ones_like(before, bool)
    mask[leaf0, mf] = False
    assert np.array_equal(after[mask], before[mask])
    # topology: a new brick far away; surviving bricks keep their materials
    changed, topo = ctx.set_voxels(gt, [[31, 31, 31, 1]])
    assert changed == 1 and topo
    _, _, _, codes1 = gt.download()
    m1 = gt.materials()
    for i, c in enumerate(codes1):
        j = np.searchsorted(codes0, c)
        if j < len(codes0) and codes0[j] == c:
            assert np.array_equal(m1[i], after[j])
        else:
            assert not m1[i].any()
