"""GPU parity of the material path (compact_surface_sparse -> tree_from_compact): trees,
per-voxel materials and compact voxel lists equal the oracle's (voxelize_surface_cpu with
owners, compact_voxels.wgsl lookup, SparseTree::from_voxels) on the same inputs."""
import numpy as np
import oracle as O
import pytest
from vxl import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _table(T, seed, short=0):
    t = np.random.default_rng(seed).integers(0, 2 ** 32, (T + 1) // 2, dtype=np.uint64)
    t = t.astype(np.uint32)
    return t[: max(len(t) - short, 0)]


def _oracle(tris, n, vs, table, g=(0, 0, 0)):
    words, owner, _ = O.voxelize_surface(tris, n, (0, 0, 0), vs, 1e-4, owner=True)
    cv = O.compact(words, owner, n, table, g)
    local = cv - np.array([g[0], g[1], g[2], 0], np.int32)
    tree, mats, dropped = O.tree_from_compact(local, n)
    assert dropped == 0
    return cv, tree, mats


def _same_tree(gt, ot):
    gn, gl, gb, gc = gt.download()
    on, ol, ob, oc = ot.export()
    assert np.array_equal(gn, on) and np.array_equal(gl, ol)
    assert np.array_equal(gb, ob) and np.array_equal(gc, oc)


def _sorted(cv):
    return cv[np.lexsort((cv[:, 0], cv[:, 1], cv[:, 2]))]


CASES = [
    ("sphere_128", lambda: W.uv_sphere(32, 16), 128, 1 / 128, 0),
    ("soup_128", lambda: W.soup(3000, seed=31, mu=np.log(5 / 128)), 128, 1 / 128, 0),
    # a table shorter than the mesh: the last triangles resolve to 0 (MISSING)
    ("soup_short_table_128", lambda: W.soup(2000, seed=32, mu=np.log(5 / 128)), 128, 1 / 128, 300),
    ("big_random_128", lambda: np.random.default_rng(4).uniform(0, 1, (40, 9)).astype(np.float32),
     128, 1 / 128, 0),
    ("terrain_slice_512", lambda: W.heightfield(64), 512, 1 / 256, 0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: c[0])
def test_voxelize_tree_materials_matches_oracle(ctx, case):
    name, make, n, vs, short = case
    tris = make()
    table = _table(len(tris), 7, short)
    _say(f"materials {name} n={n} T={len(tris)} table={len(table)}")
    cv, ot, omats = _oracle(tris, n, vs, table)
    gt = ctx.voxelize_tree_materials(tris, n, (0, 0, 0), vs, 1e-4, table)
    _same_tree(gt, ot)
    assert np.array_equal(gt.materials(), omats)
    plain = ctx.voxelize_tree(tris, n, (0, 0, 0), vs, 1e-4)
    assert gt.pairs == plain.pairs
    _same_tree(plain, ot)
    g = (-7, 300, 12)
    got = ctx.compact_surface(tris, n, (0, 0, 0), vs, 1e-4, table, g)
    want = O.compact(*O.voxelize_surface(tris, n, (0, 0, 0), vs, 1e-4, owner=True)[:2], n, table, g)
    assert np.array_equal(_sorted(got), _sorted(want))


def test_tree_from_compact_matches_oracle(ctx):
    _say("tree_from_compact: random lists with repeats and out-of-range coords")
    rng = np.random.default_rng(11)
    for n, count in ((32, 3000), (128, 40000), (512, 20000)):
        xyz = rng.integers(-3, n + 3, (count, 3))
        xyz[: count // 2] = rng.integers(0, n, (count // 2, 3)) // 3 * 3  # clustered, repeats
        mat = rng.integers(-2 ** 31, 2 ** 31, count)
        vox = np.concatenate([xyz, mat[:, None]], 1).astype(np.int32)
        vox = np.concatenate([vox, vox[: count // 10]])  # every repeat keeps the last
        vox[-(count // 10):, 3] = rng.integers(0, 2 ** 16, count // 10)
        ot, omats, odrop = O.tree_from_compact(vox, n)
        gt, gdrop = ctx.tree_from_compact(vox, n)
        assert gdrop == odrop
        _same_tree(gt, ot)
        assert np.array_equal(gt.materials(), omats)


def test_material_path_reference_kats_and_edges(ctx):
    _say("tree_from_compact reference KATs; empty inputs; a tree without materials")
    gt, dropped = ctx.tree_from_compact([(0, 0, 0, 3), (5, 5, 5, 7), (-1, 0, 0, 9),
                                         (32, 0, 0, 9), (0, 40, 0, 9)], 32)
    assert dropped == 3 and gt.leaf_count == 1
    m = gt.materials()[0]
    assert m[O.encode_brick(0, 0, 0)] == 3 and m[O.encode_brick(5, 5, 5)] == 7
    gt, _ = ctx.tree_from_compact([(4, 4, 4, 2), (4, 4, 4, 9)], 32)
    assert gt.materials()[0][O.encode_brick(4, 4, 4)] == 9
    gt, dropped = ctx.tree_from_compact(np.zeros((0, 4), np.int32), 32)
    assert dropped == 0 and gt.leaf_count == 0
    e = ctx.voxelize_tree_materials(np.zeros((0, 9), np.float32), 32, table=[1])
    assert e.leaf_count == 0
    assert ctx.compact_surface(np.zeros((0, 9), np.float32), 32).shape == (0, 4)
    plain = ctx.voxelize_tree(W.uv_sphere(16, 8), 32, (0, 0, 0), 1 / 32)
    assert not plain.materials().any()
    # no table: every voxel MISSING
    t = ctx.voxelize_tree_materials(W.uv_sphere(16, 8), 32, (0, 0, 0), 1 / 32)
    assert t.leaf_count > 0 and not t.materials().any()
