"""Oracle materials path: the compact_voxels.wgsl material lookup, compact_surface_sparse
from dense outputs, and tree_from_compact / SparseTree::from_voxels, pinned by the
reference's own tests (materials.rs, sparse.rs) and by agreement with the occupancy path."""
import numpy as np
import oracle as O
from helpers import words_to_set
from vxl import workloads as W


def _mat_at(tree, mats, x, y, z):
    _, _, _, codes = tree.export()
    code = O.morton_encode(x >> 3, y >> 3, z >> 3)
    i = np.searchsorted(codes, code)
    if i >= len(codes) or codes[i] != code:
        return 0
    return int(mats[i][O.encode_brick(x & 7, y & 7, z & 7)])


def test_packed_table_lookup_matches_material_table_for_sparse():
    """materials.rs:326-342: materials [10, MAX, 20, 10] -> globals packed as
    packed[0] = 1, packed[1] = 2 | 1 << 16 (triangle t in bits (t & 1) * 16 of word t >> 1)."""
    packed = [1, 2 | (1 << 16)]
    assert [O.resolve_material(t, packed) for t in range(4)] == [1, 0, 2, 1]
    assert O.resolve_material(4, packed) == 0            # past the table
    assert O.resolve_material(0xFFFFFFFF, packed) == 0   # no owner
    assert O.resolve_material(0, []) == 0                # empty table


def test_from_voxels_kats():
    """sparse.rs:1117-1147 (two bricks, materials carried, unoccupied reads 0) and
    :1150-1160 (a repeated coord keeps the last)."""
    pts = [(0, 0, 0, 3), (1, 0, 0, 3), (7, 7, 7, 5), (8, 0, 0, 7), (9, 2, 3, 7)]
    tree, mats, dropped = O.tree_from_compact(pts, 32)
    assert dropped == 0 and tree.leaf_count == 2
    for x, y, z, g in pts:
        assert _mat_at(tree, mats, x, y, z) == g
    assert _mat_at(tree, mats, 2, 0, 0) == 0
    occ = {(x, y, z) for x, y, z, _ in pts}
    for (x, y, z) in occ:
        assert tree.is_occupied(x, y, z)
    tree, mats, _ = O.tree_from_compact([(4, 4, 4, 2), (4, 4, 4, 9)], 32)
    assert tree.leaf_count == 1 and _mat_at(tree, mats, 4, 4, 4) == 9


def test_tree_from_compact_kats():
    """materials.rs:286-316 (out-of-range coords dropped, materials carried) and :318-331
    (material masked to its low 16 bits)."""
    vox = [(0, 0, 0, 3), (5, 5, 5, 7), (-1, 0, 0, 9), (32, 0, 0, 9), (0, 40, 0, 9)]
    tree, mats, dropped = O.tree_from_compact(vox, 32)
    assert dropped == 3
    assert tree.is_occupied(0, 0, 0) and tree.is_occupied(5, 5, 5)
    assert _mat_at(tree, mats, 0, 0, 0) == 3 and _mat_at(tree, mats, 5, 5, 5) == 7
    tree, mats, dropped = O.tree_from_compact([(1, 2, 3, np.int32(-65531))], 8)  # 0xFFFF0005
    assert dropped == 0 and _mat_at(tree, mats, 1, 2, 3) == 5


def test_compact_agrees_with_occupancy_and_owners():
    """compact lists exactly the occupied voxels, each with its owner's material, and
    tree_from_compact of that list rebuilds the same tree as the occupancy words."""
    tris = W.soup(400, seed=9, mu=np.log(6 / 32))
    words, owner, _ = O.voxelize_surface(tris, 32, (0, 0, 0), 1 / 32, 1e-4, owner=True)
    table = np.random.default_rng(2).integers(0, 2 ** 32, (len(tris) + 1) // 2, dtype=np.uint64)
    table = table.astype(np.uint32)
    g = (-3, 5, 100)
    cv = O.compact(words, owner, 32, table, g)
    got = {(x - g[0], y - g[1], z - g[2]) for x, y, z, _ in cv}
    assert got == words_to_set(words, 32)
    for x, y, z, m in cv[:500]:
        lin = (x - g[0]) + 32 * ((y - g[1]) + 32 * (z - g[2]))
        assert m == O.resolve_material(int(owner[lin]), table)
    local = cv - np.array([g[0], g[1], g[2], 0], np.int32)
    t2, mats, dropped = O.tree_from_compact(local, 32)
    assert dropped == 0
    t1 = O.Tree.build(words, 32)
    for a, b in zip(t1.export(), t2.export()):
        assert np.array_equal(a, b)
