"""GPU parity: dense occupancy and School-B trees from the HIP path are bit-exact with
the oracle (voxelize_surface_cpu -> SparseTree::build -> SchoolBBuffer::from_sparse)."""
import numpy as np
import oracle as O
import pytest
import vxl
from helpers import kats
from vxl import workloads as W

pytestmark = pytest.mark.gpu
M = kats()["voxelize_meshes"]


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _cases():
    rng = np.random.default_rng(123)
    out = [
        ("test_mesh_32", np.array(M["test_mesh_32"], np.float32), 32, (0, 0, 0), 1.0, 1e-4),
        ("hand_8", np.array([M["hand_triangle"]["tri"]], np.float32), 8, (0, 0, 0), 1.0, 1e-4),
        ("bruteforce_a_8", np.array(M["bruteforce_meshes"][0], np.float32), 8, (0, 0, 0), 1.0, 1e-4),
        ("bruteforce_b_8", np.array(M["bruteforce_meshes"][1], np.float32), 8, (0, 0, 0), 1.0, 1e-4),
        ("eps_0.5", np.array(M["eps_invariance_mesh"], np.float32), 8, (0, 0, 0), 1.0, 0.5),
        ("eps_0", np.array(M["eps_invariance_mesh"], np.float32), 8, (0, 0, 0), 1.0, 0.0),
        ("outside", np.array(M["outside_triangles"], np.float32), 8, (0, 0, 0), 1.0, 1e-4),
        ("sphere_128", W.uv_sphere(32, 16), 128, (0, 0, 0), 1 / 128, 1e-4),
        ("soup_128", W.soup(3000, seed=77, mu=np.log(4 / 128)), 128, (0, 0, 0), 1 / 128, 1e-4),
        ("offset_origin_32", (rng.uniform(-0.3, 1.3, (400, 9))).astype(np.float32), 32,
         (-0.25, 0.125, 0.3), 1 / 24.0, 1e-4),
        ("grid_space_edges_32", np.round(rng.uniform(-2, 34, (300, 9)) * 2) / 2, 32, (0, 0, 0), 1.0,
         1e-4),
        ("huge_triangle_32", np.array([[-50, -40, 16.2, 90, -30, 15.7, 10, 80, 16.9]], np.float32), 32,
         (0, 0, 0), 1.0, 1e-4),
        # large triangles: candidate windows along the normal's main axis instead of the box
        ("big_sphere_512", W.sphere_with_tris(100, seed=3), 512, (0, 0, 0), 1 / 256, 1e-4),
        ("big_random_128", rng.uniform(0.0, 1.0, (40, 9)).astype(np.float32), 128, (0, 0, 0),
         1 / 128, 1e-4),
        ("big_steep_128", np.array([[0.1, 0.1, 0.05, 0.9, 0.2, 0.95, 0.3, 0.9, 0.5],
                                    [0.05, 0.9, 0.1, 0.95, 0.1, 0.2, 0.5, 0.5, 0.99]], np.float32),
         128, (0, 0, 0), 1 / 128, 1e-4),
        # more than 2^22 candidates in one triangle's decoded box (64-bit decode): a vertex
        # with fine fractional bits makes g - c inexact for far centres, so the plane is not
        # hoisted and the whole 12.4M-cell box is enumerated
        ("wide_unhoisted_512", np.array([[0.1234567, 0.7654321, 0.3141592, 250.3, 10.7, 200.1,
                                          30.2, 245.9, 120.6]], np.float32), 512, (0, 0, 0), 1.0,
         1e-4),
        # a finite world vertex whose grid coordinate overflows to +inf (literal products,
        # 13M candidates in the clamped box)
        ("wide_literal_512", np.array([[3e38, 50.0, 50.0, 150.0, 150.0, 200.0, 150.25, 100.0,
                                        120.0]], np.float32), 512, (0, 0, 0), 0.5, 1e-4),
        # a plane crossing the grid from vertices millions of cells away (exact integers, so
        # the plane is hoisted with normal components near 1e11): the radius bound's relative
        # term makes the window wider than 15 cells, so the whole clamped box is enumerated
        ("far_vertices_128", np.array([[-1.4e6, -1.0e6, -99950.0, 1.4e6, -1.0e6, 40050.0,
                                        0.0, 1.4e6, 42050.0]], np.float32), 128, (0, 0, 0), 1.0,
         1e-4),
        ("farther_vertices_128", np.array([[-5.6e6, -4.0e6, -399950.0, 5.6e6, -4.0e6, 160050.0,
                                            0.0, 5.6e6, 168050.0]], np.float32), 128, (0, 0, 0),
         1.0, 1e-4),
    ]
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_dense_occupancy_bit_exact(ctx, case):
    name, tris, n, origin, vs, eps = case
    tris = np.asarray(tris, np.float32)
    _say(f"dense {name} n={n} T={len(tris)}")
    want, _, pairs = O.voxelize_surface(tris, n, origin, vs, eps)
    got, gpairs = ctx.voxelize_dense(tris, n, origin, vs, eps)
    assert gpairs == pairs
    assert np.array_equal(got, want), f"{int((got != want).sum())} words differ"


def _tree_equal(gt, ot):
    gn, gl, gb, gc = gt.download()
    on, ol, ob, oc = ot.export()
    if gt.leaf_count != ot.leaf_count:
        extra = sorted(set(gc.tolist()) - set(oc.tolist()))[:8]
        missing = sorted(set(oc.tolist()) - set(gc.tolist()))[:8]
        raise AssertionError(f"leaves {gt.leaf_count} vs {ot.leaf_count}; extra bricks "
                             f"{[O.morton_decode(c) for c in extra]} missing "
                             f"{[O.morton_decode(c) for c in missing]}")
    assert gt.node_count == ot.node_count
    assert np.array_equal(gl, ol), "leaf words differ"
    assert np.array_equal(gc, oc), "leaf codes differ"
    assert np.array_equal(gb, ob), "leaf bounds differ"
    assert np.array_equal(gn, on), "School-B nodes differ"


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_tree_bit_exact(ctx, case):
    name, tris, n, origin, vs, eps = case
    tris = np.asarray(tris, np.float32)
    _say(f"tree {name} n={n} T={len(tris)}")
    words, _, _ = O.voxelize_surface(tris, n, origin, vs, eps)
    ot = O.Tree.build(words, n)
    gt = ctx.voxelize_tree(tris, n, origin, vs, eps)
    _tree_equal(gt, ot)


def test_sphere_config_1_at_256_bit_exact(ctx):
    """The 16,128-triangle seeded sphere at 256^3, embedded in Resolution(512)."""
    tris = W.uv_sphere()
    n, vs = W.grid_for(256)
    _say(f"sphere C1 n={n}")
    words, _, _ = O.voxelize_surface(tris, n, (0, 0, 0), vs, 1e-4)
    got, _ = ctx.voxelize_dense(tris, n, (0, 0, 0), vs, 1e-4)
    assert np.array_equal(got, words)
    _tree_equal(ctx.voxelize_tree(tris, n, (0, 0, 0), vs, 1e-4), O.Tree.build(words, n))


def test_empty_mesh_and_k0(ctx):
    _say("empty mesh / k=0")
    t = ctx.voxelize_tree(np.zeros((0, 9), np.float32), 32)
    assert (t.leaf_count, t.node_count) == (0, 0)
    words, _ = ctx.voxelize_dense(np.zeros((0, 9), np.float32), 32)
    assert not words.any()
    tri = np.array([[1.5, 1.5, 1.5, 5.5, 1.5, 1.5, 1.5, 5.5, 1.5]], np.float32)
    w, _, _ = O.voxelize_surface(tri, 8)
    _tree_equal(ctx.voxelize_tree(tri, 8), O.Tree.build(w, 8))


def test_validation_errors(ctx):
    _say("validation")
    tri = np.zeros((1, 9), np.float32)
    with pytest.raises(vxl.VxlError) as e:
        ctx.voxelize_tree(tri, 1024)
    assert e.value.name == "VXL_NOT_REPRESENTABLE"
    with pytest.raises(vxl.VxlError) as e:
        ctx.voxelize_tree(tri, 32, voxel_size=0.0)
    assert e.value.name == "VXL_INVALID_ARG"
    with pytest.raises(vxl.VxlError) as e:
        ctx.voxelize_tree(tri, 32, eps=-1.0)
    assert e.value.name == "VXL_INVALID_ARG"
    bad = tri.copy()
    bad[0, 4] = np.inf
    with pytest.raises(vxl.VxlError) as e:
        ctx.voxelize_tree(bad, 32)
    assert e.value.name == "VXL_INVALID_ARG"


def test_repeated_calls_leave_scratch_clean(ctx):
    """The dense leaf grid is cleared by each build: calls at different n agree with fresh
    oracle results, in any order."""
    a = W.soup(2000, seed=5, mu=np.log(3 / 128))
    b = W.uv_sphere(40, 20)
    for n, tris in ((128, a), (32, b), (128, b), (512, a), (128, a)):
        vs = 1.0 / min(n, 128)
        _say(f"repeat n={n}")
        words, _, _ = O.voxelize_surface(tris, n, (0, 0, 0), vs, 1e-4)
        _tree_equal(ctx.voxelize_tree(tris, n, (0, 0, 0), vs, 1e-4), O.Tree.build(words, n))


def test_tree_from_words_fixtures(ctx):
    for kind, n, params in [("checkerboard", 32, ()), ("octant", 128, (0b0110_1001,)),
                            ("dust", 128, (4096,)), ("wire", 128, (16, 1)), ("single", 8, (5, 2, 2)),
                            ("solid", 32, ()), ("empty", 128, ())]:
        _say(f"from_words {kind} n={n}")
        words = O.fixture_words(kind, n, *params)
        _tree_equal(ctx.tree_from_words(words, n), O.Tree.build(words, n))


def test_over_16m_triangles_take_the_device_wide_scan(ctx):
    """Calls with more than 65,536 setup blocks (16,777,216 triangles) scan the candidate
    counts device-wide instead of in one workgroup. A small sphere tiled past that size
    voxelizes to the sphere's own tree, with the sphere's pair count times the tile count."""
    base = W.uv_sphere(48, 24)
    n, vs = 128, 1.0 / 128
    reps = -(-(65536 * 256 + 1) // len(base))
    tris = np.tile(base, (reps, 1))
    _say(f"tiled sphere: {len(tris)} triangles ({reps} copies of {len(base)}) at n={n}")
    words, _, pairs = O.voxelize_surface(base, n, (0, 0, 0), vs, 1e-4)
    tree = ctx.voxelize_tree(tris, n, (0, 0, 0), vs, 1e-4)
    assert tree.pairs == reps * pairs
    _tree_equal(tree, O.Tree.build(words, n))


def _affine(n):
    """A world-to-grid override that is not a scale + translation: a rotation of 30 degrees
    about z then 20 about x, an anisotropic scale and a translation into [0, n)^3, as glam's
    column-major Mat4 (16 f32)."""
    cz, sz = np.cos(np.radians(30)), np.sin(np.radians(30))
    cx, sx = np.cos(np.radians(20)), np.sin(np.radians(20))
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    a = rx @ rz @ np.diag([0.9 * n, 0.75 * n, 0.8 * n])
    t = np.array([n / 2, n / 2, n / 2]) - a @ np.array([0.5, 0.5, 0.5])
    m = np.zeros((4, 4))
    m[:3, :3], m[:3, 3], m[3, 3] = a, t, 1.0
    return m.T.reshape(16).astype(np.float32)  # column-major


def test_world_to_grid_override_bit_exact(ctx):
    """VoxelGrid::world_to_grid (core.rs:30-34,101-104): vertices go through the override with
    Mat4::transform_point3; dense words, pair count and tree equal the oracle's, single and
    batched."""
    n = 128
    mat = _affine(n)
    meshes = [W.uv_sphere(48, 24), W.soup(3000, seed=78, mu=np.log(3 / 128))]
    for i, tris in enumerate(meshes):
        _say(f"world_to_grid override mesh {i} T={len(tris)}")
        want, _, pairs = O.voxelize_surface_mat(tris, n, mat, 1e-4)
        got, gpairs = ctx.voxelize_dense(tris, n, (0, 0, 0), 1.0, 1e-4, mat=mat)
        assert gpairs == pairs
        assert np.array_equal(got, want), f"{int((got != want).sum())} words differ"
        _tree_equal(ctx.voxelize_tree(tris, n, (0, 0, 0), 1.0, 1e-4, mat=mat),
                    O.Tree.build(want, n))
    trees = ctx.voxelize_batch(meshes, [(n, (0, 0, 0), 1.0, mat), (512, (0, 0, 0), 1 / 128),
                                        ][:1] * 2)
    for tris, tr in zip(meshes, trees):
        want, _, _ = O.voxelize_surface_mat(tris, n, mat, 1e-4)
        _tree_equal(tr, O.Tree.build(want, n))
    bad = mat.copy()
    bad[5] = np.inf
    with pytest.raises(vxl.VxlError) as e:
        ctx.voxelize_tree(meshes[0], n, (0, 0, 0), 1.0, 1e-4, mat=bad)
    assert e.value.name == "VXL_INVALID_ARG"
    # with an override, a voxel size whose reciprocal overflows is not an error (core.rs:124)
    ctx.voxelize_tree(meshes[0], n, (0, 0, 0), 1e-39, 1e-4, mat=mat).free()


@pytest.mark.parametrize("waves,units", [(32, 1), (8, 4), (16, 8), (4, 64), (1, 1024)])
def test_work_unit_schedules_bit_exact(waves, units):
    """Every SAT work-distribution shape gives the oracle's words: each wave's own first unit
    alone (1 unit per wave), and units past the grid's waves handed out by the 8 interleaved
    counters (up to 1024 units per wave, so most come from the counters), on a sphere whose
    triangles split into several segments and a soup with many one-cell boxes."""
    c = vxl.Context(0)
    try:
        c.set_tuning(sat_waves_per_cu=waves, sat_units_per_wave=units)
        for name, tris, n, vs in (("sphere_256", W.uv_sphere(64, 32), 512, 1 / 256),
                                  ("soup_128", W.soup(20000, seed=91, mu=np.log(3 / 128)), 128,
                                   1 / 128)):
            _say(f"units {name} waves/cu={waves} units/wave={units}")
            want, _, pairs = O.voxelize_surface(tris, n, (0, 0, 0), vs, 1e-4)
            got, gpairs = c.voxelize_dense(tris, n, (0, 0, 0), vs, 1e-4)
            assert gpairs == pairs
            assert np.array_equal(got, want), f"{int((got != want).sum())} words differ"
    finally:
        c.close()
