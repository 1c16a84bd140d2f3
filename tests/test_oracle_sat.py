"""Pins the oracle's SAT voxelizer (reference_cpu.rs:27-200) against the reference's
analytic and brute-force tests and against an independent f64 polygon-clipping test."""
import numpy as np
import oracle as O
import pytest
from helpers import kats, overlaps_with_margin, words_to_set

HALF = (0.5, 0.5, 0.5)
M = kats()["voxelize_meshes"]


def test_sat_analytic_cases():
    for c in kats()["sat_analytic"]["cases"]:
        assert O.triangle_box_overlap(c["center"], HALF, *c["tri"]) == c["overlap"]


def test_sat_agrees_with_polygon_clipping_off_tangents():
    """A wrong sign, axis or transposed operand in the 13-axis SAT makes it disagree with
    clipping on many random cases; only near-tangent cases (within delta) may differ."""
    rng = np.random.default_rng(11)
    delta = 1e-3
    checked = hits = 0
    for _ in range(6000):
        tri = rng.uniform(-2.0, 2.0, (3, 3)).astype(np.float32)
        if rng.random() < 0.3:  # small triangles near the box
            tri = (rng.uniform(-0.8, 0.8, 3) + rng.uniform(-0.6, 0.6, (3, 3))).astype(np.float32)
        c = rng.uniform(-1.0, 1.0, 3).astype(np.float32)
        inner, outer = overlaps_with_margin(tri.astype(np.float64), c.astype(np.float64), 0.5, delta)
        got = O.triangle_box_overlap(c, HALF, tri[0], tri[1], tri[2])
        if inner:
            assert got, (tri, c)
        if not outer:
            assert not got, (tri, c)
        checked += inner or not outer
        hits += got
    assert checked > 5000 and 500 < hits < 5500


def test_cpu_voxelizes_single_triangle():
    words, _, _ = O.voxelize_surface(np.array(M["single_triangle"]), 8)
    assert words.any()


def test_epsilon_does_not_change_occupancy():
    mesh = np.array(M["eps_invariance_mesh"], np.float32)
    ref = None
    for eps in M["eps_values"]:
        w, _, _ = O.voxelize_surface(mesh, 8, eps=eps)
        if ref is None:
            ref = w
        assert np.array_equal(ref, w), eps


def _bruteforce(mesh, n):
    out = set()
    for z in range(n):
        for y in range(n):
            for x in range(n):
                c = (x + 0.5, y + 0.5, z + 0.5)
                if any(O.triangle_box_overlap(c, HALF, t[0:3], t[3:6], t[6:9]) for t in mesh):
                    out.add((x, y, z))
    return out


def test_occupancy_equals_bruteforce_sat():
    for mesh in M["bruteforce_meshes"]:
        mesh = np.array(mesh, np.float32)
        w, _, _ = O.voxelize_surface(mesh, 8)
        assert words_to_set(w, 8) == _bruteforce(mesh, 8)


def test_owner_is_lowest_triangle_index():
    tri = np.array(M["owner_triangle"], np.float32)
    mesh = np.stack([tri, tri, tri])
    w, owner, _ = O.voxelize_surface(mesh, 8, owner=True)
    covered = words_to_set(w, 8)
    assert covered
    for lin in range(512):
        v = (lin % 8, (lin // 8) % 8, lin // 64)
        assert owner[lin] == (0 if v in covered else 0xFFFFFFFF)


def test_hash_color_lcg_known_answers():
    """hash_color (reference_cpu.rs:91-99) takes the low bytes of three steps of the LCG
    x -> 1664525 x + 1013904223 (mod 2^32). Known answers from that generator's published
    sequence from seed 0 (Numerical Recipes in C, 2nd ed., sec. 7.1, "ranqd1"):
    3C6EF35F, 47502932, D1CCF6E9, AAF95334, 6252E503."""
    assert O.hash_color(0) == 0xFFE9325F            # bytes 5F, 32, E9, alpha FF
    assert O.hash_color(0x3C6EF35F) == 0xFF34E932   # seed = first value: 32, E9, 34
    assert O.hash_color(0x47502932) == 0xFF0334E9   # seed = second value: E9, 34, 03
    own = np.array([0, 0xFFFFFFFF, 0x3C6EF35F], np.uint32)
    assert list(O.owner_colors(own)) == [0xFFE9325F, 0, 0xFF34E932]


def test_fully_outside_triangles_mark_nothing():
    for tri in M["outside_triangles"]:
        w, _, pairs = O.voxelize_surface(np.array([tri], np.float32), 8)
        assert not w.any() and pairs == 0


def test_hand_triangle_marks_expected_voxels():
    h = M["hand_triangle"]
    w, _, _ = O.voxelize_surface(np.array([h["tri"]], np.float32), 8)
    s = words_to_set(w, 8)
    for v in h["occupied"]:
        assert tuple(v) in s
    for v in h["empty"]:
        assert tuple(v) not in s


def test_occupancy_is_order_invariant():
    mesh = np.array(M["test_mesh_32"], np.float32)
    a, _, _ = O.voxelize_surface(mesh, 32)
    b, _, _ = O.voxelize_surface(mesh[::-1].copy(), 32)
    assert a.any() and np.array_equal(a, b)


def test_power_of_two_grid_transform_is_exact():
    """With origin 0 and voxel size 2^-j the world->grid map is an exact scaling, so
    voxelizing world-space triangles equals voxelizing the pre-scaled ones at size 1."""
    rng = np.random.default_rng(5)
    tris = rng.uniform(0.05, 0.95, (50, 9)).astype(np.float32)
    for m in (8, 32, 128):
        a, _, pa = O.voxelize_surface(tris, m, voxel_size=1.0 / m)
        b, _, pb = O.voxelize_surface(tris * np.float32(m), m)
        assert np.array_equal(a, b) and pa == pb
    g = O.world_to_grid([0.3, 0.7, 0.1], [0, 0, 0], 1.0 / 64)
    assert g.tolist() == [np.float32(0.3) * 64, np.float32(0.7) * 64, np.float32(0.1) * 64]


def test_marked_voxels_overlap_and_candidates_count():
    rng = np.random.default_rng(6)
    for _ in range(60):
        tri = rng.uniform(0.5, 7.5, (1, 9)).astype(np.float32)
        w, _, pairs = O.voxelize_surface(tri, 8)
        for v in words_to_set(w, 8):
            c = (v[0] + 0.5, v[1] + 0.5, v[2] + 0.5)
            assert O.triangle_box_overlap(c, HALF, tri[0, :3], tri[0, 3:6], tri[0, 6:])
        g = tri.reshape(3, 3)
        lo = np.maximum(np.floor(g.min(0) - np.float32(1e-4)), 0)
        hi = np.minimum(np.floor(g.max(0) + np.float32(1e-4)), 7)
        assert pairs == int(np.prod(hi - lo + 1)) == O.count_candidates(tri, 8)


@pytest.mark.parametrize("n", [32, 128])
def test_voxelize_matches_bruteforce_on_random_soup(n):
    """The AABB scan never misses a voxel the SAT accepts (scan = exhaustive test),
    checked per triangle on its padded box grown by 2 voxels."""
    rng = np.random.default_rng(n)
    centres = rng.uniform(3, n - 3, (40, 1, 3))
    tris = (centres + rng.uniform(-2, 2, (40, 3, 3))).reshape(40, 9).astype(np.float32)
    w, _, _ = O.voxelize_surface(tris, n)
    got = words_to_set(w, n)
    want = set()
    for t in tris:
        g = t.reshape(3, 3)
        lo = np.maximum(np.floor(g.min(0)) - 2, 0).astype(int)
        hi = np.minimum(np.floor(g.max(0)) + 2, n - 1).astype(int)
        for z in range(lo[2], hi[2] + 1):
            for y in range(lo[1], hi[1] + 1):
                for x in range(lo[0], hi[0] + 1):
                    if O.triangle_box_overlap((x + .5, y + .5, z + .5), HALF, g[0], g[1], g[2]):
                        want.add((x, y, z))
    assert got == want


@pytest.mark.parametrize("threads", [2, 3, 8])
def test_threaded_scan_equals_serial(threads):
    """The z-split scan used by the large parity tests gives the serial scan's words, owners
    and pair count (a z-row dropped or taken twice would change one of them)."""
    from vxl import workloads as W
    cases = [(W.uv_sphere(32, 16), 128, 1 / 128), (W.soup(3000, seed=9, mu=np.log(5 / 128)), 128,
                                                   1 / 128),
             (W.heightfield(24), 64, 1 / 64)]
    for tris, n, vs in cases:
        w1, o1, p1 = O.voxelize_surface(tris, n, (0, 0, 0), vs, 1e-4, owner=True)
        w2, o2, p2 = O.voxelize_surface_mt(tris, n, (0, 0, 0), vs, 1e-4, owner=True,
                                           threads=threads)
        assert p1 == p2
        assert np.array_equal(w1, w2)
        assert np.array_equal(o1, o2)


def _glam_default_matrix(origin, voxel_size):
    """Mat4::from_scale(1/vs) * Mat4::from_translation(-o) as glam evaluates the product
    (column-major): the diagonal is inv, the translation column inv*(-o) (each entry one
    product plus exact zero products), the last entry 1."""
    inv = np.float32(1.0) / np.float32(voxel_size)
    m = np.zeros(16, np.float32)
    m[0] = m[5] = m[10] = inv
    for i in range(3):
        m[12 + i] = inv * np.float32(-np.float32(origin[i]))
    m[15] = 1.0
    return m


def test_world_to_grid_override_of_the_default_matrix_equals_the_default_grid():
    """A grid carrying, as its override, the matrix world_to_grid_matrix would build from its
    origin and voxel size voxelizes exactly as the grid without the override (core.rs:101-107):
    the zero entries contribute exact zeros, leaving inv*p + inv*(-o)."""
    rng = np.random.default_rng(31)
    from vxl import workloads as W
    for origin, vs, n in (((0.0, 0.0, 0.0), 1 / 64, 64), ((-0.25, 0.125, 0.3), 1 / 24.0, 32),
                          ((0.37, -0.61, 0.05), 0.0371, 32)):
        tris = (rng.uniform(-0.5, 1.5, (300, 9)) * n * vs).astype(np.float32)
        tris = np.concatenate([tris, W.uv_sphere(24, 12) * np.float32(n * vs * 0.9)])
        w1, _, p1 = O.voxelize_surface(tris, n, origin, vs, 1e-4)
        w2, _, p2 = O.voxelize_surface_mat(tris, n, _glam_default_matrix(origin, vs), 1e-4)
        assert p1 == p2 and np.array_equal(w1, w2)


def test_world_to_grid_override_association_kat():
    """transform_point3 adds x_axis*p.x, then y_axis*p.y, then z_axis*p.z, then w_axis (glam
    0.33 Mat4). With x*px = 1e8, y*py = -1e8, z*pz = 1 and w = 0.5 that order gives
    ((1e8 - 1e8) + 1) + 0.5 = 1.5 exactly, while x + (y + z) would give 0.5 (-1e8 + 1 rounds to
    -1e8 in binary32) and w first 1.0."""
    m = np.zeros(16, np.float32)
    m[0:3] = 1e8
    m[4:7] = -1e8
    m[8:11] = 1.0
    m[12:15] = 0.5
    m[15] = 1.0
    assert O.world_to_grid_mat([1.0, 1.0, 1.0], m).tolist() == [1.5, 1.5, 1.5]
