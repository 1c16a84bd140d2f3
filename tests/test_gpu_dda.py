"""GPU parity: traversal hits are identical to the f32 mirror (mirror.rs) and t agrees
with the f64 reference traversal (layout.rs) within 1e-6 relative on matching voxels."""
import numpy as np
import oracle as O
import pytest
from helpers import kats, single_voxel_school_b
from vxl import workloads as W

pytestmark = pytest.mark.gpu
T_RTOL = 1e-6  # north-star tolerance on the hit distance t


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _check(ctx, gtree, otree, rays, label, min_hits=0):
    rays = np.ascontiguousarray(rays, np.float32)
    _say(f"{label}: {len(rays)} rays")
    hits, t = ctx.traverse(gtree, rays)
    want = otree.mirror(rays)
    bad = np.any(hits != want, axis=1)
    assert not bad.any(), f"{label}: {int(bad.sum())} hit mismatches, first {np.nonzero(bad)[0][:5]}"
    assert int(hits[:, 3].sum()) >= min_hits
    ref_h, ref_t = otree.layout(rays.astype(np.float64))
    same = np.all(ref_h == hits, axis=1) & (hits[:, 3] == 1)
    err = np.abs(t[same] - ref_t[same]) / np.maximum(np.abs(ref_t[same]), 1e-30)
    zero = ref_t[same] == 0
    assert np.all(t[same][zero] == 0)
    assert err[~zero].max(initial=0) <= T_RTOL, f"{label}: max rel t error {err.max()}"
    assert np.all(t[hits[:, 3] == 0] == 0)
    return hits


FIXTURES = [("octant", 128, (0b0110_1001,)), ("checkerboard", 128, ()), ("dust", 128, (4096,)),
            ("wire", 128, (16, 1)), ("octant", 32, (0b1000_0001,)), ("solid", 32, ()),
            ("single", 8, (5, 2, 2)), ("empty", 32, ())]


@pytest.mark.parametrize("fx", FIXTURES, ids=lambda f: f"{f[0]}{f[1]}")
def test_fixture_battery(ctx, fx):
    """voxel-gpu/tests/differential.rs:202-266 battery (origins U[-4,n+4), dirs U[-1,1))."""
    kind, n, params = fx
    words = O.fixture_words(kind, n, *params)
    ot = O.Tree.build(words, n)
    gt = ctx.tree_from_words(words, n)
    _check(ctx, gt, ot, W.box_rays(20000, n), f"{kind}{n}")


def test_grazing_and_axis_aligned_rays(ctx):
    """voxel-gpu/tests/differential.rs:97-164: rays aimed at occupied voxels, jittered to
    graze, along axes and diagonals -- zero mismatches allowed."""
    n = 128
    words = O.fixture_words("dust", n, 4096)
    ot = O.Tree.build(words, n)
    gt = ctx.tree_from_words(words, n)
    rng = np.random.default_rng(0x9E37)
    occ = np.array(sorted(__import__("helpers").words_to_set(words, n)))
    targets = occ[rng.choice(len(occ), 40, replace=False)]
    axes = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, 0, 1], [0, 1, 1], [1, 1, 1]],
                    np.float64)
    rays = []
    for tgt in targets:
        for d in axes:
            aim = tgt + 0.5 + rng.uniform(-0.8, 0.8, (16, 3))
            rays.append(np.concatenate([aim - d * 50.0, np.broadcast_to(d, aim.shape)], 1))
    _check(ctx, gt, ot, np.concatenate(rays), "grazing", min_hits=100)


def test_reference_grazing_cases_2048(ctx):
    g = kats()["dda"]["grazing_cases"]
    for v, o, d in g["cases"]:
        nodes, leaves, bounds = single_voxel_school_b(g["n"], v)
        ot = O.Tree.from_school_b(g["n"], nodes, leaves, bounds)
        gt = ctx.tree_upload(g["n"], nodes, leaves, bounds)
        _check(ctx, gt, ot, np.array([o + d]), f"grazing {v}")


def test_sphere_tree_camera_and_random_rays(ctx):
    tris = W.uv_sphere()
    n, vs = W.grid_for(256)
    words, _, _ = O.voxelize_surface(tris, n, (0, 0, 0), vs, 1e-4)
    ot = O.Tree.build(words, n)
    gt = ctx.voxelize_tree(tris, n, (0, 0, 0), vs, 1e-4)
    cam = W.camera_rays(256, 480, 270)
    _check(ctx, gt, ot, cam, "sphere camera", min_hits=1000)
    _check(ctx, gt, ot, W.random_rays(100_000, 256), "sphere random", min_hits=1000)


def test_traverse_empty_batch_and_empty_tree(ctx):
    _say("empty batch / empty tree")
    e = ctx.tree_from_words(O.fixture_words("empty", 32), 32)
    hits, t = ctx.traverse(e, W.random_rays(1000, 32))
    assert not hits.any() and not t.any()
    hits, t = ctx.traverse(e, np.zeros((0, 6), np.float32))
    assert hits.shape == (0, 4)


def test_non_finite_and_degenerate_rays(ctx):
    """Directions with NaN, +-inf, zero and huge components, origins with NaN / inf:
    hits identical to the mirror (which keeps t_max = +inf for a NaN or zero component)."""
    n = 128
    words = O.fixture_words("checkerboard", n)
    ot = O.Tree.build(words, n)
    gt = ctx.tree_from_words(words, n)
    nan, inf = np.nan, np.inf
    rng = np.random.default_rng(5)
    rows = []
    specials = [nan, inf, -inf, 0.0, -0.0, 1e30, -1e30, 3e38, 1e-30, 1.0, -1.0]
    for _ in range(3000):
        o = rng.uniform(-10, n + 10, 3)
        d = rng.normal(size=3)
        for a in range(3):
            if rng.random() < 0.3:
                d[a] = specials[rng.integers(len(specials))]
        if rng.random() < 0.05:
            o[rng.integers(3)] = specials[rng.integers(3)]
        rows.append(np.concatenate([o, d]))
    rays = np.array(rows, np.float32)
    hits, _ = ctx.traverse(gt, rays)
    want = ot.mirror(rays)
    bad = np.any(hits != want, axis=1)
    assert not bad.any(), f"{int(bad.sum())} mismatches, first rays {rays[np.nonzero(bad)[0][:3]]}"


@pytest.mark.parametrize("refill", [1, 8, 64])
def test_persistent_lanes_match_mirror(ctx, refill):
    """The persistent lane schedule (idle lanes take new rays) walks every ray as the mirror
    does: box rays, sphere camera + random rays, non-finite rays and ragged batch sizes
    (1, 5, 63, 65, 4097 rays) -- a schedule changes which lane runs a ray, never its hit."""
    ctx.set_dda_refill(refill)
    try:
        n = 128
        words = O.fixture_words("dust", n, 4096)
        ot = O.Tree.build(words, n)
        gt = ctx.tree_from_words(words, n)
        rays = W.box_rays(20000, n)
        _check(ctx, gt, ot, rays, f"dust128 refill={refill}", min_hits=100)
        for R in (1, 5, 63, 65, 4097):
            _check(ctx, gt, ot, rays[:R], f"dust128 R={R} refill={refill}")
        tris = W.uv_sphere()
        n2, vs = W.grid_for(256)
        w2, _, _ = O.voxelize_surface(tris, n2, (0, 0, 0), vs, 1e-4)
        ot2 = O.Tree.build(w2, n2)
        gt2 = ctx.voxelize_tree(tris, n2, (0, 0, 0), vs, 1e-4)
        _check(ctx, gt2, ot2, W.camera_rays(256, 320, 180), f"sphere camera refill={refill}",
               min_hits=1000)
        _check(ctx, gt2, ot2, W.random_rays(50_000, 256), f"sphere random refill={refill}",
               min_hits=1000)
        odd = np.array([[5.0, 5.0, 5.0, np.nan, 1.0, 0.0], [-1.0, 3.5, 3.5, 1.0, 0.0, 0.0],
                        [3.5, -1.0, 3.5, 0.0, np.inf, 0.0], [np.inf, 1.0, 1.0, -1.0, 0.0, 0.0]],
                       np.float32)
        hits, _ = ctx.traverse(gt, odd)
        assert np.array_equal(hits, ot.mirror(odd))
    finally:
        ctx.set_dda_refill(-1)


def test_dda_refill_argument_checked(ctx):
    import vxl
    for bad in (-2, 65):
        with pytest.raises(vxl.VxlError):
            ctx.set_dda_refill(bad)


def test_walk_statistics_match_the_mirror(ctx):
    """vxl_traverse_stats counts the walk as the mirror counts it (mirror.rs:263-326 with
    the oracle's {leaf descents, node descents, visits}): sums over the batch are equal."""
    import torch
    tris = W.uv_sphere()
    n, vs = W.grid_for(256)
    words, _, _ = O.voxelize_surface(tris, n, (0, 0, 0), vs, 1e-4)
    ot = O.Tree.build(words, n)
    gt = ctx.voxelize_tree(tris, n, (0, 0, 0), vs, 1e-4)
    for label, rays in (("camera", W.camera_rays(256, 320, 180)),
                        ("random", W.random_rays(30_000, 256)),
                        ("box", W.box_rays(20_000, n))):
        _say(f"walk statistics {label}")
        rays = np.ascontiguousarray(rays, np.float32)
        d = torch.from_numpy(rays).cuda()
        got = ctx.traverse_stats(gt, d.data_ptr(), len(rays))
        _, st = ot.mirror(rays, stats=True)
        want = tuple(int(v) for v in st.astype(np.int64).sum(0))
        assert got == want, f"{label}: {got} != {want}"
