"""Pins the oracle's traversals: the f64 reference (layout.rs, dda.rs, ray.rs), the dense
tier-A oracle (oracle.rs) and the f32 mirror (mirror.rs)."""
import numpy as np
import oracle as O
import pytest
from helpers import (TIE_CASES, brute_force_first_hit, kats, single_voxel_school_b,
                     voxels_to_words, words_to_set)

D = kats()["dda"]


def _ray64(o, d):
    return np.array([list(o) + list(d)], np.float64)


def test_axis_aligned_hit_is_exact():
    k = D["axis_hit"]
    words = O.fixture_words("single", k["n"], *k["voxel"])
    h, t = O.first_hit_dense(words, k["n"], _ray64(k["origin"], k["dir"]))
    assert h[0].tolist() == k["voxel"] + [1] and abs(t[0] - k["t"]) < 1e-12
    tree = O.Tree.build(words, k["n"])
    for school in (0, 1):
        h, t = tree.layout(_ray64(k["origin"], k["dir"]), school=school)
        assert h[0].tolist() == k["voxel"] + [1] and abs(t[0] - k["t"]) < 1e-12


def test_mirror_exact_axis_hits():
    k = D["mirror_exact_axis"]
    s = k["single"]
    t = O.Tree.build(O.fixture_words("single", s["n"], *s["voxel"]), s["n"])
    ray = np.array([s["origin"] + s["dir"]], np.float32)
    assert t.mirror(ray)[0].tolist() == s["hit"] + [1]
    e = O.Tree.build(O.fixture_words("empty", 32), 32)
    assert e.mirror(ray)[0, 3] == 0
    so = k["solid"]
    t = O.Tree.build(O.fixture_words("solid", so["n"]), so["n"])
    assert t.mirror(np.array([so["origin"] + so["dir"]], np.float32))[0].tolist() == so["hit"] + [1]
    k0 = k["k0"]
    t = O.Tree.build(O.fixture_words("single", k0["n"], *k0["voxel"]), k0["n"])
    assert t.mirror(np.array([k0["origin"] + k0["dir"]], np.float32))[0].tolist() == k0["hit"] + [1]


def test_grazing_cases_keep_their_hits():
    """mirror.rs:461-492: skip-on and skip-off both return the voxel for the first three."""
    g = D["grazing_cases"]
    for v, o, d in g["cases"][: g["asserted"]]:
        nodes, leaves, bounds = single_voxel_school_b(g["n"], v)
        t = O.Tree.from_school_b(g["n"], nodes, leaves, bounds)
        ray = np.array([o + d], np.float32)
        assert t.mirror(ray, skip=False)[0].tolist() == v + [1]
        assert t.mirror(ray, skip=True)[0].tolist() == v + [1]


def test_early_skip_matches_no_skip_on_axis_aligned_rays():
    """mirror.rs:495-538."""
    v = (300, 301, 302)
    nodes, leaves, bounds = single_voxel_school_b(512, v)
    t = O.Tree.from_school_b(512, nodes, leaves, bounds)
    rng = np.random.default_rng(7)
    dirs = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1],
                     [1, 1, 0], [0, 1, 1], [1, 0, 1]], np.float64)
    rays = []
    for d in dirs:
        aim = np.array(v) + 0.5 + rng.uniform(-0.8, 0.8, (400, 3))
        rays.append(np.concatenate([aim - d * 80.0, np.broadcast_to(d, aim.shape)], 1))
    rays = np.concatenate(rays).astype(np.float32)
    assert np.array_equal(t.mirror(rays, skip=True), t.mirror(rays, skip=False))
    assert t.mirror(rays)[:, 3].sum() > 1000


@pytest.mark.parametrize("kind,params", [("checkerboard", ()), ("octant", (0b0110_1001,)),
                                         ("octant", (0b1000_0001,))])
def test_dda_matches_brute_force_slab(kind, params):
    """oracle.rs:223-251: same hit/miss and entry t (1e-6) as an independent slab test."""
    n = 8
    words = O.fixture_words(kind, n, *params)
    vox = sorted(words_to_set(words, n))
    tree = O.Tree.build(words, n)
    rng = np.random.default_rng(8)
    o = rng.uniform(-4, 12, (600, 3))
    d = rng.uniform(-1, 1, (600, 3))
    keep = np.linalg.norm(d, axis=1) > 1e-3
    rays = np.concatenate([o, d], 1)[keep]
    hd, td = O.first_hit_dense(words, n, rays)
    ha, ta = tree.layout(rays, school=0)
    hb, tb = tree.layout(rays, school=1)
    assert np.array_equal(ha, hb) and np.array_equal(ta, tb)
    assert np.array_equal(hd[:, 3], ha[:, 3])
    for i, ray in enumerate(rays):
        bf = brute_force_first_hit(vox, ray)
        assert (bf is not None) == bool(hd[i, 3])
        if bf is not None:
            assert abs(bf[1] - td[i]) < 1e-6 and abs(bf[1] - ta[i]) < 1e-6


def _box_rays(n, count, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-4, n + 4, (count, 3))
    d = rng.uniform(-1, 1, (count, 3))
    keep = np.linalg.norm(d, axis=1) >= 1e-3
    return np.concatenate([o, d], 1)[keep]


@pytest.mark.parametrize("kind,params", [("octant", (0b0110_1001,)), ("checkerboard", ())])
def test_school_a_b_and_dense_oracle_agree(kind, params):
    """school_b.rs:803-856 / sparse.rs:800-850: hierarchical == dense on voxel and t."""
    n = 128
    words = O.fixture_words(kind, n, *params)
    tree = O.Tree.build(words, n)
    rays = _box_rays(n, 3000, 9)
    hd, td = O.first_hit_dense(words, n, rays)
    ha, ta = tree.layout(rays, school=0)
    hb, tb = tree.layout(rays, school=1)
    assert np.array_equal(ha, hb) and np.array_equal(ta, tb)
    assert np.array_equal(ha, hd)
    assert np.max(np.abs(ta - td)) < 1e-6


def test_mirror_matches_f64_within_grazing_bound():
    """mirror.rs:541-590: f32 mirror vs f64 oracle disagree on < 1% of rays."""
    n = 128
    total = mism = 0
    rays = _box_rays(n, 6000, 10)
    for words in (O.sierpinski(n), O.fixture_words("checkerboard", n)):
        tree = O.Tree.build(words, n)
        hd, _ = O.first_hit_dense(words, n, rays)
        hm = tree.mirror(rays.astype(np.float32))
        # the mirror sees the f32-rounded ray; compare against the same rounding's f64 walk
        hd32, _ = O.first_hit_dense(words, n, rays.astype(np.float32).astype(np.float64))
        mism += int(np.any(hm != hd32, axis=1).sum())
        total += len(rays)
        assert np.any(hd != hd32, axis=1).sum() < len(rays) * 0.01
    assert mism / total < 0.01


@pytest.mark.parametrize("case", range(len(TIE_CASES)))
def test_walker_ties_take_the_lowest_axis(case):
    """Exact t_max ties (helpers.TIE_CASES, derived by hand from mirror.rs:87-116): the mirror,
    the f64 layout traversal (dda.rs, the same rule) and the dense oracle (oracle.rs) give the
    hand-derived voxel."""
    n, occ, ray, want = TIE_CASES[case]
    words = voxels_to_words(np.array(occ), n)
    t = O.Tree.build(words, n)
    expect = [0, 0, 0, 0] if want is None else list(want) + [1]
    assert t.mirror(np.array([ray], np.float32))[0].tolist() == expect
    for school in (0, 1):
        h, _ = t.layout(np.array([ray], np.float64), school=school)
        assert h[0].tolist() == expect
    h, _ = O.first_hit_dense(words, n, np.array([ray], np.float64))
    assert h[0].tolist() == expect
