"""GPU parity at the benchmark sizes, against the oracle.

* terrain_1024 (2,197,152 triangles at voxel size 1/1024 in Resolution(2048), k = 4): dense
  occupancy words and the School-B tree bit-exact with voxelize_surface_cpu ->
  SparseTree::build -> SchoolBBuffer::from_sparse (reference_cpu.rs:110-200,
  sparse.rs:241-305, school_b.rs:113-151);
* the 8,388,608-triangle soup at Resolution(2048): the tree bit-exact, then the full ray set
  of the traversal benchmark (1920x1080 camera rays + 2,000,000 random rays) under the
  library's default lane schedule (persistent lanes for this tree) with hits identical to the
  f32 mirror (mirror.rs:254-326) and t within 1e-6 relative of the f64 traversal
  (layout.rs:120-232); the same rays over the terrain tree;
* a many-mesh batch with every resolution x {sphere, soup}, each tree against the oracle and
  each ray segment against the mirror;
* served voxelize and ray requests against the oracle.

The oracle scans with its threaded z-split (equal to the serial scan, tests/test_oracle_sat.py)
and builds trees reading brick rows as bytes (equal to the probe build,
tests/test_oracle_tree.py), so each 2048^3 check takes seconds of CPU.
"""
import math

import numpy as np
import oracle as O
import pytest
import torch
import vxl
from vxl import workloads as W

pytestmark = pytest.mark.gpu
T_RTOL = 1e-6


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _tree_equal(gt, ot, what):
    gn, gl, gb, gc = gt.download()
    on, ol, ob, oc = ot.export()
    assert (gt.leaf_count, gt.node_count) == (ot.leaf_count, ot.node_count), what
    assert gt.level_nodes == ot.level_counts, what
    assert np.array_equal(gc, oc), f"{what}: leaf codes differ"
    assert np.array_equal(gl, ol), f"{what}: leaf words differ"
    assert np.array_equal(gb, ob), f"{what}: leaf bounds differ"
    assert np.array_equal(gn, on), f"{what}: School-B nodes differ"


def _oracle_tree(tris, n, vs):
    words, _, pairs = O.voxelize_surface_mt(tris, n, (0, 0, 0), vs, 1e-4)
    return words, pairs, O.Tree.build(words, n, rows=True)


def _check_rays(hits, t, ot, rays, label):
    want = ot.mirror(rays)
    bad = np.any(hits != want, axis=1)
    assert not bad.any(), f"{label}: {int(bad.sum())} hit mismatches, first {np.nonzero(bad)[0][:5]}"
    ref_h, ref_t = ot.layout(rays.astype(np.float64))
    same = np.all(ref_h == hits, axis=1) & (hits[:, 3] == 1)
    zero = ref_t[same] == 0
    assert np.all(t[same][zero] == 0), label
    err = np.abs(t[same][~zero] - ref_t[same][~zero]) / np.abs(ref_t[same][~zero])
    assert err.max(initial=0) <= T_RTOL, f"{label}: max rel t error {err.max()}"
    assert np.all(t[hits[:, 3] == 0] == 0), label
    _say(f"{label}: {len(rays)} rays identical, {int(hits[:, 3].sum())} hits, "
         f"{int(same.sum())} t compared (max rel err {err.max(initial=0):.2e})")


def _bench_rays(m):
    return np.ascontiguousarray(np.concatenate([W.camera_rays(m), W.random_rays(2_000_000, m)]))


@pytest.fixture(scope="module")
def terrain():
    tris = W.terrain_1024()
    n, vs = W.grid_for(1024)
    _say(f"oracle: terrain_1024 T={len(tris)} n={n}")
    words, pairs, ot = _oracle_tree(tris, n, vs)
    yield tris, n, vs, words, pairs, ot


def test_terrain_1024_dense_words_bit_exact(ctx, terrain):
    tris, n, vs, words, pairs, _ = terrain
    _say("terrain_1024 dense words")
    got, gpairs = ctx.voxelize_dense(tris, n, (0, 0, 0), vs, 1e-4)
    assert gpairs == pairs
    assert np.array_equal(got, words), f"{int((got != words).sum())} words differ"


def test_terrain_1024_tree_bit_exact(ctx, terrain):
    tris, n, vs, _, pairs, ot = terrain
    _say(f"terrain_1024 tree: {ot.leaf_count} leaves, {ot.node_count} nodes")
    gt = ctx.voxelize_tree(tris, n, (0, 0, 0), vs, 1e-4)
    assert gt.pairs == pairs
    _tree_equal(gt, ot, "terrain_1024 (host input)")
    gt.free()
    d = torch.from_numpy(tris).to("cuda")  # the benchmark's call: triangles resident in HBM
    for rep in range(2):
        gt = ctx.voxelize_tree_dev(d.data_ptr(), len(tris), n, (0, 0, 0), vs)
        _tree_equal(gt, ot, f"terrain_1024 (device input, call {rep})")
        gt.free()


def test_terrain_tree_bench_rays(ctx, terrain):
    tris, n, vs, _, _, ot = terrain
    gt = ctx.voxelize_tree(tris, n, (0, 0, 0), vs, 1e-4)
    rays = _bench_rays(1024)
    hits, t = ctx.traverse(gt, rays)
    _check_rays(hits, t, ot, rays, "terrain_1024 tree, camera + random rays")
    gt.free()


@pytest.fixture(scope="module")
def soup_2048():
    tris = W.soup(8_388_608, seed=0x5EED_0006)
    _say(f"oracle: soup_2048 T={len(tris)}")
    _, pairs, ot = _oracle_tree(tris, 2048, 1.0 / 2048)
    yield tris, pairs, ot


@pytest.fixture(scope="module")
def soup_tree(ctx, soup_2048):
    tris, pairs, ot = soup_2048
    gt = ctx.voxelize_tree(tris, 2048, (0, 0, 0), 1.0 / 2048, 1e-4)
    yield gt, pairs, ot
    gt.free()


def test_soup_2048_tree_bit_exact(soup_tree):
    gt, pairs, ot = soup_tree
    _say(f"soup_2048 tree: {ot.leaf_count} leaves, {ot.node_count} nodes, {pairs} pairs")
    assert gt.pairs == pairs
    _tree_equal(gt, ot, "soup_2048")


def test_soup_2048_bench_rays_default_schedule(ctx, soup_tree):
    """The traversal benchmark's full ray set on the 0.97 GB tree, with the schedule the
    library picks for it (persistent lanes: the tree exceeds the Infinity Cache)."""
    gt, _, ot = soup_tree
    assert gt.leaf_count * 68 + gt.node_count * 12 > 256 << 20
    rays = _bench_rays(2048)
    hits, t = ctx.traverse(gt, rays)
    _check_rays(hits, t, ot, rays, "soup_2048, camera + random rays")


def test_soup_2048_one_lane_per_ray(ctx, soup_tree):
    gt, _, ot = soup_tree
    ctx.set_dda_refill(0)
    try:
        rays = np.ascontiguousarray(_bench_rays(2048)[::7])
        hits, t = ctx.traverse(gt, rays)
        _check_rays(hits, t, ot, rays, "soup_2048 one lane per ray, every 7th ray")
    finally:
        ctx.set_dda_refill(-1)


def test_soup_2048_render(ctx, soup_tree):
    """The render path on the 0.97 GB tree (the walk pops leaf frames at their occupied
    bounds there): the oracle's image byte for byte (render.wgsl:54-91)."""
    gt, _, ot = soup_tree
    cam = W.render_camera(2048, 2048)
    _say("soup_2048 render 320x180")
    got = ctx.render(gt, cam, 320, 180)
    want = O.render(ot, cam, 320, 180, None, None)
    bad = got != want
    assert not bad.any(), f"{int(bad.sum())} pixels differ, first {np.argwhere(bad)[:3]}"


def _many_mesh_subset():
    """Every resolution of the many-mesh workload x {sphere, soup}, triangle counts drawn
    log-uniform in [100, 200,000] as there, plus two more of each kind at 1024."""
    rng = W.SplitMix64(0x5EED_0011)
    out = []
    for m in (64, 128, 256, 512, 1024):
        for kind in ("sphere", "soup"):
            u = float(rng.unit(1)[0])
            T = int(round(math.exp(math.log(100) + u * (math.log(200_000) - math.log(100)))))
            n, vs = W.grid_for(m)
            e = dict(kind=kind, T=T, m=m, n=n, voxel_size=vs, R=20_000,
                     mesh_seed=0x5EED_0100 + len(out), ray_seed=0x5EED_0200 + len(out))
            out.append(e)
    for i, kind in enumerate(("sphere", "soup", "sphere", "soup", "sphere", "soup")):
        n, vs = W.grid_for(1024 if i < 4 else 256)
        out.append(dict(kind=kind, T=[150_000, 60_000, 800, 190_000, 40_000, 120_000][i],
                        m=1024 if i < 4 else 256, n=n, voxel_size=vs, R=20_000,
                        mesh_seed=0x5EED_0300 + i, ray_seed=0x5EED_0400 + i))
    return out


def test_many_mesh_batch_against_the_oracle(ctx):
    plan = _many_mesh_subset()
    meshes = [W.batch_mesh(p) for p in plan]
    rays = [W.batch_rays(p) for p in plan]
    grids = [(p["n"], (0, 0, 0), p["voxel_size"]) for p in plan]
    _say(f"batch of {len(plan)} meshes ({sum(len(m) for m in meshes)} tris)")
    trees = ctx.voxelize_batch(meshes, grids)
    segs = ctx.traverse_batch(trees, rays)
    for p, tris, tr, r, (hits, t) in zip(plan, meshes, trees, rays, segs):
        label = f"{p['kind']} m={p['m']} T={len(tris)}"
        _, pairs, ot = _oracle_tree(tris, p["n"], p["voxel_size"])
        assert tr.pairs == pairs, label
        _tree_equal(tr, ot, label)
        _check_rays(hits, t, ot, np.ascontiguousarray(r, np.float32), label)
        tr.free()


def test_served_requests_against_the_oracle():
    """Voxelize requests of the serving workload's mesh distribution and ray queries against
    their trees, submitted together to the continuous-batching server: every tree equals the
    oracle's and every hit the mirror's."""
    _, mesh_plan, _ = W.serve_plan(duration_s=0.01)
    plan = mesh_plan[:5]
    meshes = [W.batch_mesh(p) for p in plan]
    d_mesh = [torch.from_numpy(np.ascontiguousarray(m)).to("cuda") for m in meshes]
    srv = vxl.Server(0, batching=True)
    try:
        ids = [srv.submit_voxelize(d.data_ptr(), len(m), p["n"], (0, 0, 0), p["voxel_size"])
               for d, m, p in zip(d_mesh, meshes, plan)]
        trees = [srv.wait(i)[0] for i in ids]
        oracle = []
        for p, tris, tr in zip(plan, meshes, trees):
            label = f"served {p['kind']} m={p['m']} T={len(tris)}"
            _say(label)
            _, pairs, ot = _oracle_tree(tris, p["n"], p["voxel_size"])
            _tree_equal(tr, ot, label)
            oracle.append(ot)
        rays = [np.ascontiguousarray(W.random_rays(30_000, p["m"], seed=0x5EED_0500 + i))
                for i, p in enumerate(plan)]
        d_rays = [torch.from_numpy(r).to("cuda") for r in rays]
        d_hits = [torch.zeros((len(r), 4), dtype=torch.int32, device="cuda") for r in rays]
        d_t = [torch.zeros(len(r), dtype=torch.float64, device="cuda") for r in rays]
        rids = [srv.submit_rays(tr, dr.data_ptr(), len(r), dh.data_ptr(), dt.data_ptr())
                for tr, dr, r, dh, dt in zip(trees, d_rays, rays, d_hits, d_t)]
        for rid in rids:
            srv.wait(rid)
        torch.cuda.synchronize()
        for i, (ot, r) in enumerate(zip(oracle, rays)):
            hits = d_hits[i].cpu().numpy().view(np.uint32)
            _check_rays(hits, d_t[i].cpu().numpy(), ot, r, f"served rays {i}")
        for tr in trees:
            tr.free()
    finally:
        srv.close()
