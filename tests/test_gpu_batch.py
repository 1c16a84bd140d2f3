"""GPU: batched voxelize and traversal equal per-mesh calls, tree for tree and hit for hit.

The per-mesh path (vxl_voxelize_tree / vxl_traverse) is pinned to the oracle by
test_gpu_voxelize.py and test_gpu_dda.py; here every batched result is compared with it
bit for bit, and a few batched trees directly with the oracle.
"""
import numpy as np
import oracle as O
import pytest
import vxl
from vxl import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _arrays(tree):
    nodes, leaves, bounds, codes = tree.download()
    return dict(nodes=nodes, leaves=leaves, bounds=bounds, codes=codes,
                leaf_count=tree.leaf_count, node_count=tree.node_count,
                levels=dict(tree.level_nodes))


def _assert_same_tree(a, b, what):
    A, B = _arrays(a), _arrays(b)
    assert A["leaf_count"] == B["leaf_count"], what
    assert A["node_count"] == B["node_count"], what
    assert A["levels"] == B["levels"], what
    for key in ("nodes", "leaves", "bounds", "codes"):
        assert np.array_equal(A[key], B[key]), f"{what}: {key} differ"


def _mix(seed=7):
    """Meshes of every kind the batch path sees: n = 128, 512, 2048 (batched), 8 and 32
    (per-mesh fallback), an empty mesh, meshes with no candidate voxel."""
    rng = np.random.default_rng(seed)
    out = []
    for i, m in enumerate([64, 128, 256, 512, 1024, 64, 256, 1024]):
        n, vs = W.grid_for(m)
        if i % 2 == 0:
            tris = W.sphere_with_tris(int(rng.integers(100, 6000)), seed=100 + i)
        else:
            tris = W.soup(int(rng.integers(100, 6000)), seed=200 + i, mu=np.log(2.0 / m))
        out.append((f"{'sphere' if i % 2 == 0 else 'soup'}_{m}", tris, n, vs))
    out.append(("empty_512", np.zeros((0, 9), np.float32), 512, 1 / 256))
    out.append(("outside_128", np.full((5, 9), -3.0, np.float32), 128, 1 / 64))
    out.append(("sphere_k0_8", W.uv_sphere(8, 4), 8, 1 / 8))
    out.append(("soup_k1_32", W.soup(300, seed=9, mu=np.log(2 / 32)), 32, 1 / 32))
    out.append(("sphere_2048_small", W.sphere_with_tris(120, seed=5), 2048, 1 / 1024))
    return out


def test_batch_voxelize_equals_single(ctx):
    mix = _mix()
    _say(f"voxelize batch of {len(mix)} meshes vs single calls")
    grids = [(n, (0, 0, 0), vs) for _, _, n, vs in mix]
    batch = ctx.voxelize_batch([t for _, t, _, _ in mix], grids)
    for (name, tris, n, vs), bt in zip(mix, batch):
        st = ctx.voxelize_tree(tris, n, (0, 0, 0), vs)
        assert bt.pairs == st.pairs, name
        _assert_same_tree(bt, st, name)
        st.free()
    for t in batch:
        t.free()


def test_batch_trees_match_oracle(ctx):
    mix = [m for m in _mix(seed=11) if m[2] <= 512 and len(m[1]) <= 3000][:4]
    grids = [(n, (0, 0, 0), vs) for _, _, n, vs in mix]
    batch = ctx.voxelize_batch([t for _, t, _, _ in mix], grids)
    for (name, tris, n, vs), bt in zip(mix, batch):
        _say(f"oracle {name} n={n} T={len(tris)}")
        words, _, pairs = O.voxelize_surface(tris, n, (0, 0, 0), vs, 1e-4)
        ot = O.Tree.build(words, n)
        on, ol, ob, oc = ot.export()
        gn, gl, gb, gc = bt.download()
        assert bt.pairs == pairs, name
        assert np.array_equal(gn, on) and np.array_equal(gl, ol), name
        assert np.array_equal(gb, ob) and np.array_equal(gc, oc), name
    for t in batch:
        t.free()


def test_batch_groups_and_repeat(ctx):
    """A 1 GiB budget puts each 2048^3 mesh in a group of its own; repeated calls reuse the
    pooled grid, which must be all-zero again after every group."""
    mix = [(f"soup_2048_{i}", W.soup(2000 + 500 * i, seed=300 + i, mu=np.log(2 / 1024)), 2048,
            1 / 1024) for i in range(3)]
    mix += [(f"sphere_128_{i}", W.sphere_with_tris(800, seed=400 + i), 128, 1 / 64)
            for i in range(4)]
    grids = [(n, (0, 0, 0), vs) for _, _, n, vs in mix]
    single = [ctx.voxelize_tree(t, n, (0, 0, 0), vs) for _, t, n, vs in mix]
    try:
        ctx.set_batch_budget(1 << 30)
        for rep in range(2):
            _say(f"grouped batch, repeat {rep}")
            batch = ctx.voxelize_batch([t for _, t, _, _ in mix], grids)
            for (name, *_), bt, st in zip(mix, batch, single):
                _assert_same_tree(bt, st, f"{name} rep {rep}")
            # free in an order unrelated to the group layout
            for t in batch[::-2] + batch[-2::-2]:
                t.free()
    finally:
        ctx.set_batch_budget(16 << 30)
    for t in single:
        t.free()


def test_batch_traverse_equals_single(ctx):
    mix = _mix(seed=13)
    grids = [(n, (0, 0, 0), vs) for _, _, n, vs in mix]
    trees = ctx.voxelize_batch([t for _, t, _, _ in mix], grids)
    rays = []
    for i, (name, _, n, vs) in enumerate(mix):
        m = max(int(round(1 / vs)), 1)
        R = [1000, 0, 4097, 333, 1, 2500, 64, 256, 700, 50, 900, 1200, 3000][i % 13]
        rays.append(W.random_rays(R, m, seed=500 + i) if R else np.zeros((0, 6), np.float32))
    want = [ctx.traverse(tr, r) if len(r) else None for tr, r in zip(trees, rays)]
    try:
        # one lane per ray, then persistent lanes over the run table (runs of 256 rays of one
        # segment; a wave's lanes walk several trees at once), refilled at 1, 8 and 64 idle lanes
        for refill in (0, 1, 8, 64):
            _say(f"traverse batch of {len(trees)} segments, "
                 f"{sum(len(r) for r in rays)} rays, batch refill {refill}")
            ctx.set_batch_refill(refill)
            got = ctx.traverse_batch(trees, rays)
            for (name, *_), r, w, (h, t) in zip(mix, rays, want, got):
                if len(r) == 0:
                    continue
                assert np.array_equal(h, w[0]), f"{name} refill {refill}"
                assert np.array_equal(t, w[1]), f"{name} refill {refill}"
        for bad in (-1, 65):
            with pytest.raises(vxl.VxlError):
                ctx.set_batch_refill(bad)
    finally:
        ctx.set_batch_refill(0)
    for t in trees:
        t.free()


def test_batch_traverse_dev_block_sizes(ctx):
    import torch
    mix = _mix(seed=17)[:6]
    grids = [(n, (0, 0, 0), vs) for _, _, n, vs in mix]
    trees = ctx.voxelize_batch([t for _, t, _, _ in mix], grids)
    rays = [W.random_rays(3000 + 111 * i, int(round(1 / vs)), seed=600 + i)
            for i, (_, _, n, vs) in enumerate(mix)]
    off = np.zeros(len(rays) + 1, np.uint64)
    off[1:] = np.cumsum([len(r) for r in rays])
    allr = np.concatenate(rays)
    want = [ctx.traverse(tr, r) for tr, r in zip(trees, rays)]
    dr = torch.from_numpy(allr).cuda()
    try:
        for b in (64, 128, 256):
            _say(f"traverse_batch_dev dda_block={b}")
            ctx.set_tuning(dda_block=b)
            dh = torch.zeros((len(allr), 4), dtype=torch.int32, device="cuda")
            dt = torch.zeros(len(allr), dtype=torch.float64, device="cuda")
            ctx.traverse_batch_dev(trees, dr.data_ptr(), off, dh.data_ptr(), dt.data_ptr())
            torch.cuda.synchronize()
            h = dh.cpu().numpy().view(np.uint32)
            t = dt.cpu().numpy()
            for i, (wh, wt) in enumerate(want):
                a, z = int(off[i]), int(off[i + 1])
                assert np.array_equal(h[a:z], wh) and np.array_equal(t[a:z], wt), (b, i)
    finally:
        ctx.set_tuning(dda_block=64)
    for t in trees:
        t.free()


def test_merge_words_keep_results(ctx):
    """The SAT's merge tile (words merged in LDS per wave before the flush) changes no tree,
    from a flush after every insert to the largest tile; out-of-range values are refused."""
    mix = [m for m in _mix(seed=23) if 128 <= m[2] <= 2048][:4]
    grids = [(n, (0, 0, 0), vs) for _, _, n, vs in mix]
    tris = [t for _, t, _, _ in mix]
    ref = ctx.voxelize_batch(tris, grids)
    try:
        for words in (1, 16, 63):
            _say(f"merge words {words}")
            ctx.set_merge_words(words)
            got = ctx.voxelize_batch(tris, grids)
            for (name, *_), a, b in zip(mix, ref, got):
                assert a.pairs == b.pairs, name
                _assert_same_tree(a, b, f"{name} merge {words}")
            single = ctx.voxelize_tree(tris[0], mix[0][2], (0, 0, 0), mix[0][3])
            _assert_same_tree(single, ref[0], f"{mix[0][0]} single, merge {words}")
            for t in got + [single]:
                t.free()
        for bad in (0, 65, -1):
            with pytest.raises(vxl.VxlError):
                ctx.set_merge_words(bad)
    finally:
        ctx.set_merge_words(64)
    for t in ref:
        t.free()


def test_autotune_keeps_results(ctx):
    import torch
    mix = _mix(seed=19)
    grids = [(n, (0, 0, 0), vs) for _, _, n, vs in mix]
    tris = [t for _, t, _, _ in mix]
    off = np.zeros(len(tris) + 1, np.uint64)
    off[1:] = np.cumsum([len(t) for t in tris])
    dtri = torch.from_numpy(np.concatenate(tris)).cuda()
    rays = [W.random_rays(2000, int(round(1 / vs)), seed=700 + i)
            for i, (_, _, n, vs) in enumerate(mix)]
    roff = np.zeros(len(rays) + 1, np.uint64)
    roff[1:] = np.cumsum([len(r) for r in rays])
    dr = torch.from_numpy(np.concatenate(rays)).cuda()
    dh = torch.zeros((len(dr), 4), dtype=torch.int32, device="cuda")
    dt = torch.zeros(len(dr), dtype=torch.float64, device="cuda")
    before = ctx.voxelize_batch(tris, grids)
    _say("autotune")
    try:
        res = ctx.autotune_batch(dtri.data_ptr(), off, grids, dr.data_ptr(), roff,
                                 dh.data_ptr(), dt.data_ptr())
        assert res["sat_waves_per_cu"] in (16, 20, 24, 28, 32)
        assert res["sat_units_per_wave"] in (1, 2, 4)
        assert res["dda_block"] in (64, 128, 256)
        assert res["sat_merge_words"] in (16, 32, 48, 64)
        assert res["dda_batch_refill"] in (0, 8, 16, 32)
        assert res["voxelize_ms"] > 0 and res["traverse_ms"] > 0
        after = ctx.voxelize_batch_dev(dtri.data_ptr(), off, grids)
        torch.cuda.synchronize()
        for (name, *_), a, b in zip(mix, before, after):
            _assert_same_tree(a, b, name)
    finally:
        ctx.set_tuning(sat_waves_per_cu=28, sat_units_per_wave=1, dda_block=64)
        ctx.set_merge_words(64)
        ctx.set_batch_refill(0)
    for t in before + after:
        t.free()


def test_batch_trees_outlive_context():
    c = vxl.Context(0)
    mix = _mix(seed=23)[:3]
    trees = c.voxelize_batch([t for _, t, _, _ in mix], [(n, (0, 0, 0), vs) for _, _, n, vs in mix])
    trees[1].free()
    c.close()  # releases the group's shared allocation once
    for t in trees:
        t.free()


def test_batch_rejects_bad_input(ctx):
    with pytest.raises(vxl.VxlError) as e:
        ctx.voxelize_batch([W.uv_sphere(8, 4)], [(100, (0, 0, 0), 1 / 50)])
    assert e.value.name == "VXL_NOT_REPRESENTABLE"
    with pytest.raises(vxl.VxlError) as e:
        ctx.voxelize_batch([W.uv_sphere(8, 4)], [(128, (0, 0, 0), 1 / 64)], eps=-1.0)
    assert e.value.name == "VXL_INVALID_ARG"
    with pytest.raises(vxl.VxlError):
        ctx.set_batch_budget(1 << 20)


def test_traverse_graph_replays_equal_eager(ctx):
    """The per-tree traversals captured in a HIP graph: every replay (with eager work and
    synchronisations between replays, outputs cleared before each) gives the eager hits."""
    import torch
    print("[case] traverse graph replays", flush=True)
    trees, rays = [], []
    for i, (m, kind) in enumerate(((32, "sphere"), (128, "soup"), (256, "sphere"), (64, "soup"),
                                   (512, "soup"), (128, "sphere"))):
        n, vs = W.grid_for(m)
        tris = W.uv_sphere(24 + 8 * i, 12 + 4 * i) if kind == "sphere" else \
            W.soup(2000 + 500 * i, seed=40 + i, mu=np.log(4 / m))
        trees.append(ctx.voxelize_tree(tris, n, (0, 0, 0), vs))
        rays.append(W.random_rays(1000 + 3000 * i, m, seed=50 + i))
    off = np.zeros(len(rays) + 1, np.uint64)
    off[1:] = np.cumsum([len(r) for r in rays])
    R = int(off[-1])
    d_rays = torch.from_numpy(np.concatenate(rays)).cuda()
    h_eager = torch.zeros((R, 4), dtype=torch.int32, device="cuda")
    t_eager = torch.zeros(R, dtype=torch.float64, device="cuda")
    for g, tr in enumerate(trees):
        a, z = int(off[g]), int(off[g + 1])
        ctx.traverse_dev(tr, d_rays.data_ptr() + 24 * a, z - a, h_eager.data_ptr() + 16 * a,
                         t_eager.data_ptr() + 8 * a)
    torch.cuda.synchronize()
    h = torch.zeros_like(h_eager)
    t = torch.zeros_like(t_eager)
    graph = ctx.traverse_graph(trees, d_rays.data_ptr(), off, h.data_ptr(), t.data_ptr())
    for rep in range(3):
        h.fill_(-1)
        t.fill_(-1.0)
        torch.cuda.synchronize()
        graph.launch()
        ctx.traverse_dev(trees[0], d_rays.data_ptr(), int(off[1]), h_eager.data_ptr(),
                         t_eager.data_ptr())  # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(h, h_eager) and torch.equal(t, t_eager), f"replay {rep}"
    graph.free()


def test_batch_over_16m_triangles_equals_single(ctx):
    """A batch whose meshes need more than 65,536 setup blocks scans the candidate counts
    device-wide; its trees still equal the per-mesh calls."""
    base = W.uv_sphere(48, 24)
    reps = -(-(65536 * 256 + 1) // len(base))
    mix = [("tiled_sphere_128", np.tile(base, (reps, 1)), 128, 1 / 128),
           ("soup_512", W.soup(3000, seed=31, mu=np.log(2.0 / 256)), 512, 1 / 256)]
    _say(f"batch with {sum(len(t) for _, t, _, _ in mix)} triangles vs single calls")
    batch = ctx.voxelize_batch([t for _, t, _, _ in mix], [(n, (0, 0, 0), vs) for _, _, n, vs in mix])
    for (name, tris, n, vs), bt in zip(mix, batch):
        _assert_same_tree(bt, ctx.voxelize_tree(tris, n, (0, 0, 0), vs, 1e-4), name)
