"""GPU: the whole batch as one HIP graph (vxl_batch_graph_*): voxelize every mesh, build its
tree with the sizes left on the device, walk its rays -- replay after replay equal to the
eager batch calls on the same triangles and rays, tree for tree and hit for hit.

The eager batch path is pinned to the per-mesh path and the oracle by test_gpu_batch.py;
here one graph tree is also compared with the oracle directly (reference_cpu.rs:110-200 ->
sparse.rs:241-305 -> school_b.rs:113-151). Replays run with other work and host
synchronisations between them, on new triangle contents, after a replay whose trees did not
fit the arena (which must leave the pooled grid clean), and each is checked on its own.
"""
import numpy as np
import oracle as O
import pytest
import torch
import vxl
from vxl import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _same_tree(a, b, what):
    assert (a.leaf_count, a.node_count) == (b.leaf_count, b.node_count), what
    assert dict(a.level_nodes) == dict(b.level_nodes), what
    for x, y, key in zip(a.download(), b.download(), ("nodes", "leaves", "bounds", "codes")):
        assert np.array_equal(x, y), f"{what}: {key} differ"


def _batch(seed):
    """Meshes of uneven size at n = 128, 512 and 2048 (one empty), with uneven ray counts
    (one mesh with none)."""
    specs = [
        ("sphere_256", W.uv_sphere(), 256, 20_000),
        ("soup_64", W.soup(3000, seed=seed + 1, mu=np.log(2 / 64)), 64, 5000),
        ("empty_64", np.zeros((0, 9), np.float32), 64, 3000),
        ("soup_1024", W.soup(20_000, seed=seed + 2, mu=np.log(2 / 1024)), 1024, 12_000),
        ("sphere_64", W.uv_sphere(32, 16), 64, 700),
        ("soup_256", W.soup(8000, seed=seed + 3, mu=np.log(2 / 256)), 256, 0),
    ]
    out = []
    for i, (name, tris, m, R) in enumerate(specs):
        n, vs = W.grid_for(m)
        rays = W.random_rays(R, m, seed=seed + 10 + i) if R else np.zeros((0, 6), np.float32)
        out.append((name, np.ascontiguousarray(tris, np.float32), n, vs, m,
                    np.ascontiguousarray(rays, np.float32)))
    return out


def _offsets(sizes):
    off = np.zeros(len(sizes) + 1, np.uint64)
    off[1:] = np.cumsum(sizes)
    return off


class _Buffers:
    """Device triangles, rays and result buffers of a batch."""

    def __init__(self, batch):
        self.names = [b[0] for b in batch]
        self.grids = [(b[2], (0, 0, 0), b[3]) for b in batch]
        self.tri_off = _offsets([len(b[1]) for b in batch])
        self.ray_off = _offsets([len(b[5]) for b in batch])
        self.host_tris = np.concatenate([b[1] for b in batch])
        self.d_tris = torch.from_numpy(self.host_tris.copy()).cuda()
        self.d_rays = torch.from_numpy(np.concatenate([b[5] for b in batch])).cuda()
        R = int(self.ray_off[-1])
        self.d_hits = torch.zeros((R, 4), dtype=torch.int32, device="cuda")
        self.d_t = torch.zeros(R, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()

    def set_tris(self, tris):
        self.d_tris.copy_(torch.from_numpy(np.ascontiguousarray(tris, np.float32)))
        torch.cuda.synchronize()

    def graph(self, ctx, headroom=0.0):
        return ctx.batch_graph(self.d_tris.data_ptr(), self.tri_off, self.grids,
                               self.d_rays.data_ptr(), self.ray_off, self.d_hits.data_ptr(),
                               self.d_t.data_ptr(), headroom=headroom)

    def eager(self, ctx):
        """The eager batch calls on the current contents: (trees, hits, t)."""
        trees = ctx.voxelize_batch_dev(self.d_tris.data_ptr(), self.tri_off, self.grids)
        R = int(self.ray_off[-1])
        hits = torch.zeros((R, 4), dtype=torch.int32, device="cuda")
        t = torch.zeros(R, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        ctx.traverse_batch_dev(trees, self.d_rays.data_ptr(), self.ray_off, hits.data_ptr(),
                               t.data_ptr())
        torch.cuda.synchronize()
        return trees, hits.cpu().numpy(), t.cpu().numpy()

    def results(self):
        torch.cuda.synchronize()
        return self.d_hits.cpu().numpy(), self.d_t.cpu().numpy()


def _check_replay(graph, buf, ref, what):
    ref_trees, ref_hits, ref_t = ref
    trees, pairs = graph.trees()
    hits, t = buf.results()
    try:
        for name, a, b in zip(buf.names, trees, ref_trees):
            assert a.pairs == b.pairs, f"{what} {name}: pairs"
            _same_tree(a, b, f"{what} {name}")
        assert np.array_equal(hits, ref_hits), f"{what}: hits differ"
        assert np.array_equal(t, ref_t), f"{what}: t differs"
    finally:
        for tr in trees:
            tr.free()


def _shifted(batch, frac):
    """Every mesh moved by frac of a voxel along x (same triangle counts, other trees)."""
    out = []
    for name, tris, n, vs, m, rays in batch:
        moved = tris.copy()
        moved[:, 0::3] += np.float32(frac * vs)
        out.append(moved)
    return np.concatenate(out)


def test_graph_replays_equal_eager_batch(ctx):
    batch = _batch(0x5EED_0A00)
    buf = _Buffers(batch)
    _say("eager batch on the first contents")
    ref0 = buf.eager(ctx)
    _say("create the graph")
    g = buf.graph(ctx)
    try:
        info = g.info()
        _say(f"graph: {info}")
        assert info["groups"] >= 1 and info["arena_bytes"] > 0
        _say("replay 1 (first contents)")
        g.launch()
        _check_replay(g, buf, ref0, "replay 1")
        # one graph tree against the oracle directly
        trees, _ = g.trees()
        try:
            name, tris, n, vs, m, rays = batch[4]
            words, _, pairs = O.voxelize_surface(tris, n, (0, 0, 0), vs, 1e-4)
            ot = O.Tree.build(words, n)
            assert trees[4].pairs == pairs
            for x, y, key in zip(trees[4].download(), ot.export(),
                                 ("nodes", "leaves", "bounds", "codes")):
                assert np.array_equal(x, y), f"{name}: {key} differ from the oracle"
            hits, _ = buf.results()
            a, b = int(buf.ray_off[4]), int(buf.ray_off[5])
            assert np.array_equal(hits[a:b].view(np.uint32), ot.mirror(rays)), name
        finally:
            for tr in trees:
                tr.free()

        _say("replay 2 (every mesh moved by 0.37 voxel), eager work between")
        buf.set_tris(_shifted(batch, 0.37))
        ref1 = buf.eager(ctx)
        g.launch()
        _check_replay(g, buf, ref1, "replay 2")
        for tr in ref1[0]:
            tr.free()

        _say("replay 3 and 4 (first contents again, back to back)")
        buf.set_tris(buf.host_tris)
        g.launch()
        g.launch()
        _check_replay(g, buf, ref0, "replay 4")

        _say("replay 5 on the context's stream, 6 on another stream right after it")
        side = torch.cuda.Stream()
        buf.d_hits.zero_()
        torch.cuda.synchronize()
        g.launch()
        g.launch(side.cuda_stream)
        _check_replay(g, buf, ref0, "replay 6 (other stream)")
    finally:
        g.free()
        for tr in ref0[0]:
            tr.free()


def test_graph_persistent_traversal_equals_eager(ctx):
    """The graph captured with the persistent batched traversal (runs of 256 rays of one
    segment, lanes refilled at 8 idle) gives the eager batch calls' trees and hits."""
    batch = _batch(0x5EED_0C00)
    buf = _Buffers(batch)
    ctx.set_batch_refill(8)
    try:
        ref = buf.eager(ctx)
        g = buf.graph(ctx)
        try:
            for i in range(2):
                _say(f"persistent-lane graph, replay {i + 1}")
                g.launch()
                _check_replay(g, buf, ref, f"persistent replay {i + 1}")
        finally:
            g.free()
        ctx.set_batch_refill(0)
        plain = buf.eager(ctx)
        assert np.array_equal(plain[1], ref[1]) and np.array_equal(plain[2], ref[2])
        for tr in plain[0]:
            tr.free()
    finally:
        ctx.set_batch_refill(0)
        for tr in ref[0]:
            tr.free()


def test_graph_overflow_reports_and_recovers(ctx):
    batch = [b for b in _batch(0x5EED_0B00) if b[0] in ("soup_256", "sphere_64", "soup_64")]
    buf = _Buffers(batch)
    ref = buf.eager(ctx)
    _say("graph with no headroom")
    g = buf.graph(ctx, headroom=1.0)
    try:
        g.launch()
        _check_replay(g, buf, ref, "fits")
        _say("soup triangles scaled 3x about their centroids: more leaves than the arena holds")
        big = []
        for name, tris, n, vs, m, rays in batch:
            t9 = tris.reshape(-1, 3, 3)
            c = t9.mean(axis=1, keepdims=True)
            big.append(((t9 - c) * (3.0 if name.startswith("soup") else 1.0) + c).reshape(-1, 9))
        buf.set_tris(np.concatenate(big).astype(np.float32))
        g.launch()
        with pytest.raises(vxl.VxlError) as e:
            g.trees()
        assert e.value.status == 3, e.value
        hits, _ = buf.results()
        assert not hits[:, 3].any(), "an overflowed group reports misses"
        _say("first contents again: the replay is whole")
        buf.set_tris(buf.host_tris)
        g.launch()
        _check_replay(g, buf, ref, "after overflow")
    finally:
        g.free()
        for tr in ref[0]:
            tr.free()


def test_graph_rejects_meshes_outside_the_batch_range(ctx):
    tris = torch.from_numpy(W.uv_sphere(8, 4)).cuda()
    torch.cuda.synchronize()
    with pytest.raises(vxl.VxlError) as e:
        ctx.batch_graph(tris.data_ptr(), [0, len(tris)], [(32, (0, 0, 0), 1 / 32)])
    assert e.value.status == 2
