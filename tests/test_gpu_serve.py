"""GPU: the request loop. Every result served in a mix (continuous batching, both lanes
overlapping, admission under a byte budget) equals the result of the same request alone
through the per-request entry points; the plain mode serves the same results."""
import numpy as np
import pytest
import torch
import vxl
from vxl import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)


class Fixture:
    """Resident trees, a ray pool per tree, a mesh pool, and the alone results."""

    def __init__(self, ctx, seed=1):
        rng = np.random.default_rng(seed)
        self.trees, self.ms = [], []
        for i, m in enumerate([64, 128, 256, 512]):
            n, vs = W.grid_for(m)
            tris = (W.sphere_with_tris(3000, seed=10 + i) if i % 2 == 0
                    else W.soup(4000, seed=20 + i, mu=np.log(2.0 / m)))
            self.trees.append(ctx.voxelize_tree(tris, n, (0, 0, 0), vs))
            self.ms.append(m)
        self.meshes = []
        for i, m in enumerate([64, 128, 256, 512, 1024, 8, 32]):
            n, vs = W.grid_for(m) if m >= 64 else (m, 1.0 / m)
            tris = (W.sphere_with_tris(int(rng.integers(100, 5000)), seed=30 + i) if i % 2 == 0
                    else W.soup(int(rng.integers(100, 5000)), seed=40 + i, mu=np.log(2.0 / m)))
            self.meshes.append((tris, n, vs))
        self.meshes.append((np.zeros((0, 9), np.float32), 512, 1 / 256))
        # requests: (kind, index, R or None)
        self.reqs = []
        for j in range(60):
            if j % 3 == 2:
                self.reqs.append(("vox", j % len(self.meshes), None))
            else:
                self.reqs.append(("rays", j % len(self.trees), int(rng.integers(1, 20000))))
        self.rays = [W.random_rays(20000, m, seed=50 + i) for i, m in enumerate(self.ms)]
        self.d_rays = [torch.from_numpy(r).cuda() for r in self.rays]
        self.d_tris = [torch.from_numpy(np.ascontiguousarray(t, np.float32)).cuda()
                       if len(t) else torch.zeros(9, device="cuda") for t, _, _ in self.meshes]
        self.alone_rays = [ctx.traverse(tr, r) for tr, r in zip(self.trees, self.rays)]
        self.alone_trees = [ctx.voxelize_tree(t, n, (0, 0, 0), vs) for t, n, vs in self.meshes]

    def run(self, server):
        R_total = sum(r[2] for r in self.reqs if r[0] == "rays")
        hits = torch.zeros((R_total, 4), dtype=torch.int32, device="cuda")
        t = torch.zeros(R_total, dtype=torch.float64, device="cuda")
        ids, at = [], 0
        for kind, i, R in self.reqs:
            if kind == "rays":
                rid = server.submit_rays(self.trees[i], self.d_rays[i].data_ptr(), R,
                                         hits.data_ptr() + 16 * at, t.data_ptr() + 8 * at)
                ids.append((rid, kind, i, R, at))
                at += R
            else:
                tris, n, vs = self.meshes[i]
                rid = server.submit_voxelize(self.d_tris[i].data_ptr(), len(tris), n, (0, 0, 0),
                                             vs)
                ids.append((rid, kind, i, None, None))
        out = []
        for rid, kind, i, R, a in ids:
            tree, pairs, lat = server.wait(rid)
            assert lat >= 0
            out.append((kind, i, R, a, tree, pairs))
        torch.cuda.synchronize()
        return out, hits.cpu().numpy().view(np.uint32), t.cpu().numpy()

    def check(self, out, hits, t):
        bad = []
        for pos, (kind, i, R, a, tree, pairs) in enumerate(out):
            if kind == "rays":
                wh, wt = self.alone_rays[i]
                if not (np.array_equal(hits[a:a + R], wh[:R]) and np.array_equal(t[a:a + R], wt[:R])):
                    bad.append(("rays", pos, i, R))
            else:
                want = self.alone_trees[i]
                same = (tree.leaf_count, tree.node_count) == (want.leaf_count, want.node_count)
                same = same and all(np.array_equal(x, y)
                                    for x, y in zip(tree.download(), want.download()))
                if pairs != want.pairs or not same:
                    bad.append(("vox", pos, i, pairs, want.pairs, "tree same" if same else
                                "tree differs"))
                tree.free()
        assert not bad, bad


@pytest.fixture(scope="module")
def fx(ctx):
    return Fixture(ctx)


@pytest.mark.parametrize("batching,refill", [(True, 0), (True, 16), (False, 0)],
                         ids=["batching", "batching_persistent_lanes", "plain"])
def test_served_results_equal_alone(fx, batching, refill):
    _say(f"server batching={batching} batch_refill={refill}: {len(fx.reqs)} requests")
    srv = vxl.Server(0, batching=batching, batch_refill=refill)
    out, hits, t = fx.run(srv)
    fx.check(out, hits, t)
    st = srv.stats()
    assert st["completed"] == len(fx.reqs) and st["failed"] == 0
    if batching:
        assert st["batches"] < len(fx.reqs)  # requests shared launches
    else:
        assert st["batches"] == len(fx.reqs)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    srv.close()


def test_server_rejects_bad_refill():
    with pytest.raises(vxl.VxlError):
        vxl.Server(0, batch_refill=65)


def test_admission_under_tight_budget(fx):
    """A budget that fits about one large request at a time: every request still completes,
    with identical results, and the held + in-flight footprint never exceeds the budget."""
    budget = 1_200_000_000  # one 2048^3 voxelize (1 GiB grid region) at a time
    srv = vxl.Server(0, budget_bytes=budget, batching=True)
    srv.hold(4 << 20)
    _say("tight budget")
    out, hits, t = fx.run(srv)
    fx.check(out, hits, t)
    st = srv.stats()
    assert st["completed"] == len(fx.reqs)
    assert st["peak_bytes"] <= budget
    srv.hold(-(4 << 20))
    # a request that can never fit is refused at submit
    big = vxl.Server(0, budget_bytes=1 << 19, batching=True)
    spare = torch.zeros((20000, 4), dtype=torch.int32, device="cuda")
    with pytest.raises(vxl.VxlError) as e:
        big.submit_rays(fx.trees[0], fx.d_rays[0].data_ptr(), 20000, spare.data_ptr(), 0)
    assert e.value.name == "VXL_STORAGE_EXCEEDED"
    big.close()
    srv.close()


def test_destroy_drains_pending(fx):
    srv = vxl.Server(0, batching=True)
    hits = torch.zeros((5000, 4), dtype=torch.int32, device="cuda")
    for _ in range(8):
        srv.submit_rays(fx.trees[1], fx.d_rays[1].data_ptr(), 5000, hits.data_ptr(), 0)
    tris, n, vs = fx.meshes[2]
    srv.submit_voxelize(fx.d_tris[2].data_ptr(), len(tris), n, (0, 0, 0), vs)
    srv.close()  # serves all nine, releases the uncollected tree
    torch.cuda.synchronize()
    wh, _ = fx.alone_rays[1]
    assert np.array_equal(hits.cpu().numpy().view(np.uint32), wh[:5000])


def test_submit_validation(fx):
    srv = vxl.Server(0, batching=True)
    with pytest.raises(vxl.VxlError) as e:
        srv.submit_voxelize(fx.d_tris[0].data_ptr(), 10, 100, (0, 0, 0), 0.01)
    assert e.value.name == "VXL_NOT_REPRESENTABLE"
    with pytest.raises(vxl.VxlError) as e:
        srv.submit_voxelize(fx.d_tris[0].data_ptr(), 10, 128, (0, 0, 0), 0.01, eps=-1)
    assert e.value.name == "VXL_INVALID_ARG"
    with pytest.raises(vxl.VxlError):
        srv.wait(987654321)
    srv.close()


def test_one_wait_per_request(fx):
    """Two threads waiting on one request: one collects it, the other is refused with
    VXL_INVALID_ARG (the request is claimed, or already retired), and a wait after
    retirement is refused the same way."""
    import threading
    srv = vxl.Server(0, batching=True)
    R = 200_000
    hits = torch.zeros((R, 4), dtype=torch.int32, device="cuda")
    rays = fx.d_rays[3][:R] if fx.d_rays[3].shape[0] >= R else fx.d_rays[3]
    R = rays.shape[0]
    _say(f"double wait, {R} rays")
    rid = srv.submit_rays(fx.trees[3], rays.data_ptr(), R, hits.data_ptr(), 0)
    results = []

    def waiter():
        try:
            srv.wait(rid, keep_tree=False)
            results.append("ok")
        except vxl.VxlError as e:
            results.append(e.name)

    ths = [threading.Thread(target=waiter) for _ in range(2)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    assert sorted(results) == ["VXL_INVALID_ARG", "ok"], results
    with pytest.raises(vxl.VxlError) as e:
        srv.wait(rid)
    assert e.value.name == "VXL_INVALID_ARG"
    srv.close()
