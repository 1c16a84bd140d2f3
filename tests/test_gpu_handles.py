"""Host-side refusals of handles and structures the device path must not read: trees whose
context is gone, and uploaded School-B arrays whose child indices leave the arrays. None of
these cases launches a kernel."""
import numpy as np
import oracle as O
import pytest
import torch
import vxl
from vxl import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _refused(fn):
    with pytest.raises(vxl.VxlError) as e:
        fn()
    assert e.value.name == "VXL_INVALID_ARG", e.value


def test_detached_tree_is_refused(ctx):
    _say("detached tree")
    other = vxl.Context(0)
    tree = other.voxelize_tree(W.uv_sphere(16, 8), 32, (0, 0, 0), 1 / 32)
    assert tree.leaf_count > 0
    other.close()  # releases the tree's memory and detaches the handle
    tree._refresh()
    assert (tree.leaf_count, tree.node_count, tree.device_bytes) == (0, 0, 0)
    rays = W.random_rays(64, 32)
    _refused(lambda: ctx.traverse(tree, rays))
    d_rays = torch.from_numpy(rays).to("cuda")
    d_hits = torch.zeros((64, 4), dtype=torch.int32, device="cuda")
    _refused(lambda: ctx.traverse_dev(tree, d_rays.data_ptr(), 64, d_hits.data_ptr()))
    _refused(lambda: ctx.traverse_batch([tree], [rays]))
    _refused(lambda: ctx.traverse_stats(tree, d_rays.data_ptr(), 64))
    _refused(lambda: ctx.render(tree, W.render_camera(32, 32, 8, 8), 8, 8))
    _refused(lambda: ctx.set_voxels(tree, np.array([[1, 1, 1, 1]])))
    _refused(lambda: tree.download())
    srv = vxl.Server(0)
    _refused(lambda: srv.submit_rays(tree, d_rays.data_ptr(), 64, d_hits.data_ptr()))
    srv.close()
    tree.free()


def test_collected_server_tree_dies_with_the_server(ctx):
    _say("server tree after destroy")
    tris = W.uv_sphere(16, 8)
    d = torch.from_numpy(tris).to("cuda")
    srv = vxl.Server(0)
    tree, _, _ = srv.wait(srv.submit_voxelize(d.data_ptr(), len(tris), 32, (0, 0, 0), 1 / 32))
    assert tree.leaf_count > 0
    srv.close()
    _refused(lambda: ctx.traverse(tree, W.random_rays(64, 32)))
    tree.free()


def _fixture_school_b(n=128):
    words = O.fixture_words("dust", n, 4096)
    nodes, leaves, bounds, _ = O.Tree.build(words, n).export()
    return nodes, leaves, bounds


def test_upload_checks_child_ranges(ctx):
    n = 128  # k = 2: root (level 3) -> level-2 nodes -> leaves
    nodes, leaves, bounds = _fixture_school_b(n)
    _say(f"upload checks: {len(nodes)} nodes, {len(leaves)} leaves")
    ok = ctx.tree_upload(n, nodes, leaves, bounds)
    assert ok.leaf_count == len(leaves)
    ok.free()
    bad = nodes.copy()
    bad[-1, 2] = len(leaves)  # a level-2 node whose children start past the leaves
    _refused(lambda: ctx.tree_upload(n, bad, leaves, bounds))
    bad = nodes.copy()
    bad[0, 2] = len(nodes) - 1  # the root's children run past the node array
    _refused(lambda: ctx.tree_upload(n, bad, leaves, bounds))
    _refused(lambda: ctx.tree_upload(n, nodes, leaves[:-1], bounds[:-1]))


def test_upload_rejects_a_node_reached_at_two_levels(ctx):
    n = 512  # k = 3: levels 4, 3, 2
    nodes, leaves, bounds = _fixture_school_b(n)
    root_kids = bin(int(nodes[0, 0]) | int(nodes[0, 1]) << 32).count("1")
    assert nodes[0, 2] == 1 and len(nodes) > 1 + root_kids
    _say("node at two levels")
    bad = nodes.copy()
    bad[1, 2] = 0  # a level-3 node whose children include the root
    _refused(lambda: ctx.tree_upload(n, bad, leaves, bounds))


def test_fresh_contexts_read_their_own_build_sizes():
    """A tree build waits for a sequence number its size kernel writes to mapped pinned memory.
    Pinned memory freed by a destroyed context can be handed to the next one with its contents,
    so a new context must not take an earlier context's number for its own (it would build
    from that context's sizes). Contexts and servers made one after another, each building one
    tree of a different mesh, all give the oracle's trees."""
    cases = [(W.uv_sphere(16, 8), 32, 1 / 32), (W.soup(806, seed=45, mu=np.log(2 / 8)), 8, 1 / 8),
             (W.uv_sphere(24, 12), 128, 1 / 128), (W.soup(500, seed=46, mu=np.log(2 / 32)), 32, 1 / 32)]
    want = []
    for tris, n, vs in cases:
        words, _, pairs = O.voxelize_surface(tris, n, (0, 0, 0), vs, 1e-4)
        want.append((pairs, O.Tree.build(words, n).export()))
    for rep in range(2):
        for (tris, n, vs), (pairs, arrays) in zip(cases, want):
            _say(f"fresh context {rep}: T={len(tris)} n={n}")
            c = vxl.Context(0)
            tree = c.voxelize_tree(tris, n, (0, 0, 0), vs, 1e-4)
            assert tree.pairs == pairs
            assert all(np.array_equal(a, b) for a, b in zip(tree.download(), arrays))
            tree.free()
            c.close()
            d = torch.from_numpy(tris).to("cuda")
            srv = vxl.Server(0)
            tree, got_pairs, _ = srv.wait(srv.submit_voxelize(d.data_ptr(), len(tris), n, (0, 0, 0), vs))
            assert got_pairs == pairs
            assert all(np.array_equal(a, b) for a, b in zip(tree.download(), arrays))
            tree.free()
            srv.close()
