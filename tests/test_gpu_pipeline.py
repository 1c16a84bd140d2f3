"""Back-to-back single-tree voxelize calls on device triangles, as a caller streaming meshes
makes them: each call's tree build reads its sizes back while the previous call's leaf
gather may still run, and the context's scratch (records, plan, level arrays, the size
words) is reused call after call. Every tree must equal the oracle's
(reference_cpu.rs:110-200 -> sparse.rs:241-305 -> school_b.rs:113-151) whatever came before
it: growing and shrinking meshes (the scratch grows), other calls in between, trees freed or
kept, and trees read back only after every call was issued."""
import numpy as np
import oracle as O
import pytest
import torch
import vxl
from vxl import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _oracle(tris, n, vs):
    words, _, pairs = O.voxelize_surface(tris, n, (0, 0, 0), vs, 1e-4)
    return pairs, O.Tree.build(words, n)


def _equal(gt, pairs, ot, what):
    assert gt.pairs == pairs, what
    gn, gl, gb, gc = gt.download()
    on, ol, ob, oc = ot.export()
    assert (gt.leaf_count, gt.node_count) == (ot.leaf_count, ot.node_count), what
    assert np.array_equal(gc, oc), f"{what}: leaf codes differ"
    assert np.array_equal(gl, ol), f"{what}: leaf words differ"
    assert np.array_equal(gb, ob), f"{what}: bounds differ"
    assert np.array_equal(gn, on), f"{what}: nodes differ"


def _meshes():
    n1, vs1 = W.grid_for(256)
    n2, vs2 = W.grid_for(64)
    return {
        "sphere_256": (W.uv_sphere(), n1, vs1),
        "soup_64": (W.soup(3000, seed=0x5EED_0901), n2, vs2),
        "soup_256": (W.soup(40_000, seed=0x5EED_0902), n1, vs1),
        "sphere_64": (W.uv_sphere(32, 16), n2, vs2),
    }


@pytest.fixture(scope="module")
def cases():
    out = {}
    for name, (tris, n, vs) in _meshes().items():
        pairs, ot = _oracle(tris, n, vs)
        out[name] = (torch.from_numpy(np.ascontiguousarray(tris)).to("cuda"), len(tris), n, vs,
                     pairs, ot)
    yield out


def _build(ctx, case):
    d, T, n, vs, _, _ = case
    return ctx.voxelize_tree_dev(d.data_ptr(), T, n, (0, 0, 0), vs)


def test_back_to_back_calls_equal_the_oracle(cases):
    ctx = vxl.Context(0)
    try:
        order = ["soup_64", "sphere_256", "sphere_256", "soup_256", "sphere_64", "soup_256",
                 "sphere_256", "soup_64", "soup_64", "sphere_256"]
        kept = []
        for i, name in enumerate(order):
            _say(f"call {i}: {name}")
            tr = _build(ctx, cases[name])
            kept.append((name, tr))
        # trees read back only after every call was issued
        for i, (name, tr) in enumerate(kept):
            _, _, _, _, pairs, ot = cases[name]
            _equal(tr, pairs, ot, f"call {i} ({name})")
            tr.free()
    finally:
        ctx.close()


def test_other_calls_in_between(cases):
    ctx = vxl.Context(0)
    try:
        name = "sphere_256"
        d, T, n, vs, pairs, ot = cases[name]
        rays = np.ascontiguousarray(W.random_rays(20_000, 256, seed=0x5EED_0903))
        want = ot.mirror(rays)
        for i in range(3):
            _say(f"round {i}: voxelize, dense, traverse, voxelize")
            t1 = _build(ctx, cases[name])
            words, p = ctx.voxelize_dense(W.uv_sphere(), n, (0, 0, 0), vs, 1e-4)
            assert p == pairs
            hits, _ = ctx.traverse(t1, rays)
            assert np.array_equal(hits, want)
            t2 = _build(ctx, cases["soup_256"])
            t3 = _build(ctx, cases[name])
            _equal(t1, pairs, ot, f"round {i} first")
            _equal(t3, pairs, ot, f"round {i} third")
            _equal(t2, cases["soup_256"][4], cases["soup_256"][5], f"round {i} soup")
            for t in (t1, t2, t3):
                t.free()
    finally:
        ctx.close()
