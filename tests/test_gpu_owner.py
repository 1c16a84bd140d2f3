"""GPU parity of the store_owner / store_color outputs (reference_cpu.rs:118-129,163-169,
173-186): per-voxel lowest covering triangle index and its hash colour, identical to the
oracle on the same inputs, together with the occupancy words."""
import numpy as np
import oracle as O
import pytest
from helpers import kats
from vxl import workloads as W

pytestmark = pytest.mark.gpu
M = kats()["voxelize_meshes"]
NONE = 0xFFFFFFFF


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _cases():
    rng = np.random.default_rng(321)
    tri = np.array(M["owner_triangle"], np.float32)
    soup = W.soup(4000, seed=91, mu=np.log(6 / 128))
    return [
        ("owner_kat_8", np.stack([tri, tri, tri]), 8, 1.0),
        ("sphere_128", W.uv_sphere(32, 16), 128, 1 / 128),
        ("soup_128", soup, 128, 1 / 128),
        # every triangle twice, second copy later: owners are the first copies' indices
        ("soup_dup_128", np.concatenate([soup[:500], soup[:500]]), 128, 1 / 128),
        ("big_random_128", rng.uniform(0.0, 1.0, (30, 9)).astype(np.float32), 128, 1 / 128),
        # two triangles, one with a 12.4M-cell unhoisted box (64-bit decode), overlapping a
        # small one listed after it: owners of the shared voxels are the wide triangle's
        ("wide_unhoisted_512", np.array([[0.1234567, 0.7654321, 0.3141592, 250.3, 10.7, 200.1,
                                          30.2, 245.9, 120.6],
                                         [95.5, 95.5, 106.25, 110.5, 98.5, 128.25, 97.5, 112.5,
                                          121.25]],
                                        np.float32), 512, 1.0),
    ]


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_owner_and_color_match_oracle(ctx, case):
    name, tris, n, vs = case
    _say(f"owner {name} n={n} T={len(tris)}")
    want_w, want_o, pairs = O.voxelize_surface(tris, n, (0, 0, 0), vs, 1e-4, owner=True)
    w, own, col, gpairs = ctx.voxelize_dense_owner(tris, n, (0, 0, 0), vs, 1e-4, owner=True,
                                                   color=True)
    assert gpairs == pairs
    assert np.array_equal(w, want_w)
    bad = own != want_o
    assert not bad.any(), f"{int(bad.sum())} owners differ, first voxels {np.nonzero(bad)[0][:5]}"
    assert np.array_equal(col, O.owner_colors(want_o))
    # occupancy and ownership agree: owned <=> occupied
    occ = np.unpackbits(w.view(np.uint8), bitorder="little")[: n ** 3].astype(bool)
    assert np.array_equal(occ, own != NONE)
    if name == "soup_dup_128":
        assert own[own != NONE].max() < 500


def test_color_alone_and_owner_alone(ctx):
    tris = W.uv_sphere(32, 16)
    _say("colour without owner array; owner without colour")
    _, want_o, _ = O.voxelize_surface(tris, 128, (0, 0, 0), 1 / 128, 1e-4, owner=True)
    _, own, col, _ = ctx.voxelize_dense_owner(tris, 128, (0, 0, 0), 1 / 128, 1e-4, owner=False,
                                              color=True)
    assert own is None and np.array_equal(col, O.owner_colors(want_o))
    _, own, col, _ = ctx.voxelize_dense_owner(tris, 128, (0, 0, 0), 1 / 128, 1e-4, owner=True,
                                              color=False)
    assert col is None and np.array_equal(own, want_o)


def test_owner_empty_mesh_and_repeat(ctx):
    _say("empty mesh owners; repeated calls start from no owner")
    w, own, col, pairs = ctx.voxelize_dense_owner(np.zeros((0, 9), np.float32), 32, color=True)
    assert pairs == 0 and not w.any() and np.all(own == NONE) and not col.any()
    a = W.soup(300, seed=5, mu=np.log(4 / 32))
    b = W.soup(300, seed=6, mu=np.log(4 / 32))
    ctx.voxelize_dense_owner(a, 32, (0, 0, 0), 1 / 32)
    _, own_b, _, _ = ctx.voxelize_dense_owner(b, 32, (0, 0, 0), 1 / 32)
    _, want_b, _ = O.voxelize_surface(b, 32, (0, 0, 0), 1 / 32, 1e-4, owner=True)
    assert np.array_equal(own_b, want_b)
