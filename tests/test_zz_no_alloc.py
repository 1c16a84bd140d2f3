"""The allocation-free traversal entry point (vxl_traverse_dev_noalloc) against the one it
varies (vxl_traverse_dev), through the C ABI: two calls in a row, the second smaller, for
each lane schedule; hits identical and t bit-identical. Device free memory is unchanged by
the call on a fresh context."""
import ctypes

import numpy as np
import oracle as O
import pytest
import torch
import vxl
from vxl import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _free_bytes():
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime already loaded (by soname)
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _run(fn, tree, d_rays, R):
    hits = torch.full((R, 4), -7, dtype=torch.int32, device="cuda")
    t = torch.full((R,), -7.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    fn(tree, d_rays.data_ptr(), R, hits.data_ptr(), t.data_ptr())
    torch.cuda.synchronize()
    return hits.cpu().numpy(), t.cpu().numpy()


@pytest.mark.parametrize("refill", [8, 0, -1])
def test_noalloc_equals_traverse_dev(refill):
    ctx = vxl.Context(0)
    try:
        ctx.set_dda_refill(refill)
        tris = W.soup(20_000, seed=0x5EED_0600, mu=np.log(3 / 512))
        tree = ctx.voxelize_tree(tris, 512, (0, 0, 0), 1 / 512, 1e-4)
        rays = np.ascontiguousarray(np.concatenate([W.camera_rays(512, 160, 90),
                                                    W.random_rays(40_000, 512)]))
        d_rays = torch.from_numpy(rays).to("cuda")
        words, _, _ = O.voxelize_surface(tris, 512, (0, 0, 0), 1 / 512, 1e-4)
        want_mirror = O.Tree.build(words, 512).mirror(rays)
        for R in (len(rays), 1000):  # the second call is smaller: stale counter state shows
            _say(f"refill={refill} R={R}")
            h0, t0 = _run(ctx.traverse_dev, tree, d_rays, R)
            h1, t1 = _run(ctx.traverse_dev_noalloc, tree, d_rays, R)
            assert np.array_equal(h0, h1)
            assert np.array_equal(t0.view(np.uint64), t1.view(np.uint64))
            assert np.array_equal(h1.view(np.uint32), want_mirror[:R])
        tree.free()
    finally:
        ctx.close()


def test_noalloc_allocates_no_device_memory():
    warm = vxl.Context(0)
    warm.set_dda_refill(8)
    tris = W.uv_sphere(48, 24)
    tree = warm.voxelize_tree(tris, 128, (0, 0, 0), 1 / 128, 1e-4)
    rays = np.ascontiguousarray(W.random_rays(20_000, 128))
    d_rays = torch.from_numpy(rays).to("cuda")
    hits = torch.zeros((len(rays), 4), dtype=torch.int32, device="cuda")
    t = torch.zeros(len(rays), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _say("warm-up call")
    warm.traverse_dev_noalloc(tree, d_rays.data_ptr(), len(rays), hits.data_ptr(), t.data_ptr())
    torch.cuda.synchronize()
    fresh = vxl.Context(0)  # no scratch of its own yet
    fresh.set_dda_refill(8)
    torch.cuda.synchronize()
    before = _free_bytes()
    for _ in range(3):
        fresh.traverse_dev_noalloc(tree, d_rays.data_ptr(), len(rays), hits.data_ptr(),
                                   t.data_ptr())
    torch.cuda.synchronize()
    after = _free_bytes()
    _say(f"free device memory {before} -> {after}")
    assert after == before
    with pytest.raises(vxl.VxlError) as e:
        fresh.traverse_dev_noalloc(tree, d_rays.data_ptr(), (1 << 40) + 1, hits.data_ptr(), 0)
    assert e.value.name == "VXL_STORAGE_EXCEEDED"
    fresh.close()
    tree.free()
    warm.close()
