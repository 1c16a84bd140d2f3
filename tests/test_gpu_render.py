"""GPU parity of the render path (render.wgsl:54-91): device-generated camera rays, the
walk and the shading give the oracle's image byte for byte."""
import math
import numpy as np
import oracle as O
import pytest
from vxl import workloads as W

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)


def _cam(n, eye, target, fov=60.0, aspect=16 / 9):
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    up = [0.0, 0.0, 1.0] if abs(f[2]) < 0.99 else [0.0, 1.0, 0.0]
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    return O.camera(eye, f, r, u, math.tan(math.radians(fov / 2)), aspect, n)


def _check(ctx, gt, ot, cam, w, h, mats=None, colors=(), label=""):
    _say(f"render {label} {w}x{h}")
    got = ctx.render(gt, cam, w, h, colors)
    want = O.render(ot, cam, w, h, mats, np.asarray(colors, np.uint32) if len(colors) else None)
    bad = got != want
    assert not bad.any(), f"{label}: {int(bad.sum())} pixels differ, first {np.argwhere(bad)[:3]}"
    return got


def test_render_sphere_with_and_without_materials(ctx):
    tris = W.uv_sphere()
    n, vs = W.grid_for(256)
    words, _, _ = O.voxelize_surface(tris, n, (0, 0, 0), vs, 1e-4)
    ot = O.Tree.build(words, n)
    gt = ctx.voxelize_tree(tris, n, (0, 0, 0), vs, 1e-4)
    c = (128.0, 128.0, 128.0)
    for eye, fov in (((-300.0, 100.0, 500.0), 60.0), ((600.0, 600.0, 20.0), 30.0),
                     ((130.0, 120.0, 128.0), 90.0)):  # the last one inside the sphere
        img = _check(ctx, gt, ot, _cam(n, eye, c, fov), 157, 93, label=f"sphere eye={eye}")
    assert len(np.unique(img)) > 100
    rng = np.random.default_rng(3)
    table = rng.integers(0, 2 ** 32, (len(tris) + 1) // 2, dtype=np.uint64).astype(np.uint32)
    table &= 0x001F001F  # global ids 0..31 (0 = position shading)
    colors = rng.integers(0, 2 ** 32, 24, dtype=np.uint64).astype(np.uint32)  # 24..31: past the table
    gm = ctx.voxelize_tree_materials(tris, n, (0, 0, 0), vs, 1e-4, table)
    _check(ctx, gm, ot, _cam(n, (-300.0, 100.0, 500.0), c), 160, 90, gm.materials(), colors,
           "sphere materials")


def test_render_fixtures_and_sizes(ctx):
    for kind, n, params in (("dust", 128, (4096,)), ("checkerboard", 32, ()), ("wire", 128, (16, 1)),
                            ("empty", 32, ()), ("single", 8, (5, 2, 2))):
        words = O.fixture_words(kind, n, *params)
        ot = O.Tree.build(words, n)
        gt = ctx.tree_from_words(words, n)
        cam = _cam(n, (-0.7 * n, 0.3 * n, 1.4 * n), (n / 2, n / 2, n / 2))
        for w, h in ((64, 64), (13, 7), (1, 1)):
            _check(ctx, gt, ot, cam, w, h, label=f"{kind}{n}")


def test_render_argument_checks(ctx):
    import vxl
    words = O.fixture_words("dust", 32, 64)
    gt = ctx.tree_from_words(words, 32)
    with pytest.raises(vxl.VxlError):
        ctx.render(gt, _cam(64, (0, 0, 100), (16, 16, 16)), 8, 8)  # n differs from the tree
    assert ctx.render(gt, _cam(32, (0, 0, 100), (16, 16, 16)), 0, 0).shape == (0, 0)
