"""Oracle render (render.wgsl:54-91): camera rays and shading checked against what the
shader's formulas fix -- unit directions, the centre of the image along forward, the sky
gradient's rgba8 values, position shading of a solid grid."""
import math
import numpy as np
import oracle as O


def _cam(n, eye, target, fov=60.0, aspect=16 / 9):
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, [0.0, 0.0, 1.0] if abs(f[2]) < 0.99 else [0.0, 1.0, 0.0])
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    return O.camera(eye, f, r, u, math.tan(math.radians(fov / 2)), aspect, n)


def test_camera_rays_are_unit_and_centred():
    cam = _cam(32, (-20.0, 5.0, 40.0), (16, 16, 16))
    w, h = 33, 17  # odd sizes: the middle pixel's ndc is exactly 0
    rays = O.camera_rays(cam, w, h)
    d = rays[:, 3:].astype(np.float64)
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 3e-7
    mid = rays[(h // 2) * w + w // 2, 3:]
    assert np.allclose(mid, cam[4:7], atol=1e-7)
    assert np.all(rays[:, :3] == cam[0:3])
    # left-right mirror pixels differ only by the sign of their right component
    a, b = rays[3 * w + 2, 3:], rays[3 * w + (w - 1 - 2), 3:]
    ra, rb = float(np.dot(a, cam[8:11])), float(np.dot(b, cam[8:11]))
    assert ra < 0 < rb and abs(ra + rb) < 1e-6


def test_sky_gradient_and_position_shading():
    empty = O.Tree.build(O.fixture_words("empty", 32), 32)
    cam = _cam(32, (-20.0, 5.0, 40.0), (16, 16, 16))
    img = O.render(empty, cam, 16, 9)
    # row 0: t = 0 -> rgba8(0.08, 0.10, 0.16, 1) = (20, 26, 41, 255)
    assert np.all(img[0] == (20 | (26 << 8) | (41 << 16) | (255 << 24)))
    # row 8 of 9: t = 8/9
    t = np.float32(8) / np.float32(9)
    want = [int(np.rint(np.float32(v) * np.float32(255))) for v in
            (np.float32(0.08) * (np.float32(1) - t), np.float32(0.10) * (np.float32(1) - t),
             np.float32(0.16) + np.float32(0.12) * t)]
    assert np.all(img[8] == (want[0] | (want[1] << 8) | (want[2] << 16) | (255 << 24)))
    solid = O.Tree.build(O.fixture_words("solid", 32), 32)
    cam = _cam(32, (16.5, 16.5, 80.0), (16.5, 16.5, 0.0), fov=10.0, aspect=1.0)
    img = O.render(solid, cam, 5, 5)
    # looking straight down -z at the top face: the hit voxels are (x, y, 31) with x, y
    # near 16, shaded by position
    px = img[2, 2]
    assert (px >> 16) & 255 == int(np.rint(np.float32(31) / np.float32(32) * 255))
    assert abs(int(px & 255) - 16 * 255 // 32) <= 8
