"""Seeded synthetic inputs for the voxelize and traversal workloads.

Every generator draws from splitmix64 (Steele, Lea & Flood 2014), the PRNG the
reference's own tests use (voxel-core/src/mirror.rs:343-354), with doubles drawn
as (x >> 11) * 2^-53 (mirror.rs:351-354). Meshes are produced in world space
inside the unit cube, as f32 triangle soups (T x 9, the memory layout of
``MeshInput.triangles: Vec<[Vec3; 3]>``, voxelizer/src/core.rs:321-333).

Resolution mapping: the reference grid only admits n = 8*4^k
(voxel-core/src/resolution.rs:76-118). A requested m that is not of that form is
voxelized with voxel size 1/m and origin 0 inside the next representable cube
n = 2m (64->128, 256->512, 1024->2048); the occupied region is [0, m)^3.
"""
from __future__ import annotations

import math

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


class SplitMix64:
    """Vectorised splitmix64 stream: value i is mix(seed + (i+1) * golden)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            idx = np.arange(self.count + 1, self.count + 1 + n, dtype=np.uint64)
            z = self.state + idx * GOLDEN
        self.count += n
        return _mix(z)

    def unit(self, n: int) -> np.ndarray:
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))

    def normal(self, n: int) -> np.ndarray:
        u1 = self.unit(n)
        u2 = self.unit(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)

    def sphere_dirs(self, n: int) -> np.ndarray:
        v = np.stack([self.normal(n), self.normal(n), self.normal(n)], axis=1)
        norm = np.linalg.norm(v, axis=1, keepdims=True)
        norm[norm == 0] = 1.0
        return v / norm


def grid_for(m: int):
    """(n, voxel_size) for a requested resolution m: n = m when 8*4^k, else 2m."""
    n = m
    while True:
        q = n // 8
        if n % 8 == 0 and q & (q - 1) == 0 and (q.bit_length() - 1) % 2 == 0:
            break
        n *= 2
    return n, 1.0 / m


def uv_sphere(segments=128, rings=64, radius=0.4, center=(0.5, 0.5, 0.5), noise=0.02,
              seed=0x5EED_0001) -> np.ndarray:
    """UV sphere with pole fans: 2*segments + (rings-2)*segments*2 triangles.

    Each vertex radius is scaled by (1 + U[-noise, noise]); shared vertices are
    perturbed once so the surface stays watertight.
    """
    rng = SplitMix64(seed)
    c = np.asarray(center, np.float64)
    theta = np.pi * np.arange(1, rings) / rings
    phi = 2.0 * np.pi * np.arange(segments) / segments
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    ring = np.stack([st * np.cos(phi)[None, :], st * np.sin(phi)[None, :],
                     np.broadcast_to(ct, (rings - 1, segments))], axis=-1).reshape(-1, 3)
    dirs = np.concatenate([[[0.0, 0.0, 1.0]], ring, [[0.0, 0.0, -1.0]]])
    scale = 1.0 + noise * (2.0 * rng.unit(dirs.shape[0]) - 1.0)
    verts = c + radius * scale[:, None] * dirs
    north, south = 0, dirs.shape[0] - 1

    def vid(j, i):
        return 1 + j * segments + (i % segments)

    i = np.arange(segments)
    tris = [np.stack([np.full(segments, north), vid(0, i), vid(0, i + 1)], 1)]
    for j in range(rings - 2):
        a, b, c2, d = vid(j, i), vid(j, i + 1), vid(j + 1, i), vid(j + 1, i + 1)
        tris.append(np.stack([a, c2, b], 1))
        tris.append(np.stack([b, c2, d], 1))
    tris.append(np.stack([np.full(segments, south), vid(rings - 2, i + 1), vid(rings - 2, i)], 1))
    idx = np.concatenate(tris)
    return verts[idx].reshape(-1, 9).astype(np.float32)


def sphere_with_tris(target: int, seed: int) -> np.ndarray:
    """A UV sphere whose triangle count is close to ``target`` (rings = segments / 2)."""
    seg = max(4, int(round(math.sqrt(max(target, 8)))))
    seg += seg % 2
    return uv_sphere(segments=seg, rings=max(3, seg // 2), seed=seed)


def value_noise(x: np.ndarray, y: np.ndarray, octaves=6, base_freq=4, seed=0x5EED_0002):
    """Octave value noise in [0, 1]: splitmix64 lattice values, smoothstep blend,
    frequency x2 and amplitude x1/2 per octave, normalised by the amplitude sum."""
    total = np.zeros_like(x)
    amp, norm = 1.0, 0.0
    for o in range(octaves):
        f = base_freq << o
        fx, fy = x * f, y * f
        ix, iy = np.floor(fx).astype(np.int64), np.floor(fy).astype(np.int64)
        tx, ty = fx - ix, fy - iy
        sx, sy = tx * tx * (3 - 2 * tx), ty * ty * (3 - 2 * ty)

        def lattice(a, b):
            key = (np.uint64(seed) ^ (np.uint64(o) << np.uint64(48))
                   ^ (a.astype(np.uint64) << np.uint64(24)) ^ b.astype(np.uint64))
            with np.errstate(over="ignore"):
                z = _mix(key + GOLDEN)
            return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))

        v00, v10 = lattice(ix, iy), lattice(ix + 1, iy)
        v01, v11 = lattice(ix, iy + 1), lattice(ix + 1, iy + 1)
        top = v00 + (v10 - v00) * sx
        bot = v01 + (v11 - v01) * sx
        total += amp * (top + (bot - top) * sy)
        norm += amp
        amp *= 0.5
    return total / norm


def heightfield(cells=1024, max_height=0.5, seed=0x5EED_0002) -> np.ndarray:
    """cells x cells heightfield over [0,1]^2, 2 triangles per cell (2*cells^2)."""
    g = np.arange(cells + 1, dtype=np.float64) / cells
    X, Y = np.meshgrid(g, g, indexing="xy")
    Z = max_height * value_noise(X, Y, seed=seed)
    V = np.stack([X, Y, Z], -1)
    v00, v10 = V[:-1, :-1], V[:-1, 1:]
    v01, v11 = V[1:, :-1], V[1:, 1:]
    t1 = np.stack([v00, v10, v11], 2).reshape(-1, 9)
    t2 = np.stack([v00, v11, v01], 2).reshape(-1, 9)
    out = np.empty((t1.shape[0] * 2, 9), np.float32)
    out[0::2] = t1
    out[1::2] = t2
    return out


def soup(count: int, seed=0x5EED_0003, mu=math.log(2.0 / 1024.0), sigma=0.5) -> np.ndarray:
    """Random triangle soup: centres U[0,1]^3, each vertex offset by a log-normal
    length (mu, sigma) along a uniform direction on S^2."""
    rng = SplitMix64(seed)
    centre = np.stack([rng.unit(count), rng.unit(count), rng.unit(count)], 1)
    out = np.empty((count, 9), np.float64)
    for v in range(3):
        length = np.exp(mu + sigma * rng.normal(count))
        out[:, 3 * v:3 * v + 3] = centre + length[:, None] * rng.sphere_dirs(count)
    return out.astype(np.float32)


def terrain_1024(seed=0x5EED_0002) -> np.ndarray:
    """The 1024^3 voxelize workload: 1024^2-cell heightfield (2,097,152 triangles)
    plus a 100,000-triangle soup (2,197,152 triangles)."""
    return np.concatenate([heightfield(1024, seed=seed), soup(100_000, seed=seed + 1)])


def camera_basis(fov_deg=60.0, radius=1.5, seed=0x5EED_0004):
    """The bench camera in unit-cube space: eye on a seeded point of the sphere of radius
    1.5 about the cube centre, looking at it; (eye, forward, right, up, tan(fov/2))."""
    rng = SplitMix64(seed)
    centre = np.array([0.5, 0.5, 0.5])
    eye = centre + radius * rng.sphere_dirs(1)[0]
    fwd = centre - eye
    fwd /= np.linalg.norm(fwd)
    up_w = np.array([0.0, 0.0, 1.0]) if abs(fwd[2]) < 0.99 else np.array([0.0, 1.0, 0.0])
    right = np.cross(fwd, up_w)
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    return eye, fwd, right, up, math.tan(math.radians(fov_deg) / 2)


def render_camera(m: int, n: int, width=1920, height=1080, fov_deg=60.0, radius=1.5,
                  seed=0x5EED_0004) -> np.ndarray:
    """The bench camera as the render path's 16 f32 (eye in voxel units, tan, forward,
    aspect, right, n, up, pad); the same eye and basis as camera_rays."""
    eye, fwd, right, up, tan = camera_basis(fov_deg, radius, seed)
    c = np.zeros(16, np.float32)
    c[0:3], c[3], c[4:7], c[7] = eye * m, tan, fwd, width / height
    c[8:11], c[11], c[12:15] = right, n, up
    return c


def camera_rays(m: int, width=1920, height=1080, fov_deg=60.0, radius=1.5,
                seed=0x5EED_0004) -> np.ndarray:
    """Pinhole primary rays in voxel space: eye on a seeded point of the sphere of
    radius 1.5 about the cube centre, looking at it; pixel directions follow the
    reference's camera (voxel-gpu/shaders/render.wgsl:66-70)."""
    eye, fwd, right, up, tan = camera_basis(fov_deg, radius, seed)
    aspect = width / height
    px = ((np.arange(width) + 0.5) / width * 2.0 - 1.0) * tan * aspect
    py = (1.0 - (np.arange(height) + 0.5) / height * 2.0) * tan
    PX, PY = np.meshgrid(px, py, indexing="xy")
    d = fwd[None, None, :] + right[None, None, :] * PX[..., None] + up[None, None, :] * PY[..., None]
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    rays = np.empty((height * width, 6), np.float32)
    rays[:, :3] = (eye * m).astype(np.float32)
    rays[:, 3:] = d.reshape(-1, 3).astype(np.float32)
    return rays


def random_rays(count: int, m: int, seed=0x5EED_0005) -> np.ndarray:
    """Rays with origins uniform in [0,m)^3 and directions uniform on S^2."""
    rng = SplitMix64(seed)
    o = np.stack([rng.unit(count), rng.unit(count), rng.unit(count)], 1) * m
    d = rng.sphere_dirs(count)
    return np.concatenate([o, d], 1).astype(np.float32)


def box_rays(count: int, n: int, seed=0x1357_9BDF_2468_ACE0) -> np.ndarray:
    """The reference differential's ray battery (voxel-gpu/tests/differential.rs:207-224):
    origins U[-4, n+4)^3, directions U[-1,1)^3 with |d| >= 1e-3, as f64."""
    rng = SplitMix64(seed)
    out = []
    while sum(len(o) for o in out) < count:
        u = rng.unit(6 * count).reshape(count, 6)
        o = u[:, :3] * (n + 8.0) - 4.0
        d = u[:, 3:] * 2.0 - 1.0
        keep = np.linalg.norm(d, axis=1) >= 1e-3
        out.append(np.concatenate([o, d], 1)[keep])
    return np.concatenate(out)[:count]


def _log_uniform_int(rng: SplitMix64, lo: int, hi: int) -> int:
    u = float(rng.unit(1)[0])
    return int(round(math.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))))


def batch_plan(count=512, seed=0x5EED_0007, tris=(100, 200_000), rays=(1_000, 1_000_000),
               resolutions=(64, 128, 256, 512, 1024)):
    """The many-mesh workload: per mesh a triangle count log-uniform in ``tris``, a seeded
    coin between a UV sphere of about that many triangles and a soup (edge lengths
    log-normal about 2 voxels of the mesh's resolution), a resolution m uniform over
    ``resolutions`` (voxel size 1/m, in Resolution grid_for(m)), and a ray batch of size
    log-uniform in ``rays`` with origins U[0,m)^3 and directions uniform on S^2.
    Returns one dict per mesh (kind, T, m, n, voxel_size, R, seeds), no arrays."""
    rng = SplitMix64(seed)
    plan = []
    for i in range(count):
        T = _log_uniform_int(rng, *tris)
        sphere = bool(int(rng.u64(1)[0]) & 1)
        m = resolutions[int(rng.u64(1)[0] % np.uint64(len(resolutions)))]
        R = _log_uniform_int(rng, *rays)
        n, vs = grid_for(m)
        plan.append(dict(kind="sphere" if sphere else "soup", T=T, m=m, n=n, voxel_size=vs, R=R,
                         mesh_seed=seed + 1000 + i, ray_seed=seed + 100_000 + i))
    return plan


def batch_mesh(entry) -> np.ndarray:
    if entry["kind"] == "sphere":
        return sphere_with_tris(entry["T"], seed=entry["mesh_seed"])
    return soup(entry["T"], seed=entry["mesh_seed"], mu=math.log(2.0 / entry["m"]))


def batch_rays(entry) -> np.ndarray:
    return random_rays(entry["R"], entry["m"], seed=entry["ray_seed"])


def serve_plan(duration_s=2.0, rate=2000.0, ray_share=0.7, trees=64, seed=0x5EED_0008,
               ray_range=(1_000, 500_000), tree_resolutions=(256, 512, 1024), mesh_pool=64):
    """The serving workload: Poisson arrivals at ``rate`` per second for ``duration_s``;
    each request is a ray query (probability ``ray_share``: log-uniform ray count in
    ``ray_range`` against one of ``trees`` resident trees at resolutions
    ``tree_resolutions``) or a voxelize request (a mesh of the many-mesh distribution,
    drawn from a pool of ``mesh_pool``). Returns (tree_plan, mesh_plan, requests) where a
    request is (t_arrival_s, "rays", tree_index, R, ray_offset) or
    (t_arrival_s, "vox", mesh_index)."""
    rng = SplitMix64(seed)
    tree_plan = [e for e in batch_plan(trees * 4, seed=seed + 1, resolutions=tree_resolutions)
                 ][:trees]
    mesh_plan = batch_plan(mesh_pool, seed=seed + 2)
    reqs = []
    t = 0.0
    while True:
        t += -math.log1p(-float(rng.unit(1)[0])) / rate
        if t >= duration_s:
            break
        if float(rng.unit(1)[0]) < ray_share:
            ti = int(rng.u64(1)[0] % np.uint64(trees))
            R = _log_uniform_int(rng, *ray_range)
            off = int(float(rng.unit(1)[0]) * (ray_range[1] * 2 - R))
            reqs.append((t, "rays", ti, R, off))
        else:
            reqs.append((t, "vox", int(rng.u64(1)[0] % np.uint64(mesh_pool))))
    return tree_plan, mesh_plan, reqs
