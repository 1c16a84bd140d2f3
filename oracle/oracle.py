"""ctypes wrapper over liboracle.so, the CPU restatement of the reference path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the CPU baseline. The
product library (friendly-journey_amd/libvxl.so) never imports or loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_f32p = C.POINTER(C.c_float)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.vo_morton_encode.restype = C.c_uint64
        L.vo_morton_encode.argtypes = [C.c_uint32] * 3
        L.vo_morton_decode.argtypes = [C.c_uint64, _u32p]
        L.vo_encode_brick.restype = C.c_uint32
        L.vo_encode_brick.argtypes = [C.c_uint32] * 3
        L.vo_child_bit.restype = C.c_uint32
        L.vo_child_bit.argtypes = [C.c_uint32] * 3
        L.vo_child_slot.restype = C.c_uint32
        L.vo_child_slot.argtypes = [C.c_uint32] * 4
        L.vo_resolution.restype = C.c_int
        L.vo_resolution.argtypes = [C.c_uint32, _u32p, _u32p, _u32p]
        L.vo_bounds_pack.restype = C.c_uint32
        L.vo_bounds_pack.argtypes = [_u32p, _u32p]
        L.vo_bounds_unpack.argtypes = [C.c_uint32, _u32p, _u32p]
        L.vo_leaf_occupied_bounds.restype = C.c_uint32
        L.vo_leaf_occupied_bounds.argtypes = [_u64p]
        L.vo_triangle_box_overlap.restype = C.c_int
        L.vo_triangle_box_overlap.argtypes = [_f32p] * 5
        L.vo_world_to_grid.argtypes = [_f32p, _f32p, C.c_float, _f32p]
        L.vo_voxelize_surface.restype = C.c_uint64
        L.vo_voxelize_surface.argtypes = [_f32p, C.c_uint64, C.c_uint32, _f32p, C.c_float,
                                          C.c_float, _u32p, _u32p]
        L.vo_world_to_grid_mat.argtypes = [_f32p, _f32p, _f32p]
        L.vo_voxelize_surface_mat.restype = C.c_uint64
        L.vo_voxelize_surface_mat.argtypes = [_f32p, C.c_uint64, C.c_uint32, _f32p, C.c_float,
                                              _u32p, _u32p]
        L.vo_voxelize_surface_mt.restype = C.c_uint64
        L.vo_voxelize_surface_mt.argtypes = [_f32p, C.c_uint64, C.c_uint32, _f32p, C.c_float,
                                             C.c_float, _u32p, _u32p, C.c_int]
        L.vo_hash_color.restype = C.c_uint32
        L.vo_hash_color.argtypes = [C.c_uint32]
        L.vo_owner_colors.restype = None
        L.vo_owner_colors.argtypes = [_u32p, C.c_uint64, _u32p]
        L.vo_resolve_material.restype = C.c_uint32
        L.vo_resolve_material.argtypes = [C.c_uint32, _u32p, C.c_uint32]
        L.vo_compact.restype = C.c_uint64
        L.vo_compact.argtypes = [_u32p, _u32p, C.c_uint32, _u32p, C.c_uint32, _i32p, _i32p]
        L.vo_compact_to_words.restype = C.c_uint64
        L.vo_compact_to_words.argtypes = [_i32p, C.c_uint64, C.c_uint32, _u32p]
        L.vo_compact_materials.restype = None
        L.vo_compact_materials.argtypes = [_i32p, C.c_uint64, C.c_uint32, _u64p, C.c_uint64,
                                           C.POINTER(C.c_uint16)]
        L.vo_camera_ray.restype = None
        L.vo_camera_ray.argtypes = [_f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _f32p]
        L.vo_render.restype = None
        L.vo_render.argtypes = [C.c_void_p, _f32p, C.c_uint32, C.c_uint32,
                                C.POINTER(C.c_uint16), _u32p, C.c_uint32, _u32p, C.c_int]
        _f64p = C.POINTER(C.c_double)
        L.vo_noise_grad.restype = C.c_double
        L.vo_noise_grad.argtypes = [C.c_uint32, C.c_double, C.c_double, C.c_double]
        L.vo_noise_fade.restype = C.c_double
        L.vo_noise_fade.argtypes = [C.c_double]
        L.vo_perlin3.restype = C.c_double
        L.vo_perlin3.argtypes = [C.c_uint32, C.c_double, C.c_double, C.c_double]
        L.vo_noise_value.restype = C.c_double
        L.vo_noise_value.argtypes = [_f64p, _u32p, C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.c_uint32]
        L.vo_noise_words.restype = C.c_uint64
        L.vo_noise_words.argtypes = [_f64p, _u32p, C.c_uint32, _u32p]
        L.vo_count_candidates.restype = C.c_uint64
        L.vo_count_candidates.argtypes = [_f32p, C.c_uint64, C.c_uint32, _f32p, C.c_float,
                                          C.c_float]
        L.vo_tree_build.restype = C.c_void_p
        L.vo_tree_build.argtypes = [_u32p, C.c_uint32, C.c_int]
        L.vo_tree_stage_seconds.argtypes = [C.c_void_p, _f64p]
        L.vo_tree_build_rows.restype = C.c_void_p
        L.vo_tree_build_rows.argtypes = [_u32p, C.c_uint32, C.c_int]
        L.vo_tree_from_school_b.restype = C.c_void_p
        L.vo_tree_from_school_b.argtypes = [C.c_uint32, _u32p, C.c_uint64, _u64p, C.c_uint64,
                                            _u32p]
        L.vo_tree_free.argtypes = [C.c_void_p]
        L.vo_tree_counts.argtypes = [C.c_void_p, _u64p]
        L.vo_tree_export.argtypes = [C.c_void_p, _u32p, _u64p, _u32p, _u64p]
        L.vo_tree_export_level.argtypes = [C.c_void_p, C.c_uint32, _u32p]
        L.vo_tree_is_occupied.restype = C.c_int
        L.vo_tree_is_occupied.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]
        L.vo_mirror_traverse.argtypes = [C.c_void_p, _f32p, C.c_uint64, _u32p, C.c_int, _u32p,
                                         C.c_int]
        L.vo_layout_traverse.argtypes = [C.c_void_p, _f64p, C.c_uint64, _u32p, _f64p, C.c_int,
                                         C.c_int]
        L.vo_first_hit_dense.argtypes = [_u32p, C.c_uint32, _f64p, C.c_uint64, _u32p, _f64p]
        L.vo_fixture_words.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                       _u32p]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


# ---- small integer helpers -------------------------------------------------
def morton_encode(x, y, z) -> int:
    return int(lib().vo_morton_encode(x, y, z))


def morton_decode(code):
    out = np.zeros(3, np.uint32)
    lib().vo_morton_decode(code, _p(out, _u32p))
    return tuple(int(v) for v in out)


def encode_brick(x, y, z) -> int:
    return int(lib().vo_encode_brick(x, y, z))


def child_bit(cx, cy, cz) -> int:
    return int(lib().vo_child_bit(cx, cy, cz))


def child_slot(mask: int, base: int, bit: int) -> int:
    return int(lib().vo_child_slot(mask & 0xFFFFFFFF, mask >> 32, base, bit))


def resolution(n: int):
    """Returns ('ok', k) / ('too_small',) / ('too_large',) / ('not_representable', below, above)."""
    k, b, a = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
    rc = lib().vo_resolution(n, C.byref(k), C.byref(b), C.byref(a))
    if rc == 0:
        return ("ok", k.value)
    if rc == 1:
        return ("too_small",)
    if rc == 2:
        return ("too_large",)
    return ("not_representable", b.value, a.value)


def bounds_pack(mn, mx) -> int:
    a = np.asarray(mn, np.uint32)
    b = np.asarray(mx, np.uint32)
    return int(lib().vo_bounds_pack(_p(a, _u32p), _p(b, _u32p)))


def bounds_unpack(p: int):
    a = np.zeros(3, np.uint32)
    b = np.zeros(3, np.uint32)
    lib().vo_bounds_unpack(p, _p(a, _u32p), _p(b, _u32p))
    return list(map(int, a)), list(map(int, b))


def leaf_bounds(words8) -> int:
    w = np.ascontiguousarray(words8, np.uint64)
    return int(lib().vo_leaf_occupied_bounds(_p(w, _u64p)))


# ---- SAT / voxelize --------------------------------------------------------
def triangle_box_overlap(center, half, v0, v1, v2) -> bool:
    arrs = [np.ascontiguousarray(x, np.float32) for x in (center, half, v0, v1, v2)]
    return bool(lib().vo_triangle_box_overlap(*[_p(a, _f32p) for a in arrs]))


def world_to_grid(p, origin, voxel_size):
    pa = np.ascontiguousarray(p, np.float32)
    oa = np.ascontiguousarray(origin, np.float32)
    out = np.zeros(3, np.float32)
    lib().vo_world_to_grid(_p(pa, _f32p), _p(oa, _f32p), voxel_size, _p(out, _f32p))
    return out


def voxelize_surface(tris, n, origin=(0.0, 0.0, 0.0), voxel_size=1.0, eps=1e-4, owner=False):
    """voxelize_surface_cpu: returns (x-major u32 words, owner or None, candidate pairs)."""
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    o = np.ascontiguousarray(origin, np.float32)
    nv = n ** 3
    words = np.zeros((nv + 31) // 32, np.uint32)
    own = np.zeros(nv, np.uint32) if owner else None
    pairs = lib().vo_voxelize_surface(_p(t, _f32p), t.shape[0], n, _p(o, _f32p), voxel_size, eps,
                                      _p(words, _u32p),
                                      _p(own, _u32p) if owner else None)
    return words, own, int(pairs)


def world_to_grid_mat(p, mat):
    """Mat4::transform_point3 of p by the column-major 16-float matrix mat."""
    pa = np.ascontiguousarray(p, np.float32)
    m = np.ascontiguousarray(mat, np.float32).reshape(16)
    out = np.zeros(3, np.float32)
    lib().vo_world_to_grid_mat(_p(pa, _f32p), _p(m, _f32p), _p(out, _f32p))
    return out


def voxelize_surface_mat(tris, n, mat, eps=1e-4, owner=False):
    """voxelize_surface_cpu on a grid with the world_to_grid override mat (column-major)."""
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    m = np.ascontiguousarray(mat, np.float32).reshape(16)
    nv = n ** 3
    words = np.zeros((nv + 31) // 32, np.uint32)
    own = np.zeros(nv, np.uint32) if owner else None
    pairs = lib().vo_voxelize_surface_mat(_p(t, _f32p), t.shape[0], n, _p(m, _f32p), eps,
                                          _p(words, _u32p), _p(own, _u32p) if owner else None)
    return words, own, int(pairs)


def voxelize_surface_mt(tris, n, origin=(0.0, 0.0, 0.0), voxel_size=1.0, eps=1e-4,
                        owner=False, threads=None):
    """voxelize_surface with the scan split over threads by z (same words, owners and pair
    count as the serial restatement; for checks at the large grids)."""
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    o = np.ascontiguousarray(origin, np.float32)
    nv = n ** 3
    words = np.zeros((nv + 31) // 32, np.uint32)
    own = np.zeros(nv, np.uint32) if owner else None
    pairs = lib().vo_voxelize_surface_mt(_p(t, _f32p), t.shape[0], n, _p(o, _f32p), voxel_size,
                                         eps, _p(words, _u32p),
                                         _p(own, _u32p) if owner else None,
                                         threads or _threads())
    return words, own, int(pairs)


def hash_color(i: int) -> int:
    return int(lib().vo_hash_color(i))


def owner_colors(owner: np.ndarray) -> np.ndarray:
    """Colour per voxel from its owner (reference_cpu.rs:173-186), 0 where unowned."""
    own = np.ascontiguousarray(owner, np.uint32)
    col = np.zeros_like(own)
    lib().vo_owner_colors(_p(own, _u32p), own.size, _p(col, _u32p))
    return col


def resolve_material(tri: int, table) -> int:
    t = np.ascontiguousarray(table, np.uint32)
    return int(lib().vo_resolve_material(tri, _p(t, _u32p), t.size))


def compact(words, owner, n, table, g_origin=(0, 0, 0)) -> np.ndarray:
    """compact_surface_sparse from dense outputs: (count, 4) i32 rows (vx, vy, vz, material),
    x fastest."""
    w = np.ascontiguousarray(words, np.uint32)
    o = np.ascontiguousarray(owner, np.uint32)
    t = np.ascontiguousarray(table, np.uint32)
    g = np.ascontiguousarray(g_origin, np.int32)
    c = lib().vo_compact(_p(w, _u32p), _p(o, _u32p), n, _p(t, _u32p), t.size, _p(g, _i32p), None)
    out = np.zeros((c, 4), np.int32)
    lib().vo_compact(_p(w, _u32p), _p(o, _u32p), n, _p(t, _u32p), t.size, _p(g, _i32p),
                     _p(out, _i32p))
    return out


def tree_from_compact(vox, n, threads=1):
    """tree_from_compact: (Tree over the in-range voxels, materials (L, 512) u16, dropped)."""
    v = np.ascontiguousarray(vox, np.int32).reshape(-1, 4)
    words = np.zeros((n ** 3 + 31) // 32, np.uint32)
    dropped = lib().vo_compact_to_words(_p(v, _i32p), v.shape[0], n, _p(words, _u32p))
    tree = Tree.build(words, n, threads=threads)
    _, _, _, codes = tree.export()
    codes = np.ascontiguousarray(codes, np.uint64)
    mats = np.zeros((len(codes), 512), np.uint16)
    lib().vo_compact_materials(_p(v, _i32p), v.shape[0], n, _p(codes, _u64p), len(codes),
                               mats.ctypes.data_as(C.POINTER(C.c_uint16)))
    return tree, mats, int(dropped)


def camera(eye, forward, right, up, tan, aspect, n) -> np.ndarray:
    """render.wgsl's Camera as 16 f32: eye, tan, forward, aspect, right, n, up, pad."""
    c = np.zeros(16, np.float32)
    c[0:3], c[3], c[4:7], c[7], c[8:11], c[11], c[12:15] = eye, tan, forward, aspect, right, n, up
    return c


def camera_rays(cam, width, height) -> np.ndarray:
    c = np.ascontiguousarray(cam, np.float32)
    out = np.zeros((height * width, 6), np.float32)
    for y in range(height):
        for x in range(width):
            lib().vo_camera_ray(_p(c, _f32p), width, height, x, y,
                                _p(out[y * width + x], _f32p))
    return out


def render(tree, cam, width, height, mats=None, colors=None) -> np.ndarray:
    """render_main per pixel: (height, width) u32 RGBA8 (R in the low byte)."""
    c = np.ascontiguousarray(cam, np.float32)
    out = np.zeros(height * width, np.uint32)
    m = None if mats is None else np.ascontiguousarray(mats, np.uint16)
    col = np.ascontiguousarray(colors if colors is not None else [], np.uint32)
    lib().vo_render(tree._h, _p(c, _f32p), width, height,
                    None if m is None else m.ctypes.data_as(C.POINTER(C.c_uint16)),
                    _p(col, _u32p), col.size, _p(out, _u32p), _threads())
    return out.reshape(height, width)


# NoiseField presets (voxel-core/src/fixtures.rs:284-314)
NOISE_PERLIN = dict(seed=0x5EED1234, frequency=5.0, octaves=5, lacunarity=2.0, gain=0.5,
                    ridged=False, warp=0.0, threshold=0.12)
NOISE_CAVES = dict(seed=0xCA5E5EED, frequency=4.0, octaves=5, lacunarity=2.0, gain=0.5,
                   ridged=True, warp=0.45, threshold=0.48)


def _noise_args(field):
    f = np.array([field["frequency"], field["lacunarity"], field["gain"], field["warp"],
                  field["threshold"]], np.float64)
    u = np.array([field["seed"], field["octaves"], 1 if field["ridged"] else 0], np.uint32)
    return f, u


def noise_grad(h, x, y, z) -> float:
    return float(lib().vo_noise_grad(h, x, y, z))


def noise_fade(t) -> float:
    return float(lib().vo_noise_fade(t))


def perlin3(seed, x, y, z) -> float:
    return float(lib().vo_perlin3(seed, x, y, z))


def noise_value(field, n, x, y, z) -> float:
    f, u = _noise_args(field)
    return float(lib().vo_noise_value(f.ctypes.data_as(C.POINTER(C.c_double)), _p(u, _u32p), n,
                                      x, y, z))


def noise_words(field, n):
    """NoiseField::is_occupied as x-major words, and the occupied count."""
    f, u = _noise_args(field)
    words = np.zeros((n ** 3 + 31) // 32, np.uint32)
    c = lib().vo_noise_words(f.ctypes.data_as(C.POINTER(C.c_double)), _p(u, _u32p), n,
                             _p(words, _u32p))
    return words, int(c)


def count_candidates(tris, n, origin=(0.0, 0.0, 0.0), voxel_size=1.0, eps=1e-4) -> int:
    t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
    o = np.ascontiguousarray(origin, np.float32)
    return int(lib().vo_count_candidates(_p(t, _f32p), t.shape[0], n, _p(o, _f32p), voxel_size,
                                         eps))


def fixture_words(kind: str, n: int, *params) -> np.ndarray:
    kinds = {"empty": 0, "solid": 1, "single": 2, "checkerboard": 3, "octant": 4, "wire": 5,
             "dust": 6}
    p = list(params) + [0, 0, 0]
    words = np.zeros((n ** 3 + 31) // 32, np.uint32)
    lib().vo_fixture_words(kinds[kind], n, p[0], p[1], p[2], _p(words, _u32p))
    return words


def sierpinski(n):
    return fixture_words("octant", n, 0b0110_1001)


def cantor(n):
    return fixture_words("octant", n, 0b1000_0001)


def words_get(words, n, x, y, z) -> bool:
    i = x + n * (y + n * z)
    return bool((int(words[i >> 5]) >> (i & 31)) & 1)


# ---- sparse tree / School-B -------------------------------------------------
class Tree:
    """SparseTree (School A) + SchoolBBuffer (School B) restated on the CPU."""

    def __init__(self, handle, n):
        self._h = handle
        self.n = n
        c = np.zeros(18, np.uint64)
        lib().vo_tree_counts(self._h, _p(c, _u64p))
        self.leaf_count = int(c[0])
        self.node_count = int(c[1])
        self.level_counts = {L: int(c[2 + L]) for L in range(16) if c[2 + L]}

    @classmethod
    def build(cls, words, n, threads=None, rows=False):
        """SparseTree::build + from_sparse; rows=True reads each 8-voxel brick row as one
        byte of the words (the same tree, faster at 2048^3: for checks, not the baseline)."""
        w = np.ascontiguousarray(words, np.uint32)
        fn = lib().vo_tree_build_rows if rows else lib().vo_tree_build
        h = fn(_p(w, _u32p), n, threads or _threads())
        if not h:
            raise ValueError(f"resolution {n} is not 8*4^k")
        return cls(h, n)

    @classmethod
    def from_school_b(cls, n, nodes, leaves, bounds):
        nd = np.ascontiguousarray(nodes, np.uint32).reshape(-1, 3)
        lv = np.ascontiguousarray(leaves, np.uint64).reshape(-1, 8)
        bd = np.ascontiguousarray(bounds, np.uint32)
        h = lib().vo_tree_from_school_b(n, _p(nd, _u32p), nd.shape[0], _p(lv, _u64p),
                                        lv.shape[0], _p(bd, _u32p))
        return cls(h, n)

    def stage_seconds(self):
        """(SparseTree::build s, from_sparse s) of the build that made this tree."""
        out = np.zeros(2, np.float64)
        lib().vo_tree_stage_seconds(self._h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return float(out[0]), float(out[1])

    def export(self):
        nodes = np.zeros((max(self.node_count, 0), 3), np.uint32)
        leaves = np.zeros((self.leaf_count, 8), np.uint64)
        bounds = np.zeros(self.leaf_count, np.uint32)
        codes = np.zeros(self.leaf_count, np.uint64)
        lib().vo_tree_export(self._h, _p(nodes, _u32p), _p(leaves, _u64p), _p(bounds, _u32p),
                             _p(codes, _u64p))
        return nodes, leaves, bounds, codes

    def level(self, L):
        out = np.zeros((self.level_counts.get(L, 0), 3), np.uint32)
        if out.shape[0]:
            lib().vo_tree_export_level(self._h, L, _p(out, _u32p))
        return out

    def is_occupied(self, x, y, z) -> bool:
        return bool(lib().vo_tree_is_occupied(self._h, x, y, z))

    def mirror(self, rays_f32, skip=True, stats=False, threads=None):
        r = np.ascontiguousarray(rays_f32, np.float32).reshape(-1, 6)
        hits = np.zeros((r.shape[0], 4), np.uint32)
        st = np.zeros((r.shape[0], 3), np.uint32) if stats else None
        lib().vo_mirror_traverse(self._h, _p(r, _f32p), r.shape[0], _p(hits, _u32p),
                                 1 if skip else 0, _p(st, _u32p) if stats else None,
                                 threads or _threads())
        return (hits, st) if stats else hits

    def layout(self, rays_f64, school=1, threads=None):
        r = np.ascontiguousarray(rays_f64, np.float64).reshape(-1, 6)
        hits = np.zeros((r.shape[0], 4), np.uint32)
        t = np.zeros(r.shape[0], np.float64)
        lib().vo_layout_traverse(self._h, _p(r, _f64p), r.shape[0], _p(hits, _u32p),
                                 _p(t, _f64p), school, threads or _threads())
        return hits, t

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.vo_tree_free(self._h)
            self._h = None


def first_hit_dense(words, n, rays_f64):
    w = np.ascontiguousarray(words, np.uint32)
    r = np.ascontiguousarray(rays_f64, np.float64).reshape(-1, 6)
    hits = np.zeros((r.shape[0], 4), np.uint32)
    t = np.zeros(r.shape[0], np.float64)
    lib().vo_first_hit_dense(_p(w, _u32p), n, _p(r, _f64p), r.shape[0], _p(hits, _u32p),
                             _p(t, _f64p))
    return hits, t
