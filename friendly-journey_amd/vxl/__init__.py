"""Python binding of libvxl.so, the MI355X voxelize + sparse-tree + DDA path.

This is a thin ctypes layer over the C ABI in include/vxl.h. It mirrors the
reference's GPU boundary:

* ``GpuVoxelizer.voxelize_surface`` (voxelizer/src/gpu/dense.rs:23-58) ->
  :meth:`GpuVoxelizer.voxelize_surface` (dense x-major occupancy words) and
  :meth:`GpuVoxelizer.voxelize_tree` (device-resident School-B tree);
* ``GpuTraverser::new`` / ``traverse`` (voxel-gpu/src/traverse.rs:70,154) ->
  :class:`GpuTraverser`.

There is no CPU fallback: if libvxl.so is missing or no GPU is present the calls
raise, they never compute on the host.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libvxl.so")

STATUS = {
    0: "VXL_OK", 1: "VXL_NO_DEVICE", 2: "VXL_INVALID_ARG", 3: "VXL_STORAGE_EXCEEDED",
    4: "VXL_WORKGROUPS_EXCEEDED", 5: "VXL_OOM", 6: "VXL_BUSY", 7: "VXL_DEVICE_ERROR",
    8: "VXL_NOT_REPRESENTABLE",
}

# Every symbol include/vxl.h declares.
EXPORTS = [
    "vxl_ctx_create", "vxl_ctx_destroy", "vxl_ctx_stream", "vxl_last_error", "vxl_status_str",
    "vxl_voxelize_dense", "vxl_voxelize_dense_dev", "vxl_voxelize_dense_owner",
    "vxl_voxelize_dense_owner_dev", "vxl_voxelize_tree", "vxl_voxelize_tree_dev",
    "vxl_voxelize_tree_materials", "vxl_voxelize_tree_materials_dev", "vxl_compact_surface",
    "vxl_tree_from_compact", "vxl_tree_materials_download", "vxl_render", "vxl_render_dev",
    "vxl_tree_set_voxels", "vxl_generate_noise_tree", "vxl_traverse_stats",
    "vxl_traverse_graph_create", "vxl_graph_launch", "vxl_graph_destroy",
    "vxl_batch_graph_create", "vxl_batch_graph_launch", "vxl_batch_graph_trees",
    "vxl_batch_graph_info", "vxl_batch_graph_destroy",
    "vxl_tree_from_words", "vxl_tree_info_get", "vxl_tree_download", "vxl_tree_upload",
    "vxl_tree_free", "vxl_traverse", "vxl_traverse_dev", "vxl_traverse_dev_noalloc",
    "vxl_set_timing", "vxl_last_timing",
    "vxl_set_tuning", "vxl_set_merge_words", "vxl_set_batch_refill", "vxl_set_dda_refill", "vxl_voxelize_batch", "vxl_voxelize_batch_dev", "vxl_traverse_batch",
    "vxl_traverse_batch_dev", "vxl_set_batch_budget", "vxl_autotune_batch",
    "vxl_server_create", "vxl_server_destroy", "vxl_server_hold", "vxl_server_submit_rays",
    "vxl_server_submit_voxelize", "vxl_server_poll", "vxl_server_wait", "vxl_server_stats_get",
    "vxl_server_reset_stats",
]


class VxlError(RuntimeError):
    """A non-OK vxl_status, with the library's message (the Rust error enums' Display)."""

    def __init__(self, status: int, message: str):
        self.status = status
        self.name = STATUS.get(status, str(status))
        super().__init__(f"{self.name}: {message}")


class Grid(C.Structure):
    _fields_ = [("n", C.c_uint32), ("origin", C.c_float * 3), ("voxel_size", C.c_float),
                ("has_world_to_grid", C.c_uint32), ("world_to_grid", C.c_float * 16)]


class TreeInfo(C.Structure):
    _fields_ = [("n", C.c_uint32), ("k", C.c_uint32), ("node_count", C.c_uint64),
                ("leaf_count", C.c_uint64), ("level_nodes", C.c_uint64 * 16),
                ("device_bytes", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("sat_waves_per_cu", C.c_int), ("sat_units_per_wave", C.c_int),
                ("dda_block", C.c_int), ("voxelize_ms", C.c_float), ("traverse_ms", C.c_float),
                ("sat_merge_words", C.c_int), ("dda_batch_refill", C.c_int)]


class ServerConfig(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("batching", C.c_uint32),
                ("max_batch_requests", C.c_uint32), ("max_batch_rays", C.c_uint64),
                ("max_batch_tris", C.c_uint64), ("batch_refill", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("batches", C.c_uint64), ("mean_batch_requests", C.c_double),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
                ("max_ms", C.c_double), ("ray_p50_ms", C.c_double), ("ray_p99_ms", C.c_double),
                ("vox_p50_ms", C.c_double), ("vox_p99_ms", C.c_double),
                ("throughput_rps", C.c_double), ("peak_bytes", C.c_uint64)]


_lib = None
_vp = C.c_void_p
_u64 = C.c_uint64
_pu64 = C.POINTER(C.c_uint64)


def load(path: str = LIB_PATH):
    """Loads libvxl.so; raises FileNotFoundError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: run __graft_entry__.build()")
    L = C.CDLL(path)
    L.vxl_ctx_create.argtypes = [C.c_int, C.POINTER(_vp)]
    L.vxl_ctx_destroy.argtypes = [_vp]
    L.vxl_ctx_destroy.restype = None
    L.vxl_ctx_stream.argtypes = [_vp]
    L.vxl_ctx_stream.restype = _vp
    L.vxl_last_error.restype = C.c_char_p
    L.vxl_status_str.restype = C.c_char_p
    L.vxl_status_str.argtypes = [C.c_int]
    L.vxl_voxelize_dense.argtypes = [_vp, _vp, _u64, C.POINTER(Grid), C.c_float, _vp, _pu64]
    L.vxl_voxelize_dense_dev.argtypes = [_vp, _vp, _u64, C.POINTER(Grid), C.c_float, _vp, _pu64,
                                         _vp]
    L.vxl_voxelize_dense_owner.argtypes = [_vp, _vp, _u64, C.POINTER(Grid), C.c_float, _vp, _vp,
                                           _vp, _pu64]
    L.vxl_voxelize_dense_owner_dev.argtypes = [_vp, _vp, _u64, C.POINTER(Grid), C.c_float, _vp,
                                               _vp, _vp, _pu64, _vp]
    L.vxl_voxelize_tree_materials.argtypes = [_vp, _vp, _u64, C.POINTER(Grid), C.c_float, _vp,
                                              C.c_uint32, C.POINTER(_vp), _pu64]
    L.vxl_voxelize_tree_materials_dev.argtypes = [_vp, _vp, _u64, C.POINTER(Grid), C.c_float,
                                                  _vp, C.c_uint32, C.POINTER(_vp), _pu64, _vp]
    L.vxl_traverse_stats.argtypes = [_vp, _vp, _vp, _u64, _pu64]
    L.vxl_traverse_graph_create.argtypes = [_vp, C.POINTER(_vp), C.c_uint32, _vp, _pu64, _vp,
                                            _vp, C.POINTER(_vp)]
    L.vxl_graph_launch.argtypes = [_vp, _vp]
    L.vxl_graph_destroy.argtypes = [_vp]
    L.vxl_graph_destroy.restype = None
    L.vxl_batch_graph_create.argtypes = [_vp, _vp, _pu64, C.c_uint32, _vp, C.c_float, _vp,
                                         _pu64, _vp, _vp, C.c_double, C.POINTER(_vp)]
    L.vxl_batch_graph_launch.argtypes = [_vp, _vp]
    L.vxl_batch_graph_trees.argtypes = [_vp, _vp, _pu64]
    L.vxl_batch_graph_info.argtypes = [_vp, _pu64, C.POINTER(C.c_uint32), _pu64]
    L.vxl_batch_graph_destroy.argtypes = [_vp]
    L.vxl_batch_graph_destroy.restype = None
    L.vxl_generate_noise_tree.argtypes = [_vp, C.c_uint32, C.POINTER(Noise), C.POINTER(_vp)]
    L.vxl_tree_set_voxels.argtypes = [_vp, _vp, _vp, _u64, _pu64, C.POINTER(C.c_int)]
    L.vxl_render.argtypes = [_vp, _vp, C.POINTER(Camera), _vp, C.c_uint32, _vp]
    L.vxl_render_dev.argtypes = [_vp, _vp, C.POINTER(Camera), _vp, C.c_uint32, _vp, _vp]
    L.vxl_compact_surface.argtypes = [_vp, _vp, _u64, C.POINTER(Grid), C.c_float, _vp,
                                      C.c_uint32, _vp, _vp, _u64, _pu64]
    L.vxl_tree_from_compact.argtypes = [_vp, C.c_uint32, _vp, _u64, C.POINTER(_vp), _pu64]
    L.vxl_tree_materials_download.argtypes = [_vp, _vp]
    L.vxl_voxelize_tree.argtypes = [_vp, _vp, _u64, C.POINTER(Grid), C.c_float, C.POINTER(_vp),
                                    _pu64]
    L.vxl_voxelize_tree_dev.argtypes = [_vp, _vp, _u64, C.POINTER(Grid), C.c_float,
                                        C.POINTER(_vp), _pu64, _vp]
    L.vxl_tree_from_words.argtypes = [_vp, _vp, C.c_uint32, C.POINTER(_vp)]
    L.vxl_tree_info_get.argtypes = [_vp, C.POINTER(TreeInfo)]
    L.vxl_tree_download.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.vxl_tree_upload.argtypes = [_vp, C.c_uint32, _vp, _u64, _vp, _vp, _u64, C.POINTER(_vp)]
    L.vxl_tree_free.argtypes = [_vp]
    L.vxl_tree_free.restype = None
    L.vxl_traverse.argtypes = [_vp, _vp, _vp, _u64, _vp, _vp]
    L.vxl_traverse_dev.argtypes = [_vp, _vp, _vp, _u64, _vp, _vp, _vp]
    L.vxl_traverse_dev_noalloc.argtypes = [_vp, _vp, _vp, _u64, _vp, _vp, _vp]
    L.vxl_set_timing.argtypes = [_vp, C.c_int]
    L.vxl_set_timing.restype = None
    L.vxl_last_timing.argtypes = [_vp, C.POINTER(C.c_float)]
    L.vxl_last_timing.restype = None
    L.vxl_set_tuning.argtypes = [_vp, C.c_int, C.c_int, C.c_int]
    L.vxl_set_dda_refill.argtypes = [_vp, C.c_int]
    L.vxl_set_merge_words.argtypes = [_vp, C.c_int]
    L.vxl_set_batch_refill.argtypes = [_vp, C.c_int]
    L.vxl_voxelize_batch.argtypes = [_vp, _vp, _vp, C.c_uint32, _vp, C.c_float, _vp, _vp]
    L.vxl_voxelize_batch_dev.argtypes = [_vp, _vp, _vp, C.c_uint32, _vp, C.c_float, _vp, _vp,
                                         _vp]
    L.vxl_traverse_batch.argtypes = [_vp, _vp, C.c_uint32, _vp, _vp, _vp, _vp]
    L.vxl_traverse_batch_dev.argtypes = [_vp, _vp, C.c_uint32, _vp, _vp, _vp, _vp, _vp]
    L.vxl_set_batch_budget.argtypes = [_vp, _u64]
    L.vxl_autotune_batch.argtypes = [_vp, _vp, _vp, C.c_uint32, _vp, C.c_float, _vp, _vp, _vp,
                                     _vp, C.POINTER(Tuning)]
    L.vxl_server_create.argtypes = [C.c_int, C.POINTER(ServerConfig), C.POINTER(_vp)]
    L.vxl_server_destroy.argtypes = [_vp]
    L.vxl_server_destroy.restype = None
    L.vxl_server_hold.argtypes = [_vp, C.c_int64]
    L.vxl_server_submit_rays.argtypes = [_vp, _vp, _vp, _u64, _vp, _vp, _pu64]
    L.vxl_server_submit_voxelize.argtypes = [_vp, _vp, _u64, C.POINTER(Grid), C.c_float, _pu64]
    L.vxl_server_poll.argtypes = [_vp, _u64, C.POINTER(C.c_int)]
    L.vxl_server_wait.argtypes = [_vp, _u64, C.POINTER(_vp), _pu64, C.POINTER(C.c_double)]
    L.vxl_server_stats_get.argtypes = [_vp, C.POINTER(ServerStats)]
    L.vxl_server_reset_stats.argtypes = [_vp]
    for name in ("vxl_server_create", "vxl_server_hold", "vxl_server_submit_rays",
                 "vxl_server_submit_voxelize", "vxl_server_poll", "vxl_server_wait",
                 "vxl_server_stats_get", "vxl_server_reset_stats"):
        getattr(L, name).restype = C.c_int
    for name in ("vxl_ctx_create", "vxl_voxelize_dense", "vxl_voxelize_dense_dev",
                 "vxl_voxelize_dense_owner", "vxl_voxelize_dense_owner_dev",
                 "vxl_voxelize_tree_materials", "vxl_voxelize_tree_materials_dev",
                 "vxl_compact_surface", "vxl_tree_from_compact", "vxl_tree_materials_download",
                 "vxl_render", "vxl_render_dev", "vxl_tree_set_voxels",
                 "vxl_generate_noise_tree", "vxl_traverse_stats",
                 "vxl_traverse_graph_create", "vxl_graph_launch",
                 "vxl_batch_graph_create", "vxl_batch_graph_launch", "vxl_batch_graph_trees",
                 "vxl_batch_graph_info",
                 "vxl_voxelize_tree", "vxl_voxelize_tree_dev", "vxl_tree_from_words",
                 "vxl_tree_info_get", "vxl_tree_download", "vxl_tree_upload", "vxl_traverse",
                 "vxl_traverse_dev", "vxl_traverse_dev_noalloc", "vxl_set_tuning", "vxl_set_merge_words",
                 "vxl_set_batch_refill", "vxl_set_dda_refill", "vxl_voxelize_batch",
                 "vxl_voxelize_batch_dev", "vxl_traverse_batch", "vxl_traverse_batch_dev",
                 "vxl_set_batch_budget", "vxl_autotune_batch"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise VxlError(rc, (load().vxl_last_error() or b"").decode())


class Noise(C.Structure):
    """vxl_noise: NoiseField's parameters (voxel-core fixtures.rs:259-279)."""
    _fields_ = [("seed", C.c_uint32), ("octaves", C.c_uint32), ("ridged", C.c_uint32),
                ("pad", C.c_uint32), ("frequency", C.c_double), ("lacunarity", C.c_double),
                ("gain", C.c_double), ("warp", C.c_double), ("threshold", C.c_double)]


# NoiseField presets (voxel-core fixtures.rs:284-314)
NOISE_PERLIN = dict(seed=0x5EED1234, frequency=5.0, octaves=5, lacunarity=2.0, gain=0.5,
                    ridged=False, warp=0.0, threshold=0.12)
NOISE_CAVES = dict(seed=0xCA5E5EED, frequency=4.0, octaves=5, lacunarity=2.0, gain=0.5,
                   ridged=True, warp=0.45, threshold=0.48)


class TraverseGraph:
    """A replayable HIP graph of per-tree traversals (keeps its trees alive)."""

    def __init__(self, handle: int, trees):
        self._h = handle
        self._trees = list(trees)

    def launch(self, stream: int = 0):
        _check(load().vxl_graph_launch(_vp(self._h), _vp(stream)))

    def free(self):
        if self._h:
            load().vxl_graph_destroy(_vp(self._h))
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BatchGraph:
    """A whole batch (voxelize every mesh, build its tree, walk its rays) as one HIP graph
    with no host round trip; launch() replays it, trees() copies the last replay's trees."""

    def __init__(self, handle: int, ctx, count: int):
        self._h = handle
        self._ctx = ctx
        self.count = count

    def launch(self, stream: int = 0):
        _check(load().vxl_batch_graph_launch(_vp(self._h), _vp(stream)))

    def trees(self, want_trees=True):
        """Waits for the last launch; returns (trees or None, pairs). Raises VxlError
        (VXL_STORAGE_EXCEEDED) when a group's trees did not fit the arena."""
        out = (_vp * self.count)() if want_trees else None
        pairs = np.zeros(self.count, np.uint64)
        _check(load().vxl_batch_graph_trees(_vp(self._h), out, pairs.ctypes.data_as(_pu64)))
        if not want_trees:
            return None, pairs
        trees = []
        for i in range(self.count):
            t = Tree(out[i], self._ctx)
            t.pairs = int(pairs[i])
            trees.append(t)
        return trees, pairs

    def info(self):
        a, g, n = C.c_uint64(0), C.c_uint32(0), C.c_uint64(0)
        _check(load().vxl_batch_graph_info(_vp(self._h), C.byref(a), C.byref(g), C.byref(n)))
        return dict(arena_bytes=int(a.value), groups=int(g.value), graph_nodes=int(n.value))

    def free(self):
        if self._h:
            load().vxl_batch_graph_destroy(_vp(self._h))
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Camera(C.Structure):
    """vxl_camera: render.wgsl's Camera (eye in grid units, tan of half the vertical field
    of view, unit basis, aspect, grid size n) plus the image size."""
    _fields_ = [("eye", C.c_float * 3), ("tan_half_fov", C.c_float), ("forward", C.c_float * 3),
                ("aspect", C.c_float), ("right", C.c_float * 3), ("n", C.c_float),
                ("up", C.c_float * 3), ("pad0", C.c_float), ("width", C.c_uint32),
                ("height", C.c_uint32), ("pad1", C.c_uint32), ("pad2", C.c_uint32)]


def make_camera(cam16, width: int, height: int) -> Camera:
    """From 16 f32 (eye, tan, forward, aspect, right, n, up, pad) and the image size."""
    c = np.ascontiguousarray(cam16, np.float32)
    cam = Camera()
    cam.eye[:] = c[0:3].tolist()
    cam.tan_half_fov = float(c[3])
    cam.forward[:] = c[4:7].tolist()
    cam.aspect = float(c[7])
    cam.right[:] = c[8:11].tolist()
    cam.n = float(c[11])
    cam.up[:] = c[12:15].tolist()
    cam.width, cam.height = width, height
    return cam


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_vp)


def make_grid(n: int, origin=(0.0, 0.0, 0.0), voxel_size: float = 1.0, mat=None) -> Grid:
    """vxl_grid; mat (optional): the world_to_grid override as 16 floats, column-major."""
    g = Grid()
    g.n = n
    for i in range(3):
        g.origin[i] = origin[i]
    g.voxel_size = voxel_size
    if mat is not None:
        m = np.asarray(mat, np.float32).reshape(16)
        g.has_world_to_grid = 1
        for i in range(16):
            g.world_to_grid[i] = float(m[i])
    return g


class Tree:
    """A device-resident School-B tree (nodes, Morton-sorted leaves, leaf bounds)."""

    def __init__(self, handle: int, ctx: "Context"):
        self._h = handle
        self.ctx = ctx
        self._refresh()

    def _refresh(self):
        info = TreeInfo()
        _check(load().vxl_tree_info_get(_vp(self._h), C.byref(info)))
        self.n = info.n
        self.k = info.k
        self.node_count = int(info.node_count)
        self.leaf_count = int(info.leaf_count)
        self.level_nodes = {L: int(info.level_nodes[L]) for L in range(16) if info.level_nodes[L]}
        self.device_bytes = int(info.device_bytes)

    @property
    def handle(self):
        return _vp(self._h)

    def download(self):
        """(nodes [N,3] u32, leaves [L,8] u64, bounds [L] u32, codes [L] u64)."""
        nodes = np.zeros((self.node_count, 3), np.uint32)
        leaves = np.zeros((self.leaf_count, 8), np.uint64)
        bounds = np.zeros(self.leaf_count, np.uint32)
        codes = np.zeros(self.leaf_count, np.uint64)
        _check(load().vxl_tree_download(self.handle, _ptr(nodes), _ptr(leaves), _ptr(bounds),
                                        _ptr(codes)))
        return nodes, leaves, bounds, codes

    def materials(self):
        """[L, 512] u16 per-voxel materials in intra-brick Morton order (zeros for a tree
        made without materials)."""
        mats = np.zeros((self.leaf_count, 512), np.uint16)
        _check(load().vxl_tree_materials_download(self.handle, _ptr(mats)))
        return mats

    def free(self):
        if self._h:
            load().vxl_tree_free(_vp(self._h))
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """A device + stream + grow-only scratch (vxl_ctx)."""

    def __init__(self, device: int = 0):
        h = _vp()
        _check(load().vxl_ctx_create(device, C.byref(h)))
        self._h = h.value

    @property
    def handle(self):
        return _vp(self._h)

    @property
    def stream(self) -> int:
        return load().vxl_ctx_stream(self.handle) or 0

    def close(self):
        if self._h:
            load().vxl_ctx_destroy(_vp(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- voxelize ------------------------------------------------------------------------
    def voxelize_dense(self, tris, n, origin=(0.0, 0.0, 0.0), voxel_size=1.0, eps=1e-4, mat=None):
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        words = np.zeros((n ** 3 + 31) // 32, np.uint32)
        pairs = C.c_uint64(0)
        g = make_grid(n, origin, voxel_size, mat)
        _check(load().vxl_voxelize_dense(self.handle, _ptr(t), t.shape[0], C.byref(g), eps,
                                         _ptr(words), C.byref(pairs)))
        return words, int(pairs.value)

    def voxelize_dense_owner(self, tris, n, origin=(0.0, 0.0, 0.0), voxel_size=1.0, eps=1e-4,
                             owner=True, color=False):
        """voxelize_surface with store_owner / store_color: (words, owner or None, color or
        None, pairs); owner and color are n^3 u32, x-major."""
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        words = np.zeros((n ** 3 + 31) // 32, np.uint32)
        own = np.zeros(n ** 3, np.uint32) if owner else None
        col = np.zeros(n ** 3, np.uint32) if color else None
        pairs = C.c_uint64(0)
        g = make_grid(n, origin, voxel_size)
        _check(load().vxl_voxelize_dense_owner(
            self.handle, _ptr(t), t.shape[0], C.byref(g), eps, _ptr(words),
            _ptr(own) if owner else None, _ptr(col) if color else None, C.byref(pairs)))
        return words, own, col, int(pairs.value)

    def voxelize_tree(self, tris, n, origin=(0.0, 0.0, 0.0), voxel_size=1.0, eps=1e-4, mat=None):
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        out = _vp()
        pairs = C.c_uint64(0)
        g = make_grid(n, origin, voxel_size, mat)
        _check(load().vxl_voxelize_tree(self.handle, _ptr(t), t.shape[0], C.byref(g), eps,
                                        C.byref(out), C.byref(pairs)))
        tree = Tree(out.value, self)
        tree.pairs = int(pairs.value)
        return tree

    def voxelize_tree_dev(self, d_tris: int, T: int, n, origin=(0.0, 0.0, 0.0), voxel_size=1.0,
                          eps=1e-4, stream: int = 0):
        out = _vp()
        pairs = C.c_uint64(0)
        g = make_grid(n, origin, voxel_size)
        _check(load().vxl_voxelize_tree_dev(self.handle, _vp(d_tris), T, C.byref(g), eps,
                                            C.byref(out), C.byref(pairs), _vp(stream)))
        tree = Tree(out.value, self)
        tree.pairs = int(pairs.value)
        return tree

    def voxelize_tree_materials(self, tris, n, origin=(0.0, 0.0, 0.0), voxel_size=1.0,
                                eps=1e-4, table=()):
        """Voxelize into a tree carrying per-voxel materials from a per-triangle table of
        two u16 per u32."""
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        tb = np.ascontiguousarray(table, np.uint32)
        out = _vp()
        pairs = C.c_uint64(0)
        g = make_grid(n, origin, voxel_size)
        _check(load().vxl_voxelize_tree_materials(self.handle, _ptr(t), t.shape[0], C.byref(g),
                                                  eps, _ptr(tb) if tb.size else None, tb.size,
                                                  C.byref(out), C.byref(pairs)))
        tree = Tree(out.value, self)
        tree.pairs = int(pairs.value)
        return tree

    def voxelize_tree_materials_dev(self, d_tris: int, T: int, n, origin=(0.0, 0.0, 0.0),
                                    voxel_size=1.0, eps=1e-4, d_table: int = 0,
                                    table_len: int = 0, stream: int = 0):
        out = _vp()
        pairs = C.c_uint64(0)
        g = make_grid(n, origin, voxel_size)
        _check(load().vxl_voxelize_tree_materials_dev(
            self.handle, _vp(d_tris), T, C.byref(g), eps, _vp(d_table) if d_table else None,
            table_len, C.byref(out), C.byref(pairs), _vp(stream)))
        tree = Tree(out.value, self)
        tree.pairs = int(pairs.value)
        return tree

    def compact_surface(self, tris, n, origin=(0.0, 0.0, 0.0), voxel_size=1.0, eps=1e-4,
                        table=(), g_origin=(0, 0, 0)):
        """compact_surface_sparse: [count, 4] i32 rows (vx, vy, vz, material)."""
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        tb = np.ascontiguousarray(table, np.uint32)
        go = np.ascontiguousarray(g_origin, np.int32)
        g = make_grid(n, origin, voxel_size)
        count = C.c_uint64(0)
        args = (self.handle, _ptr(t), t.shape[0], C.byref(g), eps,
                _ptr(tb) if tb.size else None, tb.size, _ptr(go))
        _check(load().vxl_compact_surface(*args, None, 0, C.byref(count)))
        out = np.zeros((int(count.value), 4), np.int32)
        if out.shape[0]:
            _check(load().vxl_compact_surface(*args, _ptr(out), out.shape[0], C.byref(count)))
        return out

    def traverse_graph(self, trees, d_rays: int, ray_offsets, d_hits: int, d_t: int = 0):
        """The per-tree traversals captured in a HIP graph: returns a TraverseGraph whose
        launch() replays them."""
        arr = (_vp * len(trees))(*[t._h for t in trees])
        off = np.ascontiguousarray(ray_offsets, np.uint64)
        out = _vp()
        _check(load().vxl_traverse_graph_create(self.handle, arr, len(trees), _vp(d_rays),
                                                off.ctypes.data_as(_pu64), _vp(d_hits),
                                                _vp(d_t) if d_t else None, C.byref(out)))
        return TraverseGraph(out.value, trees)

    def traverse_stats(self, tree: "Tree", d_rays: int, R: int):
        """Sums over the rays of (leaf descents, node descents, cell visits) of the walk."""
        out = (C.c_uint64 * 3)()
        _check(load().vxl_traverse_stats(self.handle, tree.handle, _vp(d_rays), R, out))
        return int(out[0]), int(out[1]), int(out[2])

    def generate_noise_tree(self, n: int, field: dict):
        """The NoiseField fixture's tree (fields as NOISE_PERLIN / NOISE_CAVES)."""
        f = Noise(seed=field["seed"], octaves=field["octaves"], ridged=1 if field["ridged"] else 0,
                  frequency=field["frequency"], lacunarity=field["lacunarity"],
                  gain=field["gain"], warp=field["warp"], threshold=field["threshold"])
        out = _vp()
        _check(load().vxl_generate_noise_tree(self.handle, n, C.byref(f), C.byref(out)))
        return Tree(out.value, self)

    def set_voxels(self, tree: "Tree", edits):
        """SparseTree::set_voxel over rows (x, y, z, occupied), in order: (voxels changed,
        whether bricks appeared or emptied and the tree was rebuilt). Updates tree's counts."""
        e = np.ascontiguousarray(edits, np.uint32).reshape(-1, 4)
        changed = C.c_uint64(0)
        topo = C.c_int(0)
        _check(load().vxl_tree_set_voxels(self.handle, tree.handle, _ptr(e) if e.size else None,
                                          e.shape[0], C.byref(changed), C.byref(topo)))
        tree._refresh()
        return int(changed.value), bool(topo.value)

    def render(self, tree: "Tree", cam16, width: int, height: int, colors=()):
        """render_main: (height, width) u32 RGBA8 (R in the low byte)."""
        cam = make_camera(cam16, width, height)
        col = np.ascontiguousarray(colors, np.uint32)
        img = np.zeros(height * width, np.uint32)
        _check(load().vxl_render(self.handle, tree.handle, C.byref(cam),
                                 _ptr(col) if col.size else None, col.size, _ptr(img)))
        return img.reshape(height, width)

    def render_dev(self, tree: "Tree", cam16, width: int, height: int, d_rgba: int,
                   d_colors: int = 0, ncolors: int = 0, stream: int = 0):
        cam = make_camera(cam16, width, height)
        _check(load().vxl_render_dev(self.handle, tree.handle, C.byref(cam),
                                     _vp(d_colors) if d_colors else None, ncolors, _vp(d_rgba),
                                     _vp(stream)))

    def tree_from_compact(self, voxels, n):
        """tree_from_compact: (Tree with materials, number of voxels dropped)."""
        v = np.ascontiguousarray(voxels, np.int32).reshape(-1, 4)
        out = _vp()
        dropped = C.c_uint64(0)
        _check(load().vxl_tree_from_compact(self.handle, n, _ptr(v) if v.size else None,
                                            v.shape[0], C.byref(out), C.byref(dropped)))
        return Tree(out.value, self), int(dropped.value)

    def tree_from_words(self, words, n):
        w = np.ascontiguousarray(words, np.uint32)
        out = _vp()
        _check(load().vxl_tree_from_words(self.handle, _ptr(w), n, C.byref(out)))
        return Tree(out.value, self)

    def tree_upload(self, n, nodes, leaves, bounds):
        nd = np.ascontiguousarray(nodes, np.uint32).reshape(-1, 3)
        lv = np.ascontiguousarray(leaves, np.uint64).reshape(-1, 8)
        bd = np.ascontiguousarray(bounds, np.uint32)
        out = _vp()
        _check(load().vxl_tree_upload(self.handle, n, _ptr(nd), nd.shape[0], _ptr(lv), _ptr(bd),
                                      lv.shape[0], C.byref(out)))
        return Tree(out.value, self)

    # -- traversal -----------------------------------------------------------------------
    def traverse(self, tree: Tree, rays, want_t=True):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        hits = np.zeros((r.shape[0], 4), np.uint32)
        t = np.zeros(r.shape[0], np.float64) if want_t else None
        _check(load().vxl_traverse(self.handle, tree.handle, _ptr(r), r.shape[0], _ptr(hits),
                                   _ptr(t) if want_t else None))
        return (hits, t) if want_t else hits

    def traverse_dev(self, tree: Tree, d_rays: int, R: int, d_hits: int, d_t: int = 0,
                     stream: int = 0):
        _check(load().vxl_traverse_dev(self.handle, tree.handle, _vp(d_rays), R, _vp(d_hits),
                                       _vp(d_t) if d_t else None, _vp(stream)))

    def traverse_dev_noalloc(self, tree: Tree, d_rays: int, R: int, d_hits: int, d_t: int = 0,
                             stream: int = 0):
        """vxl_traverse_dev_noalloc: the same traversal with no device allocation."""
        _check(load().vxl_traverse_dev_noalloc(self.handle, tree.handle, _vp(d_rays), R,
                                               _vp(d_hits), _vp(d_t) if d_t else None,
                                               _vp(stream)))

    # -- timing / tuning -----------------------------------------------------------------
    def set_timing(self, on: bool):
        load().vxl_set_timing(self.handle, 1 if on else 0)

    def last_timing(self):
        out = (C.c_float * 8)()
        load().vxl_last_timing(self.handle, out)
        return list(out)

    def set_tuning(self, sat_waves_per_cu=0, sat_units_per_wave=0, dda_block=0):
        _check(load().vxl_set_tuning(self.handle, sat_waves_per_cu, sat_units_per_wave, dda_block))

    def set_merge_words(self, words: int):
        """The SAT's merge tile: leaf-grid words a wave merges in LDS before writing them
        out, in [1, 64]."""
        _check(load().vxl_set_merge_words(self.handle, int(words)))

    def set_batch_refill(self, refill_min: int):
        """Lane schedule of batched traversals: 0 one lane per ray, 1..64 persistent lanes."""
        _check(load().vxl_set_batch_refill(self.handle, int(refill_min)))

    def set_dda_refill(self, refill_min: int):
        """Lane schedule of traverse / traverse_dev: refill_min in [1, 64] runs persistent
        waves whose idle lanes take new rays once refill_min are idle; 0 is one lane per
        ray; -1 (the default) picks by tree size. The results do not depend on it."""
        _check(load().vxl_set_dda_refill(self.handle, refill_min))

    # -- batches ---------------------------------------------------------------------------
    @staticmethod
    def _grids(grids):
        arr = (Grid * max(len(grids), 1))()
        for i, g in enumerate(grids):  # (n, origin, voxel_size[, world_to_grid])
            arr[i] = make_grid(*g)
        return arr

    def _batch_trees(self, out, pairs, count):
        trees = []
        for i in range(count):
            t = Tree(out[i], self)
            t.pairs = int(pairs[i])
            trees.append(t)
        return trees

    def voxelize_batch(self, meshes, grids, eps=1e-4):
        """meshes: list of (T_i, 9) f32 arrays; grids: list of (n, origin, voxel_size)."""
        count = len(meshes)
        off = np.zeros(count + 1, np.uint64)
        off[1:] = np.cumsum([len(m) for m in meshes]) if count else []
        tris = (np.ascontiguousarray(np.concatenate([np.asarray(m, np.float32).reshape(-1, 9)
                                                     for m in meshes]), np.float32)
                if count else np.zeros((0, 9), np.float32))
        out = (_vp * max(count, 1))()
        pairs = np.zeros(max(count, 1), np.uint64)
        _check(load().vxl_voxelize_batch(self.handle, _ptr(tris), _ptr(off), count,
                                         self._grids(grids), eps, out, _ptr(pairs)))
        return self._batch_trees(out, pairs, count)

    def voxelize_batch_dev(self, d_tris: int, tri_offsets, grids, eps=1e-4, stream: int = 0):
        off = np.ascontiguousarray(tri_offsets, np.uint64)
        count = len(off) - 1
        out = (_vp * max(count, 1))()
        pairs = np.zeros(max(count, 1), np.uint64)
        _check(load().vxl_voxelize_batch_dev(self.handle, _vp(d_tris), _ptr(off), count,
                                             self._grids(grids), eps, out, _ptr(pairs),
                                             _vp(stream)))
        return self._batch_trees(out, pairs, count)

    def traverse_batch(self, trees, rays_list, want_t=True):
        """rays_list[i] (R_i x 6 f32) against trees[i]; returns per-segment (hits, t)."""
        count = len(trees)
        off = np.zeros(count + 1, np.uint64)
        off[1:] = np.cumsum([len(r) for r in rays_list]) if count else []
        rays = (np.ascontiguousarray(np.concatenate([np.asarray(r, np.float32).reshape(-1, 6)
                                                     for r in rays_list]), np.float32)
                if count else np.zeros((0, 6), np.float32))
        R = rays.shape[0]
        hits = np.zeros((R, 4), np.uint32)
        t = np.zeros(R, np.float64) if want_t else None
        handles = (_vp * max(count, 1))(*[tr._h for tr in trees])
        _check(load().vxl_traverse_batch(self.handle, handles, count, _ptr(rays), _ptr(off),
                                         _ptr(hits), _ptr(t) if want_t else None))
        segs = [(int(off[i]), int(off[i + 1])) for i in range(count)]
        return [(hits[a:b], t[a:b] if want_t else None) for a, b in segs]

    def traverse_batch_dev(self, trees, d_rays: int, ray_offsets, d_hits: int, d_t: int = 0,
                           stream: int = 0):
        off = np.ascontiguousarray(ray_offsets, np.uint64)
        count = len(trees)
        handles = (_vp * max(count, 1))(*[tr._h for tr in trees])
        _check(load().vxl_traverse_batch_dev(self.handle, handles, count, _vp(d_rays), _ptr(off),
                                             _vp(d_hits), _vp(d_t) if d_t else None,
                                             _vp(stream)))

    def batch_graph(self, d_tris: int, tri_offsets, grids, d_rays: int = 0, ray_offsets=None,
                    d_hits: int = 0, d_t: int = 0, eps=1e-4, headroom=0.0):
        """vxl_batch_graph_create: the batch (meshes at tri_offsets of d_tris, rays at
        ray_offsets of d_rays against them) captured in one HIP graph."""
        off = np.ascontiguousarray(tri_offsets, np.uint64)
        count = len(off) - 1
        roff = np.ascontiguousarray(ray_offsets, np.uint64) if ray_offsets is not None else None
        out = _vp()
        _check(load().vxl_batch_graph_create(
            self.handle, _vp(d_tris) if d_tris else None, off.ctypes.data_as(_pu64), count,
            self._grids(grids), eps, _vp(d_rays) if d_rays else None,
            roff.ctypes.data_as(_pu64) if roff is not None else None,
            _vp(d_hits) if d_hits else None, _vp(d_t) if d_t else None, float(headroom),
            C.byref(out)))
        return BatchGraph(out.value, self, count)

    def set_batch_budget(self, nbytes: int):
        _check(load().vxl_set_batch_budget(self.handle, int(nbytes)))

    def autotune_batch(self, d_tris: int, tri_offsets, grids, d_rays: int = 0, ray_offsets=None,
                       d_hits: int = 0, d_t: int = 0, eps=1e-4):
        off = np.ascontiguousarray(tri_offsets, np.uint64)
        roff = np.ascontiguousarray(ray_offsets, np.uint64) if ray_offsets is not None else None
        res = Tuning()
        _check(load().vxl_autotune_batch(self.handle, _vp(d_tris), _ptr(off), len(off) - 1,
                                         self._grids(grids), eps, _vp(d_rays) if d_rays else None,
                                         _ptr(roff) if roff is not None else None,
                                         _vp(d_hits) if d_hits else None,
                                         _vp(d_t) if d_t else None, C.byref(res)))
        return dict(sat_waves_per_cu=res.sat_waves_per_cu,
                    sat_units_per_wave=res.sat_units_per_wave,
                    sat_merge_words=res.sat_merge_words, dda_block=res.dda_block,
                    dda_batch_refill=res.dda_batch_refill,
                    voxelize_ms=res.voxelize_ms, traverse_ms=res.traverse_ms)


class Server:
    """The request loop (continuous batching of ray queries and voxelize requests)."""

    def __init__(self, device: int = 0, budget_bytes: int = 0, batching: bool = True,
                 max_batch_requests: int = 0, max_batch_rays: int = 0, max_batch_tris: int = 0,
                 batch_refill: int = 0):
        cfg = ServerConfig(budget_bytes, 1 if batching else 0, max_batch_requests,
                           max_batch_rays, max_batch_tris, batch_refill)
        h = _vp()
        _check(load().vxl_server_create(device, C.byref(cfg), C.byref(h)))
        self._h = h.value
        self._ctx = _ServerLaneOwner(self)

    def close(self):
        if self._h:
            load().vxl_server_destroy(_vp(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def hold(self, delta_bytes: int):
        _check(load().vxl_server_hold(_vp(self._h), int(delta_bytes)))

    def submit_rays(self, tree: "Tree", d_rays: int, R: int, d_hits: int, d_t: int = 0) -> int:
        rid = C.c_uint64(0)
        _check(load().vxl_server_submit_rays(_vp(self._h), tree.handle, _vp(d_rays), R,
                                             _vp(d_hits), _vp(d_t) if d_t else None,
                                             C.byref(rid)))
        return int(rid.value)

    def submit_voxelize(self, d_tris: int, T: int, n: int, origin=(0.0, 0.0, 0.0),
                        voxel_size=1.0, eps=1e-4) -> int:
        rid = C.c_uint64(0)
        g = make_grid(n, origin, voxel_size)
        _check(load().vxl_server_submit_voxelize(_vp(self._h), _vp(d_tris), T, C.byref(g), eps,
                                                 C.byref(rid)))
        return int(rid.value)

    def poll(self, rid: int) -> bool:
        d = C.c_int(0)
        _check(load().vxl_server_poll(_vp(self._h), rid, C.byref(d)))
        return bool(d.value)

    def wait(self, rid: int, keep_tree: bool = True):
        """(tree or None, pairs, latency_ms); the tree is freed unless keep_tree."""
        out = _vp()
        pairs = C.c_uint64(0)
        lat = C.c_double(0)
        _check(load().vxl_server_wait(_vp(self._h), rid, C.byref(out) if keep_tree else None,
                                      C.byref(pairs), C.byref(lat)))
        tree = Tree(out.value, self._ctx) if out.value else None
        return tree, int(pairs.value), float(lat.value)

    def stats(self) -> dict:
        st = ServerStats()
        _check(load().vxl_server_stats_get(_vp(self._h), C.byref(st)))
        return {name: getattr(st, name) for name, _ in ServerStats._fields_}

    def reset_stats(self):
        _check(load().vxl_server_reset_stats(_vp(self._h)))


class _ServerLaneOwner:
    """Stands in for the Context a Tree keeps a reference to: the server's lane owns it."""

    def __init__(self, server: Server):
        self.server = server


# ---- reference-shaped API (voxelizer::GpuVoxelizer, voxel_gpu::GpuTraverser) --------------
@dataclass
class VoxelizeOpts:
    """voxelizer/src/core.rs:465-482 (owner/colour channels are not on this path)."""
    epsilon: float = 1e-4


@dataclass
class VoxelGrid:
    """voxelizer/src/core.rs:25-35."""
    n: int
    origin_world: tuple = (0.0, 0.0, 0.0)
    voxel_size: float = 1.0


@dataclass
class VoxelizationOutput:
    """voxelizer/src/core.rs:519-528: occupancy words + dispatch counters."""
    occupancy: np.ndarray
    triangles: int
    voxels: int
    pairs: int = 0


@dataclass
class GpuVoxelizer:
    ctx: Context = field(default_factory=Context)

    def voxelize_surface(self, triangles, grid: VoxelGrid, opts: VoxelizeOpts = VoxelizeOpts()):
        words, pairs = self.ctx.voxelize_dense(triangles, grid.n, grid.origin_world,
                                               grid.voxel_size, opts.epsilon)
        t = np.asarray(triangles).reshape(-1, 9).shape[0]
        return VoxelizationOutput(words, t, grid.n ** 3, pairs)

    def voxelize_tree(self, triangles, grid: VoxelGrid, opts: VoxelizeOpts = VoxelizeOpts()):
        return self.ctx.voxelize_tree(triangles, grid.n, grid.origin_world, grid.voxel_size,
                                      opts.epsilon)


class GpuTraverser:
    """Uploads a School-B structure once, then traverses batches of rays."""

    def __init__(self, ctx: Context, structure):
        self.ctx = ctx
        if isinstance(structure, Tree):
            self.tree = structure
        else:
            n, nodes, leaves, bounds = structure
            self.tree = ctx.tree_upload(n, nodes, leaves, bounds)

    def traverse(self, rays):
        """List of (x, y, z) or None per ray, like Vec<Option<VoxelCoord>>."""
        hits = self.ctx.traverse(self.tree, rays, want_t=False)
        return [tuple(map(int, h[:3])) if h[3] == 1 else None for h in hits]


__all__ = ["Context", "Tree", "Grid", "Tuning", "Server", "VxlError", "load", "make_grid", "EXPORTS", "GpuVoxelizer",
           "GpuTraverser", "VoxelGrid", "VoxelizeOpts", "VoxelizationOutput", "LIB_PATH"]
