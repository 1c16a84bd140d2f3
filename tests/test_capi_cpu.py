"""CPU-side checks of the product library and host logic (no GPU compute here)."""
import ctypes
import os
import re

import numpy as np
import pytest
import vxl
from vxl import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "vxl.h")).read()
    return sorted(set(re.findall(r"\b(vxl_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(vxl.LIB_PATH)
    declared = _declared()
    assert declared and sorted(vxl.EXPORTS) == declared
    for name in declared:
        assert hasattr(lib, name), name


def test_library_has_no_cpu_fallback_symbols():
    # the product library links nothing from the oracle
    lib = ctypes.CDLL(vxl.LIB_PATH)
    for name in ("vo_voxelize_surface", "vo_mirror_traverse", "vo_tree_build"):
        assert not hasattr(lib, name)


def test_status_strings():
    lib = vxl.load()
    assert lib.vxl_status_str(0) == b"ok"
    assert lib.vxl_status_str(8) == b"resolution not representable"


def test_uv_sphere_counts_and_watertight_radius():
    t = W.uv_sphere()
    assert t.shape == (16128, 9)
    r = np.linalg.norm(t.reshape(-1, 3).astype(np.float64) - 0.5, axis=1)
    assert r.min() >= 0.4 * 0.98 - 1e-6 and r.max() <= 0.4 * 1.02 + 1e-6


def test_heightfield_and_soup_shapes():
    h = W.heightfield(64)
    assert h.shape == (2 * 64 * 64, 9)
    z = h.reshape(-1, 3)[:, 2]
    assert z.min() >= 0.0 and z.max() <= 0.5
    s = W.soup(1000)
    assert s.shape == (1000, 9)
    e = np.linalg.norm(s[:, 0:3] - s[:, 3:6], axis=1)
    assert 0.5 / 1024 < np.median(e) < 16 / 1024


def test_rays_are_unit_and_seeded():
    a = W.camera_rays(256, 64, 36)
    b = W.camera_rays(256, 64, 36)
    assert np.array_equal(a, b)
    assert np.allclose(np.linalg.norm(a[:, 3:], axis=1), 1.0, atol=1e-5)
    r = W.random_rays(1000, 2048)
    assert r[:, :3].min() >= 0 and r[:, :3].max() < 2048


def test_grid_mapping():
    assert W.grid_for(256) == (512, 1 / 256)
    assert W.grid_for(1024) == (2048, 1 / 1024)
    assert W.grid_for(512) == (512, 1 / 512)
    assert W.grid_for(64) == (128, 1 / 64)


def test_splitmix_matches_scalar_definition():
    s = W.SplitMix64(0x0DDB1A5E5EEDF00D)
    v = s.u64(3)
    st = 0x0DDB1A5E5EEDF00D
    M = (1 << 64) - 1
    for i in range(3):
        st = (st + 0x9E3779B97F4A7C15) & M
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        assert int(v[i]) == z ^ (z >> 31)
