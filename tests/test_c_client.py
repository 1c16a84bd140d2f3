"""The C boundary used from C: tests/c/client.c includes include/vxl.h only and links
libvxl.so (no Python, no ctypes) -- it compiles and links here, and runs on a GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "friendly-journey_amd")


def _build(tmp_path):
    exe = os.path.join(str(tmp_path), "client")
    cmd = ["gcc", "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c", "client.c"), "-L", LIBDIR, "-lvxl",
           f"-Wl,-rpath,{LIBDIR}", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_client_compiles_and_links(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libvxl.so")):
        pytest.skip("libvxl.so not built")
    _build(tmp_path)


@pytest.mark.gpu
def test_c_client_runs(tmp_path):
    print("[case] C client", flush=True)
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "client ok" in r.stdout
