"""CPU: the candidate window k_tri_setup gives a large triangle (voxelize.hip, setup_one)
holds every cell whose binary32 plane test passes (reference_cpu.rs:76-82).

The window formula is restated here in binary32, operation for operation as the setup
kernel and the k_sat decode evaluate it: window start floor(k0 + ka a + kb b), W cells,
hw = R / |n_c| + 1/8, W = ceil(2 hw + 1.95). Every cell of the bounding box of random large
triangles (thin slabs along an axis included) is tested by the plane test as the kernel
evaluates it (hoisted normal, |n . v0| <= rp); each passing cell must lie in its column's
window. Coordinates stay below 2^14 here, where the kernel uses the 1/8-cell slack."""
import math

import numpy as np

f32 = np.float32


def _window_violations(trials, seed):
    rng = np.random.default_rng(seed)
    bad = passed = windows = 0
    for trial in range(trials):
        scale = rng.choice([10, 30, 100, 300, 1000])
        off = rng.choice([scale + 2, 4000, 15000])
        base = rng.uniform(off, off + 500, 3)
        g = (base + rng.uniform(-scale, scale, (3, 3))).astype(f32).reshape(-1)
        if trial % 3 == 0:  # near-planar along an axis: thin slabs
            ax = rng.integers(0, 3)
            g[ax::3] = (g[ax] + rng.uniform(-0.5, 0.5, 3)).astype(f32)
        e0 = (g[3:6] - g[0:3]).astype(f32)
        e1 = (g[6:9] - g[3:6]).astype(f32)
        n = np.array([f32(e0[1] * e1[2]) - f32(e1[1] * e0[2]),
                      f32(e0[2] * e1[0]) - f32(e1[2] * e0[0]),
                      f32(e0[0] * e1[1]) - f32(e1[0] * e0[1])], f32)
        rp = f32(f32(f32(0.5) * abs(n[0])) + f32(f32(0.5) * abs(n[1]))) + f32(f32(0.5) * abs(n[2]))
        mn = [max(math.floor(min(g[i], g[3 + i], g[6 + i]) - 1e-4), 0) for i in range(3)]
        mx = [math.floor(max(g[i], g[3 + i], g[6 + i]) + 1e-4) for i in range(3)]
        ext = [mx[i] - mn[i] + 1 for i in range(3)]
        an = np.abs(n)
        c = 1 if (an[1] > an[0] and an[1] >= an[2]) else (2 if (an[2] > an[0] and an[2] > an[1]) else 0)
        if ext[c] < 6 or n[c] == 0:
            continue
        a = 0 if c != 0 else 1
        b = 1 if c == 2 else 2
        S = f32(0)
        for i in range(3):
            lo = f32(f32(mn[i]) + f32(0.5)) - g[i]
            hi = f32(f32(mx[i]) + f32(0.5)) - g[i]
            S = f32(S + f32(abs(n[i]) * max(abs(lo), abs(hi))))
        R = f32(f32(rp * f32(1 + 1e-4)) + f32(f32(1e-4) * S))
        hw = f32(f32(R / abs(n[c])) + f32(0.125))
        wd = f32(math.ceil(f32(f32(f32(2) * hw) + f32(1.95))))
        if wd > 15:
            continue
        W = int(wd)
        na, nb, nc = n[a], n[b], n[c]
        ka = f32(-na / nc)
        kb = f32(-nb / nc)
        t = f32(f32(na * f32(f32(0.5) - g[a])) + f32(nb * f32(f32(0.5) - g[b])))
        k0 = f32(f32(f32(g[c] - f32(t / nc)) - hw) - f32(0.5))
        windows += 1
        for _ in range(40):
            ca = int(rng.integers(mn[a], mx[a] + 1))
            cb = int(rng.integers(mn[b], mx[b] + 1))
            start = int(math.floor(f32(f32(k0 + f32(ka * f32(ca))) + f32(kb * f32(cb)))))
            cen = np.zeros((ext[c], 3), f32)
            cen[:, a] = ca
            cen[:, b] = cb
            cen[:, c] = np.arange(mn[c], mx[c] + 1)
            v0 = (g[0:3] - (cen + f32(0.5))).astype(f32)
            dd = ((v0[:, 0] * n[0]).astype(f32) + (v0[:, 1] * n[1]).astype(f32)).astype(f32)
            dd = (dd + (v0[:, 2] * n[2]).astype(f32)).astype(f32)
            ok = ~(np.abs(dd) > rp)
            z = np.arange(mn[c], mx[c] + 1)[ok]
            passed += len(z)
            bad += int(np.sum((z < start) | (z > start + W - 1)))
    return windows, passed, bad


def test_every_passing_cell_lies_in_its_window():
    windows, passed, bad = _window_violations(300, seed=7)
    assert windows > 100 and passed > 5000
    assert bad == 0, f"{bad} passing cells outside their window"
