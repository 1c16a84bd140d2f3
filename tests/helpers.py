"""Independent checks used to pin the oracle: none of these share code or formulas
with the oracle's restatement of the reference."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def kats():
    with open(os.path.join(GOLDEN, "reference_kats.json")) as f:
        return json.load(f)


# ---- geometry: triangle / box overlap by polygon clipping in f64 ---------------------------
def clip_triangle_box(tri, lo, hi):
    """Sutherland-Hodgman clip of a triangle (3x3) against the box [lo, hi]; returns the
    clipped polygon (possibly empty). Exact-enough f64 reference for non-tangent cases."""
    poly = [np.asarray(v, np.float64) for v in tri]
    for axis in range(3):
        for bound, keep_ge in ((lo[axis], True), (hi[axis], False)):
            out = []
            m = len(poly)
            for i in range(m):
                a, b = poly[i], poly[(i + 1) % m]
                da = a[axis] - bound if keep_ge else bound - a[axis]
                db = b[axis] - bound if keep_ge else bound - b[axis]
                if da >= 0:
                    out.append(a)
                if (da >= 0) != (db >= 0):
                    t = da / (da - db)
                    out.append(a + t * (b - a))
            poly = out
            if not poly:
                return poly
    return poly


def overlaps_with_margin(tri, center, half, delta):
    """(inner, outer): clip against the box shrunk by delta / grown by delta."""
    c = np.asarray(center, np.float64)
    inner = len(clip_triangle_box(tri, c - (half - delta), c + (half - delta))) > 0
    outer = len(clip_triangle_box(tri, c - (half + delta), c + (half + delta))) > 0
    return inner, outer


# ---- School-B reading without the oracle ----------------------------------------------------
def morton9(x, y, z):
    r = 0
    for b in range(3):
        r |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
    return r


def school_b_occupied(n, nodes, leaves):
    """Set of occupied voxels by descending every School-B node (children addressed by
    popcount rank from child_base; level-2 child_base indexes the leaf array)."""
    k = 0
    while 8 * 4 ** k < n:
        k += 1
    out = set()
    if len(leaves) == 0:
        return out

    def emit_leaf(li, org):
        words = [int(w) for w in leaves[li]]
        for x in range(8):
            for y in range(8):
                for z in range(8):
                    m = morton9(x, y, z)
                    if (words[m >> 6] >> (m & 63)) & 1:
                        out.add((org[0] + x, org[1] + y, org[2] + z))

    def walk(idx, level, org):
        lo, hi, base = (int(v) for v in nodes[idx])
        mask = lo | (hi << 32)
        size = 1 << (2 * (level - 1) + 1)
        rank = 0
        for bit in range(64):
            if not (mask >> bit) & 1:
                continue
            cx = (bit & 1) | ((bit >> 3) & 1) << 1
            cy = ((bit >> 1) & 1) | ((bit >> 4) & 1) << 1
            cz = ((bit >> 2) & 1) | ((bit >> 5) & 1) << 1
            corg = (org[0] + cx * size, org[1] + cy * size, org[2] + cz * size)
            if level == 2:
                emit_leaf(base + rank, corg)
            else:
                walk(base + rank, level - 1, corg)
            rank += 1

    if k == 0:
        emit_leaf(0, (0, 0, 0))
    else:
        walk(0, k + 1, (0, 0, 0))
    return out


def words_to_set(words, n):
    w = np.asarray(words, np.uint32)
    bits = np.unpackbits(w.view(np.uint8), bitorder="little")[: n ** 3]
    idx = np.nonzero(bits)[0]
    x = idx % n
    y = (idx // n) % n
    z = idx // (n * n)
    return set(zip(x.tolist(), y.tolist(), z.tolist()))


def set_to_words(voxels, n):
    words = np.zeros((n ** 3 + 31) // 32, np.uint32)
    for (x, y, z) in voxels:
        i = x + n * (y + n * z)
        words[i >> 5] |= np.uint32(1 << (i & 31))
    return words


def single_voxel_school_b(n, v):
    """School-B arrays of a tree holding exactly voxel v, built by hand: one node per level
    on the root-to-leaf path, each with a single child bit."""
    k = 0
    while 8 * 4 ** k < n:
        k += 1
    bx, by, bz = v[0] >> 3, v[1] >> 3, v[2] >> 3
    nodes = []
    for i, level in enumerate(range(k + 1, 1, -1)):
        sh = 2 * (level - 2)
        cx, cy, cz = (bx >> sh) & 3, (by >> sh) & 3, (bz >> sh) & 3
        bit = (cx & 1) | (cy & 1) << 1 | (cz & 1) << 2 | (cx & 2) << 2 | (cy & 2) << 3 | (cz & 2) << 4
        mask = 1 << bit
        base = 0 if level == 2 else i + 1
        nodes.append([mask & 0xFFFFFFFF, mask >> 32, base])
    leaf = [0] * 8
    m = morton9(v[0] & 7, v[1] & 7, v[2] & 7)
    leaf[m >> 6] |= 1 << (m & 63)
    lx, ly, lz = v[0] & 7, v[1] & 7, v[2] & 7
    bounds = lx | ly << 3 | lz << 6 | lx << 9 | ly << 12 | lz << 15
    return (np.array(nodes, np.uint32).reshape(-1, 3), np.array([leaf], np.uint64),
            np.array([bounds], np.uint32))


# ---- brute-force first hit (oracle.rs:147-174 style, written independently) ---------------
def brute_force_first_hit(voxels, ray):
    o = np.asarray(ray[:3], np.float64)
    d = np.asarray(ray[3:], np.float64)
    best = None
    for v in voxels:
        lo = np.asarray(v, np.float64)
        hi = lo + 1.0
        tn, tf = -np.inf, np.inf
        ok = True
        for a in range(3):
            if d[a] == 0.0:
                if o[a] < lo[a] or o[a] > hi[a]:
                    ok = False
                    break
            else:
                t1, t2 = (lo[a] - o[a]) / d[a], (hi[a] - o[a]) / d[a]
                tn, tf = max(tn, min(t1, t2)), min(tf, max(t1, t2))
        if ok and tn <= tf and tf >= 0:
            t = max(tn, 0.0)
            if best is None or t < best[1]:
                best = (tuple(v), t)
    return best
