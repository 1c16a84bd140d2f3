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


# ---- School-B positions by closed form --------------------------------------------------------
def _popcount64(a):
    a = np.asarray(a, np.uint64)
    c = np.zeros(a.shape, np.int64)
    for b in range(64):
        c += ((a >> np.uint64(b)) & np.uint64(1)).astype(np.int64)
    return c


def school_b_from_levels(k, levels):
    """The School-B node array of School-A levels (levels[L]: N_L x 3 u32 rows mask_lo,
    mask_hi, child_base for L = 2..k+1), by arithmetic instead of a recursive walk.

    Parent-first depth-first order (school_b.rs:68-98) as closed forms: below(v), the node
    entries a node's subtree adds under it, bottom-up (0 at level 2, else its child count plus
    its children's below); then top-down, the root at 0 with its child block at 1, a node's
    j-th child at block(v) + j, and that child's own block after v's whole block and the
    subtrees of its earlier siblings: block(c_j) = block(v) + m(v) + sum_{i<j} below(c_i).
    Level-2 nodes keep their School-A leaf base."""
    top = k + 1
    mask = {L: levels[L][:, 0].astype(np.uint64) | (levels[L][:, 1].astype(np.uint64) << np.uint64(32))
            for L in range(2, top + 1)}
    m = {L: _popcount64(mask[L]) for L in mask}
    cb = {L: levels[L][:, 2].astype(np.int64) for L in mask}
    below = {2: np.zeros(len(levels[2]), np.int64)}
    prefix = {}
    for L in range(3, top + 1):
        prefix[L - 1] = np.concatenate([[0], np.cumsum(below[L - 1])])
        below[L] = m[L] + prefix[L - 1][cb[L] + m[L]] - prefix[L - 1][cb[L]]
    assert len(levels[top]) == 1
    pos = {top: np.array([0])}
    block = {top: np.array([1])}
    for L in range(top, 2, -1):
        nchild = len(levels[L - 1])
        c = np.arange(nchild)
        parent = np.searchsorted(cb[L], c, side="right") - 1
        j = c - cb[L][parent]
        pos[L - 1] = block[L][parent] + j
        block[L - 1] = (block[L][parent] + m[L][parent] +
                        prefix[L - 1][c] - prefix[L - 1][cb[L][parent]])
    total = sum(len(levels[L]) for L in range(2, top + 1))
    out = np.zeros((total, 3), np.uint32)
    for L in range(2, top + 1):
        out[pos[L], 0] = levels[L][:, 0]
        out[pos[L], 1] = levels[L][:, 1]
        out[pos[L], 2] = levels[L][:, 2] if L == 2 else block[L]
    return out


def voxels_to_words(vox, n):
    """x-major u32 occupancy words of a voxel array (N x 3), vectorised."""
    v = np.asarray(vox, np.uint64)
    lin = v[:, 0] + np.uint64(n) * (v[:, 1] + np.uint64(n) * v[:, 2])
    words = np.zeros((n ** 3 + 31) // 32, np.uint32)
    np.bitwise_or.at(words, (lin >> np.uint64(5)).astype(np.int64),
                     (np.uint32(1) << (lin & np.uint64(31)).astype(np.uint32)))
    return words


# ---- exact ties in the walker's step: expected hits derived by hand ----------------------------
# Walker::step (mirror.rs:87-116) takes the axis of the smallest t_max, the LOWEST axis on a tie,
# and a frame whose chosen axis leaves it is exhausted (the parent steps next, with the same
# rule). Rays start at voxel centres with directions of 0/+-1 components, so every t_max is a
# half-integer computed exactly in binary32 and binary64 and ties are exact. Each case: grid n,
# occupied voxels, ray (origin + direction), expected hit voxel (None: no hit). Derivations:
#  * (6.5,6.5,2.5) + (1,1,0): cells (6,6,2) (7,6,2) (7,7,2); at t = 1.5 the leaf frame (x first)
#    and the root frame (x first again, cells of 8) both tie, so the ray enters brick (1,0,0) at
#    (8,8,2.5), clamped into it as (8,7,2); (7,8,2) is never visited. A highest-axis rule would
#    enter brick (0,1,0) at (7,8,2) instead.
#  * (2.5,6.5,6.5) + (0,1,1): y before z: (2,7,6) (2,7,7) (2,8,7); (2,7,8) never.
#  * (6.5,6.5,6.5) + (1,1,1): (7,6,6) (7,7,6) (7,7,7) (8,7,7); (6,7,6) (6,6,7) (7,8,7) (7,7,8) never.
#  * (9.5,6.5,2.5) + (-1,1,0): (8,6,2) (8,7,2); at t = 1.5 x first leaves brick (1,0,0) for
#    (0,0,0), entered at (8,8) clamped to (7,7,2); then y leaves that brick for (0,1,0) at
#    (7,8,2); (8,8,2) never.
#  * n = 128, (30.5,30.5,2.5) + (1,1,0): the tie at (32,32) is resolved x first in the leaf, the
#    level-2 frame (cells of 8) and the root (cells of 32): (31,31,2) then (32,31,2); (31,32,2)
#    never.
TIE_CASES = [
    (32, [(8, 7, 2)], (6.5, 6.5, 2.5, 1, 1, 0), (8, 7, 2)),
    (32, [(7, 8, 2)], (6.5, 6.5, 2.5, 1, 1, 0), None),
    (32, [(7, 8, 2), (8, 7, 2)], (6.5, 6.5, 2.5, 1, 1, 0), (8, 7, 2)),
    (32, [(7, 6, 2)], (6.5, 6.5, 2.5, 1, 1, 0), (7, 6, 2)),
    (32, [(6, 7, 2)], (6.5, 6.5, 2.5, 1, 1, 0), None),
    (32, [(2, 8, 7)], (2.5, 6.5, 6.5, 0, 1, 1), (2, 8, 7)),
    (32, [(2, 7, 8)], (2.5, 6.5, 6.5, 0, 1, 1), None),
    (32, [(7, 6, 6)], (6.5, 6.5, 6.5, 1, 1, 1), (7, 6, 6)),
    (32, [(6, 7, 6), (6, 6, 7)], (6.5, 6.5, 6.5, 1, 1, 1), None),
    (32, [(8, 7, 7)], (6.5, 6.5, 6.5, 1, 1, 1), (8, 7, 7)),
    (32, [(7, 8, 7), (7, 7, 8)], (6.5, 6.5, 6.5, 1, 1, 1), None),
    (32, [(7, 7, 2)], (9.5, 6.5, 2.5, -1, 1, 0), (7, 7, 2)),
    (32, [(8, 8, 2)], (9.5, 6.5, 2.5, -1, 1, 0), None),
    (32, [(7, 8, 2), (8, 8, 2)], (9.5, 6.5, 2.5, -1, 1, 0), (7, 8, 2)),
    (128, [(32, 31, 2)], (30.5, 30.5, 2.5, 1, 1, 0), (32, 31, 2)),
    (128, [(31, 32, 2)], (30.5, 30.5, 2.5, 1, 1, 0), None),
]
