"""Pins the oracle's integer layer (Morton codes, Resolution, GpuNode ranks, LeafBounds)
against the reference's known-answer tests and brute force."""
import numpy as np
import oracle as O
from helpers import kats, morton9


def test_morton_known_values():
    for (x, y, z), code in kats()["morton_encode"]["cases"]:
        assert O.morton_encode(x, y, z) == code


def test_morton_round_trip_and_bit_interleave():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        x, y, z = (int(v) for v in rng.integers(0, 1 << 21, 3))
        code = O.morton_encode(x, y, z)
        assert O.morton_decode(code) == (x, y, z)
        # independent definition: bit b of x goes to 3b, of y to 3b+1, of z to 3b+2
        ref = 0
        for b in range(21):
            ref |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
        assert code == ref


def test_morton_monotone_in_x():
    rng = np.random.default_rng(2)
    for _ in range(500):
        x, y, z = (int(v) for v in rng.integers(0, (1 << 21) - 1, 3))
        assert O.morton_encode(x + 1, y, z) > O.morton_encode(x, y, z)


def test_encode_brick_is_a_bijection():
    seen = set()
    for x in range(8):
        for y in range(8):
            for z in range(8):
                i = O.encode_brick(x, y, z)
                assert i == morton9(x, y, z) and i < 512
                seen.add(i)
    assert len(seen) == 512


def test_child_bit_matches_morton():
    for cz in range(4):
        for cy in range(4):
            for cx in range(4):
                assert O.child_bit(cx, cy, cz) == O.morton_encode(cx, cy, cz) < 64


def test_resolution_brackets():
    k = kats()["resolution"]
    for n, kk in k["ok"]:
        assert O.resolution(n) == ("ok", kk)
    for n, below, above in k["not_representable"]:
        assert O.resolution(n) == ("not_representable", below, above)
    for n in k["too_small"]:
        assert O.resolution(n) == ("too_small",)
    assert O.resolution((1 << 31) + 8) == ("too_large",)
    assert O.resolution(1 << 31) == ("ok", 14)


def test_child_slot_known_answers():
    g = kats()["gpu_node"]
    m = int(g["lo_hi"]["mask"], 16)
    assert m & 0xFFFFFFFF == int(g["lo_hi"]["lo"], 16) and m >> 32 == int(g["lo_hi"]["hi"], 16)
    for case in g["child_slot"]:
        mask = sum(1 << b for b in case["bits"])
        for bit, slot in case["slots"]:
            assert O.child_slot(mask, case["base"], bit) == slot
        for bit in case["absent"]:
            assert not (mask >> bit) & 1


def test_child_slot_is_rank_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(300):
        mask = int(rng.integers(0, 1 << 63)) | (int(rng.integers(0, 2)) << 63)
        base = int(rng.integers(0, 1 << 20))
        for bit in range(64):
            assert O.child_slot(mask, base, bit) == base + bin(mask & ((1 << bit) - 1)).count("1")


def _leaf_with(voxels):
    w = [0] * 8
    for (x, y, z) in voxels:
        i = morton9(x, y, z)
        w[i >> 6] |= 1 << (i & 63)
    return np.array(w, np.uint64)


def test_leaf_bounds_known_answers():
    for case in kats()["leaf_bounds"]["cases"]:
        p = O.leaf_bounds(_leaf_with(case["set"]))
        assert O.bounds_unpack(p) == (case["min"], case["max"])


def test_bounds_pack_layout():
    for mn, mx in kats()["bounds_pack_layout"]["cases"]:
        p = O.bounds_pack(mn, mx)
        # the literal traversal.wgsl shift/mask sequence
        assert [p & 7, (p >> 3) & 7, (p >> 6) & 7] == mn
        assert [(p >> 9) & 7, (p >> 12) & 7, (p >> 15) & 7] == mx
        assert (p >> 18) & 1 == 0


def test_leaf_bounds_brute_force_and_full_rule():
    rng = np.random.default_rng(4)
    for _ in range(400):
        cnt = int(rng.integers(1, 100))
        vox = {tuple(int(v) for v in rng.integers(0, 8, 3)) for _ in range(cnt)}
        p = O.leaf_bounds(_leaf_with(vox))
        if len(vox) > 64:  # leaf.rs:188-191 BOUNDS_SCAN_LIMIT
            assert O.bounds_unpack(p) == ([0, 0, 0], [7, 7, 7])
        else:
            a = np.array(sorted(vox))
            assert O.bounds_unpack(p) == (a.min(0).tolist(), a.max(0).tolist())
