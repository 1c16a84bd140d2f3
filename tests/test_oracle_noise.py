"""Oracle noise fixtures (voxel-core noise.rs, fixtures.rs NoiseField) in f64: Ken Perlin's
improved-noise gradient table, the quintic fade, zero on the lattice, and the reference's
own tests (deterministic, seed-sensitive, bounded, preset fill fractions)."""
import numpy as np
import oracle as O


def test_perlin_gradient_table():
    """Perlin, "Improving Noise" (SIGGRAPH 2002), reference grad(): the 12 edge directions
    of the cube plus 4 repeats -- h: (u, v) = x+y, -x+y, x-y, -x-y, x+z, -x+z, x-z, -x-z,
    y+z, -y+z, y-z, -y-z, y+x, -y+z, y-x, -y-z."""
    x, y, z = 0.75, -1.5, 2.25
    want = [x + y, -x + y, x - y, -x - y, x + z, -x + z, x - z, -x - z,
            y + z, -y + z, y - z, -y - z, y + x, -y + z, y - x, -y - z]
    assert [O.noise_grad(h, x, y, z) for h in range(16)] == want
    assert O.noise_grad(16 + 5, x, y, z) == want[5]  # only the low 4 bits select


def test_fade_and_lattice():
    assert O.noise_fade(0.0) == 0.0 and O.noise_fade(1.0) == 1.0 and O.noise_fade(0.5) == 0.5
    # noise.rs:233-237: gradient noise vanishes on the integer lattice
    for p in ((3.0, -2.0, 5.0), (0.0, 0.0, 0.0), (-7.0, 11.0, 1.0)):
        assert abs(O.perlin3(99, *p)) < 1e-12


def test_reference_noise_properties():
    """noise.rs:178-230 and fixtures.rs:383-430."""
    assert O.perlin3(42, 3.25, -7.5, 12.125) != O.perlin3(43, 3.25, -7.5, 12.125)
    v = np.array([O.perlin3(7, xi * 0.37, yi * 0.37, zi * 0.37)
                  for zi in range(20) for yi in range(20) for xi in range(20)])
    assert v.min() > -1.2 and v.max() < 1.2 and abs(v.mean()) < 0.15
    assert abs(O.perlin3(1, 4.2, 1.1, -3.7) - O.perlin3(1, 4.2 + 1e-4, 1.1, -3.7)) < 1e-2
    _, cp = O.noise_words(O.NOISE_PERLIN, 128)
    _, cc = O.noise_words(O.NOISE_CAVES, 128)
    assert 0.15 <= cp / 128 ** 3 < 0.40 and 0.12 <= cc / 128 ** 3 < 0.35
    a, _ = O.noise_words(O.NOISE_PERLIN, 32)
    b, _ = O.noise_words(dict(O.NOISE_PERLIN, seed=O.NOISE_PERLIN["seed"] ^ 0xABCD), 32)
    assert int(np.unpackbits((a ^ b).view(np.uint8)).sum()) > 32 ** 3 // 20
