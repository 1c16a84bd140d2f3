"""GPU parity of the noise fixtures (generate_noise_tree over NoiseField): the device tree
is the oracle's SparseTree::build of the f64 CPU field, bit for bit."""
import numpy as np
import oracle as O
import pytest
import vxl

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[case] {msg}", flush=True)


FIELDS = [("perlin", O.NOISE_PERLIN), ("caves", O.NOISE_CAVES),
          ("custom", dict(seed=0x1234, frequency=9.5, octaves=3, lacunarity=2.5, gain=0.6,
                          ridged=True, warp=1.25, threshold=0.05))]


@pytest.mark.parametrize("field", FIELDS, ids=lambda f: f[0])
@pytest.mark.parametrize("n", [32, 128])
def test_noise_tree_matches_oracle(ctx, field, n):
    name, f = field
    _say(f"noise {name} n={n}")
    words, count = O.noise_words(f, n)
    ot = O.Tree.build(words, n)
    gt = ctx.generate_noise_tree(n, f)
    gn, gl, gb, gc = gt.download()
    on, ol, ob, oc = ot.export()
    assert np.array_equal(gc, oc) and np.array_equal(gl, ol)
    assert np.array_equal(gb, ob) and np.array_equal(gn, on)
    assert int(np.unpackbits(gl.view(np.uint8)).sum()) == count


def test_noise_presets_match_binding_presets(ctx):
    assert vxl.NOISE_PERLIN == O.NOISE_PERLIN and vxl.NOISE_CAVES == O.NOISE_CAVES
    with pytest.raises(vxl.VxlError):
        ctx.generate_noise_tree(64, O.NOISE_PERLIN)  # not 8 * 4^k
