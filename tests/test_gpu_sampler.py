"""The random-world sampler on the GPU (lh_sample_random, lh_seed_random) against its CPU
restatement: parameter sets and seeded states bit for bit; a sampled world steps.

Bar: bit-identical kernels (every field, float bits) and A; GA, S, V cleared."""
import numpy as np
import pytest

from lenia_amd import workloads as WL
from lenia_amd.engine import Engine, sample_random
from oracle import sampler_oracle as S

pytestmark = pytest.mark.gpu


def _bits(x):
    return np.float32(x).view(np.uint32)


@pytest.mark.parametrize("channels", [1, 3, 4])
def test_sampled_sets_match_restatement(channels):
    got = sample_random(99, 1_000_000, 1500, channels)
    ref = S.sample_random(99, 1_000_000, 1500, channels)
    assert len(got) == len(ref)
    for j, (g, r) in enumerate(zip(got, ref)):
        assert len(g) == len(r), j
        for a, b in zip(g, r):
            for f in ("c0", "c1", "radius", "beta_size"):
                assert a[f] == b[f], (j, f)
            for f in ("m", "s", "h", "r_a"):
                assert _bits(a[f]) == _bits(b[f]), (j, f)
            assert [_bits(v) for v in a["b"]] == [_bits(v) for v in b["b"]], j


def test_seeded_worlds_match_restatement_and_step():
    shapes = [(32, 32, 32), (24, 24, 24), (44, 24, 44)]
    seed, first_set = 1234, 500
    sets = S.sample_random(seed, first_set, len(shapes), 3)
    eng = Engine()
    for shape, ks in zip(shapes, sets):
        eng.add_world(shape, 3, ks, WL.base_cfg())
    eng.commit()
    for i in range(len(shapes)):
        eng.set(i, "GA", np.ones(3 * np.prod(shapes[i]), np.float32))  # cleared by the seeding
    eng.step(3)
    eng.seed_random(0, len(shapes), seed, first_set)
    for i, shape in enumerate(shapes):
        N = int(np.prod(shape))
        a = eng.get(i, "A")
        assert np.array_equal(a.view(np.uint32), S.seed_state(seed, first_set + i, shape, 3).view(np.uint32)), i
        assert not eng.get(i, "GA").any() and not eng.get(i, "S").any() and not eng.get(i, "V").any()
        assert eng.steps(i) == 0 and a.size == 3 * N
    eng.step(4)
    eng.synchronize()
    assert np.isfinite(eng.mass()).all()
    eng.close()
