"""The random-world sampler's CPU restatement (oracle/sampler_oracle.py) pinned by what
fixes it independently: splitmix64's published outputs, generate_random's value ranges
(lenia.cpp:372-412) and a direct double-precision evaluation of the blob sum
(seed_grid_gaussian lenia.cpp:305-334).  No GPU."""
import math

import numpy as np

from oracle import sampler_oracle as S


def test_draws_are_splitmix64():
    # splitmix64 from state 1234567 (and the next two states, + golden gamma) outputs
    # 6457827717110365317, 3203168211198807973, 9817491932198370423; the draw of
    # (seed, set, slot) is the output for state seed*G + set*256 + slot, so set 4822,
    # slot 135 (4822*256 + 135 = 1234567) with seeds 0, 1, 2 walks that sequence.
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423]
    assert [int(S.u32(seed, 4822, 135)) for seed in range(3)] == [w >> 32 for w in want]


def test_generate_random_ranges():
    sets = S.sample_random(7, 0, 3000, 3)
    nbs, counts, betas, c_used = set(), set(), set(), set()
    for ks in sets:
        counts.add(len(ks))
        assert 1 <= len(ks) <= 10
        nb = ks[0]["radius"]
        nbs.add(nb)
        for i, k in enumerate(ks):
            assert k["radius"] == nb and 1 <= nb <= 8
            betas.add(k["beta_size"])
            assert len(k["b"]) == k["beta_size"]
            for b in k["b"]:
                assert round(b * 100) in range(100) and np.float32(b) == np.float32(round(b * 100)) / np.float32(100)
            assert 0.1 <= k["m"] < 0.6
            extra = k["s"] - k["m"] / 10
            assert -1e-6 <= extra < 0.2
            assert k["h"] == 1.0 and k["r_a"] == 1.0
            if i == 0:
                assert k["c0"] == 0 and k["c1"] == 0
            c_used.add((k["c0"], k["c1"]))
    assert nbs == set(range(1, 9)) and counts == set(range(1, 11)) and betas == {1, 2}
    assert c_used == {(a, b) for a in range(3) for b in range(3)}
    # a single channel never draws a channel index
    assert all(k["c0"] == 0 and k["c1"] == 0 for ks in S.sample_random(7, 0, 200, 1) for k in ks)


def test_seed_state_matches_direct_evaluation():
    shape, C = (16, 12, 20), 2
    for sset in (3, 11, 40):
        a = S.seed_state(5, sset, shape, C)
        W, H, D = shape
        N = W * H * D
        assert a.shape == (C * N,) and np.array_equal(a[:N], a[N:])
        r = 1 + int(S.u32(5, sset, 0) % 8)
        cen = [(int(S.u32(5, sset, 128 + 3 * b) % W), int(S.u32(5, sset, 129 + 3 * b) % H),
                int(S.u32(5, sset, 130 + 3 * b) % D)) for b in range(5)]
        rng = np.random.default_rng(sset)
        for i in list(rng.integers(0, N, 200)) + [cx + W * (cy + H * cz) for cx, cy, cz in cen]:
            x, y, z = i % W, (i // W) % H, i // (W * H)
            ref = sum(math.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (2 * r * r)) for cx, cy, cz in cen)
            assert abs(float(a[i]) - ref) <= 4e-7 * max(ref, 1e-30) + 1e-38, (sset, i, float(a[i]), ref)
        for cx, cy, cz in cen:  # a blob centre holds exp(0) = 1 plus the others
            assert a[cx + W * (cy + H * cz)] >= 1.0
