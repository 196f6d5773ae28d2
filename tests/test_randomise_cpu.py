"""Host-side permutation generation of the randomise drivers (no GPU)."""
import numpy as np

from tfce_mediation_amd import randomise


def test_exchange_blocks_permute_within_blocks():
    from tfce_mediation_amd import randomise
    blocks = np.array(["a"] * 4 + ["b"] * 4 + ["c"] * 4)
    perms = randomise.permutation_indices(12, range(5), seed=1, blocks=blocks)
    for p in perms:
        assert sorted(p) == list(range(12))
        # equal blocks: each output block of 4 is one whole input block
        for j in range(3):
            seg = set(p[4 * j:4 * j + 4] // 4)
            assert len(seg) == 1


def test_permutation_indices_follow_mmr_lr_seeding():
    # tm_func.py:147-149: np.random.seed(perm_number + perm_seed) then a permutation
    p = randomise.permutation_indices(20, [3, 7], seed=100)
    for row, num in zip(p, [3, 7]):
        np.random.seed(num + 100)
        assert list(row) == list(np.random.permutation(list(range(20))))
