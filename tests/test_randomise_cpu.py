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


def test_batch_for_fills_four_items_per_compute_unit():
    """Permutations per batch: multiples of 64 in [256, 1024] giving four TFCE
    items per compute unit (256 CUs: 256 for 2 surfaces x +/-t, 512 for one
    surface with +/-t or two with +z, 1024 cap for one image per permutation)."""
    from tfce_mediation_amd.tfce import batch_for
    assert batch_for(256, 4) == 256
    assert batch_for(256, 2) == 512
    assert batch_for(256, 1) == 1024
    assert batch_for(256, 100) == 256
    assert batch_for(256, 3) == 384  # 342 rounded up to a multiple of 64
    assert batch_for(80, 1) == 320
    for cus in (64, 80, 256, 304):
        for items in range(1, 40):
            b = batch_for(cus, items)
            assert b % 64 == 0 and 256 <= b <= 1024
            assert b == 1024 or b * items >= 4 * cus
            assert b == 256 or (b - 64) * items < 4 * cus
