"""The autotuner's key groups (host logic, no GPU): every searched key is a
tuning key with candidate values, the fresh-frame and re-inserted-frame
groups do not overlap, and together they cover every candidate key."""
from bonxai_amd import autotune


def test_key_groups_cover_candidates():
    fresh, steady = set(autotune.FRESH_KEYS), set(autotune.STEADY_KEYS)
    assert not fresh & steady
    assert fresh | steady | {"small_ray_blocks_per_cu"} >= set(autotune.CANDIDATES)
    assert (fresh | steady) <= set(autotune.CANDIDATES)
    for k, values in autotune.CANDIDATES.items():
        assert values and len(set(values)) == len(values), k
