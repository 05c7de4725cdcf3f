"""clstack_mask (the fifth output under verbose bit 2 with ref_clmatrix):
libclm's host FP64 clm_clstack_mask against the oracle's restatement of
cryo_clmatrix_cpu.m:567-617, and against what the geometry fixes: lines
equal to the reference match; every line 4 degrees off still matches and 6
degrees off does not (max_angle is 5 degrees); both lines on their
antipodes match, one alone does not; an unset line (0) matches only a
reference line within 5 degrees of index 0."""
import numpy as np
import pytest

from aspire_amd.clmatrix import clstack_mask

T = 360
L = T // 2


def ref_pairs(K, rng):
    ref = np.zeros((K, K))
    k1, k2 = np.triu_indices(K, 1)
    ref[k1, k2] = rng.integers(1, L + 1, k1.size)
    ref[k2, k1] = rng.integers(1, T + 1, k1.size)
    return ref, k1, k2


def wrap(x):
    return (x - 1) % T + 1  # line indices live in [1, T]; index T+1 is index 1


def test_geometry_cases():
    rng = np.random.default_rng(1)
    K = 30
    ref, k1, k2 = ref_pairs(K, rng)
    # each degree is one index at n_theta = 360 (pi / L per index)
    for delta, want in ((0, 1), (4, 1), (-4, 1), (6, 0), (-6, 0), (L, 1)):
        est = ref.copy()
        est[k1, k2] = wrap(ref[k1, k2] + delta)
        est[k2, k1] = wrap(ref[k2, k1] + delta)
        mask, n = clstack_mask(est, ref, T)
        assert n == (k1.size if want else 0), (delta, n)
        assert np.all(mask[k1, k2] == want) and np.all(mask[k2, k1] == 0)
    # only one of the two lines on its antipode: no match
    est = ref.copy()
    est[k1, k2] = wrap(ref[k1, k2] + L)
    mask, n = clstack_mask(est, ref, T)
    assert n == 0


def test_matches_oracle_restatement(oracle):
    rng = np.random.default_rng(2)
    K = 25
    ref, k1, k2 = ref_pairs(K, rng)
    est = ref.copy()
    # offsets around both thresholds, antipodes and unset lines
    est[k1, k2] = wrap(ref[k1, k2] + rng.choice([-6, -5, -4, 0, 4, 5, 6, L, L + 5, L + 6], k1.size))
    est[k2, k1] = wrap(ref[k2, k1] + rng.choice([-6, -5, -4, 0, 4, 5, 6, L, L + 5, L + 6], k1.size))
    est[k1[::7], k2[::7]] = 0
    got, n = clstack_mask(est, ref, T)
    want, m = oracle.clstack_mask(est, ref, T)
    assert n == m and np.array_equal(got, want)
    assert 0 < n < k1.size


def test_rejects_bad_shapes():
    with pytest.raises(ValueError):
        clstack_mask(np.zeros((3, 3)), np.zeros((4, 4)), T)
