"""Host parts of cryo_estimate_shifts (no GPU):

* the shift equations as cryo_estimate_shifts.m:206-222 writes them (the
  second line flipped by its index, libclm ``clm_shift_equations_csc_ex``
  with ``CLM_FLIP_BY_INDEX``) equal the oracle's, for all pairs and for a
  subset of them, at n_theta = 150 where the index rule and cryo_clmatrix's
  angle rule part;
"""
import numpy as np
import pytest
import scipy.sparse as sp

from aspire_amd import _lib
from aspire_amd.clmatrix import shift_equations
from oracle import clm_oracle as co
from oracle import shifts_oracle as so


@pytest.mark.parametrize("subset", [False, True])
def test_equations_by_index_match_oracle(subset):
    K, n_theta = 40, 150
    L = n_theta // 2
    rng = np.random.default_rng(11)
    P = K * (K - 1) // 2
    cl1 = rng.integers(1, L + 1, P).astype(np.int32)
    cl2 = rng.integers(1, 2 * L + 1, P).astype(np.int32)
    cl2[::7] = L + 1  # the index rule flips these, the angle rule does not (L fl(pi/L) < pi)
    cl2[3::7] = L
    shift = rng.integers(-4, 5, P).astype(np.float64)
    k1, k2 = co.all_pairs(K)
    sel = None
    if subset:
        sel = (rng.uniform(size=P) < 0.3).astype(np.uint8)
        keep = sel.astype(bool)
    else:
        keep = np.ones(P, bool)
    eqs = shift_equations(K, n_theta, cl1, cl2, shift, sel, _lib.CLM_FLIP_BY_INDEX)
    A, b = so.system(K, n_theta, cl1[keep], cl2[keep], shift[keep], flip_by_index=True, k1=k1[keep], k2=k2[keep])
    ref = sp.hstack([A, sp.csr_matrix(b.reshape(-1, 1))], format="csc")
    ref.eliminate_zeros()
    assert eqs.shape == ref.shape == (int(keep.sum()), 2 * K + 1)
    assert np.array_equal(eqs.indptr, ref.indptr) and np.array_equal(eqs.indices, ref.indices)
    np.testing.assert_allclose(eqs.data, ref.data, rtol=0, atol=4e-16)
    if subset:
        # the selected pairs' records alone give the same matrix
        c = shift_equations(K, n_theta, cl1[keep], cl2[keep], shift[keep], sel,
                            _lib.CLM_FLIP_BY_INDEX | _lib.CLM_COMPACT_RECORDS)
        assert np.array_equal(c.indptr, eqs.indptr) and np.array_equal(c.indices, eqs.indices)
        assert np.array_equal(c.data, eqs.data)
        with pytest.raises(ValueError):
            shift_equations(K, n_theta, cl1, cl2, shift, sel, _lib.CLM_COMPACT_RECORDS)
    # the angle rule (flags 0) reverses exactly the rows with cl2 = L + 1
    e0 = shift_equations(K, n_theta, cl1, cl2, shift, sel, 0).toarray()[:, :-1]
    e1 = eqs.toarray()[:, :-1]
    rev = cl2[keep] == L + 1
    assert rev.any()
    np.testing.assert_allclose(e0[rev], -e1[rev], rtol=0, atol=4e-16)
    assert np.array_equal(e0[~rev], e1[~rev])
