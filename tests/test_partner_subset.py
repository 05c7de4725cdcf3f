"""Partner subset (cryo_clmatrix_gpu_1.m:150-151,346-350) and the outputs of
a subset search (:809-827,853-858), host side: the draw's counts, order and
uniformity, and the subset equations against the full equations' rows."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

from aspire_amd import _lib
from aspire_amd.clmatrix import assemble, partner_subset


def rows_of(sel, K):
    """Per k1, the selected partners k2 (0-based), from the equation-order bytes."""
    out, e = [], 0
    for k1 in range(K - 1):
        m = K - 1 - k1
        out.append([k1 + 1 + i for i in range(m) if sel[e + i]])
        e += m
    return out


@pytest.mark.parametrize("K,NK", [(1, 3), (2, 1), (9, 3), (30, 7), (30, 29), (30, 40), (17, 0), (17, -1)])
def test_counts_per_image(K, NK):
    sel, n = partner_subset(K, NK, seed=11)
    assert sel.shape == (max(K * (K - 1) // 2, 0),)
    assert set(np.unique(sel)) <= {0, 1}
    per = rows_of(sel, K)
    want = [K - 1 - k1 if NK < 0 else min(K - 1 - k1, NK) for k1 in range(K - 1)]
    assert [len(r) for r in per] == want
    assert n == sum(want) == int(sel.sum())


def test_seed_determines_draw():
    a, _ = partner_subset(60, 5, seed=1)
    b, _ = partner_subset(60, 5, seed=1)
    c, _ = partner_subset(60, 5, seed=2)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)


def test_rejects_bad_nk():
    with pytest.raises(_lib.ClmError):
        partner_subset(10, -2)


def test_every_subset_equally_likely():
    # row k1 = 0 of K = 7 chooses 2 of 6 later images: 15 subsets, each with
    # probability 1/15 (randperm's first n2 entries, sorted).  Chi-square over
    # 15,000 seeds (14 degrees of freedom; 36.1 is the 0.1% point).
    K, NK, n = 7, 2, 15000
    counts = {c: 0 for c in itertools.combinations(range(1, 7), 2)}
    for s in range(n):
        sel, _ = partner_subset(K, NK, seed=s)
        counts[tuple(np.flatnonzero(sel[:6]) + 1)] += 1
    exp = n / len(counts)
    chi2 = sum((v - exp) ** 2 / exp for v in counts.values())
    assert chi2 < 36.1, counts


def _records(K, seed, L):
    rng = np.random.default_rng(seed)
    P = K * (K - 1) // 2
    rec = dict(cl1=rng.integers(1, L + 1, P).astype(np.int32),
               cl2=rng.integers(1, 2 * L + 1, P).astype(np.int32),
               j=rng.integers(0, 5, P).astype(np.int32),
               corr=rng.uniform(-0.2, 0.9, P),
               shift=rng.choice([0.0, -1.5, 2.0, 0.5], P))
    return rec


@pytest.mark.parametrize("K,NK", [(9, 3), (25, 4), (25, -1), (6, 0)])
def test_subset_outputs_are_full_outputs_rows(K, NK):
    n_theta = 24
    rec = _records(K, 5 + K, n_theta // 2)
    sel, nsel = partner_subset(K, NK, seed=K)
    # a subset search leaves the unselected pairs at zero
    for k in rec:
        rec[k] = np.where(sel != 0, rec[k], 0 if k != "j" else -1).astype(rec[k].dtype)
    cs_f, co_f, eq_f, map_f = assemble(K, n_theta, rec)
    cs_s, co_s, eq_s, map_s = assemble(K, n_theta, rec, sel=sel)
    assert np.array_equal(cs_s, cs_f) and np.array_equal(co_s, co_f)
    assert eq_s.shape == (nsel, 2 * K + 1)
    rows = np.flatnonzero(sel)
    assert (eq_s != eq_f.tocsr()[rows]).nnz == 0
    assert eq_s.has_sorted_indices or sp.csc_matrix(eq_s).has_sorted_indices
    # the map numbers the selected pairs 1..nsel in pair order, 0 elsewhere
    want = np.zeros((K, K))
    iu = np.triu_indices(K, 1)
    num = np.zeros(K * (K - 1) // 2)
    num[rows] = np.arange(1, nsel + 1)
    want[iu] = num
    assert np.array_equal(map_s, want)
    assert np.array_equal(map_f[iu], np.arange(1, K * (K - 1) // 2 + 1))


def test_subset_equations_match_restated_loop():
    # the reference's loop (cryo_clmatrix_gpu_1.m:809-827), small K, pure Python
    K, n_theta = 7, 16
    L = n_theta // 2
    rec = _records(K, 99, L)
    sel, nsel = partner_subset(K, 3, seed=4)
    _, _, eq, _ = assemble(K, n_theta, rec, sel=sel)
    dtheta = np.pi / L
    want = np.zeros((nsel, 2 * K + 1))
    e = row = 0
    for k1 in range(K):
        for k2 in range(k1 + 1, K):
            if sel[e]:
                a = (rec["cl1"][e] - 1) * dtheta
                b = (rec["cl2"][e] - 1) * dtheta
                if b < np.pi:
                    c = [np.sin(a), np.cos(a), -np.sin(b), -np.cos(b)]
                else:
                    b -= np.pi
                    c = [-np.sin(a), -np.cos(a), -np.sin(b), -np.cos(b)]
                want[row, [2 * k1, 2 * k1 + 1, 2 * k2, 2 * k2 + 1]] = c
                want[row, 2 * K] = rec["shift"][e]
                row += 1
            e += 1
    assert np.allclose(eq.toarray(), want, rtol=0, atol=1e-15)
