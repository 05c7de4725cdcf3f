"""The sync3n oracle (oracle/sync3n_oracle.c), pinned without the reference
(whose MEX kernels need MATLAB's headers):

* geometry: relative rotations of true rotations are consistent, so every
  triangle's best match is exact (c = 0) and every score is 1: each pair's
  cumulated score is 2(N-2) and the histogram holds all 3 C(N,3) scores in
  its last bin; conjugating one side by J makes the configuration that
  undoes it the best one (1: Rij, 2: Rjk, 3: Rik);
* the mismatch values against numpy's matrix products and norms, and the
  alternatives against the configurations whose relative sign differs
  (derived here from the sign table, not retyped);
* the histogram against floor binning on noisy data;
* the per-pair sums against a per-pair brute force over the third image;
* the signs product against an explicitly built N-choose-2 square matrix,
  and on consistent data with J flips D, v = D 1 is an eigenvector with
  eigenvalue 2(N-2) (the reference's theoretical top eigenvalue 2N-4);
* the threaded split (the reference's own) agrees with one thread.
"""
import itertools

import numpy as np
import pytest

from aspire_amd import sim

J = np.diag([1.0, 1.0, -1.0])
# sign of each side pair ((ij,jk), (ik,jk), (ij,ik)) in configurations 0-3
# (0: none, 1: Rij, 2: Rjk, 3: Rik conjugated by J): a pair of sides changes
# relative sign when exactly one of its two sides is conjugated
SIDES = [(0, 1), (2, 1), (0, 2)]  # side indices (0 ij, 1 jk, 2 ik) of each side pair
FLIPPED = [None, 0, 1, 2]


def sign_table():
    t = np.ones((4, 3))
    for cfg, f in enumerate(FLIPPED):
        for q, (x, y) in enumerate(SIDES):
            if f is not None and (f == x) != (f == y):
                t[cfg, q] = -1
    return t


def pidx(N, i, j):
    return (2 * N - i - 1) * i // 2 + j - i - 1


def rel_rotations(rots, rng=None, flip=0.0, outliers=0.0):
    """R_pairs (3 x 3 x N-choose-2): R_ij = R_i^T R_j, some conjugated by J,
    some replaced by random rotations."""
    N = rots.shape[2]
    P = N * (N - 1) // 2
    R = np.zeros((3, 3, P))
    flips = np.zeros(P, bool)
    for i in range(N):
        for j in range(i + 1, N):
            e = pidx(N, i, j)
            M = rots[:, :, i].T @ rots[:, :, j]
            if rng is not None and rng.random() < outliers:
                M = sim.rand_rots(rng, 1)[:, :, 0]
            if rng is not None and rng.random() < flip:
                M = J @ M @ J
                flips[e] = True
            R[:, :, e] = M
    return R, flips


def test_consistent_triangles(oracle):
    rng = np.random.default_rng(0)
    N = 12
    R, _ = rel_rotations(sim.rand_rots(rng, N))
    cum, hist = oracle.triangles_scores(R, 100)
    assert np.allclose(cum, 2 * (N - 2), atol=1e-6)
    assert hist[-1] == 3 * N * (N - 1) * (N - 2) // 6 and hist[:-1].sum() == 0


@pytest.mark.parametrize("side", [0, 1, 2])
def test_flipped_side_selects_configuration(oracle, side):
    rng = np.random.default_rng(1 + side)
    rots = sim.rand_rots(rng, 3)
    Rij, Rjk, Rik = rots[:, :, 0].T @ rots[:, :, 1], rots[:, :, 1].T @ rots[:, :, 2], rots[:, :, 0].T @ rots[:, :, 2]
    mats = [Rij, Rjk, Rik]
    mats[side] = J @ mats[side] @ J
    b, c, s = oracle.triangle(*mats)
    assert b == side + 1
    assert c[b] < 1e-25 and np.all(np.delete(c, b) > 1e-6)


def test_mismatches_and_alternatives(oracle):
    rng = np.random.default_rng(3)
    sg = sign_table()
    for _ in range(200):
        Rij, Rjk, Rik = (sim.rand_rots(rng, 1)[:, :, 0] for _ in range(3))
        b, c, s = oracle.triangle(Rij, Rjk, Rik)
        want = np.array([np.sum((Rij @ Rjk - Rik) ** 2), np.sum((J @ Rij @ J @ Rjk - Rik) ** 2),
                         np.sum((Rij @ J @ Rjk @ J - Rik) ** 2), np.sum((Rij @ Rjk - J @ Rik @ J) ** 2)])
        assert np.allclose(c, want, rtol=1e-12, atol=1e-14)
        assert b == int(np.argmin(c))
        for q in range(3):
            alts = [cfg for cfg in range(4) if sg[cfg, q] != sg[b, q]]
            assert len(alts) == 2
            assert s[q] == pytest.approx(1 - np.sqrt(c[b] / min(c[alts[0]], c[alts[1]])), rel=1e-12, abs=1e-15)


def noisy(N, seed, flip=0.3, outliers=0.3):
    rng = np.random.default_rng(seed)
    return rel_rotations(sim.rand_rots(rng, N), rng, flip, outliers)


def test_histogram_and_sums_brute_force(oracle):
    N = 11
    R, _ = noisy(N, 5)
    cum, hist = oracle.triangles_scores(R, 20)
    scores = []
    per_pair = np.zeros(N * (N - 1) // 2)
    for i, j, k in itertools.combinations(range(N), 3):
        ij, jk, ik = pidx(N, i, j), pidx(N, j, k), pidx(N, i, k)
        _, _, s = oracle.triangle(R[:, :, ij], R[:, :, jk], R[:, :, ik])
        scores += list(s)
        per_pair[ij] += s[0] + s[2]
        per_pair[jk] += s[0] + s[1]
        per_pair[ik] += s[1] + s[2]
    scores = np.array(scores)
    bins = np.minimum(np.floor(scores * 20).astype(int), 19)
    assert np.array_equal(hist, np.bincount(bins, minlength=20).astype(float))
    assert np.allclose(cum, per_pair, rtol=1e-13)


def test_probabilities_brute_force(oracle):
    N = 9
    R, _ = noisy(N, 6)
    prm = (0.4, 2.0, 2.2, 3.0, 2.5, 0.78)
    P2, A, a, B, b, x0 = prm
    ind, arb = oracle.pairs_probabilities(R, *prm)
    want_i, want_a = np.zeros_like(ind), np.zeros_like(arb)
    for p, q in itertools.combinations(range(N), 2):
        for m in range(N):
            if m in (p, q):
                continue
            i, j, k = sorted((p, q, m))
            _, _, s = oracle.triangle(R[:, :, pidx(N, i, j)], R[:, :, pidx(N, j, k)], R[:, :, pidx(N, i, k)])
            role = [x for x, (u, v) in enumerate([(i, j), (j, k), (i, k)]) if {u, v} == {p, q}][0]
            mine = [q_ for q_, (x, y) in enumerate(SIDES) if role in (x, y)]
            for qq in mine:
                u = 1 - s[qq]
                want_i[pidx(N, p, q)] += np.log(P2 * (B * u ** b * np.exp(-b / (1 - x0) * u)) + (1 - P2) * A * u ** a)
                want_a[pidx(N, p, q)] += np.log(A * u ** a)
    assert np.allclose(ind, want_i, rtol=1e-12)
    assert np.allclose(arb, want_a, rtol=1e-12)


def explicit_signs_matrix(oracle, R, N, entries, pair_scores=None):
    P = N * (N - 1) // 2
    S = np.zeros((P, P))
    sg = sign_table()
    for i, j, k in itertools.combinations(range(N), 3):
        ij, jk, ik = pidx(N, i, j), pidx(N, j, k), pidx(N, i, k)
        b, _, s = oracle.triangle(R[:, :, ij], R[:, :, jk], R[:, :, ik])
        if entries == 0:
            e = sg[b]
        elif entries == 1:
            e = sg[b] * s
        else:
            e = sg[b] * pair_scores[ij] * pair_scores[jk] * pair_scores[ik]
        for q, (x, y) in enumerate(SIDES):
            u, v = (ij, jk, ik)[x], (ij, jk, ik)[y]
            S[u, v] += e[q]
            S[v, u] += e[q]
    return S


@pytest.mark.parametrize("entries", [0, 1, 2])
def test_signs_product_explicit_matrix(oracle, entries):
    N = 8
    R, _ = noisy(N, 7)
    P = N * (N - 1) // 2
    rng = np.random.default_rng(8)
    ps = rng.uniform(0.2, 1.0, P)
    v = rng.standard_normal((P, 3))
    got = oracle.signs_times_v(R, v, entries, ps)
    S = explicit_signs_matrix(oracle, R, N, entries, ps)
    assert np.allclose(got, S @ v, rtol=1e-12, atol=1e-12)


def test_signs_eigenvector_on_consistent_data(oracle):
    rng = np.random.default_rng(9)
    N = 10
    R, flips = rel_rotations(sim.rand_rots(rng, N), rng, flip=0.4, outliers=0.0)
    d = np.where(flips, -1.0, 1.0)[:, None]
    got = oracle.signs_times_v(R, d, 0)
    assert np.allclose(got, 2 * (N - 2) * d)


def test_threaded_split_agrees(oracle):
    R, _ = noisy(30, 10)
    c1, h1 = oracle.triangles_scores(R, 100, nthreads=1)
    c4, h4 = oracle.triangles_scores(R, 100, nthreads=4)
    assert np.array_equal(h1, h4)
    assert np.allclose(c1, c4, rtol=1e-13)
