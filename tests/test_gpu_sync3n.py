"""The sync3n triangle kernels on the GPU against the FP64 oracle
(oracle/sync3n_oracle.c, one-thread order).

Per-pair sums add a pair's triangles in ascending order of the third image
in both, with the same operations, so cumulated scores, histogram counts and
signs products are bit-identical; the log-likelihoods go through the device
and host libm (log, pow, exp), so they are compared to 1e-12 relative.
Cases: N = 3 (one triangle), tile edges (N = 17, 40, 70), J flips and
outliers, every entry kind of the signs product and 1-8 vectors, rotations
about one axis (J R J = R, so every score is 0/0: NaN, last bin), and the
power method recovering the J flips of consistent data.
"""
import sys

import numpy as np
import pytest

from aspire_amd import sim
from aspire_amd.sync3n import (jsync_power_method, pairs_probabilities_mex, signs_times_v_mex,
                               triangles_scores_mex)
from tests.test_sync3n_oracle import noisy, rel_rotations

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("N", [3, 17, 40, 70])
def test_triangles_scores(oracle, N):
    say("case sync3n scores", N)
    R, _ = noisy(N, 100 + N)
    c, h, st = triangles_scores_mex(R, 100)
    co, ho = oracle.triangles_scores(R, 100)
    assert st["triangles"] == N * (N - 1) * (N - 2) // 6
    assert np.array_equal(h, ho), (h, ho)
    assert np.array_equal(c, co), np.abs(c - co).max()
    assert h.sum() == 3 * st["triangles"]


def test_triangles_scores_bins_and_no_cum(oracle):
    say("case sync3n bins")
    R, _ = noisy(33, 7)
    for n in (1, 7, 256, 1024, 1025, 3000):  # thresholds in LDS up to 1024 intervals, global beyond
        c, h, _ = triangles_scores_mex(R, n, cum=False)
        assert c is None
        assert np.array_equal(h, oracle.triangles_scores(R, n, cum=False)[1]), n


def test_consistent_and_degenerate(oracle):
    say("case sync3n consistent/degenerate")
    rng = np.random.default_rng(2)
    N = 21
    R, _ = rel_rotations(sim.rand_rots(rng, N), rng, flip=0.3)
    c, h, _ = triangles_scores_mex(R, 100)
    assert np.array_equal(c, oracle.triangles_scores(R, 100)[0])
    assert h[-1] == 3 * N * (N - 1) * (N - 2) // 6
    # rotations about the z axis: J R J = R, all four matches equal, 0/0 scores
    th = rng.uniform(0, 2 * np.pi, N)
    rots = np.zeros((3, 3, N))
    rots[0, 0], rots[0, 1], rots[1, 0], rots[1, 1], rots[2, 2] = np.cos(th), -np.sin(th), np.sin(th), np.cos(th), 1
    Rz, _ = rel_rotations(rots)
    c, h, _ = triangles_scores_mex(Rz, 10)
    co, ho = oracle.triangles_scores(Rz, 10)
    assert np.array_equal(h, ho) and np.array_equal(c, co, equal_nan=True)


# the device evaluates both log-likelihoods from one logarithm when A, B, a,
# b > 0 and 0 <= P2 <= 1 (the first two rows), else the reference's direct
# formula (a = 0; P2 = 1 leaves only the arbitrary model's term)
PRMS = [(0.45, 1.7, 2.2, 2.9, 2.5, 0.78), (1.0, 1.7, 2.2, 2.9, 2.5, 0.78), (0.45, 1.7, 0.0, 2.9, 2.5, 0.78)]


@pytest.mark.parametrize("N,prm", [(3, PRMS[0]), (40, PRMS[0]), (40, PRMS[1]), (40, PRMS[2])])
def test_pairs_probabilities(oracle, N, prm):
    say("case sync3n probabilities", N, prm)
    R, _ = noisy(N, 200 + N)
    ind, arb, _ = pairs_probabilities_mex(R, *prm)
    io, ao = oracle.pairs_probabilities(R, *prm)
    assert np.allclose(ind, io, rtol=1e-12, atol=0) and np.allclose(arb, ao, rtol=1e-12, atol=0)


@pytest.mark.parametrize("entries,ne", [(0, 1), (1, 3), (2, 8), (0, 4)])
def test_signs_times_v(oracle, entries, ne):
    say("case sync3n signs", entries, ne)
    N = 37
    R, _ = noisy(N, 300 + ne)
    P = N * (N - 1) // 2
    rng = np.random.default_rng(ne)
    v = rng.standard_normal((P, ne))
    ps = rng.uniform(0.1, 1.0, P)
    got, _ = signs_times_v_mex(R, v, entries, ps)
    want = oracle.signs_times_v(R, v, entries, ps)
    assert np.array_equal(got, want), np.abs(got - want).max()


def test_power_method_recovers_flips():
    say("case sync3n power method")
    rng = np.random.default_rng(11)
    N = 30
    R, flips = rel_rotations(sim.rand_rots(rng, N), rng, flip=0.4)
    d = np.where(flips, -1.0, 1.0)
    J_sync, J_sig, eig, itr, dd = jsync_power_method(R, n_eigs=4, seed=3)
    say(itr, dd, eig)
    assert np.array_equal(J_sync, d) or np.array_equal(J_sync, -d)
    # the reference's stopping test (leading vector moves by <= 1e-2) leaves
    # the Rayleigh-quotient-like diagonal of R within ~1e-4 of the eigenvalues
    # 2N-4 (top) and N-4 (next), the reference's theoretical values
    assert abs(abs(eig[0]) - (2 * N - 4)) < 1e-3 * (2 * N - 4)
    assert np.all(np.abs(np.abs(eig[1:]) - (N - 4)) < 1e-3 * (N - 4))
