"""Voting synchronization matrix on the GPU (clm_syncmatrix_vote,
aspire_amd/csrc/vote.hip) against the FP64 oracle (oracle/vote_oracle.c,
itself pinned by exact geometry).

The GPU and the oracle compute each block with the same formulas; the
histogram's sums run in another order (four waves' partial sums) and the
device's transcendental functions may differ from the host's in the last
bit, so a block can only differ where the histogram peak is a numerical
near-tie.  The bar: at least 99.9% of the blocks equal to 1e-10, on true
grid lines, on lines with 20% outliers, and on the engine's own common
lines of a clean stack; S symmetric with identity diagonal blocks; fine-grid
true lines give the true relative rotations.  At K = 5000 (the common-lines
workload) a seeded sample of blocks is checked the same way.
"""
import sys

import numpy as np
import pytest

from aspire_amd import sim
from aspire_amd.vote import cryo_syncmatrix_vote, syncmatrix_vote_records
from oracle import clm_oracle as co

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def blocks_agree(Sg, So, K, tol=1e-10):
    k1, k2 = np.triu_indices(K, 1)
    d = np.array([np.abs(Sg[2 * a:2 * a + 2, 2 * b:2 * b + 2] - So[2 * a:2 * a + 2, 2 * b:2 * b + 2]).max()
                  for a, b in zip(k1, k2)])
    return float(np.mean(d <= tol)), float(d.max())


@pytest.mark.parametrize("outliers", [0.0, 0.2])
def test_vote_matches_oracle_on_grid_lines(outliers):
    say("case vote-grid", outliers)
    K, T = 60, 360
    rng = np.random.default_rng(11)
    rots = sim.rand_rots(rng, K)
    cl = sim.clmatrix_cheat(rots, T)
    k1, k2 = np.triu_indices(K, 1)
    for i in np.flatnonzero(rng.random(k1.size) < outliers):
        cl[k1[i], k2[i]] = rng.integers(1, T // 2 + 1)
        cl[k2[i], k1[i]] = rng.integers(1, T + 1)
    Sg = cryo_syncmatrix_vote(cl, T)
    So = co.syncmatrix_vote(cl, T)
    frac, worst = blocks_agree(Sg, So, K)
    say("blocks equal to 1e-10:", frac, "largest difference", worst)
    assert frac >= 0.999
    assert np.array_equal(Sg, Sg.T)
    for k in range(K):
        assert np.array_equal(Sg[2 * k:2 * k + 2, 2 * k:2 * k + 2], np.eye(2))


def test_vote_fine_grid_gives_true_rotations():
    say("case vote-fine")
    K, T = 30, 36000
    rots = sim.rand_rots(np.random.default_rng(4), K)
    cl = sim.clmatrix_cheat(rots, T)
    Sg = cryo_syncmatrix_vote(cl, T)
    k1, k2 = np.triu_indices(K, 1)
    err = max(np.abs(Sg[2 * a:2 * a + 2, 2 * b:2 * b + 2] - (rots[:, :, a] @ rots[:, :, b].T)[:2, :2]).max()
              for a, b in zip(k1, k2))
    ev = np.sort(np.abs(np.linalg.eigvalsh(Sg)))[::-1]
    say("largest block error", err, "eigenvalue 4 / 1", ev[3] / ev[0])
    assert err <= 2e-3 and ev[3] <= 1e-3 * ev[0]


def test_vote_on_engine_lines(oracle):
    from aspire_amd.clmatrix import assemble, clmatrix_records

    say("case vote-engine")
    K = 100
    st = sim.make_stack(K=K, n=65, n_r=33, n_theta=360, snr=None, seed=8)
    rec, _ = clmatrix_records(st.pf, 0, 1, precision="bf16")
    clstack, _, _, _ = assemble(K, 360, rec)
    Sg = cryo_syncmatrix_vote(clstack, 360)
    So = co.syncmatrix_vote(clstack, 360)
    frac, worst = blocks_agree(Sg, So, K)
    say("blocks equal to 1e-10:", frac, "largest difference", worst)
    assert frac >= 0.999


def test_vote_full_workload_sample():
    import torch

    from aspire_amd.clmatrix import clmatrix_records

    K, T = 5000, 360
    say("case vote-full", K)
    st = sim.make_stack(K=K, n=129, n_r=65, n_theta=T, snr=1 / 16, max_shift_2d=5.0, shift_kind="uniform",
                        seed=1003, device="cuda")
    rec, _ = clmatrix_records(st.pf, 10, 0.5, precision="bf16")
    del st
    S, stt = syncmatrix_vote_records(rec["cl1"], rec["cl2"], K, T)
    say(stt)
    assert torch.equal(S, S.T)
    cl1, cl2 = rec["cl1"].cpu().numpy(), rec["cl2"].cpu().numpy()
    k1, k2 = np.triu_indices(K, 1)
    clstack = np.zeros((K, K))
    clstack[k1, k2] = cl1
    clstack[k2, k1] = cl2
    flat = np.ascontiguousarray(clstack.T).ravel()
    rng = np.random.default_rng(0)
    ok = 0
    n = 120
    for e in rng.choice(k1.size, size=n, replace=False):
        a, b = int(k1[e]), int(k2[e])
        ref = co.vote_pair(clstack, T, a, b, _flat=flat)
        got = S[2 * a:2 * a + 2, 2 * b:2 * b + 2].cpu().numpy()
        ok += int(np.abs(got - ref).max() <= 1e-10)
    say("sampled blocks equal:", ok, "of", n)
    assert ok >= n - 1
