"""Grouped (batched) form: several uneven stacks in one device job, eager and
captured in a HIP graph.

Each stack's records must equal what the single-stack entry point returns on
the same stack (same kernels, same arithmetic) and meet the oracle parity
bar.  The graph is replayed several times, with the program's own eager work
and synchronizations between replays and the outputs overwritten before each
replay; every replay must reproduce the eager result, and a replay after the
inputs changed in place must reproduce an eager run on the new inputs.
"""
import sys

import numpy as np
import pytest

from aspire_amd import sim
from aspire_amd.clmatrix import clmatrix_grouped, clmatrix_records
from tests.gpu_helpers import assert_parity, compare, oracle_records

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(pf):
    import torch

    return torch.as_tensor(np.ascontiguousarray(pf.transpose(2, 1, 0))).cuda().permute(2, 1, 0)


@pytest.fixture(scope="module")
def stacks():
    # uneven K, two ray lengths (two kernel groups), integer and half-pixel shifts
    a = sim.make_stack(K=17, n=65, n_r=33, n_theta=360, snr=1.0, max_shift_2d=2, shift_kind="int", seed=2001)
    b = sim.make_stack(K=9, n=65, n_r=33, n_theta=360, snr=1 / 4, max_shift_2d=2, shift_kind="int", seed=2002)
    c = sim.make_stack(K=23, n=89, n_r=45, n_theta=360, snr=1 / 8, max_shift_2d=3, shift_kind="int", seed=2003)
    d = sim.make_stack(K=6, n=65, n_r=33, n_theta=360, snr=1.0, seed=2004)
    return [a.pf, b.pf, c.pf, d.pf]


MS = [2, 2, 3, 0]
SS = [1, 0.5, 1, 1]


def host(rec):
    return {k: v.cpu().numpy() for k, v in rec.items()}


def singles(pfs, precs):
    return [clmatrix_records(pf, MS[i], SS[i], precision=precs[i])[0] for i, pf in enumerate(pfs)]


@pytest.mark.parametrize("graph", [False, True])
def test_grouped_matches_single_and_oracle(oracle, stacks, graph):
    say("case grouped", "graph" if graph else "eager")
    precs = ["bf16", "fp16", "bf16", "fp64"]
    dev = [to_dev(pf) for pf in stacks]
    recs, st = clmatrix_grouped(dev, MS, SS, precs, graph=graph)
    say(st)
    assert st["pairs"] == sum(pf.shape[2] * (pf.shape[2] - 1) // 2 for pf in stacks)
    ref = singles(stacks, precs)
    for i, pf in enumerate(stacks):
        r = host(recs[i])
        for k in r:
            assert np.array_equal(r[k], ref[i][k]), (i, k)
        k1, k2 = oracle.all_pairs(pf.shape[2])
        P, ora = oracle_records(oracle, pf, MS[i], SS[i], k1, k2)
        rep = compare(oracle, P, ora, r, MS[i], SS[i], k1, k2)
        say(i, precs[i], rep["agree"], rep["n_disagree"], rep["corr_rel_max"])
        if precs[i] == "fp64":
            assert rep["n_disagree"] == 0 and np.array_equal(r["corr"], ora["best"])
        else:
            assert_parity(rep)


def test_grouped_graph_replays(stacks):
    import torch

    say("case graph-replay")
    precs = ["bf16"] * 4
    dev = [to_dev(pf) for pf in stacks]
    eager, _ = clmatrix_grouped(dev, MS, SS, precs, graph=False)
    eager = [host(r) for r in eager]
    out, _ = clmatrix_grouped(dev, MS, SS, precs, graph=True)
    for rep in range(4):
        for i in range(len(stacks)):
            got = host(out[i])
            for k in got:
                assert np.array_equal(got[k], eager[i][k]), (rep, i, k)
        # eager work of the program between replays; outputs overwritten
        x = torch.randn(512, 512, device="cuda")
        (x @ x).sum().item()
        for o in out:
            for v in o.values():
                v.fill_(-7)
        torch.cuda.synchronize()
        out, st = clmatrix_grouped(dev, MS, SS, precs, graph=True, out=out)
        say("replay", rep, st["ms_total"])
    # inputs changed in place: the replay reads the new data
    for t in dev:
        t.copy_(t.flip(2))
    torch.cuda.synchronize()
    fresh, _ = clmatrix_grouped(dev, MS, SS, precs, graph=False)
    out, _ = clmatrix_grouped(dev, MS, SS, precs, graph=True, out=out)
    for i in range(len(stacks)):
        a, b = host(out[i]), host(fresh[i])
        for k in a:
            assert np.array_equal(a[k], b[k]), ("flipped", i, k)
    assert not np.array_equal(host(fresh[0])["cl1"], eager[0]["cl1"])


def test_grouped_statistics_add_up(stacks):
    """The re-score's counters are summed per wave and added once per wave.
    A pair's re-scored cells are those within the certified margin of its
    final maximum, whichever kernel variant or job it runs in, so one
    grouped job must report the sums of the single-stack calls."""
    say("case grouped-statistics")
    precs = ["bf16", "fp16", "bf16", "bf16"]
    dev = [to_dev(pf) for pf in stacks]
    _, st = clmatrix_grouped(dev, MS, SS, precs, graph=False)
    sts = [clmatrix_records(d, MS[i], SS[i], precision=precs[i])[1] for i, d in enumerate(dev)]
    say(st)
    for k in ("pairs", "cells_rescored", "fp64_candidates", "uncertain_pairs", "refined_pairs"):
        assert st[k] == sum(x[k] for x in sts), (k, st[k], [x[k] for x in sts])
    # every certain pair re-scores at least the cell and the FP64 candidate of its maximum
    certain = st["pairs"] - st["uncertain_pairs"]
    assert st["cells_rescored"] >= certain and st["fp64_candidates"] >= certain


def test_grouped_empty_and_tiny():
    import torch

    say("case grouped-tiny")
    one = to_dev(sim.make_stack(K=1, n=65, n_r=33, n_theta=360, snr=1.0, seed=2005).pf)
    two = to_dev(sim.make_stack(K=2, n=65, n_r=33, n_theta=360, snr=1.0, seed=2006).pf)
    recs, st = clmatrix_grouped([one, two], 0, 1, "bf16", graph=True)
    assert st["pairs"] == 1 and recs[0]["cl1"].numel() == 0 and recs[1]["cl1"].numel() == 1
    r, _ = clmatrix_records(two, 0, 1, precision="bf16")
    assert torch.equal(r["cl1"], recs[1]["cl1"]) and torch.equal(r["corr"], recs[1]["corr"])
    recs, st = clmatrix_grouped([], graph=True)
    assert recs == [] and st["pairs"] == 0
