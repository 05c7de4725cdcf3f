"""The measured workloads at their full sizes, against the FP64 oracle.

* K=100 (65 px, n_r=33, no shifts, SNR 1): every one of the 4,950 pairs, in
  every precision; FP64 bit-identical.
* K=1000 (89 px, n_r=45, 13 shifts, SNR 1/8): a seeded sample of 2,000 pairs,
  every precision.
* K=5000 headline (129 px, n_r=65, 41 shifts, SNR 1/16): BF16 with the
  autotuned variant, a seeded sample of 2,000 pairs.
* 64 uneven stacks in one graph-captured grouped job: 40 seeded pairs of
  every stack.
* K=5000 learned form (shifts U[-12,12], 49-shift full grid, +-1 px window):
  the full grid against the oracle on a sample; the windowed search's
  invariants (pairs it hands back equal the full grid; pairs it settles carry
  the FP64 value of a grid entry, never above the full-grid maximum), with
  a gate that hands back a share of the pairs.

Bar: clstack agrees on >= 99.9% of the checked pairs, every disagreement is
a near tie within 1e-5 relative correlation, corr within 1e-5 relative
(sinograms/cryo_clmatrix_cpu.m:360-399 is the reference arithmetic).
"""
import sys

import numpy as np
import pytest

from aspire_amd import workloads as wl
from aspire_amd.clmatrix import autotune, clmatrix_grouped, clmatrix_records
from tests.gpu_helpers import assert_parity, compare, host_subset, oracle_sample

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check(oracle, w, pf, rec, n_sample, seed=42, exact=False):
    slots, k1, k2 = wl.sample_pairs(w.K, n_sample, seed)
    P, ora, kk1, kk2 = oracle_sample(oracle, pf, k1, k2, w.max_shift, w.shift_step)
    g = host_subset(rec, slots)
    rep = compare(oracle, P, ora, g, w.max_shift, w.shift_step, kk1, kk2)
    if exact:
        assert rep["n_disagree"] == 0, rep
        assert np.array_equal(g["corr"], ora["best"])
    else:
        assert_parity(rep)
    return rep


@pytest.fixture(scope="module")
def k100():
    return wl.K100.make()


@pytest.fixture(scope="module")
def k1000():
    return wl.K1000.make()


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp32", "fp64"])
def test_k100_every_pair(oracle, k100, prec):
    w = wl.K100
    say("case workload K=100 all pairs", prec)
    rec, st = clmatrix_records(k100.pf, w.max_shift, w.shift_step, precision=prec)
    rep = check(oracle, w, k100.pf, rec, w.pairs, exact=prec == "fp64")
    say(prec, rep["n"], rep["agree"], rep["n_disagree"], rep["corr_rel_max"], st["variant"])
    assert rep["n"] == 4950


@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp32", "fp64"])
def test_k1000_sample(oracle, k1000, prec):
    w = wl.K1000
    say("case workload K=1000 sample", prec)
    rec, st = clmatrix_records(k1000.pf, w.max_shift, w.shift_step, precision=prec)
    rep = check(oracle, w, k1000.pf, rec, 2000, exact=prec == "fp64")
    say(prec, rep["n"], rep["agree"], rep["n_disagree"], rep["corr_rel_max"], st)


def test_k5000_headline_sample(oracle):
    import torch

    w = wl.K5000
    say("case workload K=5000 headline, BF16 autotuned")
    st = w.make()
    rows, best = autotune(st.pf, w.max_shift, w.shift_step, "bf16")
    rec, stats = clmatrix_records(st.pf, w.max_shift, w.shift_step, precision="bf16")
    say(best, stats)
    assert stats["variant"] == best["variant"] and stats["pairs"] == w.pairs
    rep = check(oracle, w, st.pf, rec, 2000, seed=7)
    say(rep["n"], rep["agree"], rep["n_disagree"], rep["worst_gap_rel"], rep["corr_rel_max"])
    del st, rec
    torch.cuda.empty_cache()


def test_grouped_64_stacks_graph(oracle):
    import torch

    ws = wl.grouped()
    say("case workload 64 uneven stacks, graph-captured grouped job")
    pfs = [w.make().pf for w in ws]
    mss = [w.max_shift for w in ws]
    recs, st0 = clmatrix_grouped(pfs, mss, 1, "bf16", graph=True)
    recs, st1 = clmatrix_grouped(pfs, mss, 1, "bf16", graph=True, out=recs)  # replay
    say(st1)
    assert st1["pairs"] == sum(w.pairs for w in ws)
    n_all = agree = 0
    for j, w in enumerate(ws):
        rep = check(oracle, w, pfs[j], recs[j], 40, seed=100 + j)
        n_all += rep["n"]
        agree += rep["n"] - rep["n_disagree"]
    say("checked", n_all, "agree", agree)
    assert n_all == 64 * 40
    del pfs, recs
    torch.cuda.empty_cache()


def test_learned_workload(oracle):
    import os

    import torch

    from aspire_amd.shiftnet import WEIGHTS, ShiftNet, clmatrix_windowed

    if not os.path.exists(WEIGHTS):
        pytest.skip("trained weights not present")
    w = wl.LEARNED
    say("case workload K=5000 learned form")
    st = w.make()
    K = w.K
    full, fst = clmatrix_records(st.pf, w.max_shift, w.shift_step, precision="bf16")
    assert fst["n_eval_shifts"] == 49
    rep = check(oracle, w, st.pf, full, 600, seed=11)
    say("full grid vs oracle", rep["n"], rep["agree"], rep["n_disagree"])
    net = ShiftNet.load()
    prop = net.propose(st.pf)
    # a gate at the median window peak of the settled pairs hands back about half of them
    rec0, ws0 = clmatrix_windowed(st.pf, prop, w.max_shift, w.shift_step, 1.0, gate=-1e30, precision="bf16")
    ex0 = rec0["exhaustive"].cpu().numpy() == 1
    gate = float(np.median(rec0["corr"].cpu().numpy()[~ex0]))
    rec, ws = clmatrix_windowed(st.pf, prop, w.max_shift, w.shift_step, 1.0, gate=gate, precision="bf16")
    say("gate", gate, ws)
    assert ws["window_shifts"] == 5 and ws["full_shifts"] == 49
    assert ws["accepted"] + ws["below_gate"] + ws["off_grid"] == w.pairs
    # the gate is tested first: a pair below it is not also counted off the grid
    assert ws["below_gate"] > 0 and ws["off_grid"] <= ws0["off_grid"]
    r = {k: v.cpu().numpy() for k, v in rec.items()}
    fh = {k: v.cpu().numpy() for k, v in full.items()}
    ex = r["exhaustive"] == 1
    assert ex.sum() == ws["below_gate"] + ws["off_grid"]
    for k in ("cl1", "cl2", "j", "corr", "shift"):
        assert np.array_equal(r[k][ex], fh[k][ex]), k
    acc = np.flatnonzero(~ex)
    assert np.all(r["corr"][acc] >= gate)
    assert np.all(r["corr"][acc] <= fh["corr"][acc] * (1 + 1e-9))
    assert np.all(r["shift"][acc] == -w.max_shift + w.shift_step * r["j"][acc])
    # settled pairs carry the FP64 value of the grid entry they report
    slots = acc[np.random.default_rng(3).choice(acc.size, size=min(200, acc.size), replace=False)]
    slots.sort()
    k1, k2 = wl.pairs_of_slots(K, slots)
    Pimg, _, ii1, ii2 = oracle_sample(oracle, st.pf, k1, k2, w.max_shift, w.shift_step)
    for t, e in enumerate(slots):
        v = oracle.entry(Pimg, w.max_shift, w.shift_step, ii1[t], ii2[t], r["j"][e], r["cl1"][e] - 1,
                         r["cl2"][e] - 1)
        assert abs(v - r["corr"][e]) <= 1e-9 * abs(v), (e, v, r["corr"][e])
    del st, full, rec, rec0
    torch.cuda.empty_cache()
