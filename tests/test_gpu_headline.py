"""GPU parity at the K=5000 workload's own per-pair shape, and at the two
longest supported rays.

The per-pair shape of the headline workload is n=129 images, n_r=65
(R=64, padded ray length 64), n_theta=360 and the 41-shift grid
max_shift=10, step 0.5 (sinograms/cryo_clmatrix_cpu.m:360-399).  Every pair
of a K=64 stack of that shape (2,016 pairs) is checked against the FP64
oracle in every precision, with the BF16 kernel variant the autotuner picks
at K=5000 (3 column tiles per wave, 8 waves per workgroup: variant 308)
forced as well as chosen by the heuristic.  n_r=97 and n_r=129 cover the
padded ray lengths 96 and 128.

Bar: clstack agrees on >= 99.9% of pairs, every disagreement is a near tie
within 1e-5 relative correlation (the oracle's own value at the GPU's
choice), corr within 1e-5 relative; FP64 is bit-identical.
"""
import sys

import numpy as np
import pytest

from aspire_amd import sim
from aspire_amd.clmatrix import clmatrix_records
from tests.gpu_helpers import assert_parity, compare, oracle_records

pytestmark = pytest.mark.gpu

MS, STEP = 10.0, 0.5  # 41 shifts


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def headline(oracle):
    # the K=5000 workload's image model at K=64: 129 px, shifts U[-5,5], SNR 1/16
    st = sim.make_stack(K=64, n=129, n_r=65, n_theta=360, snr=1 / 16, max_shift_2d=5.0,
                        shift_kind="uniform", seed=1003)
    k1, k2 = oracle.all_pairs(64)
    P, ora = oracle_records(oracle, st.pf, MS, STEP, k1, k2)
    return st.pf, k1, k2, P, ora


@pytest.mark.parametrize("prec,variant", [("bf16", 308), ("bf16", 0), ("bf16", 206), ("bf16", 106), ("fp16", 308),
                                          ("fp16", 0), ("fp32", 0), ("fp64", 0)])
def test_headline_shape_all_pairs(oracle, headline, prec, variant):
    pf, k1, k2, P, ora = headline
    say("case headline-shape", prec, variant)
    rec, st = clmatrix_records(pf, MS, STEP, precision=prec, variant=variant)
    assert st["n_eval_shifts"] == 41
    if variant:
        assert st["variant"] == variant
    rep = compare(oracle, P, ora, rec, MS, STEP, k1, k2)
    say(prec, variant, st["variant"], rep["agree"], rep["n_disagree"], rep["worst_gap_rel"], rep["corr_rel_max"],
        {k: st[k] for k in ("uncertain_pairs", "refined_pairs", "fallback_pairs") if k in st})
    if prec == "fp64":
        assert rep["n_disagree"] == 0
        assert np.array_equal(rec["corr"], ora["best"])
        assert np.array_equal(rec["shift"], ora["shift"])
    else:
        assert_parity(rep)
        if prec != "fp32":
            # the winner's value is the FP64 re-score: equal to the oracle's bits
            same = (rec["cl1"] == ora["cl1"]) & (rec["cl2"] == ora["cl2"]) & (rec["j"] == ora["j"])
            assert np.array_equal(rec["corr"][same], ora["best"][same])


@pytest.mark.parametrize("n_r,K", [(97, 28), (129, 24)])
@pytest.mark.parametrize("prec", ["bf16", "fp16", "fp32", "fp64"])
def test_long_rays(oracle, n_r, K, prec):
    """Padded ray lengths 96 and 128 (n_r = 97, 129) on 129-px images with
    integer shifts in [-3,3], SNR 1/8, max_shift 4, step 1 (9 shifts)."""
    say("case long-rays", n_r, K, prec)
    st = sim.make_stack(K=K, n=129, n_r=n_r, n_theta=360, snr=1 / 8, max_shift_2d=3, shift_kind="int",
                        seed=1100 + n_r)
    k1, k2 = oracle.all_pairs(K)
    P, ora = oracle_records(oracle, st.pf, 4, 1, k1, k2)
    rec, s = clmatrix_records(st.pf, 4, 1, precision=prec)
    rep = compare(oracle, P, ora, rec, 4, 1, k1, k2)
    say(n_r, prec, s["variant"], rep["agree"], rep["n_disagree"], rep["corr_rel_max"])
    if prec == "fp64":
        assert rep["n_disagree"] == 0
        assert np.array_equal(rec["corr"], ora["best"])
    else:
        assert_parity(rep)
