"""GPU partner-subset search (clm_clmatrix_subset[_device]): the pairs of a
drawn subset (cryo_clmatrix_gpu_1.m:346-350) searched with the drop-in's
per-pair semantics.  Each selected pair's record must equal the full
search's record for that pair (the search of a pair does not depend on the
other pairs of the call), the full search's parity with the FP64 oracle
holds on them, and the left-out pairs come back as pairs with no common
line (clstack = corrstack = 0)."""
import sys

import numpy as np
import pytest

from aspire_amd import sim
from aspire_amd.clmatrix import clmatrix_records, cryo_clmatrix, cryo_clmatrix_subset, partner_subset
from tests.gpu_helpers import assert_parity, compare, oracle_records

pytestmark = pytest.mark.gpu

PRECS = ["fp16", "bf16", "fp32", "fp64"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def stack():
    return sim.make_stack(K=40, n=65, n_r=33, n_theta=360, snr=1 / 4, max_shift_2d=2, shift_kind="int",
                          seed=1041)


@pytest.mark.parametrize("prec", PRECS)
def test_subset_equals_full_search_and_oracle(oracle, stack, prec):
    pf = stack.pf
    K = pf.shape[2]
    sel, nsel = partner_subset(K, 6, seed=7)
    say("case subset K=40 NK=6", prec, "selected", nsel)
    full, _ = clmatrix_records(pf, 3, 1.0, precision=prec)
    sub, st = clmatrix_records(pf, 3, 1.0, precision=prec, sel=sel)
    say(prec, st)
    assert st["pairs"] == nsel
    on = sel != 0
    for k in ("cl1", "cl2", "j", "corr", "shift"):
        assert np.array_equal(sub[k][on], full[k][on]), k
    assert not sub["cl1"][~on].any() and not sub["cl2"][~on].any()
    assert np.all(sub["j"][~on] == -1)
    assert not sub["corr"][~on].any() and not sub["shift"][~on].any()
    # parity of the selected pairs with the oracle
    k1, k2 = oracle.all_pairs(K)
    k1, k2 = k1[on], k2[on]
    P, ora = oracle_records(oracle, pf, 3, 1.0, k1, k2)
    rec = {k: v[on] for k, v in sub.items()}
    rep = compare(oracle, P, ora, rec, 3, 1.0, k1, k2)
    say(prec, rep["agree"], rep["n_disagree"], rep["corr_rel_max"])
    if prec == "fp64":
        assert rep["n_disagree"] == 0 and np.array_equal(rec["corr"], ora["best"])
    else:
        assert_parity(rep)


def test_subset_device_tensors_equal_host(stack):
    import torch

    pf = stack.pf
    K = pf.shape[2]
    sel, _ = partner_subset(K, 9, seed=2)
    say("case subset device vs host")
    host, _ = clmatrix_records(pf, 3, 1.0, precision="bf16", sel=sel)
    dpf = torch.as_tensor(np.ascontiguousarray(pf.transpose(2, 1, 0)), device="cuda").permute(2, 1, 0)
    dsel = torch.as_tensor(sel, device="cuda")
    dev, st = clmatrix_records(dpf, 3, 1.0, precision="bf16", sel=dsel)
    for k in host:
        assert np.array_equal(dev[k].cpu().numpy(), host[k]), k
    with pytest.raises(ValueError):
        clmatrix_records(dpf, 3, 1.0, precision="bf16", sel=dsel[:-1].contiguous())


def test_subset_five_outputs(stack):
    pf = stack.pf
    K = pf.shape[2]
    say("case cryo_clmatrix_subset outputs")
    cs, co, eq, emap, mask, sel = cryo_clmatrix_subset(pf, 5, 0, 3, 1.0, seed=3)
    fcs, fco, feq, femap, _ = cryo_clmatrix(pf, -1, 0, 3, 1.0)
    on = sel != 0
    iu = np.triu_indices(K, 1)
    # selected pairs: the full search's entries; the rest zero
    assert np.array_equal(cs[iu][on], fcs[iu][on]) and not cs[iu][~on].any()
    il = (iu[1], iu[0])
    assert np.array_equal(cs[il][on], fcs[il][on]) and not cs[il][~on].any()
    assert np.array_equal(co[iu][on], fco[iu][on]) and not co[iu][~on].any()
    rows = np.flatnonzero(on)
    assert eq.shape == (rows.size, 2 * K + 1)
    assert (eq != feq.tocsr()[rows]).nnz == 0
    assert np.array_equal(emap[iu][on], np.arange(1, rows.size + 1)) and not emap[iu][~on].any()
    assert not mask.any()
    # NK = -1: every pair, the drop-in's outputs
    cs, co, eq, emap, _, sel = cryo_clmatrix_subset(pf, -1, 0, 3, 1.0)
    assert sel.all()
    assert np.array_equal(cs, fcs) and np.array_equal(co, fco) and np.array_equal(emap, femap)
    assert (eq != feq).nnz == 0


def test_subset_empty_and_repeat(stack):
    pf = stack.pf
    K = pf.shape[2]
    say("case subset empty, then two calls in a row")
    sel = np.zeros(K * (K - 1) // 2, np.uint8)
    rec, st = clmatrix_records(pf, 3, 1.0, precision="fp16", sel=sel)
    assert st["pairs"] == 0 and not rec["cl1"].any() and np.all(rec["j"] == -1)
    a, _ = partner_subset(K, 3, seed=5)
    b, _ = partner_subset(K, 12, seed=6)
    ra, _ = clmatrix_records(pf, 3, 1.0, precision="fp16", sel=a)
    rb, _ = clmatrix_records(pf, 3, 1.0, precision="fp16", sel=b)
    ra2, _ = clmatrix_records(pf, 3, 1.0, precision="fp16", sel=a)
    for k in ra:
        assert np.array_equal(ra[k], ra2[k]), k
    both = (a != 0) & (b != 0)
    for k in ra:
        assert np.array_equal(ra[k][both], rb[k][both]), k


def test_subset_at_workload_shape(oracle):
    # the K=5000 workload's per-pair shape (n_r=65, 41 shifts, SNR 1/16) on
    # K=300 with NK=15: every selected pair against the oracle (BF16 headline
    # kernel and the FP16 default)
    st = sim.make_stack(K=300, n=129, n_r=65, n_theta=360, snr=1 / 16, max_shift_2d=5, shift_kind="uniform",
                        seed=1043)
    pf = st.pf
    K = pf.shape[2]
    sel, nsel = partner_subset(K, 15, seed=9)
    k1, k2 = oracle.all_pairs(K)
    on = sel != 0
    P, ora = oracle_records(oracle, pf, 10, 0.5, k1[on], k2[on])
    for prec in ("bf16", "fp16"):
        say("case subset workload shape K=300 NK=15", prec, nsel)
        rec, s = clmatrix_records(pf, 10, 0.5, precision=prec, sel=sel)
        say(prec, s)
        rep = compare(oracle, P, ora, {k: v[on] for k, v in rec.items()}, 10, 0.5, k1[on], k2[on])
        say(prec, rep["agree"], rep["n_disagree"], rep["corr_rel_max"])
        assert_parity(rep)
