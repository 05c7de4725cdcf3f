"""Committed golden vectors (tests/golden/clmatrix_tiny.npz, written by
tests/golden/make_golden.py): tiny cryo_clmatrix cases -- random rays, a
planted common line, exact ties with an all-negative pair -- and their four
outputs.  The C oracle and the numpy restatement must keep reproducing them
(clstack and shift_equations_map exactly, corrstack to 1e-13,
shift_equations to 1e-15), and the planted and tie cases must keep their
answers fixed in closed form.
"""
import os

import numpy as np
import pytest

from oracle import clm_oracle_np as onp

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "clmatrix_tiny.npz")
CASES = ["random", "planted", "ties"]


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.mark.parametrize("name", CASES)
def test_c_oracle_reproduces_golden(oracle, gold, name):
    pf = gold[f"{name}_pf"]
    ms, st = gold[f"{name}_params"]
    cl, corr, eq, emap, mask = oracle.cryo_clmatrix_cpu(np.asfortranarray(pf), max_shift=ms, shift_step=st)
    assert np.array_equal(cl, gold[f"{name}_clstack"])
    assert np.abs(corr - gold[f"{name}_corrstack"]).max() <= 1e-13
    assert np.abs(eq.toarray() - gold[f"{name}_equations"]).max() <= 1e-15
    assert np.array_equal(emap, gold[f"{name}_map"])
    assert not mask.any()


@pytest.mark.parametrize("name", CASES)
def test_numpy_restatement_reproduces_golden(gold, name):
    ms, st = gold[f"{name}_params"]
    cl, corr, eq, emap = onp.clmatrix(gold[f"{name}_pf"], ms, st)
    assert np.array_equal(cl, gold[f"{name}_clstack"])
    assert np.abs(corr - gold[f"{name}_corrstack"]).max() <= 1e-13
    assert np.abs(eq - gold[f"{name}_equations"]).max() <= 1e-15
    assert np.array_equal(emap, gold[f"{name}_map"])


def test_golden_closed_form_answers(gold):
    l1, l2, j = gold["planted_answer"]
    cl, corr = gold["planted_clstack"], gold["planted_corrstack"]
    assert (cl[0, 2], cl[2, 0]) == (l1, l2)
    assert abs(corr[0, 2]) <= 1e-13  # corrstack = 1 - correlation, correlation 1
    ms, st = gold["planted_params"]
    e = gold["planted_map"][0, 2] - 1  # equation row of pair (0, 2), 0-based
    assert gold["planted_equations"][int(e), -1] == -ms + j * st
    cl, corr = gold["ties_clstack"], gold["ties_corrstack"]
    assert (cl[0, 1], cl[1, 0]) == (1, 1) and corr[0, 1] == 0.0
    assert cl[0, 2] == cl[2, 0] == cl[1, 2] == cl[2, 1] == 0
