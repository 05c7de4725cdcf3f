"""The committed golden vectors (tests/golden/clmatrix_tiny.npz) through the
drop-in (aspire_amd.clmatrix.cryo_clmatrix, libclm on the GPU) at every
precision.  The FP64 path and the re-scored BF16/FP16 paths must give the
golden clstack, map and equations, and corrstack to 1e-13 (their final values
are FP64 evaluations in the oracle's order); the plain FP32 path must give
the golden clstack with corrstack to 1e-6.  The planted pair and the exact
ties pin the index order, the negative pair the no-update rule.
"""
import os
import sys

import numpy as np
import pytest

from aspire_amd.clmatrix import cryo_clmatrix

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "clmatrix_tiny.npz")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("precision", ["fp64", "bf16", "fp16", "fp32"])
@pytest.mark.parametrize("name", ["random", "planted", "ties"])
def test_golden_through_the_drop_in(name, precision):
    say("case golden", name, precision)
    g = np.load(GOLD)
    pf = np.asfortranarray(g[f"{name}_pf"])
    ms, st = g[f"{name}_params"]
    cl, corr, eqs, emap, mask = cryo_clmatrix(pf, -1, 0, ms, st, precision=precision)
    eqs = eqs.toarray() if hasattr(eqs, "toarray") else np.asarray(eqs)
    assert np.array_equal(cl, g[f"{name}_clstack"])
    assert np.array_equal(emap, g[f"{name}_map"])
    assert not np.any(mask)
    tol = 1e-6 if precision == "fp32" else 1e-13
    assert np.abs(corr - g[f"{name}_corrstack"]).max() <= tol
    assert np.abs(eqs - g[f"{name}_equations"]).max() <= 1e-15
