"""What the trained shift network gains over the method without it.

The committed network (aspire_amd/data/shiftnet.npz, trained with libclm's
own BF16 kernels by ``python -m aspire_amd.shiftnet train``) is applied to a
held-out stack (seed not used in training, phantom not recentred, shifts
U[-12,12] px).  Compared with the same windowed search centred at zero shift
(no network):

* the pair 1-D shifts it predicts at the true common lines are far closer to
  the true ones (the quantity the window must contain);
* the windowed search agrees far more often with the exhaustive full-grid
  search, and finds at least as many true common lines as the full grid
  while evaluating 5 of its 49 shifts per pair.
"""
import os
import sys

import numpy as np
import pytest

from aspire_amd import sim
from aspire_amd.clmatrix import clmatrix_records
from aspire_amd.shiftnet import WEIGHTS, ShiftNet, clmatrix_windowed, pair_shift_error

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="trained weights not present")
@pytest.mark.parametrize("snr", [1.0, 1 / 16])
def test_trained_network_gain(snr):
    import torch

    say("case shiftnet-gain", snr)
    net = ShiftNet.load()
    K = 400
    st = sim.make_stack(K=K, n=129, n_r=65, n_theta=360, snr=snr, max_shift_2d=12, shift_kind="uniform",
                        seed=1205, device="cuda")
    prop = net.propose(st.pf)
    pe_net, pe_zero = pair_shift_error(prop.cpu().numpy(), st.shifts, st.rots, n_pairs=4000)
    say("pair-shift rms (px): network", pe_net, "no network", pe_zero)
    assert pe_net < 0.25 * pe_zero
    full, _ = clmatrix_records(st.pf, 12, 0.5, precision="bf16")
    full = {k: v.cpu().numpy() for k, v in full.items()}
    res = {}
    for name, p in (("net", prop), ("zero", torch.zeros_like(prop))):
        rec, ws = clmatrix_windowed(st.pf, p, 12, 0.5, window=1.0, gate=0.0, precision="bf16")
        r = {k: v.cpu().numpy() for k, v in rec.items()}
        agree = float(np.mean((r["cl1"] == full["cl1"]) & (r["cl2"] == full["cl2"]) & (r["j"] == full["j"])))
        res[name] = (agree, r)
        say(name, "agreement with the full grid", agree, ws)
    assert res["net"][0] > 2.0 * res["zero"][0]
    ref = sim.clmatrix_cheat(st.rots, 360)

    def detect(r):
        cl = np.zeros((K, K))
        k1, k2 = np.triu_indices(K, 1)
        cl[k1, k2] = r["cl1"]
        cl[k2, k1] = r["cl2"]
        return sim.comparecl(cl, ref, 360, 5)

    d_full, d_net, d_zero = detect(full), detect(res["net"][1]), detect(res["zero"][1])
    say("detection at 5 deg: full", d_full, "windowed+network", d_net, "windowed, no network", d_zero)
    assert d_net >= 0.9 * d_full
    assert d_net > d_zero
