"""cryo_clmatrix with map_filter_radius > 0 on the GPU
(clm_clmatrix_filtered, aspire_amd/csrc/clm_clmap.h) against the FP64
oracle (oracle/clm_oracle.c: clmo_pairs_filtered, itself pinned against an
FFT restatement of sinograms/cryo_average_clmap.m).

Both evaluate every map entry and the stencil in FP64 in the same order, so
the records must be identical: line indices, shift index, shift and the
filtered correlation, bit for bit.  Cases cover several radii (including the
largest, 10), odd tile remainders (L = 12, 18, 180), shifted and unshifted
stacks, complex64 and complex128 rays, host and device buffers, and the
Python mirror of cryo_clmatrix (argument 6).
"""
import sys

import numpy as np
import pytest

from aspire_amd import sim
from aspire_amd.clmatrix import clmatrix_records, cryo_clmatrix

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [
    # K, n, n_r, n_theta, max_shift, step, radius, dtype
    (8, 17, 9, 24, 2, 1, 1, "complex128"),
    (8, 17, 9, 24, 2, 1, 10, "complex128"),
    (7, 33, 17, 36, 3, 1, 2, "complex64"),
    (6, 33, 17, 36, 0, 1, 4, "complex128"),
    (5, 65, 33, 360, 1, 1, 1, "complex128"),
]


@pytest.mark.parametrize("K,n,n_r,n_theta,ms,step,d,dt", CASES)
def test_filtered_matches_oracle(oracle, K, n, n_r, n_theta, ms, step, d, dt):
    import torch

    say("case clmap", K, n, n_r, n_theta, ms, step, d, dt)
    st = sim.make_stack(K=K, n=n, n_r=n_r, n_theta=n_theta, snr=1.0, max_shift_2d=ms / 2, shift_kind="uniform",
                        seed=100 + K + d, dtype=dt)
    P, _ = oracle.preprocess(st.pf)
    ora = oracle.pairs(P, ms, step, map_filter_radius=d)
    rec, stats = clmatrix_records(st.pf, ms, step, map_filter_radius=d)
    say(stats)
    for k, ok in (("cl1", "cl1"), ("cl2", "cl2"), ("j", "j"), ("corr", "best"), ("shift", "shift")):
        assert np.array_equal(np.asarray(rec[k]), ora[ok]), k
    assert np.all(ora["best"] > 0)
    # device buffers give the same bits
    dev = torch.as_tensor(np.ascontiguousarray(st.pf.transpose(2, 1, 0))).cuda().permute(2, 1, 0)
    rd, _ = clmatrix_records(dev, ms, step, map_filter_radius=d)
    for k in ("cl1", "cl2", "j", "corr", "shift"):
        assert np.array_equal(rd[k].cpu().numpy(), np.asarray(rec[k])), k


def test_mirror_argument_six(oracle):
    say("case clmap-mirror")
    st = sim.make_stack(K=6, n=17, n_r=9, n_theta=24, snr=1.0, max_shift_2d=1, seed=5)
    clstack, corrstack, eqs, emap, mask = cryo_clmatrix(st.pf, -1, 0, 2, 1, 2)
    P, _ = oracle.preprocess(st.pf)
    ora = oracle.pairs(P, 2, 1, map_filter_radius=2)
    k1, k2 = np.triu_indices(6, 1)
    assert np.array_equal(clstack[k1, k2], ora["cl1"]) and np.array_equal(clstack[k2, k1], ora["cl2"])
    np.testing.assert_array_equal(corrstack[k1, k2], 1.0 - ora["best"])  # stored as 1 - value (:764)
    assert eqs.shape == (15, 13)


def test_radius_limits():
    from aspire_amd._lib import ClmError

    say("case clmap-limits")
    st = sim.make_stack(K=3, n=9, n_r=5, n_theta=8, snr=1.0, seed=1)
    with pytest.raises(ClmError, match="outside"):
        clmatrix_records(st.pf, 0, 1, map_filter_radius=11)
    with pytest.raises(ClmError, match="exceeds"):
        clmatrix_records(st.pf, 0, 1, map_filter_radius=5)  # 11 > 2L = 8
