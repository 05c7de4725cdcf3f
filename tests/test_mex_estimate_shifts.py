"""The cryo_estimate_shifts MEX gateway
(aspire_amd/matlab/clm_estimate_shifts_mex.c), compiled against the stand-in
MEX API (tests/mex_shim), and the host functions it uses:

* CPU: the gateway compiles with -Wall -Werror and rejects bad arguments
  with the reference's message for memoryfactor
  (cryo_estimate_shifts.m:58-60), freeing its buffers; clm_n_equations
  follows :58-94 (hand-computed cases); clm_equation_pairs draws exactly
  min(n, P) pairs, as a function of the seed, each pair about equally often.
* GPU: the gateway's est_shifts and shift_equations equal the Python
  mirror's (aspire_amd.shifts.cryo_estimate_shifts, the same library calls
  through the device entry points) for the same seed, for all pairs and
  for a memoryfactor subset, double and single pf.
"""
import numpy as np
import pytest

from aspire_amd.shifts import equation_pairs, n_equations
from test_mex_gateway import call, dense, err, load_shim, mx, sparse


@pytest.fixture(scope="module")
def mex():
    return load_shim("libclm_estimate_shifts_mex_shim.so")


def test_argument_errors(mex):
    L = mex
    rc, _ = call(L, 2, mx(L, np.ones((9, 16, 3), complex)))
    assert rc == 1 and err(L) == ("clm:arg", "pf, rotations and max_shift are required")
    pf = mx(L, np.ones((9, 16, 3), complex))
    rot = mx(L, np.stack([np.eye(3)] * 3, axis=2))
    three = mx(L, 3 * np.ones(1))
    rc, _ = call(L, 2, pf, mx(L, np.zeros((3, 3, 2))), three)
    assert rc == 1 and "rotations" in err(L)[1]
    rc, _ = call(L, 2, mx(L, np.ones((9, 15, 3), complex)), rot, three)
    assert rc == 1 and err(L) == ("clm:arg", "n_theta must be even")
    rc, _ = call(L, 2, mx(L, np.ones((9, 16, 3))), rot, three)
    assert rc == 1 and err(L) == ("clm:arg", "pf must be complex")
    for bad in (50.0, -1.0):
        rc, _ = call(L, 2, pf, rot, three, mx(L, np.ones(1)), mx(L, bad * np.ones(1)))
        assert rc == 1 and err(L) == ("clm:arg", "subsamplingfactor must be between 0 and 1 or larger than 100")
    assert L.shim_live_blocks() == 0


def test_n_equations():
    # memoryfactor <= 1: the fraction of the K(K-1)/2 pairs, at least K
    assert n_equations(10, 0.3) == 14  # ceil(90 * 0.3 / 2) = ceil(13.5)
    assert n_equations(10, 1.0) == 45
    assert n_equations(10, 0.01) == 10  # ceil(0.45) = 1, raised to K
    assert n_equations(10, 0) == 10
    # memoryfactor in MB of the dense 45 x 20 system (7200 bytes): all pairs fit
    assert n_equations(10, 100) == 45
    # K = 1000: 499500 pairs, 7.992e9 bytes dense; 1000 MB is 62500 pairs in
    # exact arithmetic, but the reference's doubles (fl(1e9/7.992e9) =
    # 0.12512512512512514, times 999000, halved: 62500.00000000001) round up
    assert n_equations(1000, 1000) == 62501
    assert n_equations(1000, 500) == 31251  # 31250.000000000004 likewise
    assert n_equations(5000, 10000) == 125000  # the default at K = 5000
    for bad in (-0.5, 1.5, 99, float("nan")):
        with pytest.raises(ValueError):
            n_equations(10, bad)


def test_equation_pairs():
    K = 30
    P = K * (K - 1) // 2
    for n in (0, 1, 100, P - 1, P, P + 5):
        s = equation_pairs(K, n, seed=3)
        assert s.shape == (P,) and s.dtype == np.uint8 and int(s.sum()) == min(n, P)
    a, b = equation_pairs(K, 100, 7), equation_pairs(K, 100, 7)
    assert np.array_equal(a, b) and not np.array_equal(a, equation_pairs(K, 100, 8))
    # every pair equally likely: over 4000 seeds each of the 435 pairs is
    # drawn 100/435 of the time (binomial sd 27 around 920)
    cnt = np.zeros(P)
    for sd in range(4000):
        cnt += equation_pairs(K, 100, sd)
    assert abs(cnt.mean() - 4000 * 100 / P) < 1e-9
    assert cnt.min() > 920 - 5 * 27 and cnt.max() < 920 + 5 * 27


@pytest.mark.gpu
@pytest.mark.parametrize("dtype,memoryfactor", [(np.complex128, 10000), (np.complex64, 10000), (np.complex128, 0.3)])
def test_gateway_outputs_equal_mirror(mex, dtype, memoryfactor):
    from aspire_amd import sim
    from aspire_amd.shifts import cryo_estimate_shifts

    L = mex
    print("case mex-estimate-shifts", np.dtype(dtype).name, memoryfactor, flush=True)
    st = sim.make_stack(K=30, n=65, n_r=33, n_theta=360, snr=None, max_shift_2d=3, shift_kind="int", seed=1401)
    pf = st.pf.astype(dtype)
    rot = np.ascontiguousarray(st.rots.transpose(1, 0, 2))
    args = [mx(L, pf), mx(L, rot), mx(L, 9 * np.ones(1)), mx(L, np.ones(1)), mx(L, memoryfactor * np.ones(1)),
            mx(L, 5 * np.ones(1))]
    rc, out = call(L, 2, *args)
    assert rc == 0, err(L)
    est_m, eqs_m = dense(L, out[0]), sparse(L, out[1])
    est, eqs = cryo_estimate_shifts(pf, rot, 9, 1.0, memoryfactor, seed=5)
    assert eqs_m.shape == eqs.shape == (min(n_equations(30, memoryfactor), 435), 61)
    assert np.array_equal(est_m, est)
    assert np.array_equal(eqs_m.indptr, eqs.indptr) and np.array_equal(eqs_m.indices, eqs.indices)
    assert np.array_equal(eqs_m.data, eqs.data)
    for a in args + out[:2]:
        L.shim_destroy(a)
    assert L.shim_live_blocks() == 0
