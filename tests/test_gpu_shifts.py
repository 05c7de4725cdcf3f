"""Device LSQR of the shift equations (clm_estimate_shifts,
aspire_amd/csrc/shifts.hip) against the CPU oracle (oracle/shifts_oracle.py,
itself pinned against scipy's lsqr and numpy's least squares).

* the iterates after a fixed number of steps (stopping tests off) agree
  with the oracle's, with and without damping;
* with the reference's settings (tol 1e-8, maxit 100) the stopping point and
  the solution agree, for an inconsistent and a consistent system;
* edge cases: b = 0, maxit = 0, line numbers outside [0, n_theta] (zero rows);
* the full K=5000 workload (common lines from the BF16 engine): the solve's
  own residual estimates agree with the residuals recomputed from x by
  torch, and the normal-equation test holds;
* a clean stack: the shifts recovered from the engine's common lines match
  the true shifts outside the 3 near-null gauge directions.
"""
import sys

import numpy as np
import pytest

from aspire_amd.shifts import estimate_shifts
from oracle import shifts_oracle as so

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_case(K, n_theta, seed, consistent=False):
    rng = np.random.default_rng(seed)
    P = K * (K - 1) // 2
    L = n_theta // 2
    cl1 = rng.integers(0, L + 1, size=P).astype(np.int32)
    cl2 = rng.integers(0, 2 * L + 1, size=P).astype(np.int32)
    shift = rng.normal(0, 2.0, size=P)
    if consistent:
        A, _ = so.system(K, n_theta, cl1, cl2, shift)
        shift = A @ rng.normal(0, 3.0, size=2 * K)
    return cl1, cl2, shift


def dev(*arrs):
    import torch

    return [torch.as_tensor(a).cuda() for a in arrs]


@pytest.mark.parametrize("iters", [1, 5, 20])
@pytest.mark.parametrize("damp", [0.0, 1.0])
def test_iterates_match_oracle(iters, damp):
    say("case lsqr-iterates", iters, damp)
    K, n_theta = 60, 36
    cl1, cl2, b = random_case(K, n_theta, seed=iters)
    A, _ = so.system(K, n_theta, cl1, cl2, b)
    xo, io = so.lsqr(A, b, tol=0.0, maxit=iters, damp=damp)
    x, st = estimate_shifts(*dev(cl1, cl2, b), K, n_theta, tol=0.0, maxit=iters, damp=damp)
    x = x.cpu().numpy().ravel()
    say(st, io)
    assert st["iterations"] == io["iterations"] == iters and st["flag"] == 1
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-10 * np.abs(xo).max())
    assert abs(st["relres"] - io["relres"]) <= 1e-9 * io["relres"]


@pytest.mark.parametrize("K,n_theta", [(300, 36), (70, 1000)])
def test_tile_edges_and_global_tables(K, n_theta):
    # K = 300: a last row block of 44 images and a partial column block
    # (64 x 256 tiles); n_theta = 1000: coefficient tables too large for
    # LDS (5 (n_theta + 2) doubles > 32 KB), read from global memory
    say("case lsqr-tiles", K, n_theta)
    cl1, cl2, b = random_case(K, n_theta, seed=K + n_theta)
    A, _ = so.system(K, n_theta, cl1, cl2, b)
    xo, io = so.lsqr(A, b, tol=0.0, maxit=12, damp=0.0)
    x, st = estimate_shifts(*dev(cl1, cl2, b), K, n_theta, tol=0.0, maxit=12, damp=0.0)
    x = x.cpu().numpy().ravel()
    say(st, io)
    assert st["iterations"] == io["iterations"] == 12
    np.testing.assert_allclose(x, xo, rtol=0, atol=1e-10 * np.abs(xo).max())


@pytest.mark.parametrize("consistent", [False, True])
def test_reference_settings_match_oracle(consistent):
    say("case lsqr-default", consistent)
    K, n_theta = 120, 360
    cl1, cl2, b = random_case(K, n_theta, seed=11, consistent=consistent)
    A, _ = so.system(K, n_theta, cl1, cl2, b)
    xo, io = so.lsqr(A, b, tol=1e-8, maxit=100)
    x, st = estimate_shifts(*dev(cl1, cl2, b), K, n_theta)
    x = x.cpu().numpy().ravel()
    say(st, io)
    assert st["flag"] == io["flag"]
    assert abs(st["iterations"] - io["iterations"]) <= 1
    if st["iterations"] == io["iterations"]:
        np.testing.assert_allclose(x, xo, rtol=0, atol=1e-9 * np.abs(xo).max())
    if consistent:
        assert st["flag"] == 0 and st["iterations"] < 100
        assert np.linalg.norm(A @ x - b) <= 1e-7 * np.linalg.norm(b)


def test_edge_cases():
    import torch

    say("case lsqr-edges")
    K, n_theta = 9, 16
    cl1, cl2, b = random_case(K, n_theta, seed=2)
    x, st = estimate_shifts(*dev(cl1, cl2, np.zeros_like(b)), K, n_theta)
    assert st["iterations"] == 0 and st["flag"] == 0 and torch.count_nonzero(x).item() == 0
    x, st = estimate_shifts(*dev(cl1, cl2, b), K, n_theta, maxit=0)
    assert st["iterations"] == 0 and st["flag"] == 1 and torch.count_nonzero(x).item() == 0
    bad = cl1.copy()
    bad[[0, 5, 7]] = [-1, 17, 1000]  # outside [0, n_theta]: zero rows
    Az, _ = so.system(K, n_theta, cl1, cl2, b)
    Az = Az.tolil()
    for r in (0, 5, 7):
        Az[r, :] = 0
    xo, io = so.lsqr(Az.tocsr(), b, tol=0.0, maxit=10)
    x, st = estimate_shifts(*dev(bad, cl2, b), K, n_theta, tol=0.0, maxit=10)
    np.testing.assert_allclose(x.cpu().numpy().ravel(), xo, rtol=0, atol=1e-10 * np.abs(xo).max())


def test_full_workload_residuals():
    import torch

    from aspire_amd import sim
    from aspire_amd.clmatrix import clmatrix_records

    K, n_theta = 5000, 360
    say("case lsqr-full", K)
    st0 = sim.make_stack(K=K, n=129, n_r=65, n_theta=n_theta, snr=1 / 16, max_shift_2d=5.0, shift_kind="uniform",
                         seed=1003, device="cuda")
    rec, _ = clmatrix_records(st0.pf, 10, 0.5, precision="bf16")
    del st0
    x, st = estimate_shifts(rec["cl1"], rec["cl2"], rec["shift"], K, n_theta)
    say(st)
    # residuals recomputed from x with torch (independent of the solver's recurrences)
    L = n_theta // 2
    k1, k2 = torch.triu_indices(K, K, 1, device="cuda")
    d = np.pi / L
    a = (rec["cl1"].double() - 1) * d
    bt = (rec["cl2"].double() - 1) * d
    flip = bt >= np.pi
    bt = torch.where(flip, bt - np.pi, bt)
    sg = torch.where(flip, -1.0, 1.0).double()
    C = torch.stack([sg * torch.sin(a), sg * torch.cos(a), -torch.sin(bt), -torch.cos(bt)], 1)
    xf = x.reshape(-1)
    Ax = C[:, 0] * xf[2 * k1] + C[:, 1] * xf[2 * k1 + 1] + C[:, 2] * xf[2 * k2] + C[:, 3] * xf[2 * k2 + 1]
    r = rec["shift"] - Ax
    relres = (r.norm() / rec["shift"].norm()).item()
    atr = torch.zeros(2 * K, dtype=torch.float64, device="cuda")
    for j, col in enumerate((2 * k1, 2 * k1 + 1, 2 * k2, 2 * k2 + 1)):
        atr.index_add_(0, col, C[:, j] * r)
    normar = (atr.norm() / (torch.sqrt((C * C).sum()) * r.norm())).item()
    say("relres", relres, "estimate", st["relres"], "normal-eq test", normar)
    assert abs(relres - st["relres"]) <= 1e-6 * relres
    assert st["iterations"] <= 100
    if st["flag"] == 0:
        # converged through the normal-equation test (b is far from the range of A)
        assert normar <= 1e-6


def test_clean_stack_recovers_shifts():
    import torch

    from aspire_amd import sim
    from aspire_amd.clmatrix import clmatrix_records

    K, n_theta = 100, 360
    say("case lsqr-clean", K)
    st0 = sim.make_stack(K=K, n=65, n_r=33, n_theta=n_theta, snr=None, max_shift_2d=3, shift_kind="int", seed=31)
    rec, _ = clmatrix_records(st0.pf, 4, 1, precision="bf16")
    rec = {k: torch.as_tensor(np.asarray(v)).cuda() for k, v in rec.items()}
    x, st = estimate_shifts(rec["cl1"], rec["cl2"], rec["shift"], K, n_theta, tol=1e-10, maxit=500)
    say(st)
    A, _ = so.system(K, n_theta, rec["cl1"].cpu().numpy(), rec["cl2"].cpu().numpy(), rec["shift"].cpu().numpy())
    _, sv, Vt = np.linalg.svd(A.toarray(), full_matrices=False)
    N = Vt[-3:]  # the three weakest directions: the shift gauge
    proj = lambda y: y - N.T @ (N @ y)  # noqa: E731
    truth = st0.shifts.reshape(-1)
    err = proj(x.cpu().numpy().reshape(-1) - truth)
    rms_err = float(np.sqrt(np.mean(err ** 2)))
    rms_truth = float(np.sqrt(np.mean(proj(truth) ** 2)))
    say("singular values (smallest 4):", sv[-4:], "rms error", rms_err, "rms of the shifts", rms_truth)
    assert rms_err <= 0.25 * rms_truth
