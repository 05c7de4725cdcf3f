"""The 2-D shift least-squares oracle (oracle/shifts_oracle.py), pinned
against independent routines:

* its LSQR iterates equal scipy.sparse.linalg.lsqr's (an independent
  implementation of the same Paige-Saunders algorithm) after the same number
  of iterations, with and without damping, early stopping disabled in both;
* run to convergence it reaches the least-squares solution
  (numpy.linalg.lstsq), and with damp = 1 the solution of
  (A'A + I) x = A'b (numpy.linalg.solve);
* the stopping rule: with tol = 1e-8 it stops before maxit on a consistent
  system and reports flag 0.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import shifts_oracle as so


def random_system(K, n_theta, seed, noise=1.0):
    rng = np.random.default_rng(seed)
    P = K * (K - 1) // 2
    L = n_theta // 2
    cl1 = rng.integers(1, L + 1, size=P).astype(np.int32)
    cl2 = rng.integers(1, 2 * L + 1, size=P).astype(np.int32)
    shift = rng.normal(0, noise, size=P)
    return so.system(K, n_theta, cl1, cl2, shift), (cl1, cl2, shift)


@pytest.mark.parametrize("iters", [1, 5, 20])
@pytest.mark.parametrize("damp", [0.0, 1.0])
def test_iterates_match_scipy(iters, damp):
    (A, b), _ = random_system(60, 36, seed=iters)  # 120 unknowns: far from converged at 20
    x, info = so.lsqr(A, b, tol=0.0, maxit=iters, damp=damp)
    ref = spla.lsqr(A, b, damp=damp, atol=0.0, btol=0.0, conlim=0.0, iter_lim=iters)
    assert info["iterations"] == iters == ref[2]
    np.testing.assert_allclose(x, ref[0], rtol=0, atol=1e-10 * np.abs(ref[0]).max())
    # residual estimates agree with scipy's (r1norm = ||b - Ax|| for damp = 0, r2norm with damping)
    rn = info["relres"] * np.linalg.norm(b)
    np.testing.assert_allclose(rn, ref[4] if damp else ref[3], rtol=1e-9)


def test_converges_to_least_squares():
    (A, b), _ = random_system(20, 36, seed=3)
    x, info = so.lsqr(A, b, tol=1e-14, maxit=2000)
    Ad = A.toarray()
    ref = np.linalg.lstsq(Ad, b, rcond=None)[0]
    np.testing.assert_allclose(x, ref, rtol=0, atol=1e-7 * np.abs(ref).max())
    xd, _ = so.lsqr(A, b, tol=1e-14, maxit=2000, damp=1.0)
    refd = np.linalg.solve(Ad.T @ Ad + np.eye(Ad.shape[1]), Ad.T @ b)
    np.testing.assert_allclose(xd, refd, rtol=0, atol=1e-9 * np.abs(refd).max())


def test_stops_on_consistent_system():
    K, n_theta = 16, 36
    (A, _), _ = random_system(K, n_theta, seed=5)
    truth = np.random.default_rng(9).normal(0, 3, size=2 * K)
    b = A @ truth
    x, info = so.lsqr(A, b, tol=1e-8, maxit=500)
    assert info["flag"] == 0 and info["iterations"] < 500
    assert np.linalg.norm(A @ x - b) <= 1e-7 * np.linalg.norm(b)


def test_equation_rows_follow_reference_branches():
    # cryo_clmatrix_cpu.m:722-741: beta >= pi flips the k1 coefficients and
    # subtracts pi from beta; columns 2k1-1, 2k1, 2k2-1, 2k2
    K, n_theta = 3, 8
    L = n_theta // 2
    cl1 = np.array([2, 1, 3], np.int32)
    cl2 = np.array([3, 6, 1], np.int32)  # pair (0,2) has beta = 5*pi/4 >= pi
    A, _ = so.system(K, n_theta, cl1, cl2, np.zeros(3))
    A = A.toarray()
    d = np.pi / L
    a, bb = (cl1[1] - 1) * d, (cl2[1] - 1) * d - np.pi
    np.testing.assert_allclose(A[1, [0, 1, 4, 5]], [-np.sin(a), -np.cos(a), -np.sin(bb), -np.cos(bb)], atol=1e-15)
    a, bb = (cl1[0] - 1) * d, (cl2[0] - 1) * d
    np.testing.assert_allclose(A[0, [0, 1, 2, 3]], [np.sin(a), np.cos(a), -np.sin(bb), -np.cos(bb)], atol=1e-15)
