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


# ---------------------------------------------------------------------------
# The per-pair loop of cryo_estimate_shifts (common lines of the rotations,
# the 1-D shift along each): pinned by the geometry the lines must satisfy
# and by the true shifts of a clean simulated stack.

def haar_rotations(K, seed):
    from scipy.spatial.transform import Rotation

    return Rotation.random(K, random_state=seed).as_matrix().transpose(1, 2, 0)  # 3 x 3 x K


def test_round_half_away_from_zero():
    x = np.array([0.0, 0.49999999999999994, 0.5, 1.5, 2.5, 3.4999999999999996, 359.5, 1e15 + 0.5])
    np.testing.assert_array_equal(so.round_half_away(x), [0, 0, 1, 2, 3, 3, 360, 1e15 + 1])


@pytest.mark.parametrize("n_theta", [3600, 150])
def test_rotation_lines_lie_in_both_image_planes(n_theta):
    # the line at angle t of image k (rotations as cryo_estimate_shifts takes
    # them) is the direction cos(t) R_k(:,1) + sin(t) R_k(:,2); the common
    # line is one direction in both planes, each index within half a grid
    # step of it, in the same orientation in both images: folding cij below
    # n_theta/2 reverses both lines (cryo_estimate_shifts.m:155-161), so
    # folding only one of them fails here
    K = 30
    R = haar_rotations(K, seed=n_theta)
    k1, k2 = np.triu_indices(K, 1)
    c1, c2 = so.lines_from_rotations(R, n_theta, k1, k2)
    L = n_theta // 2
    assert c1.min() >= 1 and c1.max() <= L and c2.min() >= 1 and c2.max() <= n_theta
    a = 2 * np.pi * (c1 - 1) / n_theta
    b = 2 * np.pi * (c2 - 1) / n_theta
    vi = np.cos(a) * R[:, 0, k1] + np.sin(a) * R[:, 1, k1]
    vj = np.cos(b) * R[:, 0, k2] + np.sin(b) * R[:, 1, k2]
    assert np.linalg.norm(vi - vj, axis=0).max() <= 2 * np.pi / n_theta + 1e-12
    # the planes' own intersection: n_k1 x n_k2 is orthogonal to both normals
    n1, n2 = R[:, 2, k1], R[:, 2, k2]
    cr = np.cross(n1.T, n2.T).T
    cr /= np.linalg.norm(cr, axis=0)
    assert np.minimum(np.linalg.norm(vi - cr, axis=0), np.linalg.norm(vi + cr, axis=0)).max() <= np.pi / n_theta + 1e-12


@pytest.fixture(scope="module")
def clean_stack():
    from aspire_amd import sim
    from oracle import clm_oracle as co

    K, n_theta = 30, 360
    st = sim.make_stack(K=K, n=65, n_r=33, n_theta=n_theta, snr=None, max_shift_2d=3, shift_kind="int", seed=31)
    P, _ = co.preprocess(st.pf)
    rot = np.ascontiguousarray(st.rots.transpose(1, 0, 2))  # estimated-rotation convention
    return K, n_theta, st, P, rot


def test_rotation_pairs_consistent_with_true_shifts(clean_stack):
    # on a clean stack the 1-D shifts of the oracle satisfy the equations of
    # the true 2-D shifts to the shift grid (step 1: |residual| <= 1/2 up to
    # interpolation); a wrong sign, swapped axes or a wrong C1/C2 choice fail
    from oracle import clm_oracle as co

    K, n_theta, st, P, rot = clean_stack
    k1, k2 = co.all_pairs(K)
    o = so.rotation_pairs(P, rot, n_theta, 9, 1.0, k1, k2)
    A, b = so.system(K, n_theta, o["cl1"], o["cl2"], o["shift"], flip_by_index=True)
    x = st.shifts.reshape(-1)
    r = A @ x - b
    assert np.mean(np.abs(r) <= 0.55) >= 0.95 and np.sqrt(np.mean(r ** 2)) <= 0.4
    for wrong in (-x, x.reshape(K, 2)[:, ::-1].ravel()):
        assert np.sqrt(np.mean((A @ wrong - b) ** 2)) > 2.0
    # the least squares recovers the shifts outside the 3-dimensional gauge
    xs, info = so.lsqr(A, b, tol=1e-8, maxit=100)
    _, _, Vt = np.linalg.svd(A.toarray(), full_matrices=False)
    N = Vt[-3:]
    err = (xs - x) - N.T @ (N @ (xs - x))
    assert np.sqrt(np.mean(err ** 2)) <= 0.2


def test_equations_by_index_flip_rule():
    # cryo_estimate_shifts flips image j's line when its index exceeds
    # n_theta/2; cryo_clmatrix tests beta < pi.  They part at cl2 = L + 1 when
    # L * fl(pi/L) rounds below pi (L = 75: n_theta = 150)
    L = 75
    n_theta = 2 * L
    d = np.pi / L
    assert L * d < np.pi
    coef, cols = so.equations_by_index(n_theta, [0, 0, 1], [1, 2, 2], [3, 3, 10], [L + 1, L, 2 * L])
    a = 2 * d
    np.testing.assert_array_equal(cols, [[1, 2, 3, 4], [1, 2, 5, 6], [3, 4, 5, 6]])
    np.testing.assert_allclose(coef[0], [-np.sin(a), -np.cos(a), -np.sin(L * d - np.pi), -np.cos(L * d - np.pi)],
                               rtol=0, atol=1e-15)
    np.testing.assert_allclose(coef[1], [np.sin(a), np.cos(a), -np.sin((L - 1) * d), -np.cos((L - 1) * d)],
                               rtol=0, atol=1e-15)
    A0, _ = so.system(3, n_theta, [3, 3, 10], [L + 1, L, 2 * L], np.zeros(3))
    A1, _ = so.system(3, n_theta, [3, 3, 10], [L + 1, L, 2 * L], np.zeros(3), flip_by_index=True)
    A0, A1 = A0.toarray(), A1.toarray()
    np.testing.assert_allclose(A0[1:], A1[1:], rtol=0, atol=1e-15)  # the rules agree elsewhere
    assert np.allclose(A0[0], -A1[0], atol=1e-15)  # and reverse the row at cl2 = L + 1
