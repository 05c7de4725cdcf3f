"""cryo_estimate_shifts on the GPU (libclm ``clm_common_lines_from_rotations``
and ``clm_estimate_shifts_ex``) against the oracle (oracle/shifts_oracle.py:
``rotation_pairs``, pinned by the planes' geometry and by the true shifts of
a clean stack; ``system(..., flip_by_index=True)`` and ``lsqr``):

* the common lines of the rotations and the 1-D shift along each equal the
  oracle's exactly (FP64, same arithmetic in the same order), at n_theta =
  360 and 150 (where the index and angle rules for the second line part),
  shift steps 1 and 0.5, with the stack given as a host array or as a
  column-major device tensor;
* a pair selection: the selected pairs' outputs equal the full call's, the
  others are cl1 = cl2 = -1 and shift 0;
* the smallest stacks (K = 2, 3) and a pair of equal rotations (Ut the
  identity up to rounding) equal the oracle too;
* the drop-in function: its equations equal the oracle's, its estimate the
  oracle's LSQR of them (tol 1e-8, maxit 100), and on a clean stack the
  true shifts outside the 3-dimensional gauge; with memoryfactor <= 1 the
  seeded subset has n_equations rows and the same agreement.
"""
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg  # noqa: F401  (sp.linalg.norm)

from aspire_amd import _lib, sim
from aspire_amd.shifts import common_lines_from_rotations, cryo_estimate_shifts, estimate_shifts, n_equations
from oracle import clm_oracle as co
from oracle import shifts_oracle as so

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_STACKS = {}


def stack(n_theta, K=40):
    key = (n_theta, K)
    if key not in _STACKS:
        st = sim.make_stack(K=K, n=65, n_r=33, n_theta=n_theta, snr=None, max_shift_2d=3, shift_kind="int",
                            seed=50 + n_theta)
        rot = np.ascontiguousarray(st.rots.transpose(1, 0, 2))  # the estimated rotations' convention
        P, _ = co.preprocess(st.pf)
        _STACKS[key] = (st, rot, P)
    return _STACKS[key]


@pytest.mark.parametrize("n_theta,max_shift,step", [(360, 9, 1.0), (150, 9, 1.0), (360, 4, 0.5)])
def test_rotation_lines_match_oracle(n_theta, max_shift, step):
    say("case rotlines", n_theta, max_shift, step)
    st, rot, P = stack(n_theta)
    K = rot.shape[2]
    k1, k2 = co.all_pairs(K)
    ora = so.rotation_pairs(P, rot, n_theta, max_shift, step, k1, k2)
    cl1, cl2, sh = common_lines_from_rotations(st.pf, rot, max_shift, step)
    assert np.array_equal(cl1.cpu().numpy(), ora["cl1"])
    assert np.array_equal(cl2.cpu().numpy(), ora["cl2"])
    assert np.array_equal(sh.cpu().numpy(), ora["shift"])
    if n_theta == 150:
        assert (ora["cl2"] == n_theta // 2 + 1).any()  # rows the two flip rules build differently


def test_rotation_lines_phases_from_global_memory():
    # 161 shifts x 64 ray points: the phase table (165 KB) exceeds the 64 KB
    # staged in LDS, and a lane walks three shifts
    say("case rotlines-global-phases")
    st = sim.make_stack(K=12, n=129, n_r=65, n_theta=360, snr=None, max_shift_2d=3, shift_kind="int", seed=77)
    rot = np.ascontiguousarray(st.rots.transpose(1, 0, 2))
    P, _ = co.preprocess(st.pf)
    k1, k2 = co.all_pairs(12)
    ora = so.rotation_pairs(P, rot, 360, 40, 0.5, k1, k2)
    cl1, cl2, sh = common_lines_from_rotations(st.pf, rot, 40, 0.5)
    assert np.array_equal(cl1.cpu().numpy(), ora["cl1"])
    assert np.array_equal(cl2.cpu().numpy(), ora["cl2"])
    assert np.array_equal(sh.cpu().numpy(), ora["shift"])


def test_rotation_lines_device_stack_and_subset():
    import torch

    say("case rotlines-subset")
    st, rot, P = stack(360)
    K = rot.shape[2]
    Pn = K * (K - 1) // 2
    full = [t.cpu().numpy() for t in common_lines_from_rotations(st.pf, rot, 9, 1.0)]
    base = torch.as_tensor(np.ascontiguousarray(st.pf.transpose(2, 1, 0))).cuda()
    dev = [t.cpu().numpy() for t in common_lines_from_rotations(base.permute(2, 1, 0), torch.as_tensor(rot), 9, 1.0)]
    for a, b in zip(full, dev):
        assert np.array_equal(a, b)
    sel = (np.random.default_rng(3).uniform(size=Pn) < 0.3).astype(np.uint8)
    sub = [t.cpu().numpy() for t in common_lines_from_rotations(st.pf, rot, 9, 1.0, sel=sel)]
    m = sel.astype(bool)
    for a, b in zip(full, sub):
        assert np.array_equal(a[m], b[m])
    assert (sub[0][~m] == -1).all() and (sub[1][~m] == -1).all() and (sub[2][~m] == 0).all()


def _oracle_estimate(P, rot, n_theta, max_shift, step, keep):
    K = rot.shape[2]
    k1, k2 = co.all_pairs(K)
    k1, k2 = k1[keep], k2[keep]
    o = so.rotation_pairs(P, rot, n_theta, max_shift, step, k1, k2)
    A, b = so.system(K, n_theta, o["cl1"], o["cl2"], o["shift"], flip_by_index=True, k1=k1, k2=k2)
    x, info = so.lsqr(A, b, tol=1e-8, maxit=100)
    return A, b, x, info


def _gauge_error(A, x, truth):
    _, _, Vt = np.linalg.svd(A.toarray(), full_matrices=False)
    N = Vt[-3:]
    d = x - truth
    return float(np.sqrt(np.mean((d - N.T @ (N @ d)) ** 2)))


@pytest.mark.parametrize("n_theta", [360, 150])
def test_cryo_estimate_shifts_clean_stack(n_theta):
    say("case estimate-shifts", n_theta)
    st, rot, P = stack(n_theta)
    K = rot.shape[2]
    Pn = K * (K - 1) // 2
    est, eqs = cryo_estimate_shifts(st.pf, rot, 9, 1.0)
    # the same steps one at a time, for the solve's statistics
    x, stats = estimate_shifts(*common_lines_from_rotations(st.pf, rot, 9, 1.0), K, n_theta,
                               flags=_lib.CLM_FLIP_BY_INDEX)
    assert np.array_equal(x.cpu().numpy(), est)
    A, b, xo, io = _oracle_estimate(P, rot, n_theta, 9, 1.0, np.ones(Pn, bool))
    ref = sp.hstack([A, sp.csr_matrix(b.reshape(-1, 1))], format="csc")
    ref.eliminate_zeros()
    assert eqs.shape == (Pn, 2 * K + 1)
    assert np.array_equal(eqs.indptr, ref.indptr) and np.array_equal(eqs.indices, ref.indices)
    np.testing.assert_allclose(eqs.data, ref.data, rtol=0, atol=4e-16)
    assert est.shape == (K, 2)
    say(stats, io)
    assert stats["flag"] == io["flag"] and abs(stats["iterations"] - io["iterations"]) <= 1
    if stats["iterations"] == io["iterations"]:
        np.testing.assert_allclose(est.ravel(), xo, rtol=0, atol=1e-9 * np.abs(xo).max())
    err = _gauge_error(A, est.ravel(), st.shifts.ravel())
    say("rms error outside the gauge", err)
    assert err <= 0.2


def test_cryo_estimate_shifts_memoryfactor():
    say("case estimate-shifts-subset")
    st, rot, P = stack(360)
    K = rot.shape[2]
    Pn = K * (K - 1) // 2
    est, eqs = cryo_estimate_shifts(st.pf, rot, 9, 1.0, memoryfactor=0.3, seed=5)
    assert eqs.shape == (n_equations(K, 0.3), 2 * K + 1)
    est2, eqs2 = cryo_estimate_shifts(st.pf, rot, 9, 1.0, memoryfactor=0.3, seed=5)
    assert np.array_equal(est, est2) and (eqs != eqs2).nnz == 0
    # the rows are the selected pairs in pair order: recover the selection
    # from the columns of each row's first image
    E = eqs.tocsr()
    k1, k2 = co.all_pairs(K)
    first = {(int(a), int(b)): e for e, (a, b) in enumerate(zip(k1, k2))}
    keep = np.zeros(Pn, bool)
    for r in range(E.shape[0]):
        cols = E.indices[E.indptr[r]:E.indptr[r + 1]]
        cols = np.sort(cols[cols < 2 * K])
        keep[first[(int(cols[0]) // 2, int(cols[-1]) // 2)]] = True
    assert keep.sum() == E.shape[0]
    A, b, xo, io = _oracle_estimate(P, rot, 360, 9, 1.0, keep)
    np.testing.assert_allclose(E[:, :-1].toarray(), A.toarray(), rtol=0, atol=4e-16)
    np.testing.assert_array_equal(E[:, -1].toarray().ravel(), b)
    sel = keep.astype(np.uint8)
    x, stats = estimate_shifts(*common_lines_from_rotations(st.pf, rot, 9, 1.0, sel=sel), K, 360,
                               flags=_lib.CLM_FLIP_BY_INDEX)
    assert np.array_equal(x.cpu().numpy(), est)
    say(stats, io)
    assert stats["flag"] == io["flag"] and abs(stats["iterations"] - io["iterations"]) <= 1
    if stats["iterations"] == io["iterations"]:
        # 234 equations in 80 unknowns, the gauge's singular values near 0.01
        # against 3.5 at the top: over 60-70 iterations the bidiagonalization
        # loses orthogonality, and the two solves (their sums in different
        # orders) take slightly different paths (their running norm estimates
        # can part in the fourth digit) to the same solution, within what the
        # stopping test (tol 1e-8) fixes of it (scipy's lsqr lands about 4e-10
        # of the largest entry from the oracle on such a subset)
        np.testing.assert_allclose(est.ravel(), xo, rtol=0, atol=1e-7 * np.abs(xo).max())
        r = b - A @ est.ravel()
        ro = b - A @ xo
        assert abs(np.linalg.norm(r) - np.linalg.norm(ro)) <= 1e-12 * np.linalg.norm(b)
        assert np.linalg.norm(A.T @ r) <= 1e-7 * sp.linalg.norm(A) * np.linalg.norm(r)


@pytest.mark.parametrize("K", [2, 3])
def test_smallest_stacks_and_repeated_rotation(K):
    # K = 2: one pair (the reference would index past its one pair, since
    # n_equations(2, .) = 2; here every pair is used); K = 3 with image 1's
    # rotation equal to image 0's (a degenerate pair: Ut is the identity up to
    # rounding)
    say("case estimate-shifts-small", K)
    st, rot, P = stack(360)
    rot = np.ascontiguousarray(rot[:, :, :K])
    if K == 3:
        rot[:, :, 1] = rot[:, :, 0]
    pf = np.ascontiguousarray(st.pf[:, :, :K])
    Pk, _ = co.preprocess(pf)
    k1, k2 = co.all_pairs(K)
    ora = so.rotation_pairs(Pk, rot, 360, 9, 1.0, k1, k2)
    cl1, cl2, sh = common_lines_from_rotations(pf, rot, 9, 1.0)
    assert np.array_equal(cl1.cpu().numpy(), ora["cl1"]) and np.array_equal(cl2.cpu().numpy(), ora["cl2"])
    assert np.array_equal(sh.cpu().numpy(), ora["shift"])
    est, eqs = cryo_estimate_shifts(pf, rot, 9, 1.0)
    assert est.shape == (K, 2) and np.isfinite(est).all() and eqs.shape == (K * (K - 1) // 2, 2 * K + 1)
