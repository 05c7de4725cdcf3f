"""Pins of the FP64 oracle (CPU only).

The reference ships no stored outputs for cryo_clmatrix, so the oracle is
pinned by checks that a plausible slip would fail:
* an independent numpy restatement on tiny inputs (indexing, tie order);
* planted common lines whose answer is fixed in closed form (C = 1 exactly
  at a known (l1, l2, shift) and both orientations);
* geometric truth: clean projections at known rotations and 2-D shifts must
  give the common lines of commonline_R and the 1-D shifts of the
  reference's own formula (cryo_clmatrix_cpu.m:667-683);
* the reference's own data (examples/p100_shifted.mat, read when present);
* exact ties (radially symmetric images) and an all-negative pair (cl = 0);
* least squares on shift_equations recovers the true 2-D shifts modulo the
  3-dimensional null space (test_cryo_estimate_shifts1.m:71-76).
"""
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp

from aspire_amd import sim
from oracle import clm_oracle_np as onp

REF_P100 = "/root/reference/examples/p100_shifted.mat"


def random_pf(rng, n_r, n_theta, K, conj_sym=True):
    L = n_theta // 2
    pf = np.empty((n_r, n_theta, K), dtype=np.complex128, order="F")
    half = rng.standard_normal((n_r, L, K)) + 1j * rng.standard_normal((n_r, L, K))
    pf[:, :L, :] = half
    pf[:, L:, :] = np.conj(half) if conj_sym else rng.standard_normal((n_r, L, K)) + 1j * rng.standard_normal((n_r, L, K))
    return pf


@pytest.mark.parametrize("n_r,n_theta,K,max_shift,step,conj_sym", [
    (9, 16, 5, 2, 1, True),
    (9, 16, 4, 1, 0.3, True),      # asymmetric shift grid
    (17, 36, 4, 0, 1, True),       # single shift
    (12, 20, 4, 1.5, 0.5, False),  # general (not conjugate-symmetric) pf
])
def test_oracle_matches_numpy_restatement(oracle, n_r, n_theta, K, max_shift, step, conj_sym):
    rng = np.random.default_rng(7)
    pf = random_pf(rng, n_r, n_theta, K, conj_sym)
    cl_c, corr_c, eq_c, map_c, mask = oracle.cryo_clmatrix_cpu(pf, max_shift=max_shift, shift_step=step)
    cl_n, corr_n, eq_n, map_n = onp.clmatrix(pf, max_shift, step)
    assert np.array_equal(cl_c, cl_n)
    np.testing.assert_allclose(corr_c, corr_n, rtol=0, atol=1e-13)
    np.testing.assert_allclose(eq_c.toarray(), eq_n, rtol=0, atol=1e-15)
    assert np.array_equal(map_c, map_n)
    assert not mask.any()


def test_oracle_preprocess_closed_form(oracle):
    """Filter, band zeroing and normalization (cryo_clmatrix_cpu.m:277-293):
    each full line has unit norm, so a conjugate-symmetric line's half has
    squared norm 1/2; rows rk = -1 are zero; rows are pf(n_r-1-r, L+l)*H/|line|."""
    rng = np.random.default_rng(1)
    n_r, T, K = 13, 20, 3
    pf = random_pf(rng, n_r, T, K)
    P, n_small = oracle.preprocess(pf)
    assert n_small == 0
    R, L = n_r - 1, T // 2
    np.testing.assert_allclose((np.abs(P) ** 2).sum(axis=2), 0.5, rtol=1e-13)
    assert np.all(P[:, :, R - 1] == 0)
    rk = np.arange(-R, R + 1)
    H = np.sqrt(np.abs(rk)) * np.exp(-rk ** 2 / (2 * (R / 4) ** 2))
    H[R - 1:R + 2] = 0
    k, l, r = 1, 4, 3
    full = np.concatenate([pf[1:, L + l, k][::-1], pf[:, l, k]]) * H
    expect = pf[n_r - 1 - r, L + l, k] * H[r] / np.linalg.norm(full)
    assert abs(P[k, l, r] - expect) < 1e-15


def plant(pf, k1, k2, l1, l2, shift, R, n_r, L, flipped=False):
    """Make line l2 of image k2 equal to line l1 of image k1 (or its
    conjugate) times the phase of `shift`, so C1 (or C2) = 1 there."""
    r = np.arange(R)
    ph = np.exp(-2j * np.pi * (r - R) * shift / (2 * R + 1))
    m = n_r - 1 - r  # pf rows of the half line
    src = pf[m, L + l1, k1]
    if flipped:
        src = np.conj(src)
    pf[m, L + l2, k2] = src * ph  # half lines are pf(n_r-1-r, L+l)*H/|line|
    pf[:, l2, k2] = np.conj(pf[:, L + l2, k2])


@pytest.mark.parametrize("flipped", [False, True])
def test_oracle_planted_common_line(oracle, flipped):
    rng = np.random.default_rng(11 + flipped)
    n_r, T, K, max_shift, step = 17, 40, 3, 3, 0.5
    R, L = n_r - 1, T // 2
    pf = random_pf(rng, n_r, T, K)
    S = oracle.n_shifts(max_shift, step)
    for trial in range(4):
        l1, l2 = rng.integers(0, L, 2)
        j = int(rng.integers(0, S))
        shift = -max_shift + j * step
        pf2 = pf.copy(order="F")
        plant(pf2, 0, 2, l1, l2, shift, R, n_r, L, flipped)
        P, _ = oracle.preprocess(pf2)
        res = oracle.pairs(P, max_shift, step, np.array([0], np.int32), np.array([2], np.int32))
        assert res["cl1"][0] == l1 + 1
        assert res["cl2"][0] == l2 + 1 + (L if flipped else 0)
        assert res["j"][0] == j
        assert res["shift"][0] == shift
        assert abs(res["best"][0] - 1.0) < 1e-13


def test_oracle_exact_ties_and_negative_pair(oracle):
    """Radially symmetric images: every line is the same vector, so all
    entries of C1 at a shift tie exactly; the reference keeps the first
    maximum in column-major order at the first shift attaining it
    (cryo_clmatrix_cpu.m:387-399) -> (cl1, cl2) = (1, 1).  An image and its
    negative with real rays correlate to -1 everywhere -> no update, cl = 0."""
    n_r, T = 11, 16
    f = np.linspace(1.0, 0.2, n_r)
    pf = np.empty((n_r, T, 3), dtype=np.complex128, order="F")
    pf[:, :, 0] = f[:, None]
    pf[:, :, 1] = (0.5 * f)[:, None]
    pf[:, :, 2] = -f[:, None]
    P, _ = oracle.preprocess(pf)
    res = oracle.pairs(P, 0, 1)
    # pairs (0,1), (0,2), (1,2)
    assert (res["cl1"][0], res["cl2"][0], res["j"][0]) == (1, 1, 0)
    assert abs(res["best"][0] - 1.0) < 1e-14
    for i in (1, 2):
        assert (res["cl1"][i], res["cl2"][i], res["j"][i]) == (0, 0, -1)
        assert res["best"][i] == 0.0
    cl, corr, eqs, emap, _ = oracle.cryo_clmatrix_cpu(pf, max_shift=0, shift_step=1)
    assert corr[0, 2] == 0 and corr[1, 2] == 0
    # cl = 0 still yields an equation (alpha = beta = -pi/L)   :724-738
    a = -math.pi / (T // 2)
    np.testing.assert_allclose(eqs[1, :6].toarray().ravel(),
                               [math.sin(a), math.cos(a), 0, 0, -math.sin(a), -math.cos(a)], atol=1e-15)


def test_oracle_nan_lines_ignored(oracle):
    """A zero line normalizes to NaN; MATLAB's max ignores NaN entries."""
    rng = np.random.default_rng(3)
    pf = random_pf(rng, 9, 16, 3)
    pf[:, 2, 1] = 0
    pf[:, 10, 1] = 0
    P, n_small = oracle.preprocess(pf)
    assert n_small == 1
    res = oracle.pairs(P, 1, 1)
    cl_n, corr_n, _, _ = onp.clmatrix(pf, 1, 1)
    k1, k2 = oracle.all_pairs(3)
    assert np.array_equal(res["cl1"], cl_n[k1, k2].astype(np.int32))
    assert np.array_equal(res["cl2"], cl_n[k2, k1].astype(np.int32))
    # line 3 of image 1 is NaN: never chosen, in either orientation
    assert res["cl2"][0] not in (3, 3 + 8)   # pair (0,1): image 1 is k2
    assert res["cl1"][2] != 3                # pair (1,2): image 1 is k1


@pytest.fixture(scope="module")
def clean_shifted_stack():
    return sim.make_stack(K=10, n=89, n_r=45, n_theta=360, snr=None, max_shift_2d=4,
                          shift_kind="int", seed=1)


def test_oracle_geometry_clean(oracle, clean_shifted_stack):
    """test_cryo_clmatrix_cpu_1.m: clean odd images with 2-D shifts up to 4
    px, 1-D search step 0.5 up to ceil(2*sqrt(2)*4) -> detection rate
    "very close to 1" at 5 degrees; the detected 1-D shifts agree with the
    reference's formula for the true shift along the detected lines."""
    st = clean_shifted_stack
    cl, corr, eqs, emap, _ = oracle.cryo_clmatrix_cpu(st.pf, max_shift=12, shift_step=0.5)
    ref = sim.clmatrix_cheat(st.rots, 360)
    assert sim.comparecl(cl, ref, 360, 5) >= 0.95
    k1, k2 = np.triu_indices(10, 1)
    rhs = eqs[:, -1].toarray().ravel()
    err = [rhs[int(emap[a, b]) - 1] - sim.true_shift_1d(int(cl[a, b]), int(cl[b, a]), 180,
                                                        st.shifts[a], st.shifts[b])
           for a, b in zip(k1, k2)]
    assert np.median(np.abs(err)) <= 0.25
    assert np.mean(np.abs(err) <= 0.5) >= 0.95


def test_oracle_shift_equations_recover_shifts(oracle, clean_shifted_stack):
    """LS on shift_equations recovers the 2-D shifts up to the 3-dim null space
    (test_cryo_estimate_shifts1.m:60-76)."""
    st = clean_shifted_stack
    cl, corr, eqs, emap, _ = oracle.cryo_clmatrix_cpu(st.pf, max_shift=12, shift_step=0.5)
    A = eqs[:, :-1].toarray()
    b = eqs[:, -1].toarray().ravel()
    s2 = np.linalg.lstsq(A, b, rcond=None)[0]
    _, S, Vt = np.linalg.svd(A, full_matrices=False)
    # the three smallest singular values belong to the projected 3-D
    # translation (exactly null only for exact common lines)
    assert S[-3] < 0.05 * S[0] and S[-4] > 3 * S[-3]
    V = Vt[:-3].T
    s1 = st.shifts.reshape(-1)
    rel = np.linalg.norm(V.T @ (s1 - s2)) / np.linalg.norm(V.T @ s1)
    assert rel < 0.1
    # flipping the sign convention would fail
    assert np.linalg.norm(V.T @ (s1 + s2)) / np.linalg.norm(V.T @ s1) > 1.0


def test_oracle_equation_rows_from_clstack(oracle):
    """shift_equations rebuilt from clstack are identical
    (test_cryo_estimate_shifts_ref.m:66-68); closed-form row values."""
    rng = np.random.default_rng(5)
    K, L = 6, 18
    k1, k2 = oracle.all_pairs(K)
    cl1 = rng.integers(0, L + 1, k1.size).astype(np.int32)
    cl2 = rng.integers(0, 2 * L + 1, k1.size).astype(np.int32)
    coef, cols = oracle.equations(L, k1, k2, cl1, cl2)
    d = math.pi / L
    for i in range(k1.size):
        a, b = (cl1[i] - 1) * d, (cl2[i] - 1) * d
        if b < math.pi:
            e = [math.sin(a), math.cos(a), -math.sin(b), -math.cos(b)]
        else:
            b -= math.pi
            e = [-math.sin(a), -math.cos(a), -math.sin(b), -math.cos(b)]
        assert np.array_equal(coef[i], e)
        assert list(cols[i]) == [2 * k1[i] + 1, 2 * k1[i] + 2, 2 * k2[i] + 1, 2 * k2[i] + 2]


@pytest.mark.skipif(not os.path.exists(REF_P100), reason="reference fixture not present")
def test_oracle_on_reference_fixture(oracle):
    """examples/p100_shifted.mat (ASPIRE's own projections, rotations and
    shifts): the projections are stored centred; detection against
    clmatrix_cheat(rots, 72) at 10 degrees (examples/abinitio_C1_1.m:30-57)
    and, with the stored shifts applied by cryo_addshifts' convention, the
    detected 1-D shifts follow the reference's formula."""
    import scipy.io as sio

    d = sio.loadmat(REF_P100)
    projs = np.transpose(d["projections"], (2, 0, 1))[:40]
    rots = d["rots"][:, :, :40]
    shifts = d["shifts"][:40]
    pf = sim.cryo_pft(projs, 33, 72)
    cl, corr, eqs, emap, _ = oracle.cryo_clmatrix_cpu(pf, max_shift=15, shift_step=1)
    assert sim.comparecl(cl, sim.clmatrix_cheat(rots, 72), 72, 10) >= 0.97
    assert np.all(np.abs(eqs[:, -1].toarray()) <= 1)
    pf2 = sim.cryo_pft(sim.add_shifts(projs, shifts), 33, 72)
    cl, corr, eqs, emap, _ = oracle.cryo_clmatrix_cpu(pf2, max_shift=15, shift_step=1)
    assert sim.comparecl(cl, sim.clmatrix_cheat(rots, 72), 72, 10) >= 0.95
    k1, k2 = np.triu_indices(40, 1)
    rhs = eqs[:, -1].toarray().ravel()
    err = np.array([rhs[int(emap[a, b]) - 1] - sim.true_shift_1d(int(cl[a, b]), int(cl[b, a]), 36,
                                                                 shifts[a], shifts[b])
                    for a, b in zip(k1, k2)])
    assert np.mean(np.abs(err) <= 1.0) >= 0.95


def test_pft_closed_form():
    """cryo_pft restatement: a single bright pixel at (x0, y0) (grid units)
    has pf(k, j) = exp(i*k*w0*(x0*sin(j*dt) + y0*cos(j*dt))), and the second
    half is the conjugate of the first (cryo_pft.m:55-75, nudft2.m:40-47)."""
    n, n_r, T = 9, 5, 12
    img = np.zeros((1, n, n))
    img[0, 6, 2] = 1.0  # x0 = 2, y0 = -2 on the centred grid
    pf = sim.cryo_pft(img, n_r, T)
    w0 = 2 * np.pi / (2 * n_r - 1)
    dt = 2 * np.pi / T
    k = np.arange(n_r)[:, None]
    j = np.arange(T // 2)[None, :]
    expect = np.exp(1j * k * w0 * (2 * np.sin(j * dt) - 2 * np.cos(j * dt)))
    np.testing.assert_allclose(pf[:, : T // 2, 0], expect, atol=1e-13)
    np.testing.assert_allclose(pf[:, T // 2:, 0], np.conj(expect), atol=1e-13)


def test_shift_theorem_closed_form():
    """cryo_addshifts convention: an integer shift (sx, sy) moves image
    content to (x+sx, y+sy) (circularly)."""
    rng = np.random.default_rng(0)
    img = rng.standard_normal((1, 9, 9))
    out = sim.add_shifts(img, np.array([[2.0, -1.0]]))
    np.testing.assert_allclose(out[0], np.roll(img[0], (2, -1), axis=(0, 1)), atol=1e-12)


def test_rotations_haar():
    rng = np.random.default_rng(2)
    R = sim.rand_rots(rng, 200)
    for k in range(0, 200, 37):
        np.testing.assert_allclose(R[:, :, k] @ R[:, :, k].T, np.eye(3), atol=1e-13)
        assert abs(np.linalg.det(R[:, :, k]) - 1) < 1e-12
    q = sim.qrand(np.random.default_rng(2), 200)
    assert np.all(q[0] >= 0)
