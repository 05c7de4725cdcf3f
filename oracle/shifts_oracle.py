"""CPU oracle of the 2-D shift least squares after cryo_clmatrix.

TEST INFRASTRUCTURE ONLY: the checker for libclm's ``clm_estimate_shifts``.
Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this module; the product path never does.

Follows sinograms/cryo_estimate_shifts.m:225-230 applied to the shift
equations that cryo_clmatrix returns (cryo_clmatrix_cpu.m:722-741,766-772):

    est_shifts = lsqr(shift_equations(:,1:end-1), shift_equations(:,end), 1.0e-8, 100);
    est_shifts = full(transpose(reshape(est_shifts, 2, n_projs)));

``lsqr`` is MATLAB's (closed source); this is the published algorithm it
implements -- C. C. Paige and M. A. Saunders, "LSQR: An algorithm for sparse
linear equations and sparse least squares", ACM TOMS 8(1):43-71, 1982 --
started from x0 = 0, with the damping term of the paper (damp = 1 is the
``eye(2*K)`` block of sinograms/test_shifts.m:24-27, without its three gauge
rows), and with the two stopping tests MATLAB documents for lsqr:
||b - A x|| <= tol ||b||, or ||A'(b - A x)|| <= tol ||A||_F ||b - A x||,
both on the algorithm's running estimates, at most maxit iterations.
The matrix rows come from the FP64 C oracle's restatement of
cryo_clmatrix_cpu.m:722-741 (``clm_oracle.equations``), or, with
``flip_by_index``, from :func:`equations_by_index` (cryo_estimate_shifts.m:197-222).

The per-pair loop of sinograms/cryo_estimate_shifts.m:142-223 (the common
line of two rotations and the 1-D shift along it) is restated by
:func:`lines_from_rotations` and :func:`rotation_pairs`: the geometry of
sinograms/commonline_R.m:22-35 in numpy (each entry of Ut = Rj.' * Ri summed
as (m=1 + m=2) + m=3; MATLAB's product does not document its order), and the
correlations C1, C2 of :190-195 as the C oracle's ``clmo_entry`` adds them
(the per-ray arithmetic the cryo_clmatrix oracle is pinned with).
"""
from __future__ import annotations

import numpy as np

from . import clm_oracle


def round_half_away(x):
    """MATLAB's round (halves away from zero) of non-negative doubles."""
    x = np.asarray(x, dtype=np.float64)
    f = np.floor(x)
    return f + ((x - f) >= 0.5)  # x - floor(x) is exact below 2^52


def lines_from_rotations(rotations, n_theta: int, k1, k2):
    """(cl1, cl2): 1-based common lines of pairs (k1, k2) as
    cryo_estimate_shifts.m:151-167 finds them: commonline_R(Ri.', Rj.', n_theta)
    (sinograms/commonline_R.m:22-35: Ut = Rj.' * Ri; atan2 of two entries
    each, shifted by pi, scaled to n_theta, rounded half away from zero, mod
    n_theta), then cij folded below n_theta/2 (cji with it, wrapped to
    [0, n_theta)) and both made 1-based.  ``rotations`` is 3 x 3 x K."""
    R = np.asarray(rotations, dtype=np.float64)
    k1 = np.asarray(k1, dtype=np.int64)
    k2 = np.asarray(k2, dtype=np.int64)
    Ri, Rj = R[:, :, k1], R[:, :, k2]

    def ut(a, b):  # (Rj.' * Ri)(a, b), 0-based
        return (Rj[0, a] * Ri[0, b] + Rj[1, a] * Ri[1, b]) + Rj[2, a] * Ri[2, b]

    PI = np.pi  # 4*atan(1.0), commonline_R.m:27
    aij = np.arctan2(ut(2, 0), -ut(2, 1)) + PI
    aji = np.arctan2(-ut(0, 2), ut(1, 2)) + PI
    cij = np.mod(round_half_away(aij / (2 * PI) * n_theta), n_theta).astype(np.int64)
    cji = np.mod(round_half_away(aji / (2 * PI) * n_theta), n_theta).astype(np.int64)
    half = n_theta // 2
    fold = cij >= half  # :155-158
    cij = np.where(fold, cij - half, cij)
    cji = np.where(fold, cji - half, cji)
    cji = np.where(cji < 0, cji + n_theta, cji)  # :159-161
    return (cij + 1).astype(np.int32), (cji + 1).astype(np.int32)


def _first_max(c):
    """MATLAB's [v, i] = max(c): NaN ignored, the first of equal maxima;
    all NaN gives (NaN, 0)."""
    best, idx = np.nan, 0
    for i, v in enumerate(c):
        if not np.isnan(v) and (np.isnan(best) or v > best):
            best, idx = v, i
    return best, idx


def rotation_pairs(P, rotations, n_theta: int, max_shift: float, shift_step: float, k1, k2):
    """cryo_estimate_shifts.m:142-204 for the pairs (k1, k2): dict of int32
    ``cl1``, ``cl2`` (1-based) and float64 ``shift`` (dx).  ``P`` are the
    filtered, normalized half rays of ``clm_oracle.preprocess`` (the
    preparation of :166-188 is cryo_clmatrix's).  C1 and C2 (:190-195) hold,
    for every shift, the correlation of the line of image k1 (conjugated for
    C2) with line cji of image k2 (cji - n_theta/2 when above n_theta/2); dx
    is the shift of the first maximum of C1 if it exceeds C2's, else of C2's
    (:197-204)."""
    K, L, R = P.shape
    if 2 * L != n_theta:
        raise ValueError("P holds n_theta/2 lines")
    cl1, cl2 = lines_from_rotations(rotations, n_theta, k1, k2)
    Pf = np.ascontiguousarray(P.astype(np.complex128)).view(np.float64).ravel()
    lib = clm_oracle.lib()
    S = clm_oracle.n_shifts(max_shift, shift_step)
    shift = np.zeros(len(cl1))
    for e, (a, b, l1, l2) in enumerate(zip(np.asarray(k1), np.asarray(k2), cl1, cl2)):
        lj = l2 - 1 if l2 <= L else l2 - 1 - L
        ent = lambda j, l2p: lib.clmo_entry(Pf, R, L, float(max_shift), float(shift_step),  # noqa: E731
                                            int(a), int(b), j, int(l1 - 1), int(l2p))
        v1, s1 = _first_max([ent(j, lj) for j in range(S)])
        v2, s2 = _first_max([ent(j, L + lj) for j in range(S)])
        shift[e] = lib.clmo_shift_value(float(max_shift), float(shift_step), s1 if v1 > v2 else s2)
    return dict(cl1=cl1, cl2=cl2, shift=shift)


def equations_by_index(n_theta: int, k1, k2, cl1, cl2):
    """Coefficients (n x 4) and 1-based columns (n x 4) of the shift equations
    as cryo_estimate_shifts.m:206-222 writes them: alpha = (cij-1) dtheta,
    beta = (cji-1) dtheta with dtheta = pi/(n_theta/2); the line of image j is
    flipped when cji > n_theta/2 (its index, :176-181), and then beta - pi is
    used with the signs of image i's coefficients reversed."""
    half = n_theta // 2
    dtheta = np.pi / half
    cl1 = np.asarray(cl1, dtype=np.int64)
    cl2 = np.asarray(cl2, dtype=np.int64)
    k1 = np.asarray(k1, dtype=np.int64)
    k2 = np.asarray(k2, dtype=np.int64)
    a = (cl1 - 1) * dtheta
    b = (cl2 - 1) * dtheta
    fl = cl2 > half
    b = np.where(fl, b - np.pi, b)
    sg = np.where(fl, -1.0, 1.0)
    coef = np.stack([sg * np.sin(a), sg * np.cos(a), -np.sin(b), -np.cos(b)], axis=1)
    cols = np.stack([2 * k1 + 1, 2 * k1 + 2, 2 * k2 + 1, 2 * k2 + 2], axis=1).astype(np.int32)
    return coef, cols


def system(K: int, n_theta: int, cl1, cl2, shift1d, flip_by_index: bool = False, k1=None, k2=None):
    """(A, b): the shift equations as a scipy CSR matrix (rows x 2K) and RHS,
    one row per pair (all K(K-1)/2 in the order k1 < k2, or the pairs given);
    ``flip_by_index`` builds the rows as cryo_estimate_shifts does."""
    import scipy.sparse as sp

    if k1 is None:
        k1, k2 = clm_oracle.all_pairs(K)
    k1 = np.asarray(k1, dtype=np.int32)
    k2 = np.asarray(k2, dtype=np.int32)
    if flip_by_index:
        coef, cols = equations_by_index(n_theta, k1, k2, cl1, cl2)
    else:
        coef, cols = clm_oracle.equations(n_theta // 2, k1, k2, cl1, cl2)
    P = k1.size
    rows = np.repeat(np.arange(P), 4)
    A = sp.csr_matrix((coef.ravel(), (rows, cols.ravel() - 1)), shape=(P, 2 * K))
    return A, np.asarray(shift1d, dtype=np.float64).copy()


def lsqr(A, b, tol: float = 1e-8, maxit: int = 100, damp: float = 0.0):
    """Paige-Saunders LSQR from x0 = 0.  Returns (x, info) with info =
    {iterations, flag (0 converged, 1 maxit reached), relres, normar, norm_a}."""
    b = np.asarray(b, dtype=np.float64)
    n = A.shape[1]
    x = np.zeros(n)
    u = b.copy()
    beta = float(np.linalg.norm(u))
    bnorm = beta
    info = dict(iterations=0, flag=0, relres=0.0, normar=0.0, norm_a=0.0)
    if beta == 0.0:
        return x, info
    u /= beta
    v = A.T @ u
    alpha = float(np.linalg.norm(v))
    if alpha == 0.0:
        info["relres"] = 1.0
        return x, info
    v /= alpha
    w = v.copy()
    phibar, rhobar = beta, alpha
    anorm = 0.0
    res2 = 0.0  # sum of psi^2: the damping part of the residual
    dampsq = damp * damp
    rnorm = beta
    arnorm = alpha * beta
    flag = 1
    itn = 0
    while itn < maxit:
        itn += 1
        # bidiagonalization
        u = A @ v - alpha * u
        beta = float(np.linalg.norm(u))
        if beta > 0.0:
            u /= beta
            anorm = float(np.sqrt(anorm * anorm + alpha * alpha + beta * beta + dampsq))
            v = A.T @ u - beta * v
            alpha = float(np.linalg.norm(v))
            if alpha > 0.0:
                v /= alpha
        # eliminate the damping parameter
        rhobar1 = float(np.sqrt(rhobar * rhobar + dampsq))
        cs1 = rhobar / rhobar1
        sn1 = damp / rhobar1
        psi = sn1 * phibar
        phibar = cs1 * phibar
        # plane rotation
        rho = float(np.sqrt(rhobar1 * rhobar1 + beta * beta))
        cs = rhobar1 / rho
        sn = beta / rho
        theta = sn * alpha
        rhobar = -cs * alpha
        phi = cs * phibar
        phibar = sn * phibar
        tau = sn * phi
        # update x and w
        t1 = phi / rho
        t2 = -theta / rho
        x = x + t1 * w
        w = v + t2 * w
        # residual estimates and stopping tests
        res2 += psi * psi
        rnorm = float(np.sqrt(phibar * phibar + res2))
        arnorm = alpha * abs(tau)
        if rnorm <= tol * bnorm or arnorm <= tol * anorm * rnorm:
            flag = 0
            break
    info.update(iterations=itn, flag=flag, relres=rnorm / bnorm,
                normar=arnorm / (anorm * rnorm) if anorm * rnorm > 0 else 0.0, norm_a=anorm)
    return x, info


def estimate_shifts(K: int, n_theta: int, cl1, cl2, shift1d, tol=1e-8, maxit=100, damp=0.0):
    """est_shifts (K x 2) of cryo_estimate_shifts.m:229-230 and the LSQR info."""
    A, b = system(K, n_theta, cl1, cl2, shift1d)
    x, info = lsqr(A, b, tol, maxit, damp)
    return x.reshape(K, 2), info
