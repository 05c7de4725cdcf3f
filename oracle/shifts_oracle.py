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
cryo_clmatrix_cpu.m:722-741 (``clm_oracle.equations``).
"""
from __future__ import annotations

import numpy as np

from . import clm_oracle


def system(K: int, n_theta: int, cl1, cl2, shift1d):
    """(A, b): the shift equations as a scipy CSR matrix (P x 2K) and RHS."""
    import scipy.sparse as sp

    k1, k2 = clm_oracle.all_pairs(K)
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
