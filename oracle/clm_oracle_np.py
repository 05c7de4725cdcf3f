"""Independent numpy restatement of ``cryo_clmatrix_cpu`` for tiny inputs.

TEST INFRASTRUCTURE ONLY (see ``clm_oracle.py``).  This second restatement is
written in the reference's own vectorised form (whole-matrix products per
shift, ``max(C(:))`` on the column-major flattening) rather than the C
oracle's loops, so an indexing slip in either shows up as a mismatch.  It is
meant for K <= ~8; it is not fast.

Follows ``sinograms/cryo_clmatrix_cpu.m``: line assembly ``:96``, filter and
normalization ``:277-293`` (``cryo_raynormalize.m:17-24``), half lines
``:295,:299,:312``, shift loop ``:360-399``, equations ``:722-741``,
finalization ``:764-772``.
"""
from __future__ import annotations

import numpy as np


def lines(pf):
    """``pf=[flipdim(pf(2:end,T/2+1:end,:),1) ; pf(:,1:T/2,:)]`` (:96)."""
    T = pf.shape[1]
    return np.concatenate([pf[1:, T // 2:, :][::-1, :, :], pf[:, : T // 2, :]], axis=0)


def normalized_lines(pf):
    G = lines(np.asarray(pf, dtype=np.complex128))
    rmax = (G.shape[0] - 1) // 2
    rk = np.arange(-rmax, rmax + 1, dtype=np.float64)
    H = np.sqrt(np.abs(rk)) * np.exp(-rk ** 2 / (2 * (rmax / 4.0) ** 2))
    G = G * H[:, None, None]
    G[rmax - 1: rmax + 2, :, :] = 0                       # rows rmax:rmax+2 (1-based)
    nr = np.linalg.norm(G, axis=0)
    return G / nr[None, :, :], rmax


def first_max_colmajor(C):
    """MATLAB ``[v,i]=max(C(:))``: first maximum, NaN ignored, 0-based index."""
    flat = C.ravel(order="F")
    ok = ~np.isnan(flat)
    if not ok.any():
        return np.nan, 0
    v = flat[ok].max()
    return v, int(np.flatnonzero(flat == v)[0])


def clmatrix(pf, max_shift, shift_step):
    G, rmax = normalized_lines(pf)
    n_theta = G.shape[1]
    K = G.shape[2]
    n_shifts = int(np.ceil(2 * max_shift / shift_step + 1))
    rk2 = np.arange(-rmax, 0, dtype=np.float64)
    clstack = np.zeros((K, K))
    corrstack = np.zeros((K, K))
    shifts_1d = np.zeros((K, K))
    rows, cols, vals, rhs = [], [], [], []
    emap = np.zeros((K, K))
    e = 0
    dtheta = np.pi / n_theta
    for k1 in range(K):
        P1 = G[:rmax, :, k1]
        P1f = np.conj(P1)
        for k2 in range(k1 + 1, K):
            P2 = G[:rmax, :, k2]
            for s in range(n_shifts):
                shift = -max_shift + s * shift_step
                ph = np.exp(-2 * np.pi * 1j * rk2 * shift / (2 * rmax + 1))
                C1 = 2 * np.real((P1 * ph[:, None]).conj().T @ P2)
                C2 = 2 * np.real((P1f * ph[:, None]).conj().T @ P2)
                sval, sidx = first_max_colmajor(np.concatenate([C1, C2], axis=1))
                if sval > corrstack[k1, k2]:
                    clstack[k1, k2] = sidx % n_theta + 1
                    clstack[k2, k1] = sidx // n_theta + 1
                    corrstack[k1, k2] = sval
                    shifts_1d[k1, k2] = shift
            a = (clstack[k1, k2] - 1) * dtheta
            b = (clstack[k2, k1] - 1) * dtheta
            if b < np.pi:
                c = [np.sin(a), np.cos(a), -np.sin(b), -np.cos(b)]
            else:
                b = b - np.pi
                c = [-np.sin(a), -np.cos(a), -np.sin(b), -np.cos(b)]
            rows += [e] * 4
            cols += [2 * k1, 2 * k1 + 1, 2 * k2, 2 * k2 + 1]
            vals += c
            rhs.append(shifts_1d[k1, k2])
            e += 1
            emap[k1, k2] = e
    nz = corrstack != 0
    corrstack[nz] = 1 - corrstack[nz]
    eqs = np.zeros((e, 2 * K + 1))
    eqs[rows, cols] = vals
    eqs[:, -1] = rhs
    return clstack, corrstack, eqs, emap
