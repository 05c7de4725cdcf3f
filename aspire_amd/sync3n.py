"""The sync3n triangle kernels on the GPU (libclm ``clm_sync3n_*``).

Mirrors of the reference's three MEX kernels of the sync3n path
(abinitio/C1/sync3n/, called from triangles_scores.m:62,
cryo_sync3n_syncmatrix_weights.m and cryo_sync3n_Jsync_power_method.m:62):

* :func:`triangles_scores_mex`  -> ``[cum_scores, scores_hist]``;
* :func:`pairs_probabilities_mex` -> ``[ln_f_ind, ln_f_arb]``;
* :func:`signs_times_v_mex` -> the J-synchronization matrix times ``v``;

and :func:`jsync_power_method`, the power method of
cryo_sync3n_Jsync_power_method.m around the last one.  ``R`` is the
3 x 3 x N-choose-2 array of relative rotations (numpy, MATLAB layout) or a
device tensor holding the same bytes (N-choose-2 x 3 x 3, C order, each
matrix stored column-major as in MATLAB).  Outputs are numpy for numpy
input, device tensors for tensor input.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _n_of(P: int) -> int:
    N = int(round((1 + np.sqrt(1 + 8 * P)) / 2))
    if N * (N - 1) // 2 != P:
        raise ValueError(f"{P} relative rotations is not N-choose-2 for any N")
    return N


def _dev_R(R, device):
    """(device tensor with R's MATLAB bytes, N, numpy input?)."""
    import torch

    if isinstance(R, torch.Tensor):
        if not R.is_cuda or R.dtype != torch.float64:
            raise ValueError("R must be a float64 device tensor")
        if R.dim() != 3 or tuple(R.shape[1:]) != (3, 3):
            raise ValueError("a device R is N-choose-2 x 3 x 3 (each matrix column-major)")
        return R.contiguous(), _n_of(R.shape[0]), False
    R = np.asarray(R, dtype=np.float64)
    if R.ndim != 3 or R.shape[:2] != (3, 3):
        raise ValueError("R must be 3 x 3 x N-choose-2")
    flat = np.ascontiguousarray(R.transpose(2, 1, 0))  # pair-major, each matrix column-major
    return torch.as_tensor(flat, device=torch.device("cuda", device)), _n_of(R.shape[2]), True


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def triangles_scores_mex(R, n_intervals: int = 100, cum: bool = True, device: int = 0):
    """(cum_scores, scores_hist, stats) of triangles_scores_mex(N, R, n_intervals, ~, cum, ~)."""
    import torch

    Rd, N, host = _dev_R(R, device)
    P = N * (N - 1) // 2
    dev = Rd.device
    c = torch.empty(P, dtype=torch.float64, device=dev) if cum else None
    h = torch.empty(int(n_intervals), dtype=torch.float64, device=dev)
    st = _lib.ClmSync3nStats()
    ctx = _lib.context(dev.index)
    torch.cuda.current_stream(dev).synchronize()
    ctx.check(_lib.load().clm_sync3n_triangles_scores(ctx.handle, N, _ptr(Rd), int(n_intervals), _ptr(c), _ptr(h),
                                                      ctypes.byref(st)))
    if host:
        return (c.cpu().numpy() if cum else None), h.cpu().numpy(), st.as_dict()
    return c, h, st.as_dict()


def pairs_probabilities_mex(R, P2, A, a, B, b, x0, device: int = 0):
    """(ln_f_ind, ln_f_arb, stats) of pairs_probabilities_mex(N, R, P2, A, a, B, b, x0, ~, ~)."""
    import torch

    Rd, N, host = _dev_R(R, device)
    P = N * (N - 1) // 2
    dev = Rd.device
    ind = torch.empty(P, dtype=torch.float64, device=dev)
    arb = torch.empty(P, dtype=torch.float64, device=dev)
    prm = (ctypes.c_double * 6)(P2, A, a, B, b, x0)
    st = _lib.ClmSync3nStats()
    ctx = _lib.context(dev.index)
    torch.cuda.current_stream(dev).synchronize()
    ctx.check(_lib.load().clm_sync3n_pairs_probabilities(ctx.handle, N, _ptr(Rd), prm, _ptr(ind), _ptr(arb),
                                                         ctypes.byref(st)))
    if host:
        return ind.cpu().numpy(), arb.cpu().numpy(), st.as_dict()
    return ind, arb, st.as_dict()


def signs_times_v_mex(R, v, scores_as_entries: int = 0, pair_scores=None, device: int = 0):
    """(v2, stats): the J-synchronization matrix times v (N-choose-2 x n_eigs)
    as signs_times_v_mex(N, R, v, n_eigs, scores_as_entries, pair_scores, ~)."""
    import torch

    Rd, N, host = _dev_R(R, device)
    P = N * (N - 1) // 2
    dev = Rd.device
    if isinstance(v, torch.Tensor):
        vd = v.to(dev, torch.float64).reshape(P, -1)
    else:
        vd = torch.as_tensor(np.asarray(v, dtype=np.float64).reshape(P, -1), device=dev)
    ne = vd.shape[1]
    vcol = vd.T.contiguous()  # column-major P x ne
    ps = None
    if int(scores_as_entries) == 2:
        if pair_scores is None:
            raise ValueError("scores_as_entries = 2 needs pair_scores")
        ps = torch.as_tensor(np.asarray(pair_scores, dtype=np.float64), device=dev) if not isinstance(
            pair_scores, torch.Tensor) else pair_scores.to(dev, torch.float64).contiguous()
    out = torch.empty((ne, P), dtype=torch.float64, device=dev)
    st = _lib.ClmSync3nStats()
    ctx = _lib.context(dev.index)
    torch.cuda.current_stream(dev).synchronize()
    ctx.check(_lib.load().clm_sync3n_signs_times_v(ctx.handle, N, _ptr(Rd), _ptr(vcol), int(ne),
                                                   int(scores_as_entries), _ptr(ps), _ptr(out), ctypes.byref(st)))
    v2 = out.T
    if host and not isinstance(v, torch.Tensor):
        return v2.cpu().numpy(), st.as_dict()
    return v2, st.as_dict()


def jsync_power_method(R, n_eigs: int = 4, scores_as_entries: bool = False, seed: int = 0, epsilon: float = 1e-2,
                       max_iterations: int = 100, device: int = 0):
    """(J_sync, J_significance, eigenvalues, iterations, dd) of
    cryo_sync3n_Jsync_power_method(Rij, n_eigs, scores_as_entries)
    (cryo_sync3n_Jsync_power_method.m:21-84): block power iterations with
    the J-synchronization matrix (:func:`signs_times_v_mex` on the device)
    and a thin QR (the device's LAPACK-style Householder QR, as MATLAB's
    qr(.,0)), from a seeded uniform start (the reference draws rand), until
    the leading vector moves by at most epsilon.  J_sync is the sign of the
    leading eigenvector: which relative rotations to conjugate by J."""
    import torch

    Rd, N, host = _dev_R(R, device)
    P = N * (N - 1) // 2
    dev = Rd.device
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    vec = torch.rand((P, int(n_eigs)), generator=g, dtype=torch.float64).to(dev)
    vec, _ = torch.linalg.qr(vec, mode="reduced")
    entries = 1 if scores_as_entries else 0
    dd, itr, eig = 1.0, 0, None
    while itr < max_iterations and dd > epsilon:
        itr += 1
        vnew, _ = signs_times_v_mex(Rd, vec, entries, device=device)
        vnew, rq = torch.linalg.qr(vnew, mode="reduced")
        dd = float(torch.linalg.vector_norm(vnew[:, 0] - vec[:, 0]))
        vec, eig = vnew, rq
    eigenvalues = torch.diagonal(eig) if eig is not None else torch.zeros(int(n_eigs), device=dev)
    lead = vec[:, 0]
    J_sync, J_sig = torch.sign(lead), torch.abs(lead)
    if host:
        return J_sync.cpu().numpy(), J_sig.cpu().numpy(), eigenvalues.cpu().numpy(), itr, dd
    return J_sync, J_sig, eigenvalues, itr, dd
