"""The voting synchronization matrix on the GPU (libclm ``clm_syncmatrix_vote``).

Mirror of ``S = cryo_syncmatrix_vote(clstack, n_theta)``
(abinitio/C1/sync2n/cryo_syncmatrix_vote.m, no reference rotations), the
step that consumes the common lines in the C1 ab-initio pipeline
(abinitio/C1/cryo_abinitio_C1_worker.m:160).  The common lines come from the
device records of :func:`aspire_amd.clmatrix.clmatrix_records` or from a
K x K clstack.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def syncmatrix_vote_records(cl1, cl2, K: int, n_theta: int, device: int | None = None):
    """(S, stats): S is a 2K x 2K float64 tensor on the device (S[i, j] is
    MATLAB's S(i+1, j+1)); cl1/cl2 are the int32 device records of the
    K(K-1)/2 pairs in the order k1 < k2."""
    import torch

    P = K * (K - 1) // 2
    for t in (cl1, cl2):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.int32 and t.numel() == P):
            raise ValueError(f"expected device int32 tensors of {P} pairs")
    cl1, cl2 = cl1.contiguous(), cl2.contiguous()
    dev = cl1.device
    Sm = torch.empty((2 * K, 2 * K), dtype=torch.float64, device=dev)  # holds S column-major
    st = _lib.ClmVoteStats()
    ctx = _lib.context(dev.index if device is None else device)
    torch.cuda.current_stream(dev).synchronize()
    ctx.check(_lib.load().clm_syncmatrix_vote(ctx.handle, int(K), int(n_theta), ctypes.c_void_p(cl1.data_ptr()),
                                              ctypes.c_void_p(cl2.data_ptr()), ctypes.c_void_p(Sm.data_ptr()),
                                              ctypes.byref(st)))
    return Sm.T, st.as_dict()


def cryo_syncmatrix_vote(clstack, n_theta: int, device: int = 0) -> np.ndarray:
    """S = cryo_syncmatrix_vote(clstack, n_theta) for a K x K clstack (numpy)."""
    import torch

    cl = np.asarray(clstack)
    K = cl.shape[0]
    if cl.ndim != 2 or cl.shape[1] != K:
        raise ValueError("clmatrix must be a square matrix")
    k1, k2 = np.triu_indices(K, 1)
    dev = torch.device("cuda", device)
    a = torch.as_tensor(cl[k1, k2].astype(np.int32), device=dev)
    b = torch.as_tensor(cl[k2, k1].astype(np.int32), device=dev)
    S, _ = syncmatrix_vote_records(a, b, K, n_theta, device)
    return S.cpu().numpy()
