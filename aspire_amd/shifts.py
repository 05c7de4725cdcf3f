"""2-D shifts from the common-lines shift equations on the GPU
(libclm ``clm_estimate_shifts``).

Mirror of the least-squares step that follows ``cryo_clmatrix``:
``est_shifts = lsqr(shift_equations(:,1:end-1), shift_equations(:,end), 1e-8, 100)``
(sinograms/cryo_estimate_shifts.m:229-230), run on the per-pair outputs that
:func:`aspire_amd.clmatrix.clmatrix_records` leaves on the device -- the
P x 2K matrix is never formed.  ``damp = 1`` adds the ``eye(2*K)`` block of
sinograms/test_shifts.m:24-27 (without its three gauge rows).
"""
from __future__ import annotations

import ctypes

from . import _lib


def estimate_shifts(cl1, cl2, shift1d, K: int, n_theta: int, tol: float = 1e-8, maxit: int = 100,
                    damp: float = 0.0, device: int | None = None):
    """(est_shifts K x 2 float64 tensor on the device, LSQR stats dict).

    cl1, cl2 (int32) and shift1d (float64) are the device tensors of P = K(K-1)/2
    pairs in the order k1 < k2 (as clmatrix_records returns them)."""
    import torch

    P = K * (K - 1) // 2
    for t, dt in ((cl1, torch.int32), (cl2, torch.int32), (shift1d, torch.float64)):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dt and t.numel() == P):
            raise ValueError(f"expected a device tensor of {P} {dt} values")
    cl1, cl2, shift1d = cl1.contiguous(), cl2.contiguous(), shift1d.contiguous()
    dev = shift1d.device
    x = torch.empty(2 * K, dtype=torch.float64, device=dev)
    st = _lib.ClmLsqrStats()
    ctx = _lib.context(dev.index if device is None else device)
    torch.cuda.current_stream(dev).synchronize()
    ctx.check(_lib.load().clm_estimate_shifts(
        ctx.handle, int(K), int(n_theta), ctypes.c_void_p(cl1.data_ptr()), ctypes.c_void_p(cl2.data_ptr()),
        ctypes.c_void_p(shift1d.data_ptr()), float(tol), int(maxit), float(damp),
        ctypes.c_void_p(x.data_ptr()), ctypes.byref(st)))
    return x.view(K, 2), st.as_dict()
