"""2-D shifts on the GPU (libclm ``clm_estimate_shifts[_ex]``,
``clm_common_lines_from_rotations``).

* :func:`estimate_shifts` -- the least-squares step
  ``est_shifts = lsqr(shift_equations(:,1:end-1), shift_equations(:,end), 1e-8, 100)``
  (sinograms/cryo_estimate_shifts.m:228), run on per-pair outputs left on the
  device (those of :func:`aspire_amd.clmatrix.clmatrix_records`, or of
  :func:`common_lines_from_rotations`) -- the P x 2K matrix is never formed.
  ``damp = 1`` adds the ``eye(2*K)`` block of sinograms/test_shifts.m:24-27
  (without its three gauge rows).
* :func:`cryo_estimate_shifts` -- the reference function
  ``[est_shifts, shift_equations] = cryo_estimate_shifts(pf, rotations,
  max_shift, shift_step, memoryfactor)``: common lines from the rotations,
  the 1-D shift along each, the equations and their least squares.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def estimate_shifts(cl1, cl2, shift1d, K: int, n_theta: int, tol: float = 1e-8, maxit: int = 100,
                    damp: float = 0.0, device: int | None = None, flags: int = 0):
    """(est_shifts K x 2 float64 tensor on the device, LSQR stats dict).

    cl1, cl2 (int32) and shift1d (float64) are the device tensors of P = K(K-1)/2
    pairs in the order k1 < k2 (as clmatrix_records returns them).  flags:
    0, or ``_lib.CLM_FLIP_BY_INDEX`` for equations built as
    cryo_estimate_shifts builds them (the second line's orientation from its
    index)."""
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
    ctx.check(_lib.load().clm_estimate_shifts_ex(
        ctx.handle, int(K), int(n_theta), ctypes.c_void_p(cl1.data_ptr()), ctypes.c_void_p(cl2.data_ptr()),
        ctypes.c_void_p(shift1d.data_ptr()), float(tol), int(maxit), float(damp), int(flags),
        ctypes.c_void_p(x.data_ptr()), ctypes.byref(st)))
    return x.view(K, 2), st.as_dict()


def _device_rotations(rotations, K, dev):
    import torch

    if isinstance(rotations, torch.Tensor):
        r = rotations.to(dev, torch.float64)
        if tuple(r.shape) != (3, 3, K):
            raise ValueError(f"rotations must be 3 x 3 x {K}")
        return r.permute(2, 1, 0).contiguous()  # R_k(r, c) at r + 3c + 9k
    r = np.asarray(rotations, dtype=np.float64)
    if r.shape != (3, 3, K):
        raise ValueError(f"rotations must be 3 x 3 x {K}")
    return torch.as_tensor(np.ascontiguousarray(r.transpose(2, 1, 0)), device=dev)


def common_lines_from_rotations(pf, rotations, max_shift, shift_step=1.0, sel=None, device: int = 0):
    """(cl1, cl2, shift1d): device tensors of the K(K-1)/2 pairs (order k1 < k2)
    -- the common lines of the rotations (1-based, cl1 <= n_theta/2) and the
    best 1-D shift along each, as the loop of cryo_estimate_shifts.m:142-223
    finds them (libclm ``clm_common_lines_from_rotations``).  ``pf`` as for
    ``cryo_clmatrix`` (numpy n_r x n_theta x K, or a column-major device
    tensor); ``rotations`` 3 x 3 x K; ``sel`` (P bytes, nonzero = use the pair)
    or None for every pair; pairs left out get cl1 = cl2 = -1 (zero rows)."""
    import torch

    from .clmatrix import _as_matlab_pf, _device_view, _params

    ctx = _lib.context(device)
    if type(pf).__module__.startswith("torch"):
        base, dt, n_r, n_theta, K = _device_view(pf)
        dev = base.device
        dpf = base
    else:
        pfm, dt = _as_matlab_pf(pf)
        n_r, n_theta, K = pfm.shape
        dev = torch.device("cuda", device)
        dpf = torch.as_tensor(np.ascontiguousarray(pfm.transpose(2, 1, 0)), device=dev)
    P = max(K * (K - 1) // 2, 0)
    rot = _device_rotations(rotations, K, dev)
    if sel is not None:
        sel = torch.as_tensor(np.ascontiguousarray(sel, dtype=np.uint8) if not isinstance(sel, torch.Tensor)
                              else sel, device=dev).to(torch.uint8).contiguous()
        if sel.numel() != P:
            raise ValueError(f"sel must hold {P} bytes")
    cl1 = torch.empty(P, dtype=torch.int32, device=dev)
    cl2 = torch.empty(P, dtype=torch.int32, device=dev)
    sh = torch.empty(P, dtype=torch.float64, device=dev)
    p = _params(n_r, n_theta, K, dt, max_shift, shift_step, "fp64", 0, 1e-3, 0)
    torch.cuda.current_stream(dev).synchronize()
    vp = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
    ctx.check(_lib.load().clm_common_lines_from_rotations(ctx.handle, ctypes.byref(p), vp(dpf), vp(rot), vp(sel),
                                                          vp(cl1), vp(cl2), vp(sh)))
    return cl1, cl2, sh


def n_equations(n_projs: int, memoryfactor: float) -> int:
    """Number of pairs cryo_estimate_shifts uses (cryo_estimate_shifts.m:58-94,
    libclm ``clm_n_equations``): all K(K-1)/2, a fraction memoryfactor of them
    (<= 1), or as many as fit in memoryfactor megabytes of a dense
    K(K-1)/2 x 2K system (>= 100); at least K."""
    n = ctypes.c_int64(0)
    if _lib.load().clm_n_equations(int(n_projs), float(memoryfactor), ctypes.byref(n)):
        raise ValueError("subsamplingfactor must be between 0 and 1 or larger than 100")
    return int(n.value)


def equation_pairs(n_projs: int, n: int, seed: int = 0) -> np.ndarray:
    """Selection (K(K-1)/2 uint8, 1 = drawn) of min(n, K(K-1)/2) pairs drawn
    uniformly without replacement (libclm ``clm_equation_pairs``; the
    reference draws them with randperm, cryo_estimate_shifts.m:130-137)."""
    P = max(int(n_projs) * (int(n_projs) - 1) // 2, 0)
    sel = np.zeros(max(P, 1), np.uint8)
    cnt = ctypes.c_int64(0)
    rc = _lib.load().clm_equation_pairs(int(n_projs), int(n), ctypes.c_uint64(seed & (2 ** 64 - 1)),
                                        sel.ctypes.data_as(ctypes.c_void_p), ctypes.byref(cnt))
    if rc:
        raise ValueError("invalid pair count")
    return sel[:P]


def cryo_estimate_shifts(pf, rotations, max_shift, shift_step=1.0, memoryfactor=10000, shifts_2d_ref=None,
                         verbose=0, *, seed=0, device: int = 0):
    """Drop-in for ``[est_shifts, shift_equations] = cryo_estimate_shifts(pf,
    rotations, max_shift, shift_step, memoryfactor)``
    (sinograms/cryo_estimate_shifts.m:1-229): est_shifts (K x 2) and the
    equations used (scipy CSC, rows x (2K+1), the right-hand side last).

    The reference picks n_equations(K, memoryfactor) pairs with ``randperm``
    (:136-137); here they are a seeded uniform sample (``equation_pairs``,
    the same draw as the MEX gateway's for the same seed), and the equations come
    in pair order (the reference's rows are in its random order; the least
    squares does not depend on it).  Per pair, the common line of the
    rotations and the 1-D shift along it (:142-223) run on the GPU
    (``common_lines_from_rotations``), then LSQR with tol 1e-8 and maxit 100
    (:228) on the device; shift_equations is assembled on the host (FP64).
    ``shifts_2d_ref`` and ``verbose`` feed only the reference's diagnostics
    and are accepted and ignored."""
    import torch

    from .clmatrix import shift_equations

    shp = tuple(pf.shape)
    n_theta, K = shp[1], (shp[2] if len(shp) == 3 else 1)
    P = K * (K - 1) // 2
    neq = min(n_equations(K, memoryfactor), P)
    sel = equation_pairs(K, neq, seed) if neq < P else None
    cl1, cl2, sh = common_lines_from_rotations(pf, rotations, max_shift, shift_step, sel, device)
    x, _ = estimate_shifts(cl1, cl2, sh, K, n_theta, tol=1e-8, maxit=100, flags=_lib.CLM_FLIP_BY_INDEX)
    flags = _lib.CLM_FLIP_BY_INDEX
    if sel is not None:
        # only the selected pairs' records leave the device
        idx = torch.as_tensor(np.flatnonzero(sel), device=cl1.device)
        cl1, cl2, sh = cl1[idx], cl2[idx], sh[idx]
        flags |= _lib.CLM_COMPACT_RECORDS
    eqs = shift_equations(K, n_theta, cl1.cpu().numpy(), cl2.cpu().numpy(), sh.cpu().numpy(), sel, flags)
    return x.cpu().numpy(), eqs
