"""Polar Fourier transform of projections on the GPU (libclm ``clm_pft``).

Mirror of ASPIRE's ``[pf, freqs] = cryo_pft(p, n_r, n_theta, precision)``
(fourier/nufft/cryo_pft.m:1-75), the step that produces the input of
``cryo_clmatrix``.  The transform is the exact non-uniform DFT that
cryo_pft's ``nufft2`` falls back to (fourier/nufft/nudft2.m:37-47), computed
in FP64 on the device (FP64 MFMA, aspire_amd/csrc/pft.hip).

* ``p``: real images, ``n x n`` or ``n x n x K``.  A numpy array (any
  memory order; host buffers go through ``clm_pft``) or a torch tensor on the
  GPU whose memory is column-major (a ``permute(2, 1, 0)`` view of a
  contiguous ``K x n x n`` tensor), which stays on the device.
* Returns ``(pf, freqs)``: ``pf`` is ``n_r x n_theta x K`` complex (complex64
  for single images, as MATLAB returns single), column-major memory, ready
  for :func:`aspire_amd.cryo_clmatrix`; ``freqs`` is ``(n_r*n_theta/2) x 2``.
* ``precision`` is accepted and ignored, as in the reference (:16-17).
* ``method``: ``"exact"`` (default) evaluates the sums exactly (FP64 MFMA,
  O(n^2) per polar point); ``"nufft"`` by a type-2 non-uniform FFT, as the
  reference's ``nufft2`` does with a NUFFT library installed
  (fourier/nufft/nufft_transform.m:108-150): O(n^2 log n) per image, the
  choice for large images.  They agree to about 1e-14 of max |pf|.

Errors follow the reference: odd ``n_theta`` (:44-46), non-square images
(nudft2.m:29-31) and non-float images (:65-67) raise ``ValueError``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def pft_freqs(n_r: int, n_theta: int) -> np.ndarray:
    """freqs of cryo_pft.m:52-58, (n_r*n_theta/2) x 2."""
    if n_theta % 2:
        raise ValueError("`n_theta` must be even.")
    M = n_r * (n_theta // 2)
    out = np.zeros(2 * M)
    rc = _lib.load().clm_pft_freqs(int(n_r), int(n_theta), out.ctypes.data_as(ctypes.c_void_p))
    if rc:
        raise ValueError(f"clm_pft_freqs failed ({rc})")
    return out.reshape(2, M).T.copy()


METHODS = {"exact": _lib.CLM_PFT_EXACT, "nufft": _lib.CLM_PFT_NUFFT}


def _params(n, K, n_r, n_theta, img_dt, pf_dt, method="exact"):
    if method not in METHODS:
        raise ValueError(f"method must be one of {sorted(METHODS)}")
    p = _lib.ClmPftParams()
    p.n, p.K, p.n_r, p.n_theta, p.img_dtype, p.pf_dtype = int(n), int(K), int(n_r), int(n_theta), img_dt, pf_dt
    p.method = METHODS[method]
    return p


def _check(shape, n_theta):
    if n_theta % 2:
        raise ValueError("`n_theta` must be even.")
    if len(shape) not in (2, 3):
        raise ValueError("Input 'im' must be of the form N1-by-N2-by-L.")
    if shape[0] != shape[1]:
        raise ValueError("Only square images supported.")


def cryo_pft(p, n_r: int, n_theta: int, precision=None, device: int | None = None, stats: dict | None = None,
             method: str = "exact"):
    """[pf, freqs] = cryo_pft(p, n_r, n_theta) on the GPU (see the module docstring)."""
    _check(tuple(p.shape), n_theta)
    freqs = pft_freqs(n_r, n_theta)
    st = _lib.ClmPftStats()
    if not isinstance(p, np.ndarray):
        import torch

        if not (isinstance(p, torch.Tensor) and p.is_cuda):
            raise TypeError("p must be a numpy array or a torch tensor on the GPU")
        if p.dtype not in (torch.float64, torch.float32):
            raise ValueError("Images data type can be 'single' or 'double' ")
        q = p.unsqueeze(2) if p.dim() == 2 else p
        n, _, K = q.shape
        base = q.permute(2, 1, 0)
        if not base.is_contiguous():
            base = base.contiguous()
        single = base.dtype == torch.float32
        dev = base.device
        out = torch.empty((K, n_theta, n_r), dtype=torch.complex64 if single else torch.complex128, device=dev)
        ctx = _lib.context(dev.index if device is None else device)
        prm = _params(n, K, n_r, n_theta, _lib.CLM_R32 if single else _lib.CLM_R64,
                      _lib.CLM_C64 if single else _lib.CLM_C128, method)
        # torch's queue must be done with the images before libclm's stream reads them
        torch.cuda.current_stream(dev).synchronize()
        ctx.check(_lib.load().clm_pft_device(ctx.handle, ctypes.byref(prm), ctypes.c_void_p(base.data_ptr()),
                                             ctypes.c_void_p(out.data_ptr()), ctypes.byref(st)))
        if stats is not None:
            stats.update(st.as_dict())
        return out.permute(2, 1, 0), freqs
    if p.dtype not in (np.float64, np.float32):
        raise ValueError("Images data type can be 'single' or 'double' ")
    q = p[:, :, None] if p.ndim == 2 else p
    n, _, K = q.shape
    single = q.dtype == np.float32
    base = np.ascontiguousarray(q.transpose(2, 1, 0))  # K x n x n: column-major n x n x K
    out = np.empty((K, n_theta, n_r), dtype=np.complex64 if single else np.complex128)
    ctx = _lib.context(0 if device is None else device)
    prm = _params(n, K, n_r, n_theta, _lib.CLM_R32 if single else _lib.CLM_R64,
                  _lib.CLM_C64 if single else _lib.CLM_C128, method)
    ctx.check(_lib.load().clm_pft(ctx.handle, ctypes.byref(prm), base.ctypes.data_as(ctypes.c_void_p),
                                  out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(st)))
    if stats is not None:
        stats.update(st.as_dict())
    return out.transpose(2, 1, 0), freqs
