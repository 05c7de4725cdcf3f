"""Host-side mirror of ASPIRE's ``cryo_clmatrix`` over libclm's C ABI.

``cryo_clmatrix(pf, NK, verbose, max_shift, shift_step)`` has the reference's
signature and returns its five outputs in the reference's form
(``sinograms/cryo_clmatrix.m:1-3``, semantics of
``sinograms/cryo_clmatrix_cpu.m``):

* ``clstack``     K x K float64, (k1,k2) = cl1 in [1, n_theta/2],
                  (k2,k1) = cl2 in [1, n_theta], 0 where not set;
* ``corrstack``   K x K float64, upper triangle = 1 - max correlation (:764);
* ``shift_equations`` scipy.sparse CSC, P x (2K+1), rows in (k1,k2) order,
                  last column = 1-D shift (:722-772);
* ``shift_equations_map`` K x K float64, equation number of (k1,k2);
* ``clstack_mask`` K x K: with verbose bit 2 and ``ref_clmatrix`` given, 1
                  where both lines are within 5 degrees of the reference's
                  (cryo_clmatrix_cpu.m:567-617,697-698), else zeros.

The pair search runs on the GPU through libclm; there is no CPU path.
"""
from __future__ import annotations

import ctypes
import logging

import numpy as np

from . import _lib

log = logging.getLogger("aspire_amd")


def _n_shifts(max_shift, shift_step):
    return int(np.ceil(2 * max_shift / shift_step + 1))


def _params(n_r, n_theta, K, dtype, max_shift, shift_step, precision, verbose, tie_margin, variant,
            flags=0):
    p = _lib.default_params()
    p.n_r, p.n_theta, p.K = int(n_r), int(n_theta), int(K)
    p.pf_dtype = dtype
    p.max_shift = float(max_shift)
    p.shift_step = float(shift_step)
    p.precision = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    p.verbose = int(verbose)
    p.tie_margin = float(tie_margin)
    p.variant = int(variant)
    p.reserved = int(flags)
    return p


def _as_matlab_pf(pf: np.ndarray):
    """Complex n_r x n_theta x K array in column-major memory."""
    pf = np.asarray(pf)
    if pf.ndim == 2:
        pf = pf[:, :, None]
    if pf.ndim != 3:
        raise ValueError("pf must be n_r x n_theta x K")
    if pf.dtype == np.complex64:
        return np.asfortranarray(pf), _lib.CLM_C64
    return np.asfortranarray(pf.astype(np.complex128, copy=False)), _lib.CLM_C128


def clmatrix_records(pf, max_shift=15, shift_step=1.0, precision="fp16", device=0,
                     tie_margin=1e-3, variant=0, verbose=0, flags=0, map_filter_radius=0, sel=None):
    """Per-pair records of the common-lines search.

    ``pf`` is a numpy array (host path) or a torch tensor on a HIP device of
    shape n_r x n_theta x K whose memory is column-major (e.g. the
    ``permute(2,1,0)`` view of a contiguous K x n_theta x n_r tensor).
    Returns ``(records, stats)``: records hold per-pair ``cl1``, ``cl2``
    (1-based), ``j`` (0-based shift index, -1 = none), ``corr`` (best
    correlation) and ``shift``, in equation order; numpy arrays for numpy
    input, device tensors for tensor input.  ``map_filter_radius`` d > 0
    filters every map with cryo_average_clmap before the search (FP64 path,
    ``clm_clmatrix_filtered``); ``corr`` is then the filtered value.
    ``sel`` (P bytes in equation order, nonzero = search the pair; a device
    tensor for device input) restricts the search to those pairs
    (``clm_clmatrix_subset``): the others come back with cl1 = cl2 = 0,
    j = -1, corr = shift = 0, and ``stats["pairs"]`` counts the selected.
    """
    if sel is not None and map_filter_radius:
        raise ValueError("the partner-subset search has no filtered-map form")
    ctx = _lib.context(device)
    stats = _lib.ClmStats()
    if type(pf).__module__.startswith("torch"):
        import torch

        base, dt, n_r, n_theta, K = _device_view(pf)
        P = max(K * (K - 1) // 2, 0)
        dev = base.device
        out = dict(cl1=torch.zeros(P, dtype=torch.int32, device=dev),
                   cl2=torch.zeros(P, dtype=torch.int32, device=dev),
                   j=torch.zeros(P, dtype=torch.int32, device=dev),
                   corr=torch.zeros(P, dtype=torch.float64, device=dev),
                   shift=torch.zeros(P, dtype=torch.float64, device=dev))
        p = _params(n_r, n_theta, K, dt, max_shift, shift_step, precision, verbose, tie_margin, variant,
                    flags)
        # torch's own queue must be done with pf before libclm's stream reads it
        torch.cuda.current_stream(dev).synchronize()
        args = (ctypes.c_void_p(base.data_ptr()),
                ctypes.c_void_p(out["cl1"].data_ptr()), ctypes.c_void_p(out["cl2"].data_ptr()),
                ctypes.c_void_p(out["j"].data_ptr()), ctypes.c_void_p(out["corr"].data_ptr()),
                ctypes.c_void_p(out["shift"].data_ptr()), ctypes.byref(stats))
        if sel is not None:
            if not (sel.is_cuda and sel.dtype == torch.uint8 and sel.is_contiguous() and sel.numel() == P
                    and sel.device == dev):
                raise ValueError(f"sel must be a contiguous uint8 tensor of {P} bytes on {dev}")
            rc = _lib.load().clm_clmatrix_subset_device(ctx.handle, ctypes.byref(p), args[0],
                                                        ctypes.c_void_p(sel.data_ptr()), *args[1:])
        elif map_filter_radius:
            rc = _lib.load().clm_clmatrix_filtered_device(ctx.handle, ctypes.byref(p), int(map_filter_radius),
                                                          *args)
        else:
            rc = _lib.load().clm_clmatrix_device(ctx.handle, ctypes.byref(p), *args)
        ctx.check(rc)
        return out, stats.as_dict()
    pfm, dt = _as_matlab_pf(pf)
    n_r, n_theta, K = pfm.shape
    P = max(K * (K - 1) // 2, 0)
    out = dict(cl1=np.zeros(P, np.int32), cl2=np.zeros(P, np.int32), j=np.zeros(P, np.int32),
               corr=np.zeros(P, np.float64), shift=np.zeros(P, np.float64))
    p = _params(n_r, n_theta, K, dt, max_shift, shift_step, precision, verbose, tie_margin, variant,
                flags)
    args = (pfm.ctypes.data_as(ctypes.c_void_p),
            out["cl1"].ctypes.data_as(ctypes.c_void_p), out["cl2"].ctypes.data_as(ctypes.c_void_p),
            out["j"].ctypes.data_as(ctypes.c_void_p), out["corr"].ctypes.data_as(ctypes.c_void_p),
            out["shift"].ctypes.data_as(ctypes.c_void_p), ctypes.byref(stats))
    if sel is not None:
        selh = np.ascontiguousarray(sel, dtype=np.uint8)
        if selh.shape != (P,):
            raise ValueError(f"sel must hold {P} bytes")
        rc = _lib.load().clm_clmatrix_subset(ctx.handle, ctypes.byref(p), args[0],
                                             selh.ctypes.data_as(ctypes.c_void_p), *args[1:])
    elif map_filter_radius:
        rc = _lib.load().clm_clmatrix_filtered(ctx.handle, ctypes.byref(p), int(map_filter_radius), *args)
    else:
        rc = _lib.load().clm_clmatrix(ctx.handle, ctypes.byref(p), *args)
    ctx.check(rc)
    return out, stats.as_dict()


def _device_view(pf):
    """(base tensor K x n_theta x n_r, dtype code, n_r, n_theta, K) of a
    column-major n_r x n_theta x K device tensor."""
    import torch

    if pf.dim() == 2:
        pf = pf.unsqueeze(2)
    n_r, n_theta, K = pf.shape
    base = pf.permute(2, 1, 0)
    if not base.is_contiguous():
        base = base.contiguous()
    if base.dtype == torch.complex64:
        dt = _lib.CLM_C64
    elif base.dtype == torch.complex128:
        dt = _lib.CLM_C128
    else:
        raise TypeError("pf must be complex64 or complex128")
    return base, dt, n_r, n_theta, K


def clmatrix_grouped(pfs, max_shift=15, shift_step=1.0, precision="fp16", device=0, tie_margin=1e-3,
                     variant=0, flags=0, graph=True, out=None):
    """Several stacks in one device job (libclm's clm_clmatrix_grouped).

    ``pfs`` is a list of device tensors (n_r x n_theta x K_i, column-major
    memory as for :func:`clmatrix_records`); ``max_shift``, ``shift_step``
    and ``precision`` are scalars or per-stack lists.  With ``graph`` the job
    is captured in a HIP graph and replayed on later calls with the same
    shapes and buffers: pass the previous call's ``out`` to reuse them.
    Returns ``(records, stats)`` with one record dict per stack.
    """
    import torch

    n = len(pfs)
    per = lambda v: list(v) if isinstance(v, (list, tuple)) else [v] * n
    ms, ss, pr = per(max_shift), per(shift_step), per(precision)
    ctx = _lib.context(device)
    params = (_lib.ClmParams * max(n, 1))()
    ios = (_lib.ClmStackIO * max(n, 1))()
    bases, outs = [], []
    for i, pf in enumerate(pfs):
        base, dt, n_r, n_theta, K = _device_view(pf)
        bases.append(base)
        P = max(K * (K - 1) // 2, 0)
        dev = base.device
        o = out[i] if out is not None else dict(
            cl1=torch.zeros(P, dtype=torch.int32, device=dev),
            cl2=torch.zeros(P, dtype=torch.int32, device=dev),
            j=torch.zeros(P, dtype=torch.int32, device=dev),
            corr=torch.zeros(P, dtype=torch.float64, device=dev),
            shift=torch.zeros(P, dtype=torch.float64, device=dev))
        outs.append(o)
        params[i] = _params(n_r, n_theta, K, dt, ms[i], ss[i], pr[i], 0, tie_margin, variant, flags)
        ios[i] = _lib.ClmStackIO(base.data_ptr(), o["cl1"].data_ptr(), o["cl2"].data_ptr(),
                                 o["j"].data_ptr(), o["corr"].data_ptr(), o["shift"].data_ptr())
    stats = _lib.ClmStats()
    torch.cuda.current_stream(torch.device("cuda", device)).synchronize()
    rc = _lib.load().clm_clmatrix_grouped(ctx.handle, n, params, ios,
                                          _lib.CLM_GROUP_GRAPH if graph else 0, ctypes.byref(stats))
    ctx.check(rc)
    return outs, stats.as_dict()


def autotune(pf, max_shift=15, shift_step=1.0, precision="fp16", device=0, sample_pairs=200000, flags=0):
    """Time every pair-kernel variant for this shape on the first
    ``sample_pairs`` pairs of ``pf`` (device tensor, column-major as for
    :func:`clmatrix_records`); later calls with ``variant=0`` on this device
    use the fastest.  Returns one dict per variant (variant = CT*100 + waves
    per workgroup, lag, ms, pairs, us_per_kpair) and the chosen one."""
    import torch

    base, dt, n_r, n_theta, K = _device_view(pf)
    ctx = _lib.context(device)
    p = _params(n_r, n_theta, K, dt, max_shift, shift_step, precision, 0, 1e-3, 0, flags)
    cap = 64
    ent = (_lib.ClmTuneEntry * cap)()
    n = ctypes.c_int32(0)
    torch.cuda.current_stream(base.device).synchronize()
    rc = _lib.load().clm_autotune(ctx.handle, ctypes.byref(p), ctypes.c_void_p(base.data_ptr()),
                                  int(sample_pairs), ent, cap, ctypes.byref(n))
    ctx.check(rc)
    rows = [dict(variant=e.variant, lag=e.lag, ms=e.ms, pairs=e.pairs,
                 us_per_kpair=1e6 * e.ms / max(e.pairs, 1)) for e in ent[:min(n.value, cap)]]
    best = min(rows, key=lambda r: r["ms"] / max(r["pairs"], 1)) if rows else None
    return rows, best


def shift_equations(K: int, n_theta: int, cl1, cl2, shift, sel=None, flags=0):
    """shift_equations (scipy CSC, rows x (2K+1), the right-hand side last)
    from per-pair lines and 1-D shifts (host FP64, libclm's
    clm_shift_equations_csc_ex): one row per pair in pair order, or per
    selected pair with ``sel`` (P bytes); ``flags`` ``_lib.CLM_FLIP_BY_INDEX``
    builds the rows as cryo_estimate_shifts does, and with
    ``_lib.CLM_COMPACT_RECORDS`` cl1, cl2 and shift hold the selected pairs'
    records only, in pair order."""
    import scipy.sparse as sp

    L = _lib.load()
    P = max(K * (K - 1) // 2, 0)
    selh = None
    if sel is not None:
        selh = np.ascontiguousarray(sel, dtype=np.uint8)
        if selh.shape != (P,):
            raise ValueError("sel must hold K(K-1)/2 bytes")
    cl1 = np.ascontiguousarray(cl1, dtype=np.int32)
    cl2 = np.ascontiguousarray(cl2, dtype=np.int32)
    shift = np.ascontiguousarray(shift, dtype=np.float64)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p) if a is not None else None
    nrow = P if selh is None else int(np.count_nonzero(selh))
    nrec = nrow if flags & _lib.CLM_COMPACT_RECORDS else P
    if selh is None and flags & _lib.CLM_COMPACT_RECORDS:
        raise ValueError("compact records need a selection")
    if not (cl1.shape == cl2.shape == shift.shape == (nrec,)):
        raise ValueError(f"expected {nrec} records")
    csc = lambda *a: L.clm_shift_equations_csc_ex(int(K), int(n_theta), vp(selh), int(flags), *a)
    nnz = ctypes.c_int64(0)
    jc = np.zeros(2 * K + 2, np.int64)
    rc = csc(vp(cl1), vp(cl2), vp(shift), vp(jc), None, None, ctypes.byref(nnz))
    if rc:
        raise _lib.ClmError(f"clm_shift_equations_csc_ex failed ({rc})")
    ir = np.zeros(nnz.value, np.int64)
    pr = np.zeros(nnz.value, np.float64)
    rc = csc(vp(cl1), vp(cl2), vp(shift), vp(jc), vp(ir), vp(pr), ctypes.byref(nnz))
    if rc:
        raise _lib.ClmError(f"clm_shift_equations_csc_ex failed ({rc})")
    if nrow < 2 ** 31 and nnz.value < 2 ** 31:
        # scipy keeps 32-bit indices when they fit; handing them over as such
        # spares its scan of the 64-bit arrays and its own converting copy
        ir, jc = ir.astype(np.int32), jc.astype(np.int32)
    return sp.csc_matrix((pr, ir, jc), shape=(nrow, 2 * K + 1))


def assemble(K: int, n_theta: int, rec, sel=None, flags=0):
    """Reference outputs from per-pair records (FP64, host; libclm's
    clm_assemble and clm_shift_equations_csc).  With ``sel`` (P bytes), the
    partner-subset form: equations for the selected pairs only, numbered in
    pair order (clm_assemble_subset, clm_shift_equations_csc_ex).  ``flags``
    (``_lib.CLM_FLIP_BY_INDEX``) builds the rows as cryo_estimate_shifts does."""
    L = _lib.load()
    if sel is not None:
        selh = np.ascontiguousarray(sel, dtype=np.uint8)
        if selh.shape != (max(K * (K - 1) // 2, 0),):
            raise ValueError("sel must hold K(K-1)/2 bytes")
    cl1 = np.ascontiguousarray(rec["cl1"], dtype=np.int32)
    cl2 = np.ascontiguousarray(rec["cl2"], dtype=np.int32)
    corr = np.ascontiguousarray(rec["corr"], dtype=np.float64)
    clstack = np.zeros((K, K), order="F")
    corrstack = np.zeros((K, K), order="F")
    emap = np.zeros((K, K), order="F")
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    if sel is None:
        rc = L.clm_assemble(int(K), vp(cl1), vp(cl2), vp(corr), vp(clstack), vp(corrstack), vp(emap))
    else:
        rc = L.clm_assemble_subset(int(K), vp(selh), vp(cl1), vp(cl2), vp(corr), vp(clstack), vp(corrstack),
                                   vp(emap))
    if rc:
        raise _lib.ClmError(f"clm_assemble failed ({rc})")
    eqs = shift_equations(K, n_theta, cl1, cl2, rec["shift"], sel, flags)
    return clstack, corrstack, eqs, emap


def cryo_clmatrix(pf, NK=-1, verbose=1, max_shift=15, shift_step=1.0, map_filter_radius=0,
                  ref_clmatrix=None, ref_shifts_2d=None, *, precision="fp16", device=0):
    """Drop-in for ``cryo_clmatrix`` (sinograms/cryo_clmatrix.m:1-3).

    Arguments and defaults follow cryo_clmatrix_cpu.m:123-152: ``NK`` is
    accepted and ignored (:123), ``max_shift`` defaults to 15, ``shift_step``
    to 1.  ``map_filter_radius`` > 0 (argument 6, :24-29) filters every map
    with cryo_average_clmap before the search (FP64 path; corrstack then
    holds 1 - the filtered value, as the reference stores it).
    ``precision`` selects the arithmetic of the unfiltered search: "fp16"
    (default), "bf16", "fp32" or "fp64".  FP16 and BF16 run at the same
    MFMA rate; FP16's rounding bound is 8x tighter, so on high-SNR data
    (class averages, broad correlation peaks) it certifies nearly every
    pair where BF16 sends many to the FP32 refinement (SNR 1, K=5000:
    452 against 5.6 M pairs, 4.3 s against 30.5 s).

    ``verbose`` is the reference's bitmask (:155-204).  Bit 1: the library
    reports the call's shape and stages through the ``aspire_amd`` logger.
    Bit 2 (detailed debugging): one report per pair (lines, correlation,
    1-D shift, and the reference line pair and match status when
    ``ref_clmatrix`` is given, :689-720), and ``clstack_mask`` marks the
    pairs whose lines are within 5 degrees of ``ref_clmatrix`` (:567-617,
    :697-698); with ``ref_shifts_2d`` also the true 1-D shift of the pair's
    estimated lines and the error (:665-685).  Bits 4 and 8 ask the
    reference for figures; they are not drawn here.
    """
    if hasattr(pf, "shape") and len(pf.shape) >= 2 and pf.shape[1] % 2:
        raise ValueError("n_theta must be even")  # cryo_clmatrix_cpu.m:84-86
    if verbose:
        log.info("Shift estimation parameters: max_shift=%g   shift_step=%g", max_shift, shift_step)
    rec, stats = clmatrix_records(pf, max_shift, shift_step, precision, device, verbose=verbose,
                                  map_filter_radius=map_filter_radius)
    if type(pf).__module__.startswith("torch"):
        rec = {k: v.cpu().numpy() for k, v in rec.items()}
        n_theta, K = pf.shape[1], (pf.shape[2] if pf.dim() == 3 else 1)
    else:
        shp = np.shape(pf)
        n_theta, K = shp[1], (shp[2] if len(shp) == 3 else 1)
    clstack, corrstack, eqs, emap = assemble(K, n_theta, rec)
    mask = np.zeros((K, K), order="F")
    have_ref = ref_clmatrix is not None and np.ndim(ref_clmatrix) == 2
    if verbose and not have_ref:
        log.info("Reference clmatrix not found")  # :166-168
    if verbose & 2 and have_ref:
        mask, matched = clstack_mask(clstack, ref_clmatrix, n_theta)
    if verbose & 2:
        _debug_report(K, n_theta, rec, clstack, corrstack, ref_clmatrix if have_ref else None, mask,
                      ref_shifts_2d if ref_shifts_2d is not None and np.ndim(ref_shifts_2d) == 2 else None)
        if have_ref:
            P = K * (K - 1) // 2
            log.info("Matched common-lines=%d/%d  (%3.1f%%)", matched, P, 100.0 * matched / max(P, 1))
    return clstack, corrstack, eqs, emap, mask


def partner_subset(K: int, NK: int, seed: int = 0):
    """(sel, count): the pairs cryo_clmatrix_gpu_1 searches
    (sinograms/cryo_clmatrix_gpu_1.m:150-151,346-350) -- image k1 with
    min(K-1-k1, NK) of its later images drawn uniformly without replacement;
    NK = -1 selects every pair.  libclm's clm_partner_subset (host); the draw
    depends on (seed, k1) only."""
    K, NK = int(K), int(NK)
    P = max(K * (K - 1) // 2, 0)
    sel = np.zeros(P, np.uint8)
    n = ctypes.c_int64(0)
    rc = _lib.load().clm_partner_subset(K, NK, ctypes.c_uint64(int(seed) & (2 ** 64 - 1)),
                                        sel.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n))
    if rc:
        raise _lib.ClmError(f"clm_partner_subset failed ({rc}): K={K}, NK={NK}")
    return sel, int(n.value)


def cryo_clmatrix_subset(pf, NK=-1, verbose=1, max_shift=15, shift_step=1.0, *, seed=0, sel=None,
                         precision="fp16", device=0):
    """The partner-subset form of ``cryo_clmatrix_gpu_1(pf, NK, verbose,
    max_shift, shift_step)`` (sinograms/cryo_clmatrix_gpu_1.m:1-24): each image
    is searched against at most ``NK`` of its later images, drawn at random
    (``partner_subset``; or the pairs of ``sel``, P bytes in equation order).
    Each selected pair takes the drop-in's search (cryo_clmatrix_cpu.m:360-399
    semantics, the same parity); the other pairs keep clstack = corrstack = 0.
    shift_equations holds one row per selected pair in pair order
    (:809-827,853-858) and shift_equations_map its 1-based number.  Returns
    the reference's five outputs (clstack_mask zeros) and the selection."""
    if hasattr(pf, "shape") and len(pf.shape) >= 2 and pf.shape[1] % 2:
        raise ValueError("n_theta must be even")
    shp = tuple(pf.shape)
    n_theta, K = shp[1], (shp[2] if len(shp) == 3 else 1)
    if sel is None:
        sel, _ = partner_subset(K, NK, seed)
    sel = np.ascontiguousarray(sel, dtype=np.uint8)
    if type(pf).__module__.startswith("torch"):
        import torch

        dsel = torch.as_tensor(sel).to(pf.device)
        rec, stats = clmatrix_records(pf, max_shift, shift_step, precision, device, verbose=verbose, sel=dsel)
        rec = {k: v.cpu().numpy() for k, v in rec.items()}
    else:
        rec, stats = clmatrix_records(pf, max_shift, shift_step, precision, device, verbose=verbose, sel=sel)
    clstack, corrstack, eqs, emap = assemble(K, n_theta, rec, sel=sel)
    return clstack, corrstack, eqs, emap, np.zeros((K, K), order="F"), sel


def clstack_mask(clstack, ref_clmatrix, n_theta):
    """(mask, matched): libclm's clm_clstack_mask (host FP64) -- 1 at (k1,k2),
    k1 < k2, where both estimated lines lie within 5 degrees of the
    reference's, or both within 5 degrees of their antipodes
    (cryo_clmatrix_cpu.m:567-617)."""
    cl = np.asfortranarray(clstack, dtype=np.float64)
    ref = np.asfortranarray(ref_clmatrix, dtype=np.float64)
    if cl.shape != ref.shape or cl.ndim != 2 or cl.shape[0] != cl.shape[1]:
        raise ValueError("clstack and ref_clmatrix must both be K x K")
    K = cl.shape[0]
    mask = np.zeros((K, K), order="F")
    matched = ctypes.c_int64(0)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = _lib.load().clm_clstack_mask(int(K), int(n_theta), vp(cl), vp(ref), vp(mask), ctypes.byref(matched))
    if rc:
        raise _lib.ClmError(f"clm_clstack_mask failed ({rc})")
    return mask, int(matched.value)


def _debug_report(K, n_theta, rec, clstack, corrstack, ref, mask, ref_shifts):
    """Per-pair reports of verbose bit 2 (cryo_clmatrix_cpu.m:689-720)."""
    L = n_theta // 2
    dtheta = np.pi / L
    e = 0
    for k1 in range(K):
        for k2 in range(k1 + 1, K):
            log.info("Finding common-line between projections [k1 k2]=[ %d %d ]", k1 + 1, k2 + 1)
            log.info("\t Common lines:")
            log.info("\t \t clstack= [ %3d %3d ]", clstack[k1, k2], clstack[k2, k1])
            if ref is not None:
                log.info("\t \t ref    = [ %3d %3d ]", ref[k1, k2], ref[k2, k1])
                log.info("\t \t status = %s (%s than 5.00 degrees)",
                         "MATCHED" if mask[k1, k2] else "NOT MATCHED", "less" if mask[k1, k2] else "more")
            log.info("\t Correlation:")
            log.info("\t \t est    =  %9.6f", corrstack[k1, k2])
            log.info("\t Shifts:")
            log.info("\t \t est  shift= %6.4f", rec["shift"][e])
            if ref_shifts is not None:
                a, b = (clstack[k1, k2] - 1) * dtheta, (clstack[k2, k1] - 1) * dtheta
                dx1, dy1 = ref_shifts[k1, 0], ref_shifts[k1, 1]
                dx2, dy2 = ref_shifts[k2, 0], ref_shifts[k2, 1]
                if b < np.pi - 1.0e-13:
                    true1d = np.sin(a) * dx1 + np.cos(a) * dy1 - np.sin(b) * dx2 - np.cos(b) * dy2
                else:
                    b -= np.pi
                    true1d = -np.sin(a) * dx1 - np.cos(a) * dy1 - np.sin(b) * dx2 - np.cos(b) * dy2
                log.info("\t \t True shift= %6.4f", true1d)
                log.info("\t \t shift err = %3.2e", rec["shift"][e] - true1d)
            e += 1
