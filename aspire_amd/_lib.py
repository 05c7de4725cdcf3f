"""ctypes binding of libclm.so (the C ABI declared in include/clm.h).

The library is loaded from this package directory.  There is no fallback:
if libclm.so is missing or cannot be loaded, every call raises.
"""
from __future__ import annotations

import ctypes
import logging
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libclm.so")

CLM_FP32, CLM_BF16_RESCORE, CLM_FP16_RESCORE, CLM_FP64_REF = 0, 1, 2, 3
CLM_C128, CLM_C64 = 0, 1
CLM_R64, CLM_R32 = 0, 1
CLM_FLIP_BY_INDEX = 1  # orientation of an equation's second line by its index
CLM_COMPACT_RECORDS = 2  # records of the selected pairs only
CLM_PFT_EXACT, CLM_PFT_NUFFT = 0, 1

PRECISIONS = {
    "fp32": CLM_FP32,
    "bf16": CLM_BF16_RESCORE,
    "fp16": CLM_FP16_RESCORE,
    "fp64": CLM_FP64_REF,
}

# every symbol include/clm.h declares
EXPORTS = (
    "clm_version", "clm_default_params", "clm_num_pairs", "clm_create",
    "clm_destroy", "clm_last_error", "clm_clmatrix", "clm_clmatrix_device", "clm_clmatrix_device_noalloc",
    "clm_clmatrix_grouped", "clm_autotune", "clm_clmatrix_windowed",
    "clm_shiftnet_nparam", "clm_shiftnet_features", "clm_shiftnet_forward",
    "clm_shiftnet_grad", "clm_shiftnet_adam", "clm_assemble", "clm_shift_equations_csc",
    "clm_pft_device", "clm_pft", "clm_pft_freqs", "clm_estimate_shifts",
    "clm_clmatrix_filtered", "clm_clmatrix_filtered_device", "clm_syncmatrix_vote",
    "clm_set_log", "clm_clstack_mask",
    "clm_sync3n_triangles_scores", "clm_sync3n_pairs_probabilities", "clm_sync3n_signs_times_v",
    "clm_clmatrix_subset_device", "clm_clmatrix_subset", "clm_partner_subset", "clm_assemble_subset",
    "clm_shift_equations_csc_subset", "clm_shift_equations_csc_ex", "clm_estimate_shifts_ex",
    "clm_common_lines_from_rotations",
    "clm_cryo_estimate_shifts",
    "clm_n_equations",
    "clm_equation_pairs",
)

# void (*)(void *user, int level, const char *msg)
LOG_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p)

CLM_GROUP_GRAPH = 1


class ClmParams(ctypes.Structure):
    _fields_ = [
        ("n_r", ctypes.c_int32),
        ("n_theta", ctypes.c_int32),
        ("K", ctypes.c_int32),
        ("pf_dtype", ctypes.c_int32),
        ("max_shift", ctypes.c_double),
        ("shift_step", ctypes.c_double),
        ("precision", ctypes.c_int32),
        ("verbose", ctypes.c_int32),
        ("tie_margin", ctypes.c_double),
        ("variant", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class ClmStackIO(ctypes.Structure):
    _fields_ = [
        ("pf", ctypes.c_void_p),
        ("cl1", ctypes.c_void_p),
        ("cl2", ctypes.c_void_p),
        ("jidx", ctypes.c_void_p),
        ("corr", ctypes.c_void_p),
        ("shift1d", ctypes.c_void_p),
    ]


class ClmTuneEntry(ctypes.Structure):
    _fields_ = [
        ("variant", ctypes.c_int32),
        ("lag", ctypes.c_int32),
        ("ms", ctypes.c_double),
        ("pairs", ctypes.c_int64),
    ]


class ClmWindowStats(ctypes.Structure):
    _fields_ = [
        ("pairs", ctypes.c_int64),
        ("accepted", ctypes.c_int64),
        ("below_gate", ctypes.c_int64),
        ("off_grid", ctypes.c_int64),
        ("window_shifts", ctypes.c_int32),
        ("full_shifts", ctypes.c_int32),
        ("ms_window", ctypes.c_double),
        ("ms_fallback", ctypes.c_double),
        ("ms_total", ctypes.c_double),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ClmStats(ctypes.Structure):
    _fields_ = [
        ("pairs", ctypes.c_int64),
        ("cells_rescored", ctypes.c_int64),
        ("uncertain_pairs", ctypes.c_int64),
        ("emitted", ctypes.c_int64),
        ("fp64_candidates", ctypes.c_int64),
        ("fallback_pairs", ctypes.c_int64),
        ("small_norm_lines", ctypes.c_int64),
        ("ms_prep", ctypes.c_double),
        ("ms_pairs", ctypes.c_double),
        ("ms_rescore", ctypes.c_double),
        ("ms_total", ctypes.c_double),
        ("variant", ctypes.c_int32),
        ("n_eval_shifts", ctypes.c_int32),
        ("refined_pairs", ctypes.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ClmPftParams(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("K", ctypes.c_int32),
        ("n_r", ctypes.c_int32),
        ("n_theta", ctypes.c_int32),
        ("img_dtype", ctypes.c_int32),
        ("pf_dtype", ctypes.c_int32),
        ("method", ctypes.c_int32),
    ]


class ClmPftStats(ctypes.Structure):
    _fields_ = [
        ("ms_tables", ctypes.c_float),
        ("ms_stage", ctypes.c_float),
        ("ms_pft", ctypes.c_float),
        ("chunks", ctypes.c_int32),
        ("ms_fft", ctypes.c_float),
        ("grid_side", ctypes.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ClmLsqrStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("flag", ctypes.c_int32),
        ("relres", ctypes.c_double),
        ("normar", ctypes.c_double),
        ("norm_a", ctypes.c_double),
        ("ms", ctypes.c_float),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ClmVoteStats(ctypes.Structure):
    _fields_ = [
        ("pairs", ctypes.c_int64),
        ("ms_build", ctypes.c_float),
        ("ms_vote", ctypes.c_float),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ClmSync3nStats(ctypes.Structure):
    _fields_ = [
        ("triangles", ctypes.c_int64),
        ("ms_layout", ctypes.c_float),
        ("ms_kernel", ctypes.c_float),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None


def load():
    """Load libclm.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m aspire_amd.build` "
            "(the HIP engine has no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32p, f64p, i64p = (ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32),
                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64))
    L.clm_version.restype = ctypes.c_int
    L.clm_version.argtypes = []
    L.clm_default_params.restype = None
    L.clm_default_params.argtypes = [ctypes.POINTER(ClmParams)]
    L.clm_num_pairs.restype = ctypes.c_int64
    L.clm_num_pairs.argtypes = [ctypes.c_int32]
    L.clm_create.restype = vp
    L.clm_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.clm_destroy.restype = None
    L.clm_destroy.argtypes = [vp]
    L.clm_last_error.restype = ctypes.c_char_p
    L.clm_last_error.argtypes = [vp]
    L.clm_clmatrix.restype = ctypes.c_int
    L.clm_clmatrix.argtypes = [vp, ctypes.POINTER(ClmParams), vp, vp, vp, vp, vp, vp,
                               ctypes.POINTER(ClmStats)]
    for name in ("clm_clmatrix_device", "clm_clmatrix_device_noalloc"):
        fn = getattr(L, name)
        fn.restype = ctypes.c_int
        fn.argtypes = [vp, ctypes.POINTER(ClmParams), vp, vp, vp, vp, vp, vp, ctypes.POINTER(ClmStats)]
    for name in ("clm_clmatrix_filtered", "clm_clmatrix_filtered_device"):
        fn = getattr(L, name)
        fn.restype = ctypes.c_int
        fn.argtypes = [vp, ctypes.POINTER(ClmParams), ctypes.c_int32, vp, vp, vp, vp, vp, vp,
                       ctypes.POINTER(ClmStats)]
    L.clm_clmatrix_grouped.restype = ctypes.c_int
    L.clm_clmatrix_grouped.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ClmParams),
                                       ctypes.POINTER(ClmStackIO), ctypes.c_int32,
                                       ctypes.POINTER(ClmStats)]
    L.clm_autotune.restype = ctypes.c_int
    L.clm_autotune.argtypes = [vp, ctypes.POINTER(ClmParams), vp, ctypes.c_int32,
                               ctypes.POINTER(ClmTuneEntry), ctypes.c_int32,
                               ctypes.POINTER(ctypes.c_int32)]
    L.clm_clmatrix_windowed.restype = ctypes.c_int
    L.clm_clmatrix_windowed.argtypes = [vp, ctypes.POINTER(ClmParams), vp, vp, ctypes.c_double, ctypes.c_double,
                                        vp, vp, vp, vp, vp, vp, ctypes.POINTER(ClmWindowStats)]
    L.clm_shiftnet_nparam.restype = ctypes.c_int
    L.clm_shiftnet_nparam.argtypes = []
    L.clm_shiftnet_features.restype = ctypes.c_int
    L.clm_shiftnet_features.argtypes = [vp, vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, vp]
    L.clm_shiftnet_forward.restype = ctypes.c_int
    L.clm_shiftnet_forward.argtypes = [vp, vp, vp, ctypes.c_int32, vp]
    L.clm_shiftnet_grad.restype = ctypes.c_int
    L.clm_shiftnet_grad.argtypes = [vp, vp, vp, vp, ctypes.c_int32, vp, vp]
    L.clm_shiftnet_adam.restype = ctypes.c_int
    L.clm_shiftnet_adam.argtypes = [vp, vp, vp, vp, vp, vp, ctypes.c_int32, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float, ctypes.c_float]
    L.clm_assemble.restype = ctypes.c_int
    L.clm_assemble.argtypes = [ctypes.c_int32, vp, vp, vp, vp, vp, vp]
    L.clm_shift_equations_csc.restype = ctypes.c_int
    L.clm_shift_equations_csc.argtypes = [ctypes.c_int32, ctypes.c_int32, vp, vp, vp, vp, vp, vp,
                                          ctypes.POINTER(ctypes.c_int64)]
    L.clm_pft_device.restype = ctypes.c_int
    L.clm_pft_device.argtypes = [vp, ctypes.POINTER(ClmPftParams), vp, vp, ctypes.POINTER(ClmPftStats)]
    L.clm_pft.restype = ctypes.c_int
    L.clm_pft.argtypes = [vp, ctypes.POINTER(ClmPftParams), vp, vp, ctypes.POINTER(ClmPftStats)]
    L.clm_estimate_shifts.restype = ctypes.c_int
    L.clm_estimate_shifts.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, vp, vp, vp, ctypes.c_double,
                                      ctypes.c_int32, ctypes.c_double, vp, ctypes.POINTER(ClmLsqrStats)]
    L.clm_syncmatrix_vote.restype = ctypes.c_int
    L.clm_syncmatrix_vote.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, vp, vp, vp, ctypes.POINTER(ClmVoteStats)]
    S3 = ctypes.POINTER(ClmSync3nStats)
    L.clm_sync3n_triangles_scores.restype = ctypes.c_int
    L.clm_sync3n_triangles_scores.argtypes = [vp, ctypes.c_int32, vp, ctypes.c_int32, vp, vp, S3]
    L.clm_sync3n_pairs_probabilities.restype = ctypes.c_int
    L.clm_sync3n_pairs_probabilities.argtypes = [vp, ctypes.c_int32, vp, ctypes.POINTER(ctypes.c_double), vp, vp, S3]
    L.clm_sync3n_signs_times_v.restype = ctypes.c_int
    L.clm_sync3n_signs_times_v.argtypes = [vp, ctypes.c_int32, vp, vp, ctypes.c_int32, ctypes.c_int32, vp, vp, S3]
    L.clm_set_log.restype = None
    L.clm_set_log.argtypes = [vp, LOG_FN, vp]
    L.clm_clstack_mask.restype = ctypes.c_int
    L.clm_clstack_mask.argtypes = [ctypes.c_int32, ctypes.c_int32, vp, vp, vp, ctypes.POINTER(ctypes.c_int64)]
    L.clm_clmatrix_subset_device.restype = ctypes.c_int
    L.clm_clmatrix_subset_device.argtypes = [vp, ctypes.POINTER(ClmParams), vp, vp, vp, vp, vp, vp, vp,
                                             ctypes.POINTER(ClmStats)]
    L.clm_clmatrix_subset.restype = ctypes.c_int
    L.clm_clmatrix_subset.argtypes = [vp, ctypes.POINTER(ClmParams), vp, vp, vp, vp, vp, vp, vp,
                                      ctypes.POINTER(ClmStats)]
    L.clm_partner_subset.restype = ctypes.c_int
    L.clm_partner_subset.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64, vp,
                                     ctypes.POINTER(ctypes.c_int64)]
    L.clm_assemble_subset.restype = ctypes.c_int
    L.clm_assemble_subset.argtypes = [ctypes.c_int32, vp, vp, vp, vp, vp, vp, vp]
    L.clm_shift_equations_csc_subset.restype = ctypes.c_int
    L.clm_shift_equations_csc_subset.argtypes = [ctypes.c_int32, ctypes.c_int32, vp, vp, vp, vp, vp, vp, vp,
                                                 ctypes.POINTER(ctypes.c_int64)]
    L.clm_shift_equations_csc_ex.restype = ctypes.c_int
    L.clm_shift_equations_csc_ex.argtypes = [ctypes.c_int32, ctypes.c_int32, vp, ctypes.c_int32, vp, vp, vp, vp,
                                             vp, vp, ctypes.POINTER(ctypes.c_int64)]
    L.clm_estimate_shifts_ex.restype = ctypes.c_int
    L.clm_estimate_shifts_ex.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, vp, vp, vp, ctypes.c_double,
                                         ctypes.c_int32, ctypes.c_double, ctypes.c_int32, vp,
                                         ctypes.POINTER(ClmLsqrStats)]
    L.clm_common_lines_from_rotations.restype = ctypes.c_int
    L.clm_common_lines_from_rotations.argtypes = [vp, ctypes.POINTER(ClmParams), vp, vp, vp, vp, vp, vp]
    L.clm_cryo_estimate_shifts.restype = ctypes.c_int
    L.clm_cryo_estimate_shifts.argtypes = [vp, ctypes.POINTER(ClmParams), vp, vp, vp, ctypes.c_double, ctypes.c_int32,
                                           vp, vp, vp, vp, ctypes.POINTER(ClmLsqrStats)]
    L.clm_n_equations.restype = ctypes.c_int
    L.clm_n_equations.argtypes = [ctypes.c_int32, ctypes.c_double, ctypes.POINTER(ctypes.c_int64)]
    L.clm_equation_pairs.restype = ctypes.c_int
    L.clm_equation_pairs.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, vp,
                                     ctypes.POINTER(ctypes.c_int64)]
    L.clm_pft_freqs.restype = ctypes.c_int
    L.clm_pft_freqs.argtypes = [ctypes.c_int32, ctypes.c_int32, vp]
    _lib = L
    return L


class ClmError(RuntimeError):
    pass


def _forward_log(user, level, msg):
    """clm_set_log callback: the library's verbose reports and warnings go to
    the ``aspire_amd`` logger."""
    text = msg.decode(errors="replace") if msg else ""
    logger = logging.getLogger("aspire_amd")
    if level:
        logger.warning(text)
    else:
        logger.info(text)


_LOG_CB = LOG_FN(_forward_log)  # one C callback object, alive as long as the module


class Context:
    """Owns one clm_ctx (one HIP stream on one device); its verbose reports
    go to the ``aspire_amd`` logger."""

    def __init__(self, device: int = 0):
        L = load()
        err = ctypes.c_int(0)
        self._h = L.clm_create(int(device), ctypes.byref(err))
        if not self._h:
            raise ClmError(f"clm_create(device={device}) failed with code {err.value}")
        self.device = device
        L.clm_set_log(self._h, _LOG_CB, None)

    @property
    def handle(self):
        return self._h

    def check(self, rc: int):
        if rc != 0:
            msg = load().clm_last_error(self._h)
            raise ClmError(f"libclm error {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "_h", None):
            load().clm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ctx_cache: dict = {}


def context(device: int = 0) -> Context:
    ctx = _ctx_cache.get(device)
    if ctx is None:
        ctx = Context(device)
        _ctx_cache[device] = ctx
    return ctx


def default_params() -> ClmParams:
    p = ClmParams()
    load().clm_default_params(ctypes.byref(p))
    return p
