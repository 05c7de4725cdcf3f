"""C ABI of libclm (CPU): the library loads, exports every symbol that
include/clm.h declares, and its host-side FP64 assembly reproduces the
reference's outputs (cryo_clmatrix_cpu.m:392-399,722-772) from per-pair
records -- checked against the oracle's assembly."""
import ctypes
import os
import re

import numpy as np
import pytest

from aspire_amd import _lib
from aspire_amd.clmatrix import assemble

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "clm.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(clm_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    L = _lib.load()
    syms = header_symbols()
    assert set(syms) == set(_lib.EXPORTS)
    for s in syms:
        assert hasattr(L, s), s


def test_version_defaults_and_sizes():
    L = _lib.load()
    assert L.clm_version() == 10000
    p = _lib.default_params()
    assert p.max_shift == 15.0 and p.shift_step == 1.0 and p.verbose == 1  # cryo_clmatrix_cpu.m:128-139
    assert p.tie_margin == 1e-3 and p.precision == _lib.CLM_FP16_RESCORE
    assert ctypes.sizeof(_lib.ClmParams) == 56
    assert L.clm_num_pairs(5000) == 12497500 and L.clm_num_pairs(1) == 0


def test_create_without_device_fails_cleanly():
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a HIP device is present")
    except ImportError:
        pass
    L = _lib.load()
    err = ctypes.c_int(0)
    h = L.clm_create(0, ctypes.byref(err))
    assert not h and err.value in (-5, -2)


# K=700: the host assembly splits its columns over threads (from K of
# about 360 up); the result must not depend on the split
@pytest.mark.parametrize("K,L_", [(7, 18), (12, 180), (700, 180)])
def test_host_assembly_matches_oracle(oracle, K, L_):
    import scipy.sparse as sp

    rng = np.random.default_rng(K)
    P = K * (K - 1) // 2
    cl1 = rng.integers(0, L_ + 1, P).astype(np.int32)
    cl2 = rng.integers(0, 2 * L_ + 1, P).astype(np.int32)
    corr = rng.uniform(-0.2, 1.0, P)
    corr[::5] = 0.0
    shift = rng.integers(-3, 4, P).astype(np.float64) * 0.5
    rec = dict(cl1=cl1, cl2=cl2, corr=corr, shift=shift)
    clstack, corrstack, eqs, emap = assemble(K, 2 * L_, rec)
    k1, k2 = oracle.all_pairs(K)
    ref_cl = np.zeros((K, K)); ref_cl[k1, k2] = cl1; ref_cl[k2, k1] = cl2
    ref_corr = np.zeros((K, K)); ref_corr[k1, k2] = corr
    nz = ref_corr != 0
    ref_corr[nz] = 1 - ref_corr[nz]
    assert np.array_equal(clstack, ref_cl)
    assert np.array_equal(corrstack, ref_corr)
    assert np.array_equal(emap[k1, k2], np.arange(1, P + 1))
    coef, cols = oracle.equations(L_, k1, k2, cl1, cl2)
    A = sp.csc_matrix((coef.ravel(), (np.repeat(np.arange(P), 4), cols.ravel() - 1)), shape=(P, 2 * K))
    ref = sp.hstack([A, sp.csc_matrix(shift.reshape(-1, 1))], format="csc")
    ref.eliminate_zeros()
    assert eqs.shape == (P, 2 * K + 1)
    assert (eqs != ref).nnz == 0
    assert eqs.nnz == ref.nnz
    assert np.array_equal(eqs.indptr, ref.indptr) and np.array_equal(eqs.indices, ref.indices)


def test_invalid_arguments_rejected_on_host():
    L = _lib.load()
    nnz = ctypes.c_int64(0)
    assert L.clm_shift_equations_csc(3, 15, None, None, None, None, None, None, ctypes.byref(nnz)) == -1
    assert L.clm_assemble(-1, None, None, None, None, None, None) == -1
