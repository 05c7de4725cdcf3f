"""The MEX gateway (aspire_amd/matlab/clm_mex.c), compiled against a
stand-in of MATLAB's MEX/MX API (tests/mex_shim) and driven through it.

CPU: the gateway compiles with -Wall -Werror, validates its arguments with
the reference's messages (cryo_clmatrix_cpu.m:84-86), reports a missing GPU
as an error rather than a crash, and frees its mxMalloc'd buffers on every
error path.  GPU: its five outputs equal the Python mirror's assembly of
libclm's records (the same library calls), for double and single pf;
verbose bit 1 prints libclm's report; verbose bit 2 with ref_clmatrix fills
clstack_mask as clm_clstack_mask does.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "tests", "mex_shim")
MX_DOUBLE, MX_SINGLE = 6, 7


@pytest.fixture(scope="module")
def mex():
    return load_shim("libclm_mex_shim.so")


def load_shim(name):
    """Build the gateways against the stand-in MEX API and load one of them."""
    subprocess.run(["make", "-s", "-C", SHIM], check=True)
    L = ctypes.CDLL(os.path.join(SHIM, name))
    vp = ctypes.c_void_p
    L.shim_create.restype = vp
    L.shim_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int64), vp]
    L.shim_call.restype = ctypes.c_int
    L.shim_call.argtypes = [ctypes.c_int, ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(vp)]
    for f in ("shim_error_id", "shim_error_msg", "shim_output"):
        getattr(L, f).restype = ctypes.c_char_p
    L.shim_live_blocks.restype = ctypes.c_long
    L.shim_ndim.argtypes = [vp]
    L.shim_dim.restype = ctypes.c_int64
    L.shim_dim.argtypes = [vp, ctypes.c_int]
    L.shim_is_sparse.argtypes = [vp]
    for f in ("shim_data", "shim_jc", "shim_ir"):
        getattr(L, f).restype = vp
        getattr(L, f).argtypes = [vp]
    L.shim_destroy.argtypes = [vp]
    return L


def mx(L, a, cls=None):
    a = np.asarray(a)
    if cls is None:
        cls = MX_SINGLE if a.dtype in (np.float32, np.complex64) else MX_DOUBLE
    cplx = np.iscomplexobj(a)
    dt = {(MX_DOUBLE, False): np.float64, (MX_DOUBLE, True): np.complex128,
          (MX_SINGLE, False): np.float32, (MX_SINGLE, True): np.complex64}[(cls, cplx)]
    a = np.asfortranarray(a.astype(dt))
    dims = (ctypes.c_int64 * max(a.ndim, 2))(*(a.shape if a.ndim >= 2 else (a.size, 1)))
    return L.shim_create(cls, int(cplx), max(a.ndim, 2), dims, a.ctypes.data_as(ctypes.c_void_p))


def call(L, nlhs, *args):
    prhs = (ctypes.c_void_p * max(len(args), 1))(*args)
    plhs = (ctypes.c_void_p * 5)()
    rc = L.shim_call(nlhs, plhs, len(args), prhs)
    return rc, [plhs[i] for i in range(5)]


def dense(L, a):
    m, n = L.shim_dim(a, 0), L.shim_dim(a, 1)
    buf = (ctypes.c_double * (m * n)).from_address(L.shim_data(a))
    return np.frombuffer(buf, dtype=np.float64).reshape((m, n), order="F").copy()


def sparse(L, a):
    import scipy.sparse as sp

    m, n = L.shim_dim(a, 0), L.shim_dim(a, 1)
    jc = np.frombuffer((ctypes.c_uint64 * (n + 1)).from_address(L.shim_jc(a)), dtype=np.uint64).astype(np.int64)
    nnz = int(jc[-1])
    ir = np.frombuffer((ctypes.c_uint64 * max(nnz, 1)).from_address(L.shim_ir(a)), dtype=np.uint64)[:nnz]
    pr = np.frombuffer((ctypes.c_double * max(nnz, 1)).from_address(L.shim_data(a)), dtype=np.float64)[:nnz]
    return sp.csc_matrix((pr.copy(), ir.astype(np.int64), jc), shape=(m, n))


def err(L):
    return L.shim_error_id().decode(), L.shim_error_msg().decode()


def test_argument_errors(mex):
    L = mex
    rc, _ = call(L, 5)
    assert rc == 1 and err(L) == ("clm:arg", "pf is required")
    real = mx(L, np.zeros((9, 16, 3)))
    rc, _ = call(L, 5, real)
    assert rc == 1 and err(L) == ("clm:arg", "pf must be complex")
    odd = mx(L, np.zeros((9, 15, 3), complex))
    rc, _ = call(L, 5, odd)
    assert rc == 1 and err(L) == ("clm:arg", "n_theta must be even")
    pf = mx(L, np.ones((9, 16, 3), complex))
    vec = mx(L, np.zeros(2))
    rc, _ = call(L, 5, pf, vec)
    assert rc == 1 and err(L)[0] == "clm:arg" and "argument 2" in err(L)[1]
    bad_ref = mx(L, np.zeros((2, 2)))
    one = mx(L, np.ones(1))
    rc, _ = call(L, 5, pf, one, one, one, one, mx(L, np.zeros(1)), bad_ref)
    assert rc == 1 and "ref_clmatrix" in err(L)[1]
    assert L.shim_live_blocks() == 0


def test_without_gpu_raises(mex):
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a HIP device is present")
    except ImportError:
        pass
    L = mex
    pf = mx(L, np.ones((9, 16, 3), complex))
    rc, _ = call(L, 5, pf)
    assert rc == 1 and err(L)[0] == "clm:nodev"
    assert L.shim_live_blocks() == 0


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_gateway_outputs_equal_mirror(mex, dtype):
    from aspire_amd import sim
    from aspire_amd.clmatrix import assemble, clmatrix_records, clstack_mask

    L = mex
    print("case mex-gateway", np.dtype(dtype).name, flush=True)
    st = sim.make_stack(K=20, n=65, n_r=33, n_theta=360, snr=1.0, max_shift_2d=2, shift_kind="int", seed=1301)
    pf = st.pf.astype(dtype)
    K = pf.shape[2]
    ref = sim.clmatrix_cheat(st.rots, 360)
    args = [mx(L, pf), mx(L, np.ones(1)), mx(L, 3 * np.ones(1)), mx(L, np.ones(1)), mx(L, np.ones(1))]
    rc, out = call(L, 5, *args)
    assert rc == 0, err(L)
    assert "pairs in" in L.shim_output().decode()  # verbose bit 1: libclm's report
    rec, _ = clmatrix_records(pf, 3, 1, precision="bf16")
    want = assemble(K, 360, rec)
    for i in (0, 1, 3):
        assert np.array_equal(dense(L, out[i]), want[i]), i
    got_eq = sparse(L, out[2])
    assert got_eq.shape == want[2].shape and (got_eq != want[2]).nnz == 0
    assert not dense(L, out[4]).any()  # no ref_clmatrix: zeros
    # verbose bit 2 with ref_clmatrix: the mask
    args2 = [mx(L, pf), mx(L, 2 * np.ones(1)), mx(L, 3 * np.ones(1)), mx(L, np.ones(1)), mx(L, np.ones(1)),
             mx(L, np.zeros(1)), mx(L, ref)]
    rc, out2 = call(L, 5, *args2)
    assert rc == 0, err(L)
    mask, matched = clstack_mask(want[0], ref, 360)
    assert np.array_equal(dense(L, out2[4]), mask)
    assert matched > K  # clean-ish data: most lines within 5 degrees
    assert f"Matched common-lines={matched}/{K * (K - 1) // 2}" in L.shim_output().decode()
    assert L.shim_live_blocks() == 0
    for a in out + out2:
        L.shim_destroy(a)
