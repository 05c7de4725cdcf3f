"""ctypes front end of the FP64 C oracle (``oracle/clm_oracle.c``).

TEST INFRASTRUCTURE ONLY: the oracle is the checker for the HIP common-lines
engine.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this module; the product path never does.

The C code restates ``sinograms/cryo_clmatrix_cpu.m`` (ASPIRE) line by line;
this module adds the MATLAB-shaped assembly of the outputs
(``cryo_clmatrix_cpu.m:764-772``) so tests can compare whole outputs.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libclm_oracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.clmo_n_shifts.restype = ctypes.c_int
        L.clmo_n_shifts.argtypes = [ctypes.c_double, ctypes.c_double]
        L.clmo_shift_value.restype = ctypes.c_double
        L.clmo_shift_value.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.clmo_preprocess.restype = ctypes.c_int64
        L.clmo_preprocess.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p]
        L.clmo_pairs.restype = None
        L.clmo_pairs.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_double, _i32p, _i32p,
                                 ctypes.c_int64, _i32p, _i32p, _i32p, _f64p, _f64p,
                                 ctypes.c_int]
        L.clmo_entry.restype = ctypes.c_double
        L.clmo_entry.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.clmo_equations.restype = None
        L.clmo_equations.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, _i32p,
                                     ctypes.c_int64, _f64p, _i32p]
        L.clmo_map.restype = None
        L.clmo_map.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                               ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p]
        L.clmo_filter_map.restype = None
        L.clmo_filter_map.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, _f64p]
        L.clmo_pairs_filtered.restype = None
        L.clmo_pairs_filtered.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_double, ctypes.c_double, ctypes.c_int, _i32p, _i32p,
                                          ctypes.c_int64, _i32p, _i32p, _i32p, _f64p, _f64p, ctypes.c_int]
        L.clmo_vote_pair.restype = None
        L.clmo_vote_pair.argtypes = [_f64p, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, _f64p]
        L.clmo_syncmatrix_vote.restype = None
        L.clmo_syncmatrix_vote.argtypes = [_f64p, ctypes.c_int, ctypes.c_double, _f64p, ctypes.c_int]
        L.clmo_max_threads.restype = ctypes.c_int
        L.s3o_triangles_scores.restype = ctypes.c_int
        L.s3o_triangles_scores.argtypes = [ctypes.c_int64, _f64p, ctypes.c_int, ctypes.c_void_p, _f64p, ctypes.c_int]
        L.s3o_pairs_probabilities.restype = ctypes.c_int
        L.s3o_pairs_probabilities.argtypes = [ctypes.c_int64, _f64p, _f64p, _f64p, _f64p, ctypes.c_int]
        L.s3o_signs_times_v.restype = ctypes.c_int
        L.s3o_signs_times_v.argtypes = [ctypes.c_int64, _f64p, _f64p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                        _f64p, ctypes.c_int]
        L.s3o_triangle.restype = ctypes.c_int
        L.s3o_triangle.argtypes = [_f64p, _f64p, _f64p, _f64p, _f64p]
        L.clmo_pft_freqs.restype = None
        L.clmo_pft_freqs.argtypes = [ctypes.c_int, ctypes.c_int, _f64p]
        L.clmo_pft.restype = None
        L.clmo_pft.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _f64p,
                               ctypes.c_int]
        _lib = L
    return _lib


def n_shifts(max_shift: float, shift_step: float) -> int:
    """``ceil(2*max_shift/shift_step+1)`` (cryo_clmatrix_cpu.m:140)."""
    return int(lib().clmo_n_shifts(float(max_shift), float(shift_step)))


def preprocess(pf: np.ndarray):
    """Half lines ``P[k, l, r]`` (complex128) of the filtered, normalized rays.

    ``pf`` is n_r x n_theta x K complex (MATLAB layout).  Returns
    ``(P, n_small)`` where ``n_small`` counts lines with norm < 1e-13.
    """
    pf = np.asarray(pf)
    if pf.ndim == 2:
        pf = pf[:, :, None]
    n_r, T, K = pf.shape
    if T % 2:
        raise ValueError("n_theta must be even")
    buf = np.ascontiguousarray(np.asfortranarray(pf.astype(np.complex128)).ravel(order="F").view(np.float64))
    L, R = T // 2, n_r - 1
    P = np.empty(2 * K * L * R, dtype=np.float64)
    n_small = lib().clmo_preprocess(buf, n_r, T, K, P)
    return P.view(np.complex128).reshape(K, L, R), int(n_small)


def all_pairs(K: int):
    k1, k2 = np.triu_indices(K, 1)
    return k1.astype(np.int32), k2.astype(np.int32)


def pairs(P: np.ndarray, max_shift: float, shift_step: float, k1=None, k2=None,
          nthreads: int = 0, map_filter_radius: int = 0):
    """Run the pair loop (cryo_clmatrix_cpu.m:296-399) on a list of pairs;
    with ``map_filter_radius`` d > 0 every map is filtered first
    (cryo_average_clmap, :383-385).

    Returns dict with int32 ``cl1``, ``cl2`` (1-based, 0 = not set),
    ``j`` (0-based shift index or -1), float64 ``best`` and ``shift``.
    """
    K, L, R = P.shape
    if k1 is None:
        k1, k2 = all_pairs(K)
    k1 = np.ascontiguousarray(k1, dtype=np.int32)
    k2 = np.ascontiguousarray(k2, dtype=np.int32)
    n = k1.size
    out = dict(cl1=np.zeros(n, np.int32), cl2=np.zeros(n, np.int32),
               j=np.zeros(n, np.int32), best=np.zeros(n), shift=np.zeros(n))
    Pf = np.ascontiguousarray(P.astype(np.complex128)).view(np.float64).ravel()
    if map_filter_radius > 0:
        lib().clmo_pairs_filtered(Pf, R, L, K, float(max_shift), float(shift_step), int(map_filter_radius),
                                  k1, k2, n, out["cl1"], out["cl2"], out["j"], out["best"], out["shift"],
                                  int(nthreads))
        return out
    lib().clmo_pairs(Pf, R, L, K, float(max_shift), float(shift_step), k1, k2, n,
                     out["cl1"], out["cl2"], out["j"], out["best"], out["shift"],
                     int(nthreads))
    return out


def corr_map(P: np.ndarray, max_shift, shift_step, k1, k2, j) -> np.ndarray:
    """C = [C1, C2] (L x 2L) of one pair at shift index j."""
    K, L, R = P.shape
    Pf = np.ascontiguousarray(P.astype(np.complex128)).view(np.float64).ravel()
    out = np.zeros(2 * L * L)
    lib().clmo_map(Pf, R, L, float(max_shift), float(shift_step), int(k1), int(k2), int(j), out)
    return out.reshape(2 * L, L).T


def filter_map(m: np.ndarray, d: int) -> np.ndarray:
    """cryo_average_clmap(m, d) (default weights) of an L x 2L map."""
    L = m.shape[0]
    flat = np.ascontiguousarray(np.asarray(m, np.float64).T).ravel()  # column-major
    out = np.zeros(2 * L * L)
    lib().clmo_filter_map(flat, L, int(d), out)
    return out.reshape(2 * L, L).T


def entry(P: np.ndarray, max_shift, shift_step, k1, k2, j, l1, l2p) -> float:
    """C(l1, l2p) at shift index j (0-based indices; l2p >= L selects C2)."""
    K, L, R = P.shape
    Pf = np.ascontiguousarray(P.astype(np.complex128)).view(np.float64).ravel()
    return float(lib().clmo_entry(Pf, R, L, float(max_shift), float(shift_step),
                                  int(k1), int(k2), int(j), int(l1), int(l2p)))


def equations(L: int, k1, k2, cl1, cl2):
    """Shift-equation coefficients (n x 4) and 1-based columns (n x 4)."""
    k1 = np.ascontiguousarray(k1, dtype=np.int32)
    k2 = np.ascontiguousarray(k2, dtype=np.int32)
    cl1 = np.ascontiguousarray(cl1, dtype=np.int32)
    cl2 = np.ascontiguousarray(cl2, dtype=np.int32)
    n = k1.size
    coef = np.zeros(4 * n)
    cols = np.zeros(4 * n, np.int32)
    lib().clmo_equations(int(L), k1, k2, cl1, cl2, n, coef, cols)
    return coef.reshape(n, 4), cols.reshape(n, 4)


def cryo_clmatrix_cpu(pf, NK=None, verbose=0, max_shift=15, shift_step=1.0, nthreads=0):
    """Whole-output oracle: ``[clstack, corrstack, shift_equations,
    shift_equations_map, clstack_mask]`` as cryo_clmatrix_cpu.m returns them
    (dense numpy / scipy.sparse CSC, MATLAB 1-based values)."""
    import scipy.sparse as sp

    P, _ = preprocess(pf)
    K, L, R = P.shape
    k1, k2 = all_pairs(K)
    res = pairs(P, max_shift, shift_step, k1, k2, nthreads)
    clstack = np.zeros((K, K))
    corrstack = np.zeros((K, K))
    clstack[k1, k2] = res["cl1"]
    clstack[k2, k1] = res["cl2"]
    corrstack[k1, k2] = res["best"]
    nz = corrstack != 0
    corrstack[nz] = 1.0 - corrstack[nz]                      # :764
    coef, cols = equations(L, k1, k2, res["cl1"], res["cl2"])
    npairs = k1.size
    rows = np.repeat(np.arange(npairs), 4)
    A = sp.csc_matrix((coef.ravel(), (rows, cols.ravel() - 1)), shape=(npairs, 2 * K))
    b = sp.csc_matrix(res["shift"].reshape(-1, 1))
    eqs = sp.hstack([A, b], format="csc")                    # :768-772
    eqs.eliminate_zeros()
    emap = np.zeros((K, K))
    emap[k1, k2] = np.arange(1, npairs + 1)
    return clstack, corrstack, eqs, emap, np.zeros((K, K))


def clstack_mask(clstack, ref_clmatrix, n_theta):
    """clstack_mask of the reference's detailed-debugging mode
    (``cryo_clmatrix_cpu.m:567-617,697-698``), in its own arithmetic: with
    n_theta := n_theta/2 (:119), alpha = 2*pi*i/(2 n_theta), a pair is
    matched when |e^{alpha(l-1)} - e^{alpha(t-1)}| <= 2 sin(max_angle/2) +
    1e-10 (max_angle = 5 degrees) for both lines, or the same with e^{i pi}
    applied to both estimated lines.  Returns (mask, matched)."""
    cl = np.asarray(clstack, dtype=np.float64)
    ref = np.asarray(ref_clmatrix, dtype=np.float64)
    K = cl.shape[0]
    L = n_theta // 2
    max_angle = 5 / 180 * np.pi
    angle_tol = 2 * np.sin(max_angle / 2) + 1.0e-10
    alpha = 2 * np.pi * 1j / (2 * L)
    PI = 4 * np.arctan(1.0)
    mask = np.zeros((K, K))
    matched = 0
    for k1 in range(K):
        for k2 in range(k1 + 1, K):
            tcl1, tcl2 = ref[k1, k2], ref[k2, k1]
            l1, l2 = cl[k1, k2], cl[k2, k1]
            d1s = abs(np.exp(alpha * (l1 - 1)) - np.exp(alpha * (tcl1 - 1)))
            d2s = abs(np.exp(alpha * (l2 - 1)) - np.exp(alpha * (tcl2 - 1)))
            d1f = abs(np.exp(alpha * (l1 - 1) + 1j * PI) - np.exp(alpha * (tcl1 - 1)))
            d2f = abs(np.exp(alpha * (l2 - 1) + 1j * PI) - np.exp(alpha * (tcl2 - 1)))
            if (d1s <= angle_tol and d2s <= angle_tol) or (d1f <= angle_tol and d2f <= angle_tol):
                mask[k1, k2] = 1
                matched += 1
    return mask, matched


# ------------------------------------------------------ polar Fourier transform
def pft_freqs(n_r: int, n_theta: int) -> np.ndarray:
    """freqs of cryo_pft.m:52-58: (n_r*n_theta/2) x 2."""
    M = n_r * (n_theta // 2)
    out = np.zeros(2 * M)
    lib().clmo_pft_freqs(n_r, n_theta, out)
    return out.reshape(2, M).T.copy()


def cryo_pft(p: np.ndarray, n_r: int, n_theta: int, nthreads: int = 0) -> np.ndarray:
    """pf (n_r x n_theta x K complex128) of real images p (n x n x K, or n x n)
    by the exact NUDFT of cryo_pft.m:49-75 / nudft2.m:37-47 (pft_oracle.c)."""
    p = np.asarray(p, dtype=np.float64)
    if p.ndim == 2:
        p = p[:, :, None]
    n, n2, K = p.shape
    if n != n2:
        raise ValueError("Only square images supported.")
    if n_theta % 2:
        raise ValueError("`n_theta` must be even.")
    flat = np.ascontiguousarray(p.transpose(2, 1, 0)).ravel()  # column-major n x n x K
    out = np.zeros(2 * n_r * n_theta * K)
    lib().clmo_pft(flat, n, K, n_r, n_theta, out, nthreads)
    pf = out.view(np.complex128).reshape(K, n_theta, n_r)
    return pf.transpose(2, 1, 0)


# ------------------------------------------------------ synchronization by voting
def _cl_colmajor(clstack):
    cl = np.asarray(clstack, dtype=np.float64)
    return np.ascontiguousarray(cl.T).ravel()  # column-major K x K


def syncmatrix_vote(clstack, L, nthreads: int = 0) -> np.ndarray:
    """S (2K x 2K) of cryo_syncmatrix_vote(clstack, L) (vote_oracle.c)."""
    K = np.shape(clstack)[0]
    S = np.zeros(4 * K * K)
    lib().clmo_syncmatrix_vote(_cl_colmajor(clstack), int(K), float(L), S, int(nthreads))
    return S.reshape(2 * K, 2 * K).T


def vote_pair(clstack, L, k1, k2, _flat=None) -> np.ndarray:
    """The 2x2 block of pair (k1, k2) (0-based)."""
    K = np.shape(clstack)[0]
    r = np.zeros(4)
    lib().clmo_vote_pair(_cl_colmajor(clstack) if _flat is None else _flat, int(K), float(L), int(k1), int(k2), r)
    return r.reshape(2, 2)


# ------------------------------------------------------ sync3n triangle kernels
def _rpairs(R):
    """R_pairs (3 x 3 x N-choose-2, MATLAB layout) as a flat column-major array."""
    R = np.asarray(R, dtype=np.float64)
    return np.ascontiguousarray(R.transpose(2, 1, 0)).ravel()


def _n_of(R):
    P = np.shape(R)[2]
    N = int(round((1 + np.sqrt(1 + 8 * P)) / 2))
    assert N * (N - 1) // 2 == P
    return N


def triangles_scores(R, n_intervals=100, nthreads=1, cum=True):
    """[cumulated_scores, scores_histogram] of triangles_scores_mex (sync3n_oracle.c)."""
    N = _n_of(R)
    P = N * (N - 1) // 2
    c = np.zeros(P) if cum else None
    h = np.zeros(n_intervals)
    lib().s3o_triangles_scores(N, _rpairs(R), int(n_intervals),
                               c.ctypes.data_as(ctypes.c_void_p) if cum else None, h, int(nthreads))
    return c, h


def pairs_probabilities(R, P2, A, a, B, b, x0, nthreads=1):
    """[ln_f_ind, ln_f_arb] of pairs_probabilities_mex (sync3n_oracle.c)."""
    N = _n_of(R)
    P = N * (N - 1) // 2
    ind, arb = np.zeros(P), np.zeros(P)
    prm = np.array([P2, A, a, B, b, x0], dtype=np.float64)
    lib().s3o_pairs_probabilities(N, _rpairs(R), prm, ind, arb, int(nthreads))
    return ind, arb


def signs_times_v(R, v, scores_as_entries=0, pair_scores=None, nthreads=1):
    """signs_matrix * v of signs_times_v_mex (sync3n_oracle.c); v is P x n_eigs."""
    N = _n_of(R)
    P = N * (N - 1) // 2
    v = np.asarray(v, dtype=np.float64).reshape(P, -1)
    ne = v.shape[1]
    vf = np.ascontiguousarray(v.T).ravel()
    out = np.zeros(P * ne)
    ps = None if pair_scores is None else np.ascontiguousarray(pair_scores, dtype=np.float64)
    lib().s3o_signs_times_v(N, _rpairs(R), vf, ne, int(scores_as_entries),
                            ps.ctypes.data_as(ctypes.c_void_p) if ps is not None else None, out, int(nthreads))
    return out.reshape(ne, P).T


def triangle(Rij, Rjk, Rik):
    """(best configuration, c[4], s[3]) of one triangle."""
    c, s = np.zeros(4), np.zeros(3)
    f = lambda M: np.ascontiguousarray(np.asarray(M, dtype=np.float64).T).ravel()
    b = lib().s3o_triangle(f(Rij), f(Rjk), f(Rik), c, s)
    return int(b), c, s
