"""Writes tests/golden/clmatrix_tiny.npz: tiny cryo_clmatrix cases with their
expected outputs (clstack, corrstack, shift_equations, shift_equations_map).

The expected outputs come from the numpy restatement (oracle/clm_oracle_np.py)
and are written only when the C oracle (oracle/clm_oracle.c) agrees with it:
clstack and the map exactly, corrstack to 1e-13, shift_equations to 1e-15.
Cases (all seeded):
* random: conjugate-symmetric random rays, K = 6, n_r = 9, n_theta = 16,
  max_shift 2, step 1 (5 shifts);
* planted: n_r = 17, n_theta = 40, K = 3, max_shift 3, step 0.5; line l2 of
  image 2 is line l1 of image 0 times the phase of grid shift j, so that
  pair's answer is (l1+1, l2+1, shift) with correlation 1;
* ties: radially symmetric images (every line is the same vector: the first
  maximum in column-major order, (1, 1) at the first shift) and an image's
  negative (correlation -1 everywhere: no update, line 0).

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import clm_oracle as co  # noqa: E402
from oracle import clm_oracle_np as onp  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "clmatrix_tiny.npz")


def random_pf(rng, n_r, n_theta, K):
    L = n_theta // 2
    pf = np.empty((n_r, n_theta, K), dtype=np.complex128, order="F")
    half = rng.standard_normal((n_r, L, K)) + 1j * rng.standard_normal((n_r, L, K))
    pf[:, :L, :] = half
    pf[:, L:, :] = np.conj(half)
    return pf


def planted_case():
    rng = np.random.default_rng(2027)
    n_r, T, K, max_shift, step = 17, 40, 3, 3.0, 0.5
    R, L = n_r - 1, T // 2
    pf = random_pf(rng, n_r, T, K)
    l1, l2, j = 5, 13, 9
    shift = -max_shift + j * step
    r = np.arange(R)
    m = n_r - 1 - r  # pf rows of the half line
    pf[m, L + l2, 2] = pf[m, L + l1, 0] * np.exp(-2j * np.pi * (r - R) * shift / (2 * R + 1))
    pf[:, l2, 2] = np.conj(pf[:, L + l2, 2])
    return pf, max_shift, step, np.array([l1 + 1, l2 + 1, j])


def ties_case():
    n_r, T = 11, 16
    f = np.linspace(1.0, 0.2, n_r)
    pf = np.empty((n_r, T, 3), dtype=np.complex128, order="F")
    pf[:, :, 0] = f[:, None]
    pf[:, :, 1] = (0.5 * f)[:, None]
    pf[:, :, 2] = -f[:, None]
    return pf, 0.0, 1.0


def main():
    co.build()
    cases = {
        "random": (random_pf(np.random.default_rng(2026), 9, 16, 6), 2.0, 1.0),
        "planted": planted_case()[:3],
        "ties": ties_case(),
    }
    out = {"planted_answer": planted_case()[3]}
    for name, (pf, ms, st) in cases.items():
        cl_n, corr_n, eq_n, map_n = onp.clmatrix(pf, ms, st)
        cl_c, corr_c, eq_c, map_c, _ = co.cryo_clmatrix_cpu(pf, max_shift=ms, shift_step=st)
        assert np.array_equal(cl_c, cl_n), name
        assert np.abs(corr_c - corr_n).max() <= 1e-13, name
        assert np.abs(eq_c.toarray() - eq_n).max() <= 1e-15, name
        assert np.array_equal(map_c, map_n), name
        out[f"{name}_pf"] = np.asarray(pf)
        out[f"{name}_params"] = np.array([ms, st])
        out[f"{name}_clstack"] = cl_n
        out[f"{name}_corrstack"] = corr_n
        out[f"{name}_equations"] = eq_n
        out[f"{name}_map"] = map_n
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
