"""The map filter of cryo_clmatrix's sixth argument in the oracle
(oracle/clm_oracle.c: clmo_filter_map, clmo_pairs_filtered), pinned against
a numpy restatement of sinograms/cryo_average_clmap.m that evaluates it the
way the .m file does -- 2-D FFTs of the extended map and of the weight
square, conjugate product, inverse FFT, the index shift by d -- a different
route from the oracle's direct circular sum:

* filtered maps agree to FFT rounding for several L and radii;
* the filtered pair loop picks the same (cl1, cl2, shift) as a brute-force
  search over those FFT-filtered maps (first maximum in column-major order,
  strict > over shifts) and stores the filtered value.
"""
import numpy as np
import pytest

from oracle import clm_oracle as co


def avg_fft(m, d):
    """cryo_average_clmap.m:20-48 with numpy's FFTs."""
    L = m.shape[0]
    W0 = 1.25 ** np.arange(21, 0, -1)[: 1 + 2 * d]
    a = np.abs(np.arange(-d, d + 1))
    W = W0[a[None, :] + a[:, None]]
    ext = np.vstack([m, m[:, np.r_[L:2 * L, 0:L]]])
    avg = np.fft.ifft2(np.fft.fft2(ext) * np.conj(np.fft.fft2(W, s=ext.shape)))
    rows = np.r_[2 * L - d:2 * L, 0:L - d]
    cols = np.r_[2 * L - d:2 * L, 0:2 * L - d]
    return avg[np.ix_(rows, cols)].real


@pytest.mark.parametrize("L,d", [(8, 1), (18, 2), (30, 3), (10, 4), (12, 10)])
def test_filter_matches_fft_restatement(L, d):
    m = np.random.default_rng(L * 10 + d).standard_normal((L, 2 * L))
    a = co.filter_map(m, d)
    b = avg_fft(m, d)
    assert a.shape == (L, 2 * L)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-14 * np.abs(b).max())


def test_filtered_pair_loop_matches_brute_force():
    from aspire_amd import sim

    st = sim.make_stack(K=6, n=17, n_r=9, n_theta=24, snr=1.0, max_shift_2d=1, seed=21)
    P, _ = co.preprocess(st.pf)
    K, L, R = P.shape
    ms, step, d = 2, 1, 2
    S = co.n_shifts(ms, step)
    k1, k2 = co.all_pairs(K)
    res = co.pairs(P, ms, step, k1, k2, map_filter_radius=d)
    for i in range(k1.size):
        best, key = 0.0, None
        for j in range(S):
            f = avg_fft(co.corr_map(P, ms, step, k1[i], k2[i], j), d)
            col_major = f.T.ravel()
            q = int(np.argmax(col_major))
            if col_major[q] > best:
                best, key = col_major[q], (j, q % L + 1, q // L + 1)
        assert key is not None
        assert (res["j"][i], res["cl1"][i], res["cl2"][i]) == key
        assert abs(res["best"][i] - best) <= 1e-12 * abs(best)
        assert res["shift"][i] == -ms + step * key[0]
