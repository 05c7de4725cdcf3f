"""The polar Fourier transform oracle (oracle/pft_oracle.c), pinned without
restating its formula:

* the theta = 0 ray is a 1-D DFT of the image's sums over its first index,
  and the theta = 90 degree ray one of its sums over the second index,
  computed here with numpy's FFT (w0 = 2*pi/(2 n_r - 1) makes the radial
  samples DFT frequencies of length 2 n_r - 1); this fixes w0, the pixel grid
  ceil(t - n/2) for odd and even n, the sign of the exponent and which image
  index goes with sin(theta) (cryo_pft.m:49-58,70; nudft2.m:40-47);
* rotating every image by 90 degrees moves every ray by n_theta/4 columns
  (all angles, both halves of pf);
* the second half of the rays is the conjugate of the first (:75), exactly;
* a unit pixel at the grid origin transforms to 1 everywhere;
* libclm's clm_pft_freqs equals the oracle's freqs (host code, no GPU).
"""
import numpy as np
import pytest

from aspire_amd import pft as gpft
from oracle import clm_oracle as co


def fft_ray(sums: np.ndarray, n: int, n_r: int) -> np.ndarray:
    """sum_t s_t exp(+2 pi i k g_t / N), k < n_r, g_t = ceil(t - n/2), N = 2 n_r - 1."""
    N = 2 * n_r - 1
    assert n <= N
    a = np.zeros(N, dtype=np.complex128)
    g = np.ceil(np.arange(n) - n / 2).astype(int)
    a[g % N] = sums
    return (N * np.fft.ifft(a))[:n_r]


@pytest.mark.parametrize("n,n_r", [(17, 9), (16, 9), (33, 17), (12, 7)])
def test_axis_rays_are_ffts(n, n_r):
    rng = np.random.default_rng(n)
    K, n_theta = 3, 36
    p = rng.standard_normal((n, n, K))
    pf = co.cryo_pft(p, n_r, n_theta)
    for k in range(K):
        # theta = 0: frequencies (0, k w0): sum over the first index u
        ref0 = fft_ray(p[:, :, k].sum(axis=0), n, n_r)
        np.testing.assert_allclose(pf[:, 0, k], ref0, rtol=0, atol=1e-12 * np.abs(ref0).max())
        # theta = 90 deg (column n_theta/4): frequencies (k w0, ~0): sum over v
        ref90 = fft_ray(p[:, :, k].sum(axis=1), n, n_r)
        np.testing.assert_allclose(pf[:, n_theta // 4, k], ref90, rtol=0, atol=1e-12 * np.abs(ref90).max())


@pytest.mark.parametrize("n,n_r,n_theta", [(17, 9, 16), (33, 17, 360), (9, 5, 8)])
def test_rotation_moves_rays(n, n_r, n_theta):
    rng = np.random.default_rng(7)
    p = rng.standard_normal((n, n, 2))
    rot = np.transpose(p[:, ::-1, :], (1, 0, 2))  # rot(a, b) = p(b, n-1-a)
    pf = co.cryo_pft(p, n_r, n_theta)
    pr = co.cryo_pft(rot, n_r, n_theta)
    shifted = np.roll(pf, -(n_theta // 4), axis=1)  # pf(:, j + n_theta/4)
    np.testing.assert_allclose(pr, shifted, rtol=0, atol=1e-12 * np.abs(pf).max())


def test_conjugate_half_and_origin_delta():
    for n in (15, 16):
        n_r, n_theta = 8, 20
        rng = np.random.default_rng(n)
        p = rng.standard_normal((n, n, 2))
        pf = co.cryo_pft(p, n_r, n_theta)
        L = n_theta // 2
        assert np.array_equal(pf[:, L:, :], np.conj(pf[:, :L, :]))
        d = np.zeros((n, n))
        c = n // 2  # g(c) = ceil(c - n/2) = 0 for odd and even n
        d[c, c] = 1.0
        pd = co.cryo_pft(d, n_r, n_theta)
        np.testing.assert_allclose(pd[:, :, 0], np.ones((n_r, n_theta)), rtol=0, atol=1e-15)


def test_library_freqs_match_oracle():
    for n_r, n_theta in ((9, 16), (65, 360), (33, 36)):
        a = gpft.pft_freqs(n_r, n_theta)
        b = co.pft_freqs(n_r, n_theta)
        assert a.shape == (n_r * n_theta // 2, 2)
        np.testing.assert_array_equal(a, b)
        # ray 0 runs along the second image axis, the first radial sample is 0
        assert np.all(a[:n_r, 0] == 0.0) and np.all(a[::n_r] == 0.0)


def test_argument_errors():
    with pytest.raises(ValueError, match="must be even"):
        gpft.cryo_pft(np.zeros((5, 5)), 3, 7)
    with pytest.raises(ValueError, match="square"):
        gpft.cryo_pft(np.zeros((5, 6)), 3, 8)
    with pytest.raises(ValueError, match="single"):
        gpft.cryo_pft(np.zeros((5, 5), dtype=np.int32), 3, 8)
