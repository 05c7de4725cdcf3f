"""The NUFFT method of the GPU polar Fourier transform (clm_pft with
CLM_PFT_NUFFT: deconvolution onto an oversampled grid, hipFFT, kernel
interpolation at the polar points) against the FP64 oracle's exact sums
(oracle/pft_oracle.c: cryo_pft.m:49-75 with nudft2.m:37-47).

The method approximates the sums; with its kernel of width 16 on a grid of
side >= 2n it is within 1e-13 of the largest |pf| (measured about 1.5e-14,
the accuracy finufft gives at the epsilon of 1e-15 the reference asks for,
nufft_initialize.m:29), on odd and even n, tiny images whose kernel window
is wider than the image, single images (complex64 output: within its
rounding), and more than one batch of grids.  The second half of the rays
is the exact conjugate of the first.  At the common-lines workload's shape
it is checked against the exact method on the GPU.
"""
import sys

import numpy as np
import pytest

from aspire_amd import pft as gpft

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(p):
    import torch

    return torch.as_tensor(np.ascontiguousarray(p.transpose(2, 1, 0))).cuda().permute(2, 1, 0)


def smooth_at_least(n):
    m = n
    while True:
        r = m
        for f in (2, 3, 5):
            while r % f == 0:
                r //= f
        if r == 1:
            return m
        m += 1


CASES = [
    # n, n_r, n_theta, K, image dtype
    (17, 9, 16, 7, np.float64),
    (16, 9, 36, 65, np.float64),
    (33, 17, 36, 70, np.float32),
    (65, 33, 360, 13, np.float64),
    (129, 65, 360, 5, np.float64),
    (5, 3, 4, 1, np.float64),
    (1, 2, 2, 3, np.float64),
]


@pytest.mark.parametrize("n,n_r,n_theta,K,dt", CASES)
def test_nufft_matches_oracle(oracle, n, n_r, n_theta, K, dt):
    say("case pft-nufft", n, n_r, n_theta, K, np.dtype(dt).name)
    rng = np.random.default_rng(2000 * n + K)
    p = rng.standard_normal((n, n, K)).astype(dt)
    ref = oracle.cryo_pft(p.astype(np.float64), n_r, n_theta)
    st = {}
    pf, freqs = gpft.cryo_pft(to_dev(p), n_r, n_theta, stats=st, method="nufft")
    got = pf.cpu().numpy()
    assert got.shape == (n_r, n_theta, K)
    assert got.dtype == (np.complex64 if dt == np.float32 else np.complex128)
    assert st["grid_side"] == smooth_at_least(max(2 * n, 32))
    scale = np.abs(ref).max()
    err = np.abs(got.astype(np.complex128) - ref).max()
    say("max abs error", err, "relative to max |pf|", err / scale, st)
    tol = 1e-6 if dt == np.float32 else 1e-13
    assert err <= tol * scale
    L = n_theta // 2
    assert np.array_equal(got[:, L:, :], np.conj(got[:, :L, :]))
    # the host-buffer form returns the device form's bits
    host, _ = gpft.cryo_pft(p, n_r, n_theta, method="nufft")
    assert np.array_equal(host, got)


def test_nufft_batches(oracle):
    # more than 2048 images: the grids go through the FFT in two batches
    say("case pft-nufft-batches")
    rng = np.random.default_rng(78)
    n, n_r, n_theta, K = 9, 5, 8, 2100
    p = rng.standard_normal((n, n, K))
    st = {}
    pf, _ = gpft.cryo_pft(to_dev(p), n_r, n_theta, stats=st, method="nufft")
    assert st["chunks"] == 2
    got = pf.cpu().numpy()
    ref = oracle.cryo_pft(p, n_r, n_theta)
    assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max()


def test_nufft_matches_exact_at_workload_shape():
    # the common-lines workload's images: n=129, n_r=65, n_theta=360
    say("case pft-nufft-workload")
    import torch

    from aspire_amd import sim

    imgs, _, _, _ = sim.make_images(K=300, n=129, snr=1 / 16, max_shift_2d=5.0, shift_kind="uniform", seed=1003,
                                    device="cuda")
    # column-major n x n x K with p(u, v, k) = imgs[k, u, v]
    imgs = imgs.transpose(1, 2).contiguous().permute(2, 1, 0)
    a, _ = gpft.cryo_pft(imgs, 65, 360)
    s1 = {}
    b, _ = gpft.cryo_pft(imgs, 65, 360, stats=s1, method="nufft")
    torch.cuda.synchronize()
    err = (a - b).abs().max().item()
    scale = a.abs().max().item()
    say("nufft vs exact", err, err / scale, s1)
    assert err <= 1e-13 * scale


def test_nufft_rejects_unknown_method():
    with pytest.raises(ValueError):
        gpft.cryo_pft(np.zeros((5, 5, 1)), 3, 4, method="fast")
