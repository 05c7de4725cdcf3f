"""GPU polar Fourier transform (clm_pft, aspire_amd/csrc/pft.hip) against
the FP64 oracle (oracle/pft_oracle.c: cryo_pft.m:49-75 with nudft2.m:37-47).

* small and workload-shaped stacks (odd and even n, K around the 64-image
  block, single and double images): the GPU's FP64 result is within 1e-12 of
  the oracle's largest value (complex64 output for single images: within
  its rounding), the second half of the rays is the exact conjugate of the
  first, and the host-buffer form returns the device form's bits;
* the full K=5000, n=129, n_r=65, n_theta=360 stack of the common-lines
  workload, checked against an independent evaluation on the GPU (the
  weight matrix materialised with torch and multiplied by rocBLAS ZGEMM,
  aspire_amd/sim.py) -- the oracle would take hours at that size.
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


CASES = [
    # n, n_r, n_theta, K, image dtype
    (17, 9, 16, 7, np.float64),
    (16, 9, 36, 65, np.float64),
    (33, 17, 36, 70, np.float32),
    (65, 33, 360, 130, np.float64),
    (129, 65, 360, 20, np.float64),
    (5, 3, 4, 1, np.float64),
]


@pytest.mark.parametrize("n,n_r,n_theta,K,dt", CASES)
def test_pft_matches_oracle(oracle, n, n_r, n_theta, K, dt):
    say("case pft", n, n_r, n_theta, K, np.dtype(dt).name)
    rng = np.random.default_rng(1000 * n + K)
    p = rng.standard_normal((n, n, K)).astype(dt)
    ref = oracle.cryo_pft(p.astype(np.float64), n_r, n_theta)
    st = {}
    pf, freqs = gpft.cryo_pft(to_dev(p), n_r, n_theta, stats=st)
    got = pf.cpu().numpy()
    assert got.shape == (n_r, n_theta, K)
    assert got.dtype == (np.complex64 if dt == np.float32 else np.complex128)
    scale = np.abs(ref).max()
    err = np.abs(got.astype(np.complex128) - ref).max()
    say("max abs error", err, "max |pf|", scale, st)
    tol = 1e-6 if dt == np.float32 else 1e-12
    assert err <= tol * scale
    L = n_theta // 2
    assert np.array_equal(got[:, L:, :], np.conj(got[:, :L, :]))
    np.testing.assert_array_equal(freqs, oracle.pft_freqs(n_r, n_theta))
    host, _ = gpft.cryo_pft(p, n_r, n_theta)
    assert np.array_equal(host, got)


def test_pft_image_chunks(oracle):
    # more than 8192 images: the stack is staged in chunks
    say("case pft-chunks")
    rng = np.random.default_rng(77)
    n, n_r, n_theta, K = 9, 5, 8, 8200
    p = rng.standard_normal((n, n, K))
    st = {}
    pf, _ = gpft.cryo_pft(to_dev(p), n_r, n_theta, stats=st)
    assert st["chunks"] == 2
    got = pf.cpu().numpy()
    ref = oracle.cryo_pft(p, n_r, n_theta)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()


def test_pft_empty_stack():
    import torch

    say("case pft-empty")
    pf, _ = gpft.cryo_pft(torch.zeros((9, 9, 0), dtype=torch.float64, device="cuda"), 5, 8)
    assert tuple(pf.shape) == (5, 8, 0)


def test_pft_full_workload_vs_zgemm():
    import torch

    from aspire_amd import sim

    n, n_r, n_theta, K = 129, 65, 360, 5000
    say("case pft-full", n, n_r, n_theta, K)
    g = torch.Generator(device="cuda").manual_seed(5)
    imgs = torch.randn((K, n, n), dtype=torch.float64, device="cuda", generator=g)  # imgs[k, u, v]
    st = {}
    pf, _ = gpft.cryo_pft(imgs.permute(1, 2, 0), n_r, n_theta, stats=st)
    say("clm_pft", st)
    err, scale = 0.0, 0.0
    for k0 in range(0, K, 1000):
        ref = sim.cryo_pft(imgs[k0:k0 + 1000], n_r, n_theta)
        d = (pf[:, :, k0:k0 + 1000] - ref).abs().max().item()
        err = max(err, d)
        scale = max(scale, ref.abs().max().item())
        del ref
    say("max abs difference", err, "max |pf|", scale)
    assert err <= 1e-11 * scale
