"""Seeded synthetic projection stacks and their polar Fourier transforms.

This is the input producer for the common-lines engine (tests, benchmark,
training data of the shift proposal).  It restates the reference's
simulation conventions:

* rotations: normalized Gaussian quaternions with q0 >= 0 then the
  quaternion-to-matrix map (common/qrand.m:16-26, common/q_to_rot.m:15-25);
* phantom: sum of isotropic Gaussians rho*exp(-|x-x0|^2/w^2) on [-1,1]^3
  (projections/simulation/cryo_gaussian_phantom_3d.m:45-49,86);
* projection: the analytic line integral of each Gaussian along the third
  row of R on the pixel grid t = (-(n-1)/2..(n-1)/2)/((n-1)/2)
  (projections/simulation/cryo_project_gaussian.m:47-53,110-144, isotropic
  case), stored transposed as cryo_gen_projections does after cryo_project
  (cryo_gen_projections.m: permute(projections,[2 1 3])): the first image
  axis runs along the second row of R, the second axis along the first row.
  This is the convention commonline_R/clmatrix_cheat describe;
* shifts: Fourier phase exp(i(wx*sx + wy*sy)), wx = -2*pi*kx/n
  (projections/simulation/cryo_addshifts.m:76-88);
* noise: white Gaussian, each image's noise scaled to unit sample std and
  then by sigma = sqrt(var(image 1)/SNR) (cryo_addnoise.m:36,61-73);
* polar Fourier transform: exact non-uniform DFT at
  freqs = (k*w0*sin(j*dtheta), k*w0*cos(j*dtheta)), w0 = 2*pi/(2n_r-1),
  pf = sum_x,y img(x,y) exp(+i(fx*x + fy*y)), and pf(:,T/2+j) = conj(pf(:,j))
  (fourier/nufft/cryo_pft.m:49-75 with nudft2.m:40-47, the library's exact
  fallback).

Every random draw comes from one numpy PCG64 generator, in the order
rotations, phantom, shifts, noise.  Heavy arithmetic (projection, NUDFT)
runs on a torch device when one is given, else in numpy.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


# ----------------------------------------------------------------- rotations
def qrand(rng: np.random.Generator, K: int) -> np.ndarray:
    """Haar-uniform unit quaternions, 4 x K, with q[0] >= 0 (qrand.m:16-26)."""
    q = rng.standard_normal((4, K))
    q = q / np.sqrt((q ** 2).sum(axis=0, keepdims=True))
    q[:, q[0] < 0] *= -1
    return q


def q_to_rot(q: np.ndarray) -> np.ndarray:
    """3 x 3 x K rotation matrices (q_to_rot.m:15-25)."""
    a, b, c, d = q
    R = np.empty((3, 3, q.shape[1]))
    R[0, 0] = a ** 2 + b ** 2 - c ** 2 - d ** 2
    R[0, 1] = 2 * b * c - 2 * a * d
    R[0, 2] = 2 * a * c + 2 * b * d
    R[1, 0] = 2 * b * c + 2 * a * d
    R[1, 1] = a ** 2 - b ** 2 + c ** 2 - d ** 2
    R[1, 2] = -2 * a * b + 2 * c * d
    R[2, 0] = -2 * a * c + 2 * b * d
    R[2, 1] = 2 * a * b + 2 * c * d
    R[2, 2] = a ** 2 - b ** 2 - c ** 2 + d ** 2
    return R


def rand_rots(rng: np.random.Generator, K: int) -> np.ndarray:
    return q_to_rot(qrand(rng, K))


# ------------------------------------------------------------------- phantom
@dataclass
class Blobs:
    centers: np.ndarray  # B x 3, in a ball of radius 0.5 of the [-1,1]^3 box
    widths: np.ndarray   # B
    amps: np.ndarray     # B


def gaussian_blobs(rng: np.random.Generator, n_blobs: int = 20, radius: float = 0.5,
                   wmin: float = 0.05, wmax: float = 0.15, amin: float = 0.5,
                   amax: float = 1.0, recenter: bool = False) -> Blobs:
    """Centers uniform in the ball (direction x radius^(1/3) rule), widths and
    amplitudes uniform.  With ``recenter`` every center is moved by the same
    vector so that the volume's center of mass (blob mass proportional to
    amplitude x width^3) is the origin: then each projection's centroid is
    its 2-D shift."""
    v = rng.standard_normal((n_blobs, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    r = radius * rng.random(n_blobs) ** (1.0 / 3.0)
    centers = v * r[:, None]
    widths = rng.uniform(wmin, wmax, n_blobs)
    amps = rng.uniform(amin, amax, n_blobs)
    if recenter:
        mass = amps * widths ** 3
        centers = centers - (mass[:, None] * centers).sum(0) / mass.sum()
    return Blobs(centers, widths, amps)


def pixel_grid(n: int, rmax: float = 1.0) -> np.ndarray:
    if n % 2 == 1:
        return (np.arange(n) - (n - 1) / 2) / ((n - 1) / 2) * rmax
    return (np.arange(n) - n / 2 + 0.5) / (n / 2) * rmax


# ---------------------------------------------------------------- backends
def _xp(device):
    if device is None:
        return np, None
    import torch

    return torch, torch.device(device)


def project(blobs: Blobs, rots: np.ndarray, n: int, device=None):
    """Analytic projections, K x n x n (index order [k, x, y]); axis x runs
    along the second row of each R, axis y along the first."""
    K = rots.shape[2]
    t = pixel_grid(n)
    alpha = rots[1, :, :].T  # K x 3: direction of image axis x
    beta = rots[0, :, :].T   # direction of image axis y
    u0 = alpha @ blobs.centers.T  # K x B
    v0 = beta @ blobs.centers.T
    w = blobs.widths
    coef = blobs.amps * w * math.sqrt(math.pi)
    xp, dev = _xp(device)
    if xp is np:
        ex = np.exp(-((t[None, :, None] - u0[:, None, :]) ** 2) / (w ** 2))  # K x n x B
        ey = np.exp(-((t[None, :, None] - v0[:, None, :]) ** 2) / (w ** 2))
        return np.einsum("kxb,b,kyb->kxy", ex, coef, ey, optimize=True)
    import torch

    tt = torch.as_tensor(t, device=dev)
    U0 = torch.as_tensor(u0, device=dev)
    V0 = torch.as_tensor(v0, device=dev)
    W2 = torch.as_tensor(w ** 2, device=dev)
    C = torch.as_tensor(coef, device=dev)
    ex = torch.exp(-((tt[None, :, None] - U0[:, None, :]) ** 2) / W2)
    ey = torch.exp(-((tt[None, :, None] - V0[:, None, :]) ** 2) / W2)
    return torch.bmm(ex * C, ey.transpose(1, 2).contiguous())


def add_shifts(imgs, shifts: np.ndarray):
    """Shift each image by (sx, sy) pixels through its Fourier phase
    (cryo_addshifts.m:76-88).  imgs: K x n x n (numpy or torch)."""
    n = imgs.shape[1]
    rng_ = np.arange(n) - (n - 1) // 2 if n % 2 else np.arange(n) - n // 2
    omega = -2 * np.pi * rng_ / n
    if isinstance(imgs, np.ndarray):
        F = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(imgs, axes=(1, 2))), axes=(1, 2))
        ph = np.exp(1j * (omega[None, :, None] * shifts[:, 0, None, None]
                          + omega[None, None, :] * shifts[:, 1, None, None]))
        return np.real(np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(F * ph, axes=(1, 2))), axes=(1, 2)))
    import torch

    dev = imgs.device
    om = torch.as_tensor(omega, device=dev)
    sh = torch.as_tensor(shifts, device=dev)
    F = torch.fft.fftshift(torch.fft.fft2(torch.fft.ifftshift(imgs, dim=(1, 2))), dim=(1, 2))
    ph = torch.exp(1j * (om[None, :, None] * sh[:, 0, None, None] + om[None, None, :] * sh[:, 1, None, None]))
    return torch.real(torch.fft.fftshift(torch.fft.ifft2(torch.fft.ifftshift(F * ph, dim=(1, 2))), dim=(1, 2)))


def add_noise(rng: np.random.Generator, imgs, snr: float):
    """White noise at the given SNR, sigma from the variance of image 1
    (cryo_addnoise.m:36,61-73)."""
    K, n, _ = imgs.shape
    first = imgs[0].reshape(-1)
    if isinstance(imgs, np.ndarray):
        sigma = math.sqrt(float(np.var(first, ddof=1)) / snr)
    else:
        sigma = math.sqrt(float(first.var(unbiased=True)) / snr)
    g = rng.standard_normal((K, n, n))
    g /= g.reshape(K, -1).std(axis=1, ddof=1)[:, None, None]
    g *= sigma
    if isinstance(imgs, np.ndarray):
        return imgs + g
    import torch

    return imgs + torch.as_tensor(g, device=imgs.device)


def pft_freqs(n_r: int, n_theta: int) -> np.ndarray:
    """(n_r * n_theta/2) x 2 sampling frequencies of cryo_pft.m:49-58."""
    omega0 = 2 * np.pi / (2 * n_r - 1)
    dtheta = 2 * np.pi / n_theta
    j = np.arange(n_theta // 2)
    k = np.arange(n_r)
    fx = (k[None, :] * omega0) * np.sin(j[:, None] * dtheta)
    fy = (k[None, :] * omega0) * np.cos(j[:, None] * dtheta)
    return np.stack([fx.ravel(), fy.ravel()], axis=1)


def cryo_pft(imgs, n_r: int, n_theta: int, chunk: int = 2048):
    """Polar Fourier transform by exact NUDFT.  imgs: K x n x n.

    Returns pf with shape (n_r, n_theta, K) whose memory is column-major
    (a transposed view of a contiguous K x n_theta x n_r array), complex128.
    """
    if n_theta % 2:
        raise ValueError("`n_theta` must be even.")
    K, n, _ = imgs.shape
    g = np.ceil(np.arange(-n / 2, n / 2)).astype(np.float64)  # nudft2.m:40
    freqs = pft_freqs(n_r, n_theta)
    M = freqs.shape[0]
    L = n_theta // 2
    if isinstance(imgs, np.ndarray):
        X = imgs.reshape(K, n * n).T.astype(np.complex128)  # (x,y) x K, x fastest? see below
        # image index [k, x, y] flattened as x*n + y; points (gx[x], gy[y])
        out = np.empty((M, K), dtype=np.complex128)
        gx = np.repeat(g, n)
        gy = np.tile(g, n)
        for m0 in range(0, M, chunk):
            f = freqs[m0:m0 + chunk]
            E = np.exp(1j * (f[:, 0:1] * gx[None, :] + f[:, 1:2] * gy[None, :]))
            out[m0:m0 + chunk] = E @ X
        half = out.reshape(L, n_r, K)  # [j, k, img]
        full = np.empty((K, n_theta, n_r), dtype=np.complex128)
        full[:, :L, :] = np.transpose(half, (2, 0, 1))
        full[:, L:, :] = np.conj(full[:, :L, :])
        return full.transpose(2, 1, 0)
    import torch

    dev = imgs.device
    X = imgs.reshape(K, n * n).to(torch.complex128).T.contiguous()
    gx = torch.as_tensor(np.repeat(g, n), device=dev)
    gy = torch.as_tensor(np.tile(g, n), device=dev)
    F = torch.as_tensor(freqs, device=dev)
    out = torch.empty((M, K), dtype=torch.complex128, device=dev)
    for m0 in range(0, M, chunk):
        f = F[m0:m0 + chunk]
        E = torch.exp(1j * (f[:, 0:1] * gx[None, :] + f[:, 1:2] * gy[None, :]))
        out[m0:m0 + chunk] = E @ X
    half = out.reshape(L, n_r, K)
    full = torch.empty((K, n_theta, n_r), dtype=torch.complex128, device=dev)
    full[:, :L, :] = half.permute(2, 0, 1)
    full[:, L:, :] = torch.conj(full[:, :L, :])
    return full.permute(2, 1, 0)


@dataclass
class Stack:
    pf: object            # n_r x n_theta x K (column-major memory)
    rots: np.ndarray      # 3 x 3 x K
    shifts: np.ndarray    # K x 2 (pixels, image axes x,y)
    blobs: Blobs
    n: int
    snr: float | None


def make_images(K: int, n: int, snr: float | None = None, max_shift_2d: float = 0.0, shift_kind: str = "int",
                seed: int = 0, n_blobs: int = 20, device=None, recenter: bool = False):
    """The images of :func:`make_stack` (K x n x n, imgs[k, u, v] = p(u, v, k)),
    with their rotations, shifts and phantom: (imgs, rots, shifts, blobs)."""
    rng = np.random.default_rng(seed)
    rots = rand_rots(rng, K)
    blobs = gaussian_blobs(rng, n_blobs, recenter=recenter)
    if max_shift_2d > 0:
        if shift_kind == "int":
            shifts = rng.integers(-int(max_shift_2d), int(max_shift_2d) + 1, size=(K, 2)).astype(np.float64)
        else:
            shifts = rng.uniform(-max_shift_2d, max_shift_2d, size=(K, 2))
    else:
        shifts = np.zeros((K, 2))
    imgs = project(blobs, rots, n, device)
    if max_shift_2d > 0:
        imgs = add_shifts(imgs, shifts)
    if snr is not None:
        imgs = add_noise(rng, imgs, snr)
    return imgs, rots, shifts, blobs


def make_stack(K: int, n: int, n_r: int, n_theta: int = 360, snr: float | None = None,
               max_shift_2d: float = 0.0, shift_kind: str = "int", seed: int = 0,
               n_blobs: int = 20, device=None, chunk_images: int = 1024,
               dtype: str = "complex128", recenter: bool = False) -> Stack:
    """A seeded stack: Haar rotations, Gaussian-blob projections, 2-D shifts
    (integer or uniform in [-max_shift_2d, max_shift_2d]), white noise at
    ``snr`` (None = clean) and the polar Fourier transform."""
    imgs, rots, shifts, blobs = make_images(K, n, snr, max_shift_2d, shift_kind, seed, n_blobs, device, recenter)
    if device is None:
        pf = cryo_pft(imgs, n_r, n_theta)
        if dtype == "complex64":
            pf = np.ascontiguousarray(pf.transpose(2, 1, 0).astype(np.complex64)).transpose(2, 1, 0)
    else:
        import torch

        parts = []
        for k0 in range(0, K, chunk_images):
            parts.append(cryo_pft(imgs[k0:k0 + chunk_images], n_r, n_theta).permute(2, 1, 0))
        base = torch.cat(parts, dim=0)
        if dtype == "complex64":
            base = base.to(torch.complex64)
        pf = base.permute(2, 1, 0)
        del imgs
    return Stack(pf=pf, rots=rots, shifts=shifts, blobs=blobs, n=n, snr=snr)


# ------------------------------------------------------------ geometric truth
def commonline_R(Ri: np.ndarray, Rj: np.ndarray, L: int):
    """0-based common-line indices (l_ij, l_ji) of two rotations
    (sinograms/commonline_R.m:22-35)."""
    Ut = Rj @ Ri.T
    aij = math.atan2(Ut[2, 0], -Ut[2, 1]) + math.pi
    aji = math.atan2(-Ut[0, 2], Ut[1, 2]) + math.pi
    lij = int(round(aij / (2 * math.pi) * L)) % L
    lji = int(round(aji / (2 * math.pi) * L)) % L
    return lij, lji


def clmatrix_cheat(rots: np.ndarray, n_theta: int) -> np.ndarray:
    """True common-lines matrix, 1-based (sinograms/clmatrix_cheat.m:22-31)."""
    K = rots.shape[2]
    cl = np.zeros((K, K))
    for k1 in range(K - 1):
        for k2 in range(k1 + 1, K):
            a, b = commonline_R(rots[:, :, k1], rots[:, :, k2], n_theta)
            cl[k1, k2] = a + 1
            cl[k2, k1] = b + 1
    return cl


def matched_lines(l1, l2, t1, t2, n_theta: int, max_angle: float) -> np.ndarray:
    """Per pair: both detected lines (l1, l2) within max_angle degrees of the
    true ones (t1, t2), same or flipped orientation (sinograms/comparecl.m:20-52)."""
    max_angle = max_angle * math.pi / 180
    half = n_theta / 2
    tol = 2 * math.sin(max_angle / 2) + 1.0e-10
    alpha = 2j * math.pi / (2 * half)
    e = lambda x: np.exp(alpha * (np.asarray(x, dtype=np.float64) - 1))
    d1s = np.abs(e(l1) - e(t1))
    d2s = np.abs(e(l2) - e(t2))
    d1f = np.abs(e(l1) * np.exp(1j * math.pi) - e(t1))
    d2f = np.abs(e(l2) * np.exp(1j * math.pi) - e(t2))
    return ((d1s <= tol) & (d2s <= tol)) | ((d1f <= tol) & (d2f <= tol))


def comparecl(clstack: np.ndarray, ref: np.ndarray, n_theta: int, max_angle: float) -> float:
    """Fraction of pairs whose detected lines are within max_angle degrees of
    the true ones, same or flipped orientation (sinograms/comparecl.m:20-52)."""
    K = clstack.shape[0]
    k1, k2 = np.triu_indices(K, 1)
    t1, t2 = ref[k1, k2], ref[k2, k1]
    use = t1 != 0
    ok = matched_lines(clstack[k1, k2], clstack[k2, k1], t1, t2, n_theta, max_angle)
    return float(ok[use].sum() / max(use.sum(), 1))


def true_shift_1d(cl1: int, cl2: int, L: int, s1, s2) -> float:
    """Reference 1-D shift of a detected common line from the true 2-D shifts
    (cryo_clmatrix_cpu.m:667-683)."""
    dtheta = math.pi / L
    a = (cl1 - 1) * dtheta
    b = (cl2 - 1) * dtheta
    if b < math.pi - 1.0e-13:
        return math.sin(a) * s1[0] + math.cos(a) * s1[1] - math.sin(b) * s2[0] - math.cos(b) * s2[1]
    b -= math.pi
    return -math.sin(a) * s1[0] - math.cos(a) * s1[1] - math.sin(b) * s2[0] - math.cos(b) * s2[1]
