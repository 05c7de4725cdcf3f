"""The measured workloads, defined once for the benchmark and the GPU tests.

Each workload is a seeded synthetic stack (:func:`aspire_amd.sim.make_stack`:
Gaussian-blob phantom, Haar rotations, 2-D shifts, white noise, exact polar
NUDFT) plus the 1-D shift grid of the common-lines search:

* ``K100``: 100 images of 65 px, no shifts, SNR 1, n_r=33, max_shift 0 (one
  shift) -- the plain CPU method's workload;
* ``K1000``: 1000 images of 89 px, integer shifts in [-4,4], SNR 1/8, n_r=45,
  max_shift 6, step 1 (13 shifts) -- precision forms against plain FP32;
* ``K5000``: 5000 images of 129 px, shifts U[-5,5], SNR 1/16, n_r=65,
  max_shift 10, step 0.5 (41 shifts), 12,497,500 pairs -- the headline;
* ``GROUPED``: 64 independent stacks, each with its own phantom (seed
  4000+j): K ~ U[50,2000], n_r in {33,45,65} (images of 2 n_r - 1 px),
  max_shift in {0,4,8}, step 1, SNR in {1,1/8,1/32}, complex64 rays;
* ``LEARNED``: 5000 images of 129 px, shifts U[-12,12], SNR 1/16, n_r=65,
  full grid max_shift 12, step 0.5 (49 shifts) against a +-1 px window.
"""
from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np


@dataclass(frozen=True)
class Workload:
    name: str
    K: int
    n: int
    n_r: int
    snr: float | None
    shifts_2d: float
    shift_kind: str
    max_shift: float
    shift_step: float
    seed: int
    n_theta: int = 360
    dtype: str = "complex128"

    @property
    def pairs(self) -> int:
        return self.K * (self.K - 1) // 2

    @property
    def n_shifts(self) -> int:
        return int(np.ceil(2 * self.max_shift / self.shift_step + 1))

    def describe(self) -> str:
        snr = "clean" if self.snr is None else f"SNR={self.snr:g}"
        return (f"K={self.K} n={self.n} n_r={self.n_r} n_theta={self.n_theta} {snr} "
                f"shifts {self.shift_kind} [-{self.shifts_2d:g},{self.shifts_2d:g}] "
                f"max_shift={self.max_shift:g} shift_step={self.shift_step:g} ({self.n_shifts} shifts)")

    def with_K(self, K: int) -> "Workload":
        return replace(self, K=int(K))

    def make(self, device="cuda"):
        """The seeded stack (:class:`aspire_amd.sim.Stack`); ``pf`` is a
        column-major n_r x n_theta x K device tensor for ``device="cuda"``."""
        from . import sim

        return sim.make_stack(K=self.K, n=self.n, n_r=self.n_r, n_theta=self.n_theta, snr=self.snr,
                              max_shift_2d=self.shifts_2d, shift_kind=self.shift_kind, seed=self.seed,
                              device=device, dtype=self.dtype)


K100 = Workload("K100", K=100, n=65, n_r=33, snr=1.0, shifts_2d=0.0, shift_kind="int", max_shift=0, shift_step=1,
                seed=1001)
K1000 = Workload("K1000", K=1000, n=89, n_r=45, snr=1 / 8, shifts_2d=4.0, shift_kind="int", max_shift=6,
                 shift_step=1, seed=1002)
K5000 = Workload("K5000", K=5000, n=129, n_r=65, snr=1 / 16, shifts_2d=5.0, shift_kind="uniform", max_shift=10,
                 shift_step=0.5, seed=1003)
LEARNED = Workload("LEARNED", K=5000, n=129, n_r=65, snr=1 / 16, shifts_2d=12.0, shift_kind="uniform",
                   max_shift=12, shift_step=0.5, seed=1005)


def grouped(n_stacks: int = 64, seed: int = 4000):
    """The uneven multi-stack batch: one :class:`Workload` per stack."""
    rng = np.random.default_rng(seed)
    out = []
    for j in range(n_stacks):
        K = int(rng.integers(50, 2001))
        n_r = int(rng.choice([33, 45, 65]))
        ms = int(rng.choice([0, 4, 8]))
        snr = float(rng.choice([1.0, 1 / 8, 1 / 32]))
        out.append(Workload(f"GROUPED[{j}]", K=K, n=2 * n_r - 1, n_r=n_r, snr=snr, shifts_2d=ms / 2,
                            shift_kind="uniform", max_shift=ms, shift_step=1, seed=seed + j, dtype="complex64"))
    return out


def sample_pairs(K: int, n: int, seed: int = 42):
    """A seeded uniform sample of ``n`` pairs (all pairs if n >= K(K-1)/2):
    (equation slots ascending, k1, k2), 0-based, k1 < k2."""
    P = K * (K - 1) // 2
    rng = np.random.default_rng(seed)
    slots = np.arange(P, dtype=np.int64) if n >= P else np.sort(rng.choice(P, size=n, replace=False))
    k1, k2 = pairs_of_slots(K, slots)
    return slots.astype(np.int64), k1, k2


def pairs_of_slots(K: int, slots):
    """(k1, k2), 0-based, of equation slots e = k1 K - k1(k1+1)/2 + (k2-k1-1)
    (the row order of cryo_clmatrix_cpu.m:722-741)."""
    slots = np.asarray(slots, dtype=np.int64)
    firsts = np.array([k * K - k * (k + 1) // 2 for k in range(K)], dtype=np.int64)
    k1 = np.searchsorted(firsts, slots, side="right") - 1
    k2 = slots - firsts[k1] + k1 + 1
    return k1.astype(np.int64), k2.astype(np.int64)
