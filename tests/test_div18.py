"""The vote histogram's division by 2 sigma^2 = 18 (cryo_voteIJ.m:101,
exp(h./(2*sigma.^2))) is computed on the device as q0 = t*RN(1/18), then
two FMAs (the remainder t - 18 q0, then q0 + remainder*RN(1/18)); the
kernel's comment proves the result equals the correctly rounded t/18.  This
checks that claim on the host against IEEE division, bit for bit: on the
kernel's own t values (angles against the 60 histogram points, formed in
the kernel's order), on values spread over every binade the kernel can
produce, and on values next to rounding midpoints of t/18."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

SRC = r"""
#include <math.h>
static double div18(double t) {
    const double R = 1.0 / 18.0;
    const double q0 = t * R;
    const double r = fma(-q0, 18.0, t);
    return fma(r, R, q0);
}
long count_mismatch(const double *t, long n) {
    long bad = 0;
    for (long i = 0; i < n; i++) {
        const double a = div18(t[i]), b = t[i] / 18.0;
        if (*(const long long *)&a != *(const long long *)&b) bad++;
    }
    return bad;
}
"""


@pytest.fixture(scope="module")
def lib():
    d = tempfile.mkdtemp()
    src, so = os.path.join(d, "div18.c"), os.path.join(d, "libdiv18.so")
    open(src, "w").write(SRC)
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fno-strict-aliasing", "-shared", "-fPIC",
                           "-o", so, src, "-lm"])
    L = ctypes.CDLL(so)
    L.count_mismatch.restype = ctypes.c_long
    L.count_mismatch.argtypes = [ctypes.c_void_p, ctypes.c_long]
    return L


def mismatches(lib, t):
    t = np.ascontiguousarray(t, dtype=np.float64)
    return lib.count_mismatch(t.ctypes.data_as(ctypes.c_void_p), t.size)


def test_kernel_values(lib):
    rng = np.random.default_rng(18)
    NT = 60
    x = np.arange(NT) * (180.0 / (NT - 1))
    x[-1] = 180.0
    a = np.concatenate([rng.uniform(0, 180, 20000), rng.uniform(0, 1e-3, 2000), x + 1e-9, x])
    A, X = np.meshgrid(a, x, indexing="ij")
    t = 2 * A * X
    t = t - A * A
    t = t - X * X
    assert mismatches(lib, t.ravel()) == 0


def test_every_binade_and_near_midpoints(lib):
    rng = np.random.default_rng(1818)
    # random mantissas over binades 2^-40 .. 2^16, both signs
    e = rng.integers(-40, 16, 2_000_000)
    t = np.ldexp(rng.uniform(1, 2, e.size), e) * rng.choice([-1.0, 1.0], e.size)
    assert mismatches(lib, t) == 0
    # t next to 18 * (a rounding midpoint of the quotient): the hardest cases
    q = np.ldexp(rng.uniform(1, 2, 500_000), rng.integers(-30, 12, 500_000))
    mid = q + np.spacing(q) / 2
    t = 18.0 * mid
    near = np.concatenate([t, np.nextafter(t, np.inf), np.nextafter(t, -np.inf), -t])
    assert mismatches(lib, near) == 0
