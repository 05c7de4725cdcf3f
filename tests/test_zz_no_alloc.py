"""The allocation-free entry point (clm_clmatrix_device_noalloc) against the
allocating one (clm_clmatrix_device), through the C ABI.

Both run the same kernels on the same inputs, so every output must be
bit-identical.  Two calls in a row (the second on a smaller stack) check
that what the first call leaves in the shared scratch array does not leak
into the second; a shape beyond the scratch array is refused before any
kernel runs, and the context keeps working afterwards.
"""
import ctypes
import sys

import numpy as np
import pytest

from aspire_amd import _lib, sim
from aspire_amd.clmatrix import _device_view, _params

pytestmark = pytest.mark.gpu

KEYS = ("cl1", "cl2", "j", "corr", "shift")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run(ctx, entry, pf, ms, step, precision):
    """One call of `entry` with device outputs; returns (host records, stats, rc)."""
    import torch

    base, dt, n_r, n_theta, K = _device_view(pf)
    P = K * (K - 1) // 2
    dev = base.device
    out = dict(cl1=torch.full((P,), -7, dtype=torch.int32, device=dev),
               cl2=torch.full((P,), -7, dtype=torch.int32, device=dev),
               j=torch.full((P,), -7, dtype=torch.int32, device=dev),
               corr=torch.full((P,), -7.0, dtype=torch.float64, device=dev),
               shift=torch.full((P,), -7.0, dtype=torch.float64, device=dev))
    p = _params(n_r, n_theta, K, dt, ms, step, precision, 0, 1e-3, 0)
    stats = _lib.ClmStats()
    torch.cuda.synchronize()
    rc = getattr(_lib.load(), entry)(ctx.handle, ctypes.byref(p), ctypes.c_void_p(base.data_ptr()),
                                     *[ctypes.c_void_p(out[k].data_ptr()) for k in KEYS], ctypes.byref(stats))
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}, stats.as_dict(), rc


def to_dev(pf):
    import torch

    return torch.as_tensor(np.ascontiguousarray(pf.transpose(2, 1, 0))).cuda().permute(2, 1, 0)


@pytest.fixture(scope="module")
def stack():
    # the K=5000 workload's per-pair shape (n=129, n_r=65, 41 shifts) at K=40
    st = sim.make_stack(K=40, n=129, n_r=65, n_theta=360, snr=1 / 16, max_shift_2d=5.0, shift_kind="uniform",
                        seed=1203)
    return to_dev(st.pf)


def same(a, b):
    return all(np.array_equal(a[k], b[k]) for k in KEYS)


def test_noalloc_matches_device_entry_two_calls(stack):
    import torch

    ctx = _lib.Context(0)
    ref_ctx = _lib.context(0)
    for prec in ("bf16", "fp32", "fp64"):
        big, small = stack, stack[:, :, :17]
        say("case noalloc", prec, "K=40 then K=17")
        ref_big, st_big, rc = run(ref_ctx, "clm_clmatrix_device", big, 10, 0.5, prec)
        assert rc == 0, _lib.load().clm_last_error(ref_ctx.handle)
        ref_small, _, rc = run(ref_ctx, "clm_clmatrix_device", small, 10, 0.5, prec)
        assert rc == 0
        free0 = torch.cuda.mem_get_info(0)[0]
        got_big, sg, rc = run(ctx, "clm_clmatrix_device_noalloc", big, 10, 0.5, prec)
        assert rc == 0, _lib.load().clm_last_error(ctx.handle)
        got_small, _, rc = run(ctx, "clm_clmatrix_device_noalloc", small, 10, 0.5, prec)
        assert rc == 0, _lib.load().clm_last_error(ctx.handle)
        free1 = torch.cuda.mem_get_info(0)[0]
        say(prec, "stats", sg, "free-memory change (MiB)", (free1 - free0) / 2**20)
        assert same(got_big, ref_big), prec
        assert same(got_small, ref_small), prec
        assert sg["variant"] == st_big["variant"] and sg["pairs"] == st_big["pairs"]
        # the allocating path's BF16 job at K=40 takes about 245 MiB of device buffers
        assert free0 - free1 < 64 << 20


def test_noalloc_rejects_beyond_scratch_and_context_survives(stack):
    import torch

    say("case noalloc-too-large")
    ctx = _lib.Context(0)
    rng = np.random.default_rng(5)
    K = 160  # A-stacks alone: 160 x 1.9 MB (BF16) + 3.8 MB (FP32 twins) > 512 MiB
    big = torch.as_tensor(rng.standard_normal((K, 360, 65)) + 1j * rng.standard_normal((K, 360, 65))).cuda()
    _, _, rc = run(ctx, "clm_clmatrix_device_noalloc", big.permute(2, 1, 0), 10, 0.5, "bf16")
    msg = _lib.load().clm_last_error(ctx.handle).decode()
    say(rc, msg)
    assert rc == -4 and "scratch" in msg
    # the same context still runs both entry points afterwards
    small = stack[:, :, :9]
    a, _, rc = run(ctx, "clm_clmatrix_device_noalloc", small, 10, 0.5, "bf16")
    assert rc == 0
    b, _, rc = run(ctx, "clm_clmatrix_device", small, 10, 0.5, "bf16")
    assert rc == 0
    assert same(a, b)
