"""Shift network kernels and the windowed search on the GPU.

* features, forward and gradient kernels against the CPU oracle
  (oracle/shiftnet_oracle.py, itself pinned against torch autograd and
  finite differences) on the same BF16 inputs and weights;
* a short training run with the project's own kernels lowers the loss;
* the windowed search: every pair it settles reports a reference-grid entry
  (its corr equals the FP64 oracle's value at the reported line pair and
  shift), every pair it hands to the exhaustive search gets exactly the
  full-grid result, and with exact proposals it agrees with the full grid
  on the pairs it settles.
"""
import sys

import numpy as np
import pytest

from aspire_amd import sim
from aspire_amd.clmatrix import clmatrix_records
from aspire_amd.shiftnet import ShiftNet, clmatrix_windowed, init_params
from oracle import shiftnet_oracle as so

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def to_dev(pf):
    import torch

    return torch.as_tensor(np.ascontiguousarray(pf.transpose(2, 1, 0))).cuda().permute(2, 1, 0)


@pytest.fixture(scope="module")
def small():
    return sim.make_stack(K=12, n=65, n_r=33, n_theta=360, snr=1 / 2, max_shift_2d=4, shift_kind="uniform",
                          seed=6001, recenter=True)


def bits(t):
    import torch

    return t.view(torch.int16).cpu().numpy().astype(np.int32) & 0xFFFF


def test_features_match_oracle(small):
    say("case shiftnet-features")
    net = ShiftNet()
    x = net.features(to_dev(small.pf))
    ref = so.bf16_bits(so.features(small.pf, quant=False)).astype(np.int32)
    got = bits(x)
    ref[ref == 0x8000] = 0  # -0 and +0 are the same value
    got[got == 0x8000] = 0
    diff = np.abs(got - ref)
    say("features differing:", int((diff != 0).sum()), "of", diff.size)
    assert diff.max() <= 1 and (diff != 0).mean() <= 1e-3


def test_forward_and_grad_match_oracle(small):
    import torch

    say("case shiftnet-forward-grad")
    theta = init_params(3)
    net = ShiftNet(params=theta)
    x = net.features(to_dev(small.pf))
    xq = so.bf16_value(bits(x).astype(np.uint16)).reshape(x.shape)
    out = net.forward(x).cpu().numpy()
    ref = so.predict(theta, xq, quant=True)
    say("forward max abs diff", float(np.abs(out - ref).max()), "scale", float(np.abs(ref).max()))
    np.testing.assert_allclose(out, ref, rtol=2e-3, atol=2e-3 * float(np.abs(ref).max()))
    tgt = np.random.default_rng(1).uniform(-4, 4, size=(x.shape[0], 2)).astype(np.float32)
    loss, g = net.loss_grad(x, torch.as_tensor(tgt, device=net.dev))
    g = g.cpu().numpy().astype(np.float64)
    rl, rg = so.loss_grad(theta, xq, tgt, quant=True)
    say("loss", loss, rl)
    assert abs(loss - rl) <= 1e-3 * rl
    off = 0
    for name, shp in so.SHAPES:
        n = int(np.prod(shp))
        a, b = g[off:off + n], rg[off:off + n]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        say(f"grad {name}: rel {rel:.2e}")
        assert rel <= 1e-2, name
        off += n


def test_training_lowers_loss():
    import torch

    say("case shiftnet-train")
    st = sim.make_stack(K=512, n=65, n_r=33, n_theta=360, snr=1.0, max_shift_2d=6, shift_kind="uniform",
                        seed=6002, device="cuda", recenter=True)
    net = ShiftNet(seed=1)
    x = net.features(st.pf)
    y = torch.as_tensor(st.shifts, dtype=torch.float32, device=net.dev)
    l0 = float(((net.forward(x[448:]) - y[448:]) ** 2).sum(1).mean())
    g = torch.Generator().manual_seed(0)
    for step in range(300):
        idx = torch.randint(0, 448, (64,), generator=g).to(net.dev)
        net.loss_grad(x.index_select(0, idx), y.index_select(0, idx))
        net.adam(lr=2e-3)
    l1 = float(((net.forward(x[448:]) - y[448:]) ** 2).sum(1).mean())
    say("held-out mean squared error (px^2):", l0, "->", l1)
    assert l1 < 0.25 * l0


@pytest.fixture(scope="module")
def shifted():
    # learned-form shape at small K: n=129, n_r=65, shifts U[-12,12], 49-shift grid
    return sim.make_stack(K=40, n=129, n_r=65, n_theta=360, snr=1 / 4, max_shift_2d=12, shift_kind="uniform",
                          seed=6003)


def window_offsets(prop, K, L, step):
    """Per-line 1-D shares a_k(l), rounded to the step (as the kernel does)."""
    th = np.arange(L) * np.pi / L
    a = np.sin(th)[None, :] * prop[:, 0:1].astype(np.float64) + np.cos(th)[None, :] * prop[:, 1:2].astype(np.float64)
    return np.rint(a / step) * step


@pytest.mark.parametrize("prec", ["bf16", "fp16"])
def test_windowed_search_properties(oracle, shifted, prec):
    """Window shifts are reference-grid shifts, so for every pair the window
    settles: (a) its corr is the FP64 value of a grid entry, (b) it never
    exceeds the full-grid maximum, (c) when the full-grid maximum lies in
    the pair's window both report the same line pair and shift.  Pairs the
    window hands back get exactly the full-grid result."""
    import torch

    say("case windowed-properties", prec)
    pf = shifted.pf
    K, L = pf.shape[2], pf.shape[1] // 2
    full, _ = clmatrix_records(pf, 12, 0.5, precision=prec)
    rng = np.random.default_rng(0)
    prop_np = (shifted.shifts + rng.normal(0, 0.3, size=shifted.shifts.shape)).astype(np.float32)
    prop = torch.as_tensor(prop_np).cuda()
    rec, ws = clmatrix_windowed(to_dev(pf), prop, 12, 0.5, window=1.0, gate=0.0, precision=prec)
    say(ws)
    w = {k: v.cpu().numpy() for k, v in rec.items()}
    assert ws["window_shifts"] == 5 and ws["full_shifts"] == 49
    assert ws["accepted"] + ws["below_gate"] + ws["off_grid"] == K * (K - 1) // 2
    ex = w["exhaustive"] == 1
    assert ex.sum() == ws["below_gate"] + ws["off_grid"]
    for k in ("cl1", "cl2", "j", "corr", "shift"):
        assert np.array_equal(w[k][ex], full[k][ex]), k
    P, _ = oracle.preprocess(pf)
    k1, k2 = np.triu_indices(K, 1)
    acc = np.flatnonzero(~ex)
    assert acc.size > 0
    for i in acc[:: max(1, acc.size // 60)]:
        v = oracle.entry(P, 12, 0.5, k1[i], k2[i], w["j"][i], w["cl1"][i] - 1, w["cl2"][i] - 1)
        assert abs(v - w["corr"][i]) <= 1e-9 * abs(v), (i, v, w["corr"][i])
        assert w["shift"][i] == -12 + 0.5 * w["j"][i]
    assert np.all(w["corr"][acc] <= full["corr"][acc] * (1 + 1e-9))
    a = window_offsets(prop_np, K, L, 0.5)
    inside = []
    for i in acc:
        l1, c2 = full["cl1"][i] - 1, full["cl2"][i] - 1
        off = a[k1[i], l1] - a[k2[i], c2] if c2 < L else -a[k1[i], l1] - a[k2[i], c2 - L]
        inside.append(abs((-12 + 0.5 * full["j"][i]) - off) <= 1.0 + 1e-9)
    inside = np.array(inside)
    same = (w["cl1"] == full["cl1"]) & (w["cl2"] == full["cl2"]) & (w["j"] == full["j"])
    say("settled:", acc.size, "full-grid max inside the window:", int(inside.sum()),
        "agree there:", float(same[acc][inside].mean()) if inside.any() else None,
        "agree elsewhere:", float(same[acc][~inside].mean()) if (~inside).any() else None)
    assert inside.sum() >= 10
    assert same[acc][inside].mean() >= 0.99
    assert not np.any(same[acc][~inside])
