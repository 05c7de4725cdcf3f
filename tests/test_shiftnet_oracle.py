"""The shift network's CPU oracle (oracle/shiftnet_oracle.py) against two
independent references: torch autograd in float64 (circular conv1d, the same
architecture written with torch's own layers) and central finite
differences.  A dropped term, a wrong roll direction or a transposed weight
in the oracle's backward fails one of them."""
import numpy as np
import pytest

from oracle import shiftnet_oracle as so

torch = pytest.importorskip("torch")


def torch_loss(theta, x, t):
    import torch.nn.functional as F

    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    parts, o = {}, 0
    for name, shp in so.SHAPES:
        n = int(np.prod(shp))
        parts[name] = th[o:o + n].reshape(shp)
        o += n
    xt = torch.tensor(x, dtype=torch.float64)
    h1 = torch.relu(F.conv1d(F.pad(xt, (so.HALF, so.HALF), mode="circular"), parts["W1"], parts["b1"]))
    h2 = torch.relu(F.conv1d(F.pad(h1, (so.HALF, so.HALF), mode="circular"), parts["W2"], parts["b2"]))
    z = torch.relu(h2.reshape(x.shape[0], -1) @ parts["F1"].T + parts["c1"])
    y = z @ parts["F2"].T + parts["c2"]
    d = y - torch.tensor(t, dtype=torch.float64) / so.SCALE
    loss = 0.5 * (d * d).sum() / x.shape[0]
    loss.backward()
    return float(loss), th.grad.numpy(), (so.SCALE * y).detach().numpy()


def random_case(seed, B=3):
    rng = np.random.default_rng(seed)
    theta = rng.standard_normal(so.NPARAM) * 0.05
    x = rng.standard_normal((B, so.CIN, so.T))
    t = rng.uniform(-12, 12, size=(B, so.NO))
    return theta, x, t


def test_parameter_count():
    assert so.NPARAM == 52570


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_matches_torch_autograd(seed):
    theta, x, t = random_case(seed)
    loss, g = so.loss_grad(theta, x, t, quant=False)
    tl, tg, ty = torch_loss(theta, x, t)
    assert abs(loss - tl) <= 1e-12 * max(1.0, abs(tl))
    np.testing.assert_allclose(so.predict(theta, x, quant=False), ty, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(g, tg, rtol=1e-9, atol=1e-12)
    # every parameter block is exercised (no block with an all-zero gradient)
    off = 0
    for name, shp in so.SHAPES:
        n = int(np.prod(shp))
        assert np.any(g[off:off + n] != 0), name
        off += n


def test_oracle_finite_differences():
    theta, x, t = random_case(5, B=2)
    _, g = so.loss_grad(theta, x, t, quant=False)
    rng = np.random.default_rng(9)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in so.SHAPES])
    idx = [int(rng.integers(offs[i], offs[i + 1])) for i in range(len(so.SHAPES))]
    h = 1e-6
    for i in idx:
        e = np.zeros_like(theta)
        e[i] = h
        lp, _ = so.loss_grad(theta + e, x, t, quant=False)
        lm, _ = so.loss_grad(theta - e, x, t, quant=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_bf16_rounding_known_values():
    # exact bf16 values map to themselves; ties round to even
    v = np.array([1.0, -2.5, 0.0, 3.140625, 65280.0], np.float32)
    assert np.array_equal(so.bf16_value(so.bf16_bits(v)), v.astype(np.float64))
    one_plus_half_ulp = np.float32(1.0) + np.float32(2.0 ** -8)      # halfway 1 -> 1 + 2^-7
    assert so.bf16_value(so.bf16_bits([one_plus_half_ulp]))[0] == 1.0
    three_half = np.float32(1.0 + 2.0 ** -7 + 2.0 ** -8)              # halfway, odd -> round up
    assert so.bf16_value(so.bf16_bits([three_half]))[0] == 1.0 + 2.0 ** -6


def test_features_closed_form():
    rng = np.random.default_rng(3)
    pf = rng.standard_normal((20, 360, 2)) + 1j * rng.standard_normal((20, 360, 2))
    x = so.features(pf, quant=False)
    assert x.shape == (2, 32, 360)
    np.testing.assert_allclose(np.mean(x ** 2, axis=(1, 2)) * 2, 1.0, rtol=1e-12)  # unit RMS per complex value
    sc = 1 / np.sqrt(np.mean(np.abs(pf[:16, :, 1]) ** 2))
    assert abs(x[1, 7, 100] - pf[3, 100, 1].imag * sc) < 1e-14
