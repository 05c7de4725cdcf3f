"""CPU oracle of the shift network (test infrastructure only: only tests/,
the smoke check and the bench's CPU baseline use it; the product path never
calls it).

Float64 numpy restatement of the network that libclm's HIP kernels
implement (aspire_amd/csrc/shiftnet.hip), forward and backward:

    h1 = relu(conv5(x;  32 -> 32) + b1)     circular, kernel 5, centred
    h2 = relu(conv5(h1; 32 -> 8)  + b2)
    z  = relu(F1 . flatten(h2) + c1)        2880 -> 16
    y  = F2 . z + c2                        16 -> 2
    loss = 1/(2B) sum_b sum_o (y_bo - t_bo / 8)^2   (t in pixels)

With ``quant=True`` the values the GPU stores in BF16 are rounded the same
way (weights, inputs, h1 and h2 after the ReLU, round to nearest even), so
the GPU's FP32 sums can be compared with this FP64 reference to FP32-level
tolerances.  Features follow the same definition as the kernel: channel
2r + c of image k at ray t is the real (c=0) or imaginary part of
pf(r, t, k) divided by the root mean square of the image's 16 x 360 values.

The oracle itself is pinned in tests/test_shiftnet_oracle.py against torch
autograd (float64, circular conv1d) and central finite differences.
"""
from __future__ import annotations

import numpy as np

NR, T = 16, 360
CIN, C1, C2, KW, H, NO = 2 * NR, 32, 8, 5, 16, 2
HALF = KW // 2
FLAT = C2 * T
SCALE = 8.0

SHAPES = [("W1", (C1, CIN, KW)), ("b1", (C1,)), ("W2", (C2, C1, KW)), ("b2", (C2,)),
          ("F1", (H, FLAT)), ("c1", (H,)), ("F2", (NO, H)), ("c2", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(theta: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(shp)
        o += n
    return out


def pack(parts: dict) -> np.ndarray:
    return np.concatenate([np.asarray(parts[name], np.float64).ravel() for name, _ in SHAPES])


# ------------------------------------------------------------------ bf16
def bf16_bits(a) -> np.ndarray:
    """Round to BF16 (nearest even, via float32) and return the 16-bit codes."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return r.astype(np.uint16)


def bf16_value(bits) -> np.ndarray:
    b = np.asarray(bits, dtype=np.uint16).astype(np.uint32) << 16
    return b.view(np.float32).astype(np.float64)


def q(a, quant=True):
    return bf16_value(bf16_bits(a)) if quant else np.asarray(a, np.float64)


# ------------------------------------------------------------------ layers
def conv(x, W, b):
    """out[n, co, t] = b[co] + sum_ci sum_j W[co, ci, j] x[n, ci, (t + j - HALF) mod T]."""
    out = np.broadcast_to(b[None, :, None], (x.shape[0], W.shape[0], x.shape[2])).copy()
    for j in range(W.shape[2]):
        xs = np.roll(x, -(j - HALF), axis=2)  # xs[..., t] = x[..., t + j - HALF]
        out += np.einsum("oc,nct->not", W[:, :, j], xs, optimize=True)
    return out


def conv_grad_input(dout, W):
    """d in[n, ci, u] = sum_co sum_j W[co, ci, j] dout[n, co, (u - j + HALF) mod T]."""
    din = np.zeros((dout.shape[0], W.shape[1], dout.shape[2]))
    for j in range(W.shape[2]):
        ds = np.roll(dout, j - HALF, axis=2)  # ds[..., u] = dout[..., u - j + HALF]
        din += np.einsum("oc,not->nct", W[:, :, j], ds, optimize=True)
    return din


def conv_grad_weight(dout, x):
    """dW[co, ci, j] = sum_n sum_t dout[n, co, t] x[n, ci, (t + j - HALF) mod T]."""
    dW = np.zeros((dout.shape[1], x.shape[1], KW))
    for j in range(KW):
        xs = np.roll(x, -(j - HALF), axis=2)
        dW[:, :, j] = np.einsum("not,nct->oc", dout, xs, optimize=True)
    return dW


# ------------------------------------------------------------------ network
def features(pf: np.ndarray, quant=True) -> np.ndarray:
    """x (K x 32 x 360) from pf (n_r x 360 x K complex)."""
    z = np.asarray(pf)[:NR, :, :]                    # r, t, k
    ms = np.mean(np.abs(z.astype(np.complex128)) ** 2, axis=(0, 1))
    sc = 1.0 / np.sqrt(ms + 1e-30)
    x = np.empty((z.shape[2], CIN, T))
    zz = np.transpose(z, (2, 0, 1)) * sc[:, None, None]  # k, r, t
    x[:, 0::2, :] = zz.real
    x[:, 1::2, :] = zz.imag
    return q(x, quant)


def forward(theta, x, quant=True):
    """Activations of a batch x (B x 32 x 360)."""
    p = unpack(q(theta, quant))
    x = q(x, quant)
    a1 = conv(x, p["W1"], p["b1"])
    h1 = q(np.maximum(a1, 0.0), quant)
    a2 = conv(h1, p["W2"], p["b2"])
    h2 = q(np.maximum(a2, 0.0), quant)
    flat = h2.reshape(h2.shape[0], FLAT)
    z = np.maximum(flat @ p["F1"].T + p["c1"], 0.0)
    y = z @ p["F2"].T + p["c2"]
    return dict(p=p, x=x, h1=h1, h2=h2, z=z, y=y)


def predict(theta, x, quant=True):
    """Proposed shifts in pixels (B x 2)."""
    return SCALE * forward(theta, x, quant)["y"]


def loss_grad(theta, x, target_px, quant=True):
    """(loss, gradient over the packed parameters)."""
    a = forward(theta, x, quant)
    p, B = a["p"], x.shape[0]
    d = a["y"] - np.asarray(target_px, np.float64) / SCALE
    loss = 0.5 * float(np.sum(d * d)) / B
    dy = d / B
    g = {}
    g["F2"] = dy.T @ a["z"]
    g["c2"] = dy.sum(axis=0)
    dz = (dy @ p["F2"]) * (a["z"] > 0)
    flat = a["h2"].reshape(B, FLAT)
    g["F1"] = dz.T @ flat
    g["c1"] = dz.sum(axis=0)
    dz2 = ((dz @ p["F1"]) * (flat > 0)).reshape(B, C2, T)
    g["W2"] = conv_grad_weight(dz2, a["h1"])
    g["b2"] = dz2.sum(axis=(0, 2))
    dz1 = conv_grad_input(dz2, p["W2"]) * (a["h1"] > 0)
    g["W1"] = conv_grad_weight(dz1, a["x"])
    g["b1"] = dz1.sum(axis=(0, 2))
    return loss, pack(g)
