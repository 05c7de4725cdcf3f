"""Learned form of the shift search: the shift-proposal network and the
windowed common-lines search built on it.

* :class:`ShiftNet` holds the network's parameters on the GPU (FP32 master
  weights, their BF16 copy, Adam moments) and drives libclm's hand-written
  HIP kernels: features, forward, gradient, Adam.
* :func:`estimate_shifts` recovers per-image 2-D shifts from the exhaustive
  search's common lines and 1-D shifts by robust least squares on the shift
  equations (the training targets; cryo_clmatrix_cpu.m:722-741 rows,
  solved as in sinograms/test_shifts.m:24-27 with iterative reweighting).
* :func:`train` fits the network; :func:`calibrate_gate` picks the window
  gate on a stack whose full-grid result is known.
* :func:`clmatrix_windowed` runs the windowed search with exhaustive
  fallback (libclm's clm_clmatrix_windowed).

The network and its kernels are described in aspire_amd/csrc/shiftnet.hip.
"""
from __future__ import annotations

import ctypes
import math
import os
import time

import numpy as np

from . import _lib
from .clmatrix import _device_view, _params

NPARAM = 52570
NR, T = 16, 360
SHAPES = [("W1", (32, 32, 5)), ("b1", (32,)), ("W2", (8, 32, 5)), ("b2", (8,)),
          ("F1", (16, 2880)), ("c1", (16,)), ("F2", (2, 16)), ("c2", (2,))]
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "shiftnet.npz")


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights (bound sqrt(6 / fan_in)), zero biases; the output
    layer is scaled by 0.1 so that initial proposals are small."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shp in SHAPES:
        if len(shp) == 1:
            parts.append(np.zeros(shp))
            continue
        fan_in = int(np.prod(shp[1:]))
        w = rng.uniform(-1, 1, size=shp) * math.sqrt(6.0 / fan_in)
        if name == "F2":
            w *= 0.1
        parts.append(w)
    theta = np.concatenate([p.ravel() for p in parts]).astype(np.float32)
    assert theta.size == NPARAM
    return theta


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class ShiftNet:
    """Shift-proposal network on one HIP device."""

    def __init__(self, params=None, device: int = 0, seed: int = 0):
        import torch

        L = _lib.load()
        if L.clm_shiftnet_nparam() != NPARAM:
            raise _lib.ClmError("libclm's shift network has a different parameter layout")
        self.device = device
        self.dev = torch.device("cuda", device)
        theta = init_params(seed) if params is None else np.asarray(params, np.float32)
        self.master = torch.as_tensor(theta.copy(), device=self.dev)
        self.w = self.master.to(torch.bfloat16)
        self.m = torch.zeros_like(self.master)
        self.v = torch.zeros_like(self.master)
        self.grad = torch.zeros_like(self.master)
        self.loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        self.steps = 0
        self.meta: dict = {}
        self.ctx = _lib.context(device)

    # --------------------------------------------------------------- kernels
    def _sync(self):
        import torch

        torch.cuda.current_stream(self.dev).synchronize()

    def features(self, pf):
        """x (K x 32 x 360, bf16) from pf (device tensor, column-major
        n_r x 360 x K as for clmatrix_records)."""
        import torch

        base, dt, n_r, n_theta, K = _device_view(pf)
        x = torch.empty((K, 2 * NR, T), dtype=torch.bfloat16, device=self.dev)
        self._sync()
        rc = _lib.load().clm_shiftnet_features(self.ctx.handle, _ptr(base), n_r, n_theta, K, dt, _ptr(x))
        self.ctx.check(rc)
        return x

    def forward(self, x, chunk: int = 8192):
        """Proposed shifts in pixels (B x 2, float32 device tensor)."""
        import torch

        x = x.contiguous()
        out = torch.empty((x.shape[0], 2), dtype=torch.float32, device=self.dev)
        self._sync()
        for b0 in range(0, x.shape[0], chunk):
            b1 = min(x.shape[0], b0 + chunk)
            rc = _lib.load().clm_shiftnet_forward(self.ctx.handle, _ptr(self.w), _ptr(x[b0:b1]), b1 - b0,
                                                  _ptr(out[b0:b1]))
            self.ctx.check(rc)
        return out

    def loss_grad(self, x, target_px):
        """(loss, gradient tensor) of one batch; the gradient is self.grad."""
        x = x.contiguous()
        t = target_px.contiguous().float()
        self._sync()
        rc = _lib.load().clm_shiftnet_grad(self.ctx.handle, _ptr(self.w), _ptr(x), _ptr(t), x.shape[0],
                                           _ptr(self.grad), _ptr(self.loss))
        self.ctx.check(rc)
        return float(self.loss.item()), self.grad

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.steps += 1
        rc = _lib.load().clm_shiftnet_adam(self.ctx.handle, _ptr(self.master), _ptr(self.w), _ptr(self.m),
                                           _ptr(self.v), _ptr(self.grad), self.steps, lr, beta1, beta2, eps)
        self.ctx.check(rc)

    def propose(self, pf):
        return self.forward(self.features(pf))

    # ---------------------------------------------------------------- files
    def params(self) -> np.ndarray:
        return self.master.detach().cpu().numpy().astype(np.float32)

    def save(self, path: str = WEIGHTS, **meta):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        m = dict(self.meta)
        m.update(meta)
        np.savez(path, params=self.params(), meta=np.array(repr(m)))

    @classmethod
    def load(cls, path: str = WEIGHTS, device: int = 0) -> "ShiftNet":
        import ast

        with np.load(path, allow_pickle=False) as f:
            net = cls(params=f["params"], device=device)
            net.meta = ast.literal_eval(str(f["meta"]))
        return net


# ------------------------------------------------------------------ targets
def equation_rows(cl1, cl2, L):
    """Coefficients (sin a, cos a, -sin b, -cos b) or the b >= pi variant of
    each pair's shift equation (cryo_clmatrix_cpu.m:722-738)."""
    a = (np.asarray(cl1, np.float64) - 1) * math.pi / L
    b = (np.asarray(cl2, np.float64) - 1) * math.pi / L
    low = b < math.pi
    bb = np.where(low, b, b - math.pi)
    s = np.where(low, 1.0, -1.0)
    return np.stack([s * np.sin(a), s * np.cos(a), -np.sin(bb), -np.cos(bb)], axis=1)


def estimate_shifts(rec, K: int, L: int, iters: int = 5, weights=None):
    """2-D shift per image (K x 2) from per-pair records by iteratively
    reweighted least squares (Cauchy weights) on the shift equations.  Each
    step solves the 2K x 2K normal equations by eigendecomposition and keeps
    the minimum-norm solution: eigenvalues below 1e-9 of the largest (the
    3-dim null space of projected 3-D translations) are dropped."""
    import scipy.sparse as sp

    cl1 = np.asarray(rec["cl1"])
    cl2 = np.asarray(rec["cl2"])
    rhs = np.asarray(rec["shift"], np.float64)
    k1, k2 = np.triu_indices(K, 1)
    ok = cl1 > 0
    k1, k2, cl1, cl2, rhs = k1[ok], k2[ok], cl1[ok], cl2[ok], rhs[ok]
    C = equation_rows(cl1, cl2, L)
    n = k1.size
    rows = np.repeat(np.arange(n), 4)
    cols = np.stack([2 * k1, 2 * k1 + 1, 2 * k2, 2 * k2 + 1], axis=1).ravel()
    A = sp.csr_matrix((C.ravel(), (rows, cols)), shape=(n, 2 * K))
    w = np.ones(n) if weights is None else np.asarray(weights, np.float64)[ok]
    x = np.zeros(2 * K)
    for _ in range(iters):
        Aw = sp.diags(w) @ A
        N = (A.T @ Aw).toarray()
        g = A.T @ (w * rhs)
        ev, V = np.linalg.eigh(N)
        keep = ev > 1e-9 * ev[-1]
        x = V[:, keep] @ ((V[:, keep].T @ g) / ev[keep])
        r = A @ x - rhs
        scale = 1.4826 * np.median(np.abs(r)) + 1e-6
        w = 1.0 / (1.0 + (r / (2.385 * scale)) ** 2)
    return x.reshape(K, 2)


# ----------------------------------------------------------------- training
def make_training_set(n_stacks=20, K=1000, seed0=5001, n=129, n_r=65, input_snrs=(1 / 16, 1 / 4, 1.0),
                      max_shift_2d=12.0, max_shift=12.0, shift_step=0.5, log=print):
    """Training pairs (features, shift targets) from n_stacks seeded stacks.

    For every seed the exhaustive search (BF16 engine, full 49-shift grid)
    runs on the noise-free stack and robust least squares on its shift
    equations recovers each image's 2-D shift: those are the targets.  The
    inputs are the same stack (same rotations, phantom and shifts) with white
    noise at each SNR of ``input_snrs``.  Phantoms are recentred, so each
    image's centroid is its shift.  Returns (x, targets, truth, per-seed LS
    rms error in px)."""
    import torch

    from . import sim
    from .clmatrix import clmatrix_records

    net = ShiftNet()
    xs, ys, truth, errs = [], [], [], []
    for s in range(n_stacks):
        t0 = time.time()
        seed = seed0 + s
        clean = sim.make_stack(K=K, n=n, n_r=n_r, n_theta=T, snr=None, max_shift_2d=max_shift_2d,
                               shift_kind="uniform", seed=seed, device="cuda", recenter=True)
        rec, _ = clmatrix_records(clean.pf, max_shift, shift_step, precision="bf16")
        rec = {k: v.cpu().numpy() for k, v in rec.items()}
        est = estimate_shifts(rec, K, T // 2)
        target = torch.as_tensor(est, dtype=torch.float32, device=net.dev)
        del clean
        for snr in input_snrs:
            st = sim.make_stack(K=K, n=n, n_r=n_r, n_theta=T, snr=snr, max_shift_2d=max_shift_2d,
                                shift_kind="uniform", seed=seed, device="cuda", recenter=True)
            xs.append(net.features(st.pf))
            ys.append(target)
            truth.append(st.shifts)
            del st
        err = float(np.sqrt(np.mean(np.sum((est - truth[-1]) ** 2, axis=1))))
        errs.append(err)
        log(f"seed {seed}: K={K}, least-squares shift rms error {err:.3f} px ({time.time() - t0:.1f}s)")
    return torch.cat(xs), torch.cat(ys), np.concatenate(truth), errs


def train(net: ShiftNet, x, y, epochs=40, batch=256, lr=1e-3, val_frac=0.1, seed=0, log=print):
    """Adam on the BF16 network with libclm's kernels.  Returns the
    per-epoch (train loss, validation rms error in px)."""
    import torch

    g = torch.Generator(device="cpu").manual_seed(seed)
    N = x.shape[0]
    perm = torch.randperm(N, generator=g).to(net.dev)
    nv = int(N * val_frac)
    vi, ti = perm[:nv], perm[nv:]
    hist = []
    for ep in range(epochs):
        order = ti[torch.randperm(ti.numel(), generator=g).to(net.dev)]
        tot, nb = 0.0, 0
        lr_ep = lr * (0.5 * (1 + math.cos(math.pi * ep / epochs)))
        for b0 in range(0, order.numel() - batch + 1, batch):
            idx = order[b0:b0 + batch]
            loss, _ = net.loss_grad(x.index_select(0, idx), y.index_select(0, idx))
            net.adam(lr=lr_ep)
            tot += loss
            nb += 1
        pv = net.forward(x.index_select(0, vi)) if nv else None
        verr = float(torch.sqrt(((pv - y.index_select(0, vi)) ** 2).sum(1).mean())) if nv else float("nan")
        hist.append((tot / max(nb, 1), verr))
        log(f"epoch {ep}: loss {tot / max(nb, 1):.5f}, validation rms {verr:.3f} px")
    return hist


def clmatrix_windowed(pf, prop, max_shift=12.0, shift_step=0.5, window=1.0, gate=0.0, precision="fp16",
                      device=0, variant=0, flags=0):
    """Windowed search with exhaustive fallback (libclm's
    clm_clmatrix_windowed).  ``prop``: K x 2 float32 device tensor of
    proposed shifts in pixels.  Returns (records, window stats); records
    hold ``exhaustive`` (uint8, 1 = the pair was searched over the full
    grid)."""
    import torch

    base, dt, n_r, n_theta, K = _device_view(pf)
    P = max(K * (K - 1) // 2, 0)
    dev = base.device
    out = dict(cl1=torch.zeros(P, dtype=torch.int32, device=dev),
               cl2=torch.zeros(P, dtype=torch.int32, device=dev),
               j=torch.zeros(P, dtype=torch.int32, device=dev),
               corr=torch.zeros(P, dtype=torch.float64, device=dev),
               shift=torch.zeros(P, dtype=torch.float64, device=dev),
               exhaustive=torch.zeros(P, dtype=torch.uint8, device=dev))
    prop = prop.contiguous().float()
    if prop.shape != (K, 2):
        raise ValueError("proposals must be K x 2")
    ctx = _lib.context(device)
    p = _params(n_r, n_theta, K, dt, max_shift, shift_step, precision, 0, 1e-3, variant, flags)
    ws = _lib.ClmWindowStats()
    torch.cuda.current_stream(dev).synchronize()
    rc = _lib.load().clm_clmatrix_windowed(ctx.handle, ctypes.byref(p), _ptr(base), _ptr(prop), float(window),
                                           float(gate), _ptr(out["cl1"]), _ptr(out["cl2"]), _ptr(out["j"]),
                                           _ptr(out["corr"]), _ptr(out["shift"]), _ptr(out["exhaustive"]),
                                           ctypes.byref(ws))
    ctx.check(rc)
    return out, ws.as_dict()


def calibrate_gate(pf, prop, full_rec, max_shift=12.0, shift_step=0.5, window=1.0, target=0.999,
                   precision="fp16", rots=None, max_detect_loss=0.005):
    """The window gate from a stack whose full-grid result ``full_rec``
    (host arrays) is known.  Returns (gate, table of candidates).

    With the true rotations ``rots`` (a simulated stack): the smallest gate
    such that the combined result (window where its peak is at or above the
    gate, full grid elsewhere) finds at least (1 - max_detect_loss) as many
    true common lines (within 5 degrees, comparecl's test) as the full grid
    alone.  Without them: the smallest gate such that at least ``target`` of
    the pairs the window settles agree with the full grid.  Table rows:
    (gate, fraction of on-grid pairs settled, agreement with the full grid
    among them, detection of the combined result / of the full grid)."""
    from . import sim

    rec, _ = clmatrix_windowed(pf, prop, max_shift, shift_step, window, gate=-1e30, precision=precision)
    w = {k: v.cpu().numpy() for k, v in rec.items()}
    same = (w["cl1"] == full_rec["cl1"]) & (w["cl2"] == full_rec["cl2"]) & (w["j"] == full_rec["j"])
    # with no gate the window settles every pair whose peak lands on the
    # grid; at gate g it settles those of them with window corr >= g, and
    # the rest are searched exhaustively (and then agree by construction)
    keep = w["exhaustive"] == 0
    corr = w["corr"][keep]
    order = np.argsort(-corr)
    n_acc = np.arange(1, corr.size + 1)
    cum_bad = np.cumsum(~same[keep][order])
    if rots is not None:
        K = rots.shape[2]
        ref = sim.clmatrix_cheat(rots, T)
        k1, k2 = np.triu_indices(K, 1)
        t1, t2 = ref[k1, k2], ref[k2, k1]
        ok_full = sim.matched_lines(full_rec["cl1"], full_rec["cl2"], t1, t2, T, 5)
        ok_win = sim.matched_lines(w["cl1"], w["cl2"], t1, t2, T, 5)
        base = float(ok_full.sum())
        gain = np.cumsum(ok_win[keep][order].astype(np.int64) - ok_full[keep][order].astype(np.int64))
        detect = (base + gain) / max(base, 1.0)
        good = detect >= 1.0 - max_detect_loss
    else:
        detect = None
        good = cum_bad <= (1 - target) * n_acc
    table = []
    for qn in (1.0, 0.99, 0.95, 0.9, 0.8, 0.7, 0.6, 0.5, 0.4, 0.3, 0.2, 0.1):
        i = int(qn * (corr.size - 1))
        table.append((float(corr[order[i]]), float(n_acc[i] / corr.size), float(1 - cum_bad[i] / n_acc[i]),
                      float(detect[i]) if detect is not None else None))
    idx = np.flatnonzero(good)
    gate = float(corr[order[idx[-1]]]) if idx.size else float(corr.max())
    return gate, table


# --------------------------------------------------------------- evaluation
def pair_shift_error(prop, shifts, rots, L=T // 2, n_pairs=20000, seed=0):
    """RMS error (px) of the pair 1-D shifts that proposals predict at the
    true common lines, against the true shifts' values there.  Both use the
    shift-equation coefficients (cryo_clmatrix_cpu.m:722-738), so a common
    3-D translation of the proposals (their null space) does not count."""
    from . import sim

    prop = np.asarray(prop, np.float64)
    K = prop.shape[0]
    rng = np.random.default_rng(seed)
    m = min(n_pairs, K * (K - 1) // 2)
    k1 = rng.integers(0, K, size=3 * m)
    k2 = rng.integers(0, K, size=3 * m)
    ok = k1 < k2
    k1, k2 = k1[ok][:m], k2[ok][:m]
    err, base = [], []
    for a, b in zip(k1, k2):
        l1, l2 = sim.commonline_R(rots[:, :, a], rots[:, :, b], 2 * L)
        c1, c2 = l1 + 1, l2 + 1
        if c1 > L:  # the reference reports l1 in [1, L]: flip both lines
            c1, c2 = c1 - L, (c2 + L - 1) % (2 * L) + 1
        st = sim.true_shift_1d(c1, c2, L, shifts[a], shifts[b])
        sp = sim.true_shift_1d(c1, c2, L, prop[a], prop[b])
        err.append(sp - st)
        base.append(st)
    err, base = np.asarray(err), np.asarray(base)
    return float(np.sqrt(np.mean(err ** 2))), float(np.sqrt(np.mean(base ** 2)))


def evaluate(net: "ShiftNet", K=5000, seed=1005, snr=1 / 16, n=129, n_r=65, max_shift=12.0, shift_step=0.5,
             window=1.0, gate=None, log=print, precision="fp16"):
    """The learned form on one stack: full-grid search against the windowed
    search with the network's proposals and with zero proposals (the
    method without the network): time, agreement with the full grid,
    detection against the true common lines, pair-shift prediction error."""
    import torch

    from . import sim
    from .clmatrix import clmatrix_records

    st = sim.make_stack(K=K, n=n, n_r=n_r, n_theta=T, snr=snr, max_shift_2d=max_shift, shift_kind="uniform",
                        seed=seed, device="cuda")
    gate = net_gate(net, snr) if gate is None else float(gate)
    P = K * (K - 1) // 2
    torch.cuda.synchronize()
    # warm both paths once (kernel code, buffers), then time
    clmatrix_records(st.pf, max_shift, shift_step, precision=precision)
    torch.cuda.synchronize()
    t = time.perf_counter()
    full, fst = clmatrix_records(st.pf, max_shift, shift_step, precision=precision)
    torch.cuda.synchronize()
    t_full = time.perf_counter() - t
    full = {k: v.cpu().numpy() for k, v in full.items()}
    prop = net.propose(st.pf)
    clmatrix_windowed(st.pf, prop, max_shift, shift_step, window, gate, precision=precision)
    torch.cuda.synchronize()
    t = time.perf_counter()
    prop = net.propose(st.pf)
    torch.cuda.synchronize()
    t_net = time.perf_counter() - t
    t = time.perf_counter()
    rec, ws = clmatrix_windowed(st.pf, prop, max_shift, shift_step, window, gate, precision=precision)
    torch.cuda.synchronize()
    t_win = time.perf_counter() - t
    rec = {k: v.cpu().numpy() for k, v in rec.items()}
    zero = torch.zeros_like(prop)
    rz, wz = clmatrix_windowed(st.pf, zero, max_shift, shift_step, window, gate, precision=precision)
    rz = {k: v.cpu().numpy() for k, v in rz.items()}

    def agree(r):
        return float(np.mean((r["cl1"] == full["cl1"]) & (r["cl2"] == full["cl2"]) & (r["j"] == full["j"])))

    nd = min(K, 1000)  # detection on the first images (all their pairs)
    ref = sim.clmatrix_cheat(st.rots[:, :, :nd], 2 * (T // 2))

    def detect(r):
        cl = np.zeros((nd, nd))
        k1, k2 = np.triu_indices(K, 1)
        m = k2 < nd
        cl[k1[m], k2[m]] = r["cl1"][m]
        cl[k2[m], k1[m]] = r["cl2"][m]
        return float(sim.comparecl(cl, ref, 2 * (T // 2), 5))

    pe_net, pe_zero = pair_shift_error(prop.cpu().numpy(), st.shifts, st.rots)
    out = dict(K=K, pairs=P, snr=snr, seed=seed, full_shifts=ws["full_shifts"], window_shifts=ws["window_shifts"],
               gate=gate, full_seconds=t_full, full_pairs_per_s=P / t_full,
               net_seconds=t_net, windowed_seconds=t_win, learned_pairs_per_s=P / (t_net + t_win),
               speedup=t_full / (t_net + t_win), precision=precision,
               settled=ws["accepted"] / P, below_gate=ws["below_gate"] / P, off_grid=ws["off_grid"] / P,
               agree_full=agree(rec), agree_full_zero_proposals=agree(rz),
               settled_zero_proposals=wz["accepted"] / P,
               detect_5deg_full=detect(full), detect_5deg_learned=detect(rec), detect_5deg_zero=detect(rz),
               pair_shift_rms_px_net=pe_net, pair_shift_rms_px_zero=pe_zero, window_stats=ws)
    log(out)
    return out


# ---------------------------------------------------------------------- CLI
def _cmd_train(args):
    import json

    t0 = time.time()
    x, y, truth, errs = make_training_set(n_stacks=args.stacks, K=args.K, seed0=args.seed0)
    t_data = time.time() - t0
    net = ShiftNet(seed=args.init_seed)
    hist = train(net, x, y, epochs=args.epochs, batch=args.batch, lr=args.lr)
    t_train = time.time() - t0 - t_data
    info = dict(stacks=args.stacks, K=args.K, seed0=args.seed0, input_snrs=[1 / 16, 1 / 4, 1.0],
                epochs=args.epochs, batch=args.batch, lr=args.lr, samples=int(x.shape[0]),
                ls_target_rms_px=float(np.mean(errs)), final_loss=hist[-1][0], val_rms_px=hist[-1][1],
                seconds_data=t_data, seconds_train=t_train, steps=net.steps)
    net.save(args.out, training=info, window=1.0)
    print(json.dumps(info))


def snr_key(snr: float) -> str:
    return "%g" % snr


def net_gate(net: "ShiftNet", snr: float) -> float:
    """The window gate stored with the weights for data of this SNR (the
    calibration of that SNR, else the default ``gate``, else 0).  Window
    peaks scale with the SNR, so a gate holds for the SNR it was calibrated
    on."""
    gates = net.meta.get("gates", {})
    if snr_key(snr) in gates:
        return float(gates[snr_key(snr)])
    return float(net.meta.get("gate", 0.0))


def _cmd_calibrate(net: "ShiftNet", args) -> dict:
    """The gate: on a held-out stack (seed outside the training seeds, images
    of the learned workload's model: 129 px, shifts U[-12,12], 49-shift full
    grid, SNR 1 by default), the smallest window peak at which the combined
    search finds at least (1 - max_detect_loss) of the true common lines the
    full grid finds (:func:`calibrate_gate` with the true rotations).
    Stored in the weights' meta under ``gates`` for that SNR; the SNR-1
    calibration is also the default ``gate``."""
    from . import sim
    from .clmatrix import clmatrix_records

    st = sim.make_stack(K=args.K, n=129, n_r=65, n_theta=T, snr=args.snr, max_shift_2d=12.0, shift_kind="uniform",
                        seed=args.seed, device="cuda")
    full, _ = clmatrix_records(st.pf, 12.0, 0.5, precision="fp16")
    full = {k: v.cpu().numpy() for k, v in full.items()}
    prop = net.propose(st.pf)
    gate, table = calibrate_gate(st.pf, prop, full, 12.0, 0.5, 1.0, target=args.target,
                                 rots=None if args.no_truth else st.rots, max_detect_loss=args.max_detect_loss)
    _, ws = clmatrix_windowed(st.pf, prop, 12.0, 0.5, 1.0, gate)
    info = dict(gate=gate, K=args.K, snr=args.snr, seed=args.seed,
                criterion="agreement >= %g" % args.target if args.no_truth else
                "detection loss <= %g" % args.max_detect_loss, table=table, window_stats=ws)
    net.meta.setdefault("gates", {})[snr_key(args.snr)] = float(gate)
    cal = net.meta.get("gate_calibration", {})
    if "gate" in cal:  # one calibration stored flat: keep it under its SNR
        cal = {snr_key(cal.get("snr", 1.0)): cal}
    cal[snr_key(args.snr)] = {k: v for k, v in info.items() if k not in ("table", "window_stats")}
    net.meta["gate_calibration"] = cal
    if args.snr == 1.0:
        net.meta["gate"] = float(gate)
    return info


def main(argv=None):
    import argparse

    ap = argparse.ArgumentParser(prog="python -m aspire_amd.shiftnet",
                                 description="Train the shift-proposal network with libclm's kernels.")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--stacks", type=int, default=20)
    t.add_argument("--K", type=int, default=1000)
    t.add_argument("--seed0", type=int, default=5001)
    t.add_argument("--init-seed", type=int, default=0)
    t.add_argument("--epochs", type=int, default=30)
    t.add_argument("--batch", type=int, default=256)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--out", default=WEIGHTS)
    c = sub.add_parser("calibrate", help="pick the window gate on a held-out stack and store it in the weights")
    c.add_argument("--weights", default=WEIGHTS)
    c.add_argument("--K", type=int, default=1000)
    c.add_argument("--snr", type=float, default=1.0)
    c.add_argument("--seed", type=int, default=5101)
    c.add_argument("--target", type=float, default=0.99, help="agreement target (without --no-truth: unused)")
    c.add_argument("--max-detect-loss", type=float, default=0.005)
    c.add_argument("--no-truth", action="store_true", help="calibrate on agreement with the full grid instead")
    e = sub.add_parser("evaluate")
    e.add_argument("--weights", default=WEIGHTS)
    e.add_argument("--K", type=int, default=5000)
    e.add_argument("--snr", type=float, default=1 / 16)
    e.add_argument("--seed", type=int, default=1005)
    e.add_argument("--gate", type=float, default=None)
    args = ap.parse_args(argv)
    if args.cmd == "train":
        _cmd_train(args)
    elif args.cmd == "calibrate":
        import json

        net = ShiftNet.load(args.weights)
        print(json.dumps(_cmd_calibrate(net, args)))
        net.save(args.weights)
    elif args.cmd == "evaluate":
        import json

        net = ShiftNet.load(args.weights)
        print(json.dumps(evaluate(net, K=args.K, snr=args.snr, seed=args.seed, gate=args.gate, log=lambda *a: None)))


if __name__ == "__main__":
    main()
