"""Shared helpers of the GPU parity tests: compare libclm's per-pair records
with the FP64 oracle on the same inputs."""
import numpy as np


def oracle_records(oracle, pf, max_shift, step, k1=None, k2=None):
    P, n_small = oracle.preprocess(pf)
    res = oracle.pairs(P, max_shift, step, k1, k2)
    return P, res


def compare(oracle, P, ora, gpu, max_shift, step, k1, k2, near_tie_rel=1e-5):
    """Agreement report.  A disagreement is acceptable only as a near tie:
    the oracle's own value at the GPU's choice is within near_tie_rel
    (relative) of the oracle's maximum."""
    n = k1.size
    g_cl1 = np.asarray(gpu["cl1"])
    g_cl2 = np.asarray(gpu["cl2"])
    g_j = np.asarray(gpu["j"])
    g_corr = np.asarray(gpu["corr"])
    g_shift = np.asarray(gpu["shift"])
    same = (g_cl1 == ora["cl1"]) & (g_cl2 == ora["cl2"]) & (g_j == ora["j"])
    bad = []
    worst = 0.0
    for i in np.flatnonzero(~same):
        best = ora["best"][i]
        if g_cl1[i] == 0 or g_j[i] < 0:
            gap = best  # GPU found nothing above 0; oracle did
            v = 0.0
        else:
            v = oracle.entry(P, max_shift, step, k1[i], k2[i], g_j[i], g_cl1[i] - 1, g_cl2[i] - 1)
            gap = best - v
        rel = gap / max(abs(best), 1e-300)
        worst = max(worst, rel)
        if rel > near_tie_rel:
            bad.append((int(k1[i]), int(k2[i]), (int(ora["cl1"][i]), int(ora["cl2"][i]), int(ora["j"][i])),
                        (int(g_cl1[i]), int(g_cl2[i]), int(g_j[i])), float(best), float(v)))
    m = same & (ora["best"] != 0)
    corr_rel = np.abs(g_corr[m] - ora["best"][m]) / np.maximum(np.abs(ora["best"][m]), 1e-300)
    shift_ok = np.all(g_shift[same] == ora["shift"][same])
    return dict(n=n, agree=float(same.mean()) if n else 1.0, n_disagree=int((~same).sum()),
                not_near_tie=bad, worst_gap_rel=float(worst),
                corr_rel_max=float(corr_rel.max()) if corr_rel.size else 0.0,
                shift_ok=bool(shift_ok))


def assert_parity(rep, min_agree=0.999, corr_rel=1e-5):
    assert not rep["not_near_tie"], rep["not_near_tie"][:5]
    assert rep["n_disagree"] <= max(1, int((1 - min_agree) * rep["n"])), rep
    assert rep["corr_rel_max"] <= corr_rel, rep
    assert rep["shift_ok"], rep


def oracle_sample(oracle, pf, k1, k2, max_shift, step):
    """The oracle on the pairs (k1, k2) of ``pf`` (numpy array or column-major
    device tensor), reading only the images those pairs use.  Returns
    (P of those images, oracle records, k1 and k2 re-indexed into P)."""
    imgs = np.unique(np.concatenate([k1, k2]))
    if type(pf).__module__.startswith("torch"):
        import torch

        sub = pf[:, :, torch.as_tensor(imgs, device=pf.device)].cpu().numpy()
    else:
        sub = np.asarray(pf)[:, :, imgs]
    sub = sub.astype(np.complex128)
    kk1 = np.searchsorted(imgs, k1).astype(np.int32)
    kk2 = np.searchsorted(imgs, k2).astype(np.int32)
    P, ora = oracle_records(oracle, sub, max_shift, step, kk1, kk2)
    return P, ora, kk1, kk2


def host_subset(rec, slots):
    """Host copies of per-pair records (numpy arrays or device tensors) at
    the given equation slots."""
    out = {}
    for k, v in rec.items():
        a = v.cpu().numpy() if hasattr(v, "cpu") else np.asarray(v)
        out[k] = a[slots]
    return out
