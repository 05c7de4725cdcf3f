"""Benchmark of the common-lines engine on its headline workload.

Workload (one "step" = one cryo_clmatrix call over the whole stack):
K=5000 projections of 129x129 from a seeded Gaussian-blob phantom, 2-D shifts
uniform in [-5,5] px, SNR=1/16, polar Fourier transform n_r=65, n_theta=360,
1-D shift search max_shift=10, shift_step=0.5 (41 shifts): 12,497,500 pairs.
Inputs are resident in HBM when the timed region starts; outputs (per-pair
records) stay in HBM.

Reports BF16-MFMA + FP32/FP64 re-score pairs/s (value), the plain FP32 GPU
path's pairs/s on the same K=5000 stack, clstack agreement with the FP64
oracle on a seeded sample of 20,000 pairs (whose oracle time is also the CPU
baseline), the roofline position of the pair kernel, and the drop-in's
end-to-end time from host buffers to the five cryo_clmatrix outputs.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pairs/s K=5000 n_r=65 41 shifts: BF16+rescore vs FP32 GPU; clstack match %"
BF16_PEAK = 2.5e15  # dense MFMA FLOP/s, MI355X


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--K", type=int, default=5000)
    ap.add_argument("--precision", default="bf16")
    ap.add_argument("--fp32-K", type=int, default=5000)
    ap.add_argument("--sample-pairs", type=int, default=20000)
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the drop-in line (host pf -> the five cryo_clmatrix outputs)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the learned-form line (windowed search with the shift network)")
    ap.add_argument("--variant", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (heuristic variant)")
    ap.add_argument("--tune-pairs", type=int, default=200000)
    ap.add_argument("--flags", type=int, default=0, help="clm_params.reserved bits (1: no lagged schedule)")
    ap.add_argument("--mode", default="bench", choices=["bench", "pairs", "copy", "grouped", "configs", "pft", "shifts", "vote",
                                                        "sync3n"],
                    help="pairs: one BF16 call on the workload and nothing else (for counter "
                         "runs); copy: a plain 1 GiB device copy (HBM byte-counter calibration); "
                         "grouped: the 64-stack uneven batch in one graph-captured grouped job; "
                         "configs: the K=100 and K=1000 workloads in every precision against the oracle; "
                         "pft: the polar Fourier transform of the K=5000 workload's images; "
                         "shifts: the LSQR solve of the K=5000 workload's shift equations; "
                         "vote: the voting synchronization matrix of the K=5000 workload's common lines; "
                         "sync3n: the sync3n triangle kernels on N=K relative rotations")
    ap.add_argument("--stacks", type=int, default=64, help="grouped mode: number of stacks")
    ap.add_argument("--grouped-sample", type=int, default=40, help="grouped mode: oracle pairs per stack")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU runs are not part of this build",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    if args.mode == "grouped":
        return grouped_bench(args)
    if args.mode == "configs":
        return configs_bench(args)
    if args.mode == "pft":
        return pft_bench(args)
    if args.mode == "shifts":
        return shifts_bench(args)
    if args.mode == "vote":
        return vote_bench(args)
    if args.mode == "sync3n":
        return sync3n_bench(args)

    if args.mode == "copy":
        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"mode": "copy", "bytes_read": src.numel() * 4, "bytes_written": dst.numel() * 4}))
        return

    from aspire_amd import sim
    from aspire_amd.clmatrix import autotune, clmatrix_records

    n, n_r, T, ms, step, snr = 129, 65, 360, 10.0, 0.5, 1 / 16
    K = args.K
    t0 = time.time()
    st = sim.make_stack(K=K, n=n, n_r=n_r, n_theta=T, snr=snr, max_shift_2d=5.0,
                        shift_kind="uniform", seed=1003, device="cuda")
    pf = st.pf
    torch.cuda.synchronize()
    log(f"data: K={K} generated in {time.time() - t0:.1f}s")
    S = int(np.ceil(2 * ms / step + 1))
    L, R = T // 2, n_r - 1
    P = K * (K - 1) // 2
    W_pair = 8.0 * S * L * L * R  # algorithmic flops per pair

    def run(prec, x, variant=0):
        return clmatrix_records(x, ms, step, precision=prec, variant=variant, flags=args.flags)

    tune = None
    if not args.no_tune and args.variant == 0:
        t = time.perf_counter()
        rows, best = autotune(pf, ms, step, args.precision, sample_pairs=args.tune_pairs, flags=args.flags)
        tune = {"sample_pairs": args.tune_pairs, "seconds": time.perf_counter() - t, "chosen": best,
                "variants": rows}
        log(f"autotune ({tune['seconds']:.1f}s): {best}")

    if args.mode == "pairs":
        rec, stats = run(args.precision, pf, args.variant)
        torch.cuda.synchronize()
        print(json.dumps({"mode": "pairs", "K": K, "stats": stats}))
        return

    for i in range(args.warmup):
        rec, stats = run(args.precision, pf, args.variant)
        log(f"warmup {i}: {stats}")
    torch.cuda.synchronize()
    times, kms, all_stats = [], [], []
    for i in range(args.steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        rec, stats = run(args.precision, pf, args.variant)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        kms.append(stats["ms_pairs"])
        all_stats.append(stats)
        log(f"step {i}: {times[-1]:.3f}s {stats}")
    sec = float(np.mean(times))
    value = P / sec
    k_ms = float(np.mean(kms))
    achieved = W_pair * P / (k_ms * 1e-3)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pair_kernel_k5000_traffic.json")
    if os.path.exists(tpath) and K == 5000:
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    # FP16 + re-score (the drop-in's default precision) on the same stack:
    # the same MFMA rate, an 8x tighter rounding bound
    fp16 = None
    rec16 = None
    if args.precision != "fp16":
        run("fp16", pf)
        torch.cuda.synchronize()
        t = time.perf_counter()
        rec16, st16 = run("fp16", pf)
        torch.cuda.synchronize()
        sec16 = time.perf_counter() - t
        fp16 = {"value": P / sec16, "unit": "pairs/s", "seconds": sec16, "ms_pairs": st16["ms_pairs"],
                "ms_rescore": st16["ms_rescore"], "variant": st16["variant"], "cells_rescored": st16["cells_rescored"],
                "uncertain_pairs": st16["uncertain_pairs"]}
        log(f"fp16: {sec16:.3f}s {st16}")

    # plain FP32 GPU path on the same stack (the first fp32-K images; all by
    # default).  The autotuner's timing launches warm its kernels; the one
    # timed call includes its own buffer growth, a few ms against ~50 s.
    Kf = min(args.fp32_K, K)
    pf_f = pf[:, :, :Kf]
    Pf = Kf * (Kf - 1) // 2
    tune32 = None
    if not args.no_tune:
        rows32, best32 = autotune(pf_f, ms, step, "fp32", sample_pairs=min(args.tune_pairs, Pf), flags=args.flags)
        tune32 = {"chosen": best32, "variants": rows32}
    else:
        _ = run("fp32", pf[:, :, :min(Kf, 200)])
    torch.cuda.synchronize()
    t = time.perf_counter()
    rec32, st32 = run("fp32", pf_f)
    torch.cuda.synchronize()
    fp32_sec = time.perf_counter() - t
    fp32_value = Pf / fp32_sec
    log(f"fp32 K={Kf}: {fp32_sec:.3f}s {st32}")
    del rec32

    # clstack agreement with the FP64 oracle on a seeded sample of pairs
    from oracle import clm_oracle

    from aspire_amd.workloads import sample_pairs

    ns = min(args.sample_pairs, P)
    slots, k1, k2 = sample_pairs(K, ns, seed=42)
    imgs = np.unique(np.concatenate([k1, k2]))
    idx = {int(v): i for i, v in enumerate(imgs)}
    sub = pf[:, :, torch.as_tensor(imgs, device=pf.device)].cpu().numpy()
    Psub, _ = clm_oracle.preprocess(sub)
    kk1 = np.array([idx[int(v)] for v in k1], np.int32)
    kk2 = np.array([idx[int(v)] for v in k2], np.int32)
    nth = clm_oracle.lib().clmo_max_threads()
    t = time.perf_counter()
    ora = clm_oracle.pairs(Psub, ms, step, kk1, kk2, nthreads=nth)
    cpu_sec = time.perf_counter() - t
    g = {k: v.cpu().numpy()[slots] for k, v in rec.items()}
    same = (g["cl1"] == ora["cl1"]) & (g["cl2"] == ora["cl2"]) & (g["j"] == ora["j"])
    near = 0
    for i in np.flatnonzero(~same):
        v = clm_oracle.entry(Psub, ms, step, kk1[i], kk2[i], g["j"][i], g["cl1"][i] - 1, g["cl2"][i] - 1)
        if ora["best"][i] - v <= 1e-5 * abs(ora["best"][i]):
            near += 1
    m = same & (ora["best"] != 0)
    corr_rel = float(np.max(np.abs(g["corr"][m] - ora["best"][m]) / np.abs(ora["best"][m]))) if m.any() else 0.0
    log(f"parity sample {ns}: agree {same.mean():.5f}, near-tie disagreements {near}/{(~same).sum()}, "
        f"corr rel {corr_rel:.2e}, oracle {cpu_sec:.1f}s on {nth} threads")
    if rec16 is not None:
        g16 = {k: v.cpu().numpy()[slots] for k, v in rec16.items()}
        s16 = (g16["cl1"] == ora["cl1"]) & (g16["cl2"] == ora["cl2"]) & (g16["j"] == ora["j"])
        m16 = s16 & (ora["best"] != 0)
        fp16["parity_sample"] = {
            "pairs": int(ns), "disagree": int((~s16).sum()),
            "corr_rel_max": float(np.max(np.abs(g16["corr"][m16] - ora["best"][m16]) / np.abs(ora["best"][m16])))
            if m16.any() else 0.0}
        del rec16

    # learned form on its own workload: K=5000, shifts U[-12,12], full grid
    # max_shift 12 step 0.5 (49 shifts) against the windowed search (5 shifts)
    # with the committed shift network
    learned = None
    from aspire_amd import shiftnet

    if not args.no_learned and os.path.exists(shiftnet.WEIGHTS):
        del st
        torch.cuda.empty_cache()
        net = shiftnet.ShiftNet.load()
        learned = shiftnet.evaluate(net, K=5000, seed=1005, snr=1 / 16, log=lambda *a: None)
        learned["workload"] = ("K=5000 n=129 n_r=65 n_theta=360 shifts U[-12,12] SNR=1/16; full grid "
                               "max_shift=12 step=0.5 (49 shifts) vs window +-1 px (5 shifts) + exhaustive fallback "
                               "below the calibrated gate or off the grid")
        log(f"learned: {learned}")
        # the same at SNR 1, where the full grid finds most true common lines
        # and its agreement with the window is a meaningful error measure
        learned["snr_1"] = shiftnet.evaluate(net, K=5000, seed=1005, snr=1.0, log=lambda *a: None)
        log(f"learned SNR 1: {learned['snr_1']}")

    cpu = None
    if not args.no_cpu:
        # 1-thread oracle on a bounded sample for the per-core figure
        n1 = max(4, int(min(40, ns)))
        t = time.perf_counter()
        clm_oracle.pairs(Psub, ms, step, kk1[:n1], kk2[:n1], nthreads=1)
        one_sec = time.perf_counter() - t
        cpu = {"value": ns / cpu_sec, "unit": "pairs/s", "cores": int(nth), "kind": "port",
               "sample": f"{ns} seeded pairs of the K=5000 workload (the parity sample), FP64 C oracle (OpenMP)",
               "seconds": cpu_sec, "value_1thread": n1 / one_sec, "sample_1thread": f"{n1} pairs"}

    dropin = None
    if not args.no_dropin:
        dropin = dropin_bench(pf, K, T, ms, step, args)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (seeded Gaussian-blob phantom projections, exact polar NUDFT)",
        "config": {"workload": "cryo_clmatrix K=5000 n=129 n_r=65 n_theta=360 max_shift=10 "
                               "shift_step=0.5 (41 shifts) SNR=1/16 shifts U[-5,5]",
                   "K": K, "pairs": P, "n_r": n_r, "n_theta": T, "n_shifts": S,
                   "precision": args.precision, "variant": all_stats[-1]["variant"]},
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": BF16_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / BF16_PEAK, "traffic": traffic,
                     "kernel": "pair_kernel (pass 1)", "kernel_ms": k_ms,
                     "flops_per_pair": W_pair, "executed_flops_per_pair": W_pair / 2,
                     "executed_frac": achieved / 2 / BF16_PEAK,
                     "note": "flops_per_pair is the reference's arithmetic (8 S L^2 R); the kernel "
                             "obtains it with half the products (X = Re'Re and Y = Im'Im give "
                             "C1(s) = 2(X+Y) and C2(-s) = 2(X-Y) at once), so frac may pass 1; "
                             "executed_frac is the MFMA work actually issued"},
        "cpu_baseline": cpu,
        "fp16": fp16,
        "fp32_gpu": {"value": fp32_value, "unit": "pairs/s", "K": Kf, "seconds": fp32_sec,
                     "ms_pairs": st32["ms_pairs"], "variant": st32["variant"], "autotune": tune32},
        "dropin": dropin,
        "autotune": tune,
        "learned": learned,
        "speedup_vs_fp32": value / fp32_value,
        "speedup_vs_fp32_same_K": Kf == K,
        "clstack_match_pct": 100.0 * float(same.mean()),
        "parity_sample": {"pairs": int(ns), "disagree": int((~same).sum()), "near_tie": int(near),
                          "corr_rel_max": corr_rel},
        "stats": all_stats[-1],
    }
    print(json.dumps(out))


def dropin_bench(pf, K, T, ms, step, args):
    """What a MATLAB caller of the drop-in pays: host pf (complex128,
    column-major, as the MEX gateway receives it) through clm_clmatrix (H2D,
    the device job, D2H of the per-pair records) and the host FP64 assembly
    of clstack, corrstack, shift_equations_map (clm_assemble) and the sparse
    shift_equations (clm_shift_equations_csc), one timed call after an
    untimed one."""
    import torch

    from aspire_amd.clmatrix import assemble, clmatrix_records

    host = pf.permute(2, 1, 0).cpu().numpy().transpose(2, 1, 0)  # n_r x T x K, column-major
    assert host.flags.f_contiguous
    rec, stats = clmatrix_records(host, ms, step, precision=args.precision, flags=args.flags)
    del rec
    t0 = time.perf_counter()
    rec, stats = clmatrix_records(host, ms, step, precision=args.precision, flags=args.flags)
    t1 = time.perf_counter()
    clstack, corrstack, eqs, emap = assemble(K, T, rec)
    t2 = time.perf_counter()
    out = {"seconds": t2 - t0, "pairs_per_s": K * (K - 1) // 2 / (t2 - t0),
           "clm_clmatrix_s": t1 - t0, "device_job_s": stats["ms_total"] * 1e-3,
           "transfers_s": (t1 - t0) - stats["ms_total"] * 1e-3, "pf_bytes": int(host.nbytes),
           "assemble_s": t2 - t1, "shift_equations_nnz": int(eqs.nnz),
           "note": "host pf -> clm_clmatrix (H2D, device job, D2H of records) -> clm_assemble + "
                   "clm_shift_equations_csc (FP64 host) -> scipy CSC"}
    log(f"drop-in: {out}")
    del clstack, corrstack, eqs, emap, rec, host
    torch.cuda.empty_cache()
    return out


def configs_bench(args):
    """K=100 (65 px, n_r=33, no shifts, SNR 1, max_shift 0) with every pair
    checked against the FP64 oracle, and K=1000 (89 px, integer shifts in
    [-4,4], SNR 1/8, n_r=45, max_shift 6, step 1: 13 shifts) with a seeded
    sample of 2000 pairs, in FP32, BF16, FP16 and FP64."""
    import torch

    from aspire_amd import sim
    from aspire_amd.clmatrix import clmatrix_records
    from oracle import clm_oracle

    cases = [dict(name="K=100 n=65 n_r=33 T=360 SNR=1 no shifts, max_shift=0", K=100, n=65, n_r=33, snr=1.0,
                  shifts=0, kind="int", ms=0, step=1, seed=1001, sample=None),
             dict(name="K=1000 n=89 n_r=45 T=360 SNR=1/8 int shifts [-4,4], max_shift=6 step=1 (13 shifts)",
                  K=1000, n=89, n_r=45, snr=1 / 8, shifts=4, kind="int", ms=6, step=1, seed=1002, sample=2000)]
    out = []
    for cs in cases:
        st = sim.make_stack(K=cs["K"], n=cs["n"], n_r=cs["n_r"], n_theta=360, snr=cs["snr"],
                            max_shift_2d=cs["shifts"], shift_kind=cs["kind"], seed=cs["seed"], device="cuda")
        pf = st.pf
        K = cs["K"]
        P = K * (K - 1) // 2
        k1, k2 = np.triu_indices(K, 1)
        sel = np.arange(P) if cs["sample"] is None else np.sort(
            np.random.default_rng(42).choice(P, size=cs["sample"], replace=False))
        host = pf.cpu().numpy()
        Pr, _ = clm_oracle.preprocess(host)
        nth = clm_oracle.lib().clmo_max_threads()
        t = time.perf_counter()
        ora = clm_oracle.pairs(Pr, cs["ms"], cs["step"], k1[sel].astype(np.int32), k2[sel].astype(np.int32),
                               nthreads=nth)
        cpu_s = time.perf_counter() - t
        row = dict(workload=cs["name"], pairs=P, checked_pairs=int(sel.size),
                   cpu_oracle_pairs_per_s=sel.size / cpu_s, cpu_threads=int(nth), precisions={})
        for prec in ("fp32", "bf16", "fp16", "fp64"):
            run = lambda: clmatrix_records(pf, cs["ms"], cs["step"], precision=prec)
            run()
            torch.cuda.synchronize()
            t = time.perf_counter()
            rec, stats = run()
            torch.cuda.synchronize()
            sec = time.perf_counter() - t
            g = {k: v.cpu().numpy()[sel] for k, v in rec.items()}
            same = (g["cl1"] == ora["cl1"]) & (g["cl2"] == ora["cl2"]) & (g["j"] == ora["j"])
            near = 0
            for i in np.flatnonzero(~same):
                v = clm_oracle.entry(Pr, cs["ms"], cs["step"], k1[sel][i], k2[sel][i], g["j"][i], g["cl1"][i] - 1,
                                     g["cl2"][i] - 1)
                if ora["best"][i] - v <= 1e-5 * abs(ora["best"][i]):
                    near += 1
            m = same & (ora["best"] != 0)
            crel = float(np.max(np.abs(g["corr"][m] - ora["best"][m]) / np.abs(ora["best"][m]))) if m.any() else 0.0
            row["precisions"][prec] = dict(seconds=sec, pairs_per_s=P / sec, clstack_match_pct=100 * float(same.mean()),
                                           near_tie_disagreements=near, corr_rel_max=crel,
                                           corr_bit_identical=bool(np.array_equal(g["corr"][same], ora["best"][same])),
                                           uncertain=stats["uncertain_pairs"], refined=stats["refined_pairs"])
            log(f"{cs['name']} {prec}: {row['precisions'][prec]}")
        if cs["K"] == 100:
            # argument 6 (map_filter_radius = 1): the filtered FP64 search, every pair
            t = time.perf_counter()
            oraf = clm_oracle.pairs(Pr, cs["ms"], cs["step"], k1[sel].astype(np.int32), k2[sel].astype(np.int32),
                                    nthreads=nth, map_filter_radius=1)
            cpu_f = time.perf_counter() - t
            clmatrix_records(pf, cs["ms"], cs["step"], map_filter_radius=1)
            torch.cuda.synchronize()
            t = time.perf_counter()
            rec, stats = clmatrix_records(pf, cs["ms"], cs["step"], map_filter_radius=1)
            torch.cuda.synchronize()
            sec = time.perf_counter() - t
            g = {k: v.cpu().numpy()[sel] for k, v in rec.items()}
            ident = all(np.array_equal(g[a_], oraf[b_]) for a_, b_ in
                        (("cl1", "cl1"), ("cl2", "cl2"), ("j", "j"), ("corr", "best"), ("shift", "shift")))
            row["map_filter_radius_1"] = dict(seconds=sec, pairs_per_s=P / sec, records_bit_identical=ident,
                                              cpu_oracle_pairs_per_s=sel.size / cpu_f)
            log(f"{cs['name']} map_filter_radius=1: {row['map_filter_radius_1']}")
        out.append(row)
    print(json.dumps({"mode": "configs", "cases": out}))


def grouped_bench(args):
    """64 independent stacks (seed 4000+j): K ~ U[50,2000], n_r in {33,45,65}
    (images of 2 n_r - 1 px), n_theta=360, max_shift in {0,4,8} with step 1,
    SNR in {1,1/8,1/32}, complex64 rays; all stacks in one graph-captured
    grouped job.  Per-stack clstack agreement with the FP64 oracle on a
    seeded sample of pairs of every stack."""
    import torch

    from aspire_amd import sim
    from aspire_amd.clmatrix import clmatrix_grouped
    from oracle import clm_oracle

    rng = np.random.default_rng(4000)
    specs = []
    for j in range(args.stacks):
        K = int(rng.integers(50, 2001))
        n_r = int(rng.choice([33, 45, 65]))
        ms = int(rng.choice([0, 4, 8]))
        snr = float(rng.choice([1.0, 1 / 8, 1 / 32]))
        specs.append((K, n_r, ms, snr))
    t0 = time.time()
    pfs = []
    for j, (K, n_r, ms, snr) in enumerate(specs):
        st = sim.make_stack(K=K, n=2 * n_r - 1, n_r=n_r, n_theta=360, snr=snr, max_shift_2d=ms / 2,
                            shift_kind="uniform", seed=4000 + j, device="cuda", dtype="complex64")
        pfs.append(st.pf)
    torch.cuda.synchronize()
    log(f"grouped: {len(pfs)} stacks generated in {time.time() - t0:.1f}s")
    mss = [s[2] for s in specs]
    P = sum(s[0] * (s[0] - 1) // 2 for s in specs)
    recs, st0 = clmatrix_grouped(pfs, mss, 1, "bf16", graph=True)       # capture + first run
    for i in range(args.warmup):
        recs, _ = clmatrix_grouped(pfs, mss, 1, "bf16", graph=True, out=recs)
    times = []
    for i in range(args.steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        recs, stats = clmatrix_grouped(pfs, mss, 1, "bf16", graph=True, out=recs)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        log(f"grouped step {i}: {times[-1]:.3f}s {stats}")
    sec = float(np.mean(times))
    # per-stack parity on a seeded sample
    srng = np.random.default_rng(42)
    per, agree_all, n_all, near_all = [], 0, 0, 0
    cpu_t, cpu_pairs = 0.0, 0
    for j, (K, n_r, ms, snr) in enumerate(specs):
        Pj = K * (K - 1) // 2
        ns = min(args.grouped_sample, Pj)
        slots = np.sort(srng.choice(Pj, size=ns, replace=False))
        firsts = np.array([k * K - k * (k + 1) // 2 for k in range(K)], dtype=np.int64)
        k1 = np.searchsorted(firsts, slots, side="right") - 1
        k2 = slots - firsts[k1] + k1 + 1
        imgs = np.unique(np.concatenate([k1, k2]))
        idx = {int(v): i for i, v in enumerate(imgs)}
        sub = pfs[j][:, :, torch.as_tensor(imgs, device=pfs[j].device)].cpu().numpy().astype(np.complex128)
        Psub, _ = clm_oracle.preprocess(sub)
        kk1 = np.array([idx[int(v)] for v in k1], np.int32)
        kk2 = np.array([idx[int(v)] for v in k2], np.int32)
        t = time.perf_counter()
        ora = clm_oracle.pairs(Psub, ms, 1, kk1, kk2, nthreads=clm_oracle.lib().clmo_max_threads())
        cpu_t += time.perf_counter() - t
        cpu_pairs += ns
        g = {k: v.cpu().numpy()[slots] for k, v in recs[j].items()}
        same = (g["cl1"] == ora["cl1"]) & (g["cl2"] == ora["cl2"]) & (g["j"] == ora["j"])
        near = 0
        for i in np.flatnonzero(~same):
            v = clm_oracle.entry(Psub, ms, 1, kk1[i], kk2[i], g["j"][i], g["cl1"][i] - 1, g["cl2"][i] - 1)
            if ora["best"][i] - v <= 1e-5 * abs(ora["best"][i]):
                near += 1
        agree_all += int(same.sum())
        near_all += near
        n_all += ns
        per.append({"K": K, "n_r": n_r, "max_shift": ms, "snr": snr, "sample": ns,
                    "agree": float(same.mean()), "near_tie_disagreements": near})
    out = {"mode": "grouped", "stacks": len(specs), "pairs": P, "seconds": sec, "pairs_per_s": P / sec,
           "graph": True, "stats": stats, "clstack_match_pct": 100.0 * agree_all / n_all,
           "sample_pairs": n_all, "near_tie_disagreements": near_all,
           "cpu_oracle_pairs_per_s": cpu_pairs / cpu_t, "per_stack": per}
    print(json.dumps(out))


def pft_bench(args):
    """cryo_pft of the K=5000 workload's images (n=129, n_r=65, n_theta=360,
    FP64) with clm_pft: images/s; the contraction kernel's FP64 MFMA roofline
    (algorithmic 4 flops per polar point, pixel and image: a complex weight
    times a real pixel); the FP64 C oracle on a bounded image sample as the
    CPU baseline and parity check; the same transform by torch (weights
    materialised, rocBLAS ZGEMM) for reference."""
    import torch

    from aspire_amd import sim
    from aspire_amd.pft import cryo_pft
    from oracle import clm_oracle

    n, n_r, T, K = 129, 65, 360, args.K
    t0 = time.time()
    imgs, _, _, _ = sim.make_images(K=K, n=n, snr=1 / 16, max_shift_2d=5.0, shift_kind="uniform", seed=1003,
                                    device="cuda")
    # column-major n x n x K with p(u, v, k) = imgs[k, u, v]
    p = imgs.transpose(1, 2).contiguous().permute(2, 1, 0)
    del imgs
    torch.cuda.synchronize()
    log(f"pft: K={K} images generated in {time.time() - t0:.1f}s")
    M = n_r * (T // 2)
    W_img = 4.0 * M * n * n  # algorithmic flops per image
    pf = None
    for i in range(args.warmup):
        pf, _ = cryo_pft(p, n_r, T)
    times, kms = [], []
    for i in range(args.steps):
        del pf
        torch.cuda.synchronize()
        st = {}
        t = time.perf_counter()
        pf, _ = cryo_pft(p, n_r, T, stats=st)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        kms.append(st["ms_pft"])
        log(f"pft step {i}: {times[-1]:.4f}s {st}")
    sec = float(np.mean(times))
    k_ms = float(np.mean(kms))
    achieved = W_img * K / (k_ms * 1e-3)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pft_kernel_k5000_traffic.json")
    if os.path.exists(tpath) and K == 5000:
        traffic = json.load(open(tpath))["hbm_bytes_per_launch"]
    Mp = (M + 127) // 128 * 128
    nup = (n + 3) // 4 * 4
    Kp = (K + 63) // 64 * 64
    # torch (rocBLAS ZGEMM over materialised weights), same images
    torch.cuda.synchronize()
    t = time.perf_counter()
    ref = sim.cryo_pft(p.permute(2, 0, 1), n_r, T)
    torch.cuda.synchronize()
    t_torch = time.perf_counter() - t
    zdiff = (pf - ref).abs().max().item() / ref.abs().max().item()
    del ref
    # CPU oracle on a bounded sample; parity on the same images
    ns = 16
    sample = np.linspace(0, K - 1, ns).astype(int)
    ps = p[:, :, torch.as_tensor(sample, device=p.device)].cpu().numpy()
    nth = clm_oracle.lib().clmo_max_threads()
    t = time.perf_counter()
    ora = clm_oracle.cryo_pft(ps, n_r, T, nthreads=nth)
    cpu_sec = time.perf_counter() - t
    got = pf[:, :, torch.as_tensor(sample, device=pf.device)].cpu().numpy()
    rel = float(np.abs(got - ora).max() / np.abs(ora).max())
    nufft = nufft_bench(args, p, pf, n_r, T, sample, ora)
    out = {"metric": "images/s: cryo_pft n=129 n_r=65 n_theta=360 (exact NUDFT, FP64)", "value": K / sec,
           "unit": "images/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * sec, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "cryo_pft of the K=5000 common-lines workload's images", "K": K, "n": n,
                      "n_r": n_r, "n_theta": T},
           "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                        "frac": achieved / 78.6e12, "traffic": traffic, "kernel": "pft_kernel", "kernel_ms": k_ms,
                        "flops_per_image": W_img, "executed_flops": 4.0 * Mp * n * nup * Kp},
           "stats": st,
           "cpu_baseline": {"value": ns / cpu_sec, "unit": "images/s", "cores": nth, "kind": "port",
                            "sample": f"{ns} images of the workload, FP64 C oracle (OpenMP)"},
           "parity_sample": {"images": ns, "max_abs_err_rel_to_max": rel},
           "torch_zgemm": {"seconds": t_torch, "images_per_s": K / t_torch, "max_diff_rel_to_max": zdiff},
           "nufft": nufft}
    print(json.dumps(out))


def nufft_bench(args, p, pf_exact, n_r, T, sample, ora):
    """cryo_pft's NUFFT method (CLM_PFT_NUFFT) on the same images: time,
    stage breakdown, difference from the exact method over all images and
    from the oracle on its sample; then both methods on 400 images of side
    359 (n_r=180), where the exact sums cost 21x more per image."""
    import torch

    from aspire_amd.pft import cryo_pft

    K = p.shape[2]

    def timed(q, nr, method, steps):
        out, _ = cryo_pft(q, nr, T, method=method)  # warm-up (FFT plans)
        ts, sts = [], []
        for _ in range(steps):
            del out
            torch.cuda.synchronize()
            st = {}
            t = time.perf_counter()
            out, _ = cryo_pft(q, nr, T, stats=st, method=method)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
            sts.append(st)
        return out, float(np.mean(ts)), sts[-1]

    pfn, sec, st = timed(p, n_r, "nufft", args.steps)
    diff = (pfn - pf_exact).abs().max().item() / pf_exact.abs().max().item()
    got = pfn[:, :, torch.as_tensor(sample, device=pfn.device)].cpu().numpy()
    rel = float(np.abs(got - ora).max() / np.abs(ora).max())
    del pfn
    log(f"pft nufft K={K}: {sec:.4f}s {st} vs exact {diff:.2e}, vs oracle {rel:.2e}")
    n2, nr2, K2 = 359, 180, 400
    g = torch.Generator(device=p.device).manual_seed(359)
    q = torch.randn((K2, n2, n2), dtype=torch.float64, device=p.device, generator=g).permute(2, 1, 0)
    a, sec_e, st_e = timed(q, nr2, "exact", 1)
    b, sec_n, st_n = timed(q, nr2, "nufft", 1)
    diff2 = (a - b).abs().max().item() / a.abs().max().item()
    del a, b, q
    log(f"pft n={n2} K={K2}: exact {sec_e:.4f}s, nufft {sec_n:.4f}s, diff {diff2:.2e}")
    return {"seconds": sec, "images_per_s": K / sec, "stats": st, "max_diff_vs_exact_rel_to_max": diff,
            "max_abs_err_vs_oracle_rel_to_max": rel, "oracle_images": len(sample),
            "large": {"n": n2, "n_r": nr2, "n_theta": T, "K": K2, "images": "standard normal pixels",
                      "exact_seconds": sec_e, "nufft_seconds": sec_n, "speedup": sec_e / sec_n,
                      "max_diff_rel_to_max": diff2, "exact_stats": st_e, "nufft_stats": st_n}}


def shifts_bench(args):
    """2-D shifts by LSQR (tol 1e-8, maxit 100: cryo_estimate_shifts.m:229)
    on the shift equations of the K=5000 workload, computed on the device
    from the BF16 engine's per-pair outputs.  HBM-bound: the algorithmic
    traffic is 20 bytes per pair per iteration (one pass forms u' = A v -
    (alpha/beta) u and A'u' together: the packed lines and u read once, u
    written once; a solver with separate A v and A'u passes moves 32).
    CPU baseline: the oracle's
    LSQR (scipy CSR products) on the same full system, a few iterations
    timed and scaled to the iterations the device ran."""
    import torch

    from aspire_amd import sim
    from aspire_amd.clmatrix import clmatrix_records
    from aspire_amd.shifts import estimate_shifts
    from oracle import shifts_oracle

    K, n_theta = args.K, 360
    st0 = sim.make_stack(K=K, n=129, n_r=65, n_theta=n_theta, snr=1 / 16, max_shift_2d=5.0, shift_kind="uniform",
                         seed=1003, device="cuda")
    rec, _ = clmatrix_records(st0.pf, 10, 0.5, precision="bf16")
    drop_in = estimate_shifts_dropin(st0, K, n_theta)
    del st0
    torch.cuda.synchronize()
    P = K * (K - 1) // 2
    for i in range(args.warmup):
        x, st = estimate_shifts(rec["cl1"], rec["cl2"], rec["shift"], K, n_theta)
    times, sts = [], []
    for i in range(args.steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        x, st = estimate_shifts(rec["cl1"], rec["cl2"], rec["shift"], K, n_theta)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        sts.append(st)
        log(f"shifts step {i}: {times[-1]:.4f}s {st}")
    sec = float(np.mean(times))
    it = st["iterations"]
    dev_ms = float(np.mean([s_["ms"] for s_ in sts]))
    achieved = 20.0 * P * it / (dev_ms * 1e-3)
    # measured HBM bytes of one solve (rocprof FETCH_SIZE/WRITE_SIZE, calibrated),
    # valid for the workload and iteration count it was measured on
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "lsqr_k5000_traffic.json")
    if os.path.exists(tpath) and K == 5000:
        tj = json.load(open(tpath))
        if tj.get("iterations") == it:
            traffic = tj["hbm_bytes_per_launch"]
    # CPU oracle: same system, a few iterations
    cl1, cl2, b = (rec[k].cpu().numpy() for k in ("cl1", "cl2", "shift"))
    A, bb = shifts_oracle.system(K, n_theta, cl1, cl2, b)
    n_cpu = 3
    t = time.perf_counter()
    xo, io = shifts_oracle.lsqr(A, bb, tol=0.0, maxit=n_cpu)
    cpu_it = (time.perf_counter() - t) / n_cpu
    xg, _ = estimate_shifts(rec["cl1"], rec["cl2"], rec["shift"], K, n_theta, tol=0.0, maxit=n_cpu)
    xg = xg.cpu().numpy().ravel()
    out = {"metric": "pairs/s: LSQR of the shift equations, K=5000 (cryo_estimate_shifts.m:229)",
           "value": P / sec, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * sec, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "shift equations of the K=5000 common-lines workload", "K": K, "pairs": P,
                      "tol": 1e-8, "maxit": 100},
           "lsqr": st,
           "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": 8000.0, "unit": "GB/s",
                        "frac": achieved / 8e12, "traffic": traffic, "kernel": "whole solve",
                        "bytes_per_pair_iteration": 20},
           "cpu_baseline": {"value": P / (cpu_it * it), "unit": "pairs/s", "cores": 1, "kind": "port",
                            "sample": f"{n_cpu} LSQR iterations of the oracle (scipy CSR) on the full system, "
                                      f"scaled to {it} iterations"},
           "parity_sample": {"iterations": n_cpu,
                             "max_abs_diff_rel_to_max": float(np.abs(xg - xo).max() / np.abs(xo).max())},
           "cryo_estimate_shifts": drop_in}
    print(json.dumps(out))


def estimate_shifts_dropin(st0, K, n_theta, max_shift=10, shift_step=0.5):
    """The whole cryo_estimate_shifts(pf, rotations, 10, 0.5) on the stack's
    true rotations (in the estimated rotations' convention), at the default
    memoryfactor (10000 MB: 125,000 of the 12.5 M pairs at K=5000), and the
    common-lines-from-rotations kernel over every pair; exact agreement of the
    kernel's lines and 1-D shifts with the oracle on a sample of pairs."""
    import torch

    from aspire_amd.shifts import common_lines_from_rotations, cryo_estimate_shifts, n_equations
    from oracle import clm_oracle, shifts_oracle

    rot = np.ascontiguousarray(st0.rots.transpose(1, 0, 2))
    cryo_estimate_shifts(st0.pf, rot, max_shift, shift_step)  # warm-up
    torch.cuda.synchronize()
    t = time.perf_counter()
    est, eqs = cryo_estimate_shifts(st0.pf, rot, max_shift, shift_step)
    sec = time.perf_counter() - t
    P = K * (K - 1) // 2
    torch.cuda.synchronize()
    t = time.perf_counter()
    cl1, cl2, sh = common_lines_from_rotations(st0.pf, rot, max_shift, shift_step)
    torch.cuda.synchronize()
    lines_sec = time.perf_counter() - t
    # parity sample: 64 seeded pairs against the oracle
    sample = np.sort(np.random.default_rng(7).choice(P, size=64, replace=False))
    k1, k2 = np.triu_indices(K, 1)
    n_img = np.unique(np.concatenate([k1[sample], k2[sample]]))
    sub = st0.pf[:, :, n_img].cpu().numpy() if isinstance(st0.pf, torch.Tensor) else st0.pf[:, :, n_img]
    Ps, _ = clm_oracle.preprocess(sub)
    pos = {int(k): i for i, k in enumerate(n_img)}
    a = np.array([pos[int(k)] for k in k1[sample]])
    b = np.array([pos[int(k)] for k in k2[sample]])
    t = time.perf_counter()
    ora = shifts_oracle.rotation_pairs(Ps, rot[:, :, n_img], n_theta, max_shift, shift_step, a, b)
    cpu_sec = time.perf_counter() - t
    c1, c2, s1 = cl1.cpu().numpy()[sample], cl2.cpu().numpy()[sample], sh.cpu().numpy()[sample]
    same = int(np.sum((c1 == ora["cl1"]) & (c2 == ora["cl2"]) & (s1 == ora["shift"])))
    truth = st0.shifts.reshape(-1)
    A = eqs[:, :-1]
    resid = np.asarray(A @ truth - eqs[:, -1].toarray().ravel())
    return {"equations": int(eqs.shape[0]), "n_equations": n_equations(K, 10000), "seconds": sec,
            "max_shift": max_shift, "shift_step": shift_step,
            "lines_all_pairs": {"pairs": P, "seconds": lines_sec, "pairs_per_s": P / lines_sec},
            "parity_sample": {"pairs": len(sample), "identical_to_oracle": same},
            "cpu_baseline": {"value": len(sample) / cpu_sec, "unit": "pairs/s", "cores": 1, "kind": "port",
                             "sample": f"the oracle's per-pair loop (numpy geometry, one FP64 C correlation "
                                       f"call per shift and orientation, one thread) on the {len(sample)} "
                                       f"sampled pairs"},
            "rms_residual_of_true_shifts": float(np.sqrt(np.mean(resid ** 2))),
            "rms_rhs": float(np.sqrt(np.mean(eqs[:, -1].toarray() ** 2)))}


def vote_bench(args):
    """S = cryo_syncmatrix_vote(clstack, 360) of the K=5000 workload's common
    lines (BF16 engine), on the device; the oracle on a seeded sample of pairs
    (one thread) as the CPU baseline and parity sample.  The work is FP64
    transcendental evaluation: the kernel evaluates 60 Gaussians per voting
    triangle (the exponential count per second is reported)."""
    import torch

    from aspire_amd import sim
    from aspire_amd.clmatrix import clmatrix_records
    from aspire_amd.vote import syncmatrix_vote_records
    from oracle import clm_oracle

    K, T = args.K, 360
    st0 = sim.make_stack(K=K, n=129, n_r=65, n_theta=T, snr=1 / 16, max_shift_2d=5.0, shift_kind="uniform",
                         seed=1003, device="cuda")
    rec, _ = clmatrix_records(st0.pf, 10, 0.5, precision="bf16")
    del st0
    torch.cuda.synchronize()
    P = K * (K - 1) // 2
    S = None
    for i in range(args.warmup):
        S, st = syncmatrix_vote_records(rec["cl1"], rec["cl2"], K, T)
        del S
    times = []
    for i in range(args.steps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        S, st = syncmatrix_vote_records(rec["cl1"], rec["cl2"], K, T)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        log(f"vote step {i}: {times[-1]:.3f}s {st}")
        if i + 1 < args.steps:
            del S
    sec = float(np.mean(times))
    cl1, cl2 = rec["cl1"].cpu().numpy(), rec["cl2"].cpu().numpy()
    k1, k2 = np.triu_indices(K, 1)
    clstack = np.zeros((K, K))
    clstack[k1, k2] = cl1
    clstack[k2, k1] = cl2
    flat = np.ascontiguousarray(clstack.T).ravel()
    ntri = int(np.count_nonzero(clstack)) // 2  # pairs with a line
    sample = np.random.default_rng(0).choice(P, size=60, replace=False)
    t = time.perf_counter()
    same = 0
    for e in sample:
        a, b = int(k1[e]), int(k2[e])
        ref = clm_oracle.vote_pair(clstack, T, a, b, _flat=flat)
        same += int(np.abs(S[2 * a:2 * a + 2, 2 * b:2 * b + 2].cpu().numpy() - ref).max() <= 1e-10)
    cpu_sec = time.perf_counter() - t
    out = {"metric": "pairs/s: cryo_syncmatrix_vote of the K=5000 workload's common lines", "value": P / sec,
           "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * sec,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "voting synchronization matrix, K=5000, n_theta=360", "K": K, "pairs": P},
           "stats": st, "pairs_with_line": ntri,
           "exp_per_s_upper_bound": 60.0 * P * (K - 2) / sec,
           "roofline": vote_roofline(K, st),
           "cpu_baseline": {"value": sample.size / cpu_sec, "unit": "pairs/s", "cores": 1, "kind": "port",
                            "sample": f"{sample.size} seeded pairs, FP64 C oracle (one thread)"},
           "parity_sample": {"pairs": int(sample.size), "blocks_equal_1e-10": same}}
    print(json.dumps(out))


def vote_roofline(K, st):
    """FP64 roofline of the vote at K=5000: its FP64 operations per vote
    as rocprof counts them (ADD + MUL + 2 FMA + TRANS wave instructions x 64,
    profiles/vote_kernel_k5000_f64.md: the reference's arithmetic, each
    triangle evaluated once, 60 Gaussians of the histogram per triangle)
    over the vote's own device time (HIP events around its kernels)."""
    if K != 5000:
        return None
    flops = 6.9e13
    ms = float(st.get("ms_vote", 0.0))
    if ms <= 0:
        return None
    achieved = flops / (ms * 1e-3)
    return {"bound": "fp64-valu", "achieved": achieved / 1e12, "peak": 78.6, "unit": "TFLOP/s",
            "frac": achieved / 78.6e12, "traffic": None, "kernel": "vote (peak_kernel + rot_kernel)",
            "kernel_ms": ms, "flops_per_vote": flops,
            "flops_basis": "FP64 operations counted by rocprof for this workload (each triangle once)"}


def sync3n_bench(args):
    """The sync3n triangle kernels (triangles_scores_mex, pairs_probabilities_mex,
    signs_times_v_mex) on the relative rotations of the K=5000 workload's
    rotations (R_ij = R_i^T R_j, 30% replaced by random rotations, 40%
    conjugated by J).  value = triangles/s of the scores kernel.  Roofline:
    FP64 vector; the reference's arithmetic is 297 flops per triangle (four
    3x3 products, four squared norms, three divisions, square roots and
    subtractions), which the per-pair kernel evaluates three times (once per
    side; rocprof counts 659 executed FP64 operations per triangle, with the
    division and square-root sequences).  CPU baseline: the oracle's split
    of the triangle loop over threads (the reference's) on the first 600
    images; parity: the GPU against the one-thread oracle on those images."""
    import torch

    from aspire_amd import sim
    from aspire_amd.sync3n import pairs_probabilities_mex, signs_times_v_mex, triangles_scores_mex
    from oracle import clm_oracle

    N = args.K
    rng = np.random.default_rng(1003)
    rots = torch.as_tensor(sim.rand_rots(rng, N).transpose(2, 0, 1).copy(), device="cuda")  # N x 3 x 3
    i, j = np.triu_indices(N, 1)
    P = i.size
    Rt = torch.empty((P, 3, 3), dtype=torch.float64, device="cuda")  # C order of R^T = MATLAB order of R
    Jm = torch.diag(torch.tensor([1.0, 1.0, -1.0], dtype=torch.float64, device="cuda"))
    gen = torch.Generator(device="cpu").manual_seed(5)
    for a in range(0, P, 1 << 22):
        ii = torch.as_tensor(i[a:a + (1 << 22)], device="cuda")
        jj = torch.as_tensor(j[a:a + (1 << 22)], device="cuda")
        M = torch.matmul(rots[ii].transpose(1, 2).contiguous(), rots[jj].contiguous())  # R_i^T R_j
        n = M.shape[0]
        out = torch.rand(n, generator=gen) < 0.3
        if out.any():
            q = torch.randn((int(out.sum()), 4), generator=gen, dtype=torch.float64)
            q = q / q.norm(dim=1, keepdim=True)
            M[out.to("cuda")] = torch.as_tensor(sim.q_to_rot(q.numpy().T).transpose(2, 0, 1).copy(), device="cuda")
        fl = (torch.rand(n, generator=gen) < 0.4).to("cuda")
        M[fl] = Jm @ M[fl] @ Jm
        Rt[a:a + n] = M.transpose(1, 2)
    del rots
    torch.cuda.synchronize()
    T = N * (N - 1) * (N - 2) // 6

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        ts, out = [], None
        for _ in range(args.steps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        return float(np.mean(ts)), out

    sec, (c, h, st) = timed(lambda: triangles_scores_mex(Rt, 100))
    log(f"sync3n scores: {sec:.3f}s {st}")
    prm = (0.5, 1.0, 2.2, 1.0, 2.5, 0.78)
    sec_p, (_, _, stp) = timed(lambda: pairs_probabilities_mex(Rt, *prm))
    log(f"sync3n probabilities: {sec_p:.3f}s {stp}")
    v = torch.rand((P, 4), dtype=torch.float64, device="cuda")
    sec_v, (_, stv) = timed(lambda: signs_times_v_mex(Rt, v, 0))
    log(f"sync3n signs x v: {sec_v:.3f}s {stv}")
    achieved = 297.0 * T / (st["ms_kernel"] * 1e-3)
    # parity and CPU baseline on the first Ns images (their pairs come first in each row)
    Ns = min(600, N)
    sel = np.flatnonzero(j < Ns)
    Rsub = Rt[torch.as_tensor(sel, device="cuda")].cpu().numpy().transpose(2, 1, 0)  # 3 x 3 x Ps
    cg, hg, _ = triangles_scores_mex(Rsub, 100)
    co, ho = clm_oracle.triangles_scores(Rsub, 100, nthreads=1)
    nth = clm_oracle.lib().clmo_max_threads()
    t = time.perf_counter()
    clm_oracle.triangles_scores(Rsub, 100, nthreads=nth)
    cpu_s = time.perf_counter() - t
    Ts = Ns * (Ns - 1) * (Ns - 2) // 6
    # measured HBM bytes of the scores kernel at N=5000 (rocprof, calibrated)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "sync3n_k5000_traffic.json")
    if os.path.exists(tpath) and N == 5000:
        traffic = json.load(open(tpath))["hbm_bytes_per_launch"]
    out = {"metric": "triangles/s: sync3n triangles_scores_mex, N=5000 relative rotations", "value": T / sec,
           "unit": "triangles/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * sec, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "sync3n on the relative rotations of the K=5000 workload's rotations "
                                  "(30% random, 40% J-conjugated)", "N": N, "pairs": P, "triangles": T,
                      "n_intervals": 100},
           "stats": st,
           "pairs_probabilities": {"seconds": sec_p, "triangles_per_s": T / sec_p, "stats": stp},
           "signs_times_v": {"seconds": sec_v, "triangles_per_s": T / sec_v, "n_eigs": 4, "stats": stv},
           "roofline": {"bound": "fp64-valu", "achieved": achieved / 1e12, "peak": 78.6, "unit": "TFLOP/s",
                        "frac": achieved / 78.6e12, "traffic": traffic, "kernel": "tri_scores_kernel",
                        "kernel_ms": st["ms_kernel"], "flops_per_triangle": 297,
                        "executed_flops_per_triangle": 659},
           "cpu_baseline": {"value": Ts / cpu_s, "unit": "triangles/s", "cores": int(nth), "kind": "port",
                            "sample": f"N={Ns} (first images), the oracle's threaded split (the reference's)"},
           "parity_sample": {"N": Ns, "hist_equal": bool(np.array_equal(hg, ho)),
                             "cum_bit_identical": bool(np.array_equal(cg, co))}}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
