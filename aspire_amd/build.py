"""Build libclm.so (HIP, gfx950) in-tree with hipcc.

``python -m aspire_amd.build`` compiles, in parallel, ``aspire_amd/csrc/
clm_api.hip`` (the engine: context, orchestration and its kernels),
``shiftnet.hip`` (the shift network), ``pft.hip`` (the polar Fourier
transform), ``shifts.hip`` (the 2-D shift least squares), ``vote.hip`` (the
voting synchronization matrix), ``sync3n.hip`` (the sync3n triangle kernels) and ``pair_variants.hip`` once per (operand
type, shift-grid symmetry), and links them into ``aspire_amd/libclm.so``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libclm.so")
CSRC = os.path.join(PKG, "csrc")
# (source, extra flags): the engine, the shift network, and the pair-kernel
# instantiations, one object per (operand type, shift-grid symmetry)
UNITS = [(os.path.join(CSRC, "clm_api.hip"), []), (os.path.join(CSRC, "shiftnet.hip"), []),
         (os.path.join(CSRC, "pft.hip"), []), (os.path.join(CSRC, "shifts.hip"), []),
         (os.path.join(CSRC, "vote.hip"), []), (os.path.join(CSRC, "sync3n.hip"), [])] + [
    (os.path.join(CSRC, "pair_variants.hip"), [f"-DCLM_VARIANT_OP={op}", f"-DCLM_VARIANT_SYM={sym}"])
    for op in ("bf16", "f16", "f32") for sym in (1, 0)]
ARCH = "gfx950"
ROCM_LIB = "/opt/rocm/lib"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # fixed per-element arithmetic in the FP64 stages (bit-parity with the
    # oracle); hot FP32 code uses explicit fmaf where a fused op is wanted
    "-ffp-contract=off",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found; libclm needs ROCm's hipcc to build")


def build(verbose: bool = False) -> str:
    """Compile libclm.so; returns its path.  Always recompiles."""
    import tempfile

    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "csrc")]
    with tempfile.TemporaryDirectory(prefix="clm_build_") as tmpdir:
        objs, procs = [], []
        for n, (src, extra) in enumerate(UNITS):
            obj = os.path.join(tmpdir, f"{n}_{os.path.basename(src)}.o")
            cmd = [hipcc(), *HIPCC_FLAGS, *extra, *inc, "-c", "-o", obj, src]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
            objs.append(obj)
        failed = None
        for cmd, pr in procs:
            out, err = pr.communicate()
            if pr.returncode != 0:
                sys.stderr.write(out + err)
                failed = failed or pr.returncode
        if failed:
            raise RuntimeError(f"hipcc failed with exit code {failed}")
        tmp = LIB + ".tmp"
        # hipFFT: the grid transforms of cryo_pft's NUFFT method
        cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
               f"-L{ROCM_LIB}", "-lhipfft", f"-Wl,-rpath,{ROCM_LIB}"]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc link failed with exit code {res.returncode}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
