"""In-tree build of the HIP engine for gfx950 (``hipcc`` cross-compiles without a GPU)."""

from __future__ import annotations

import os
import shutil
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(_HERE), "csrc", "fdtd_engine.hip")
OUT = os.path.join(_HERE, "_lib", "libmeep_fdtd.so")
#: The same engine built to flush subnormals (the reference's ``flush`` policy).
OUT_FLUSH = os.path.join(_HERE, "_lib", "libmeep_fdtd_flush.so")

#: One IEEE rounding per operation (no contraction), subnormals kept, packed FP32 off.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

#: The subnormal policy is a build-time flag (the reference's floating-point contract keeps or
#: flushes them, docs/design/floating-point.md): the flushing build runs every kernel with the
#: float32 and float64 denormal modes at flush (subnormal inputs read as zero, subnormal results
#: written as zero, sign kept: x86's FTZ|DAZ), and reports it through fdtd_subnormal_mode().
FLUSH_FLAGS = [f for f in HIPCC_FLAGS if f != "-fno-gpu-flush-denormals-to-zero"] + [
    "-Xarch_device", "-fdenormal-fp-math=preserve-sign", "-DFDTD_FLUSH_SUBNORMALS=1",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found (ROCm is required to build the engine)")
    return path


def build(verbose: bool = False) -> str:
    """Both builds, compiled side by side; returns the default (subnormals kept) library."""
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    jobs = []
    for out, flags in ((OUT, HIPCC_FLAGS), (OUT_FLUSH, FLUSH_FLAGS)):
        tmp = out + ".tmp"
        proc = subprocess.Popen([hipcc(), *flags, "-o", tmp, CSRC], stdout=subprocess.PIPE,
                                stderr=subprocess.PIPE, text=True)
        jobs.append((out, tmp, proc))
    failed = []
    for out, tmp, proc in jobs:
        _, err = proc.communicate()
        if proc.returncode != 0:
            failed.append(f"{os.path.basename(out)}: hipcc failed ({proc.returncode}):\n"
                          f"{err[-4000:]}")
            continue
        if verbose and err:
            print(err)
        os.replace(tmp, out)
    if failed:
        raise RuntimeError("\n".join(failed))
    return OUT
