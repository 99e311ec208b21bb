"""meep_gpu for MI355X: the reference package's API over a HIP time-stepping engine.

The per-step FDTD update (Yee curl, UPML auxiliaries, chi1inv constitutive step, Lorentz/Drude
ADE, source deposit, flux/DFT accumulation) runs in hand-written gfx950 kernels
(``csrc/fdtd_engine.hip``) reached through a C ABI (``include/meep_fdtd.h``). Exports are
lazy so that importing the package loads nothing until a name is used.
"""

from __future__ import annotations

import importlib

__version__ = "0.1.0"

_LAZY = {
    "Grid": ("grid", "Grid"),
    "PML": ("pml", "PML"),
    "Susceptibility": ("dispersion", "Susceptibility"),
    "LORENTZIAN": ("dispersion", "LORENTZIAN"),
    "DRUDE": ("dispersion", "DRUDE"),
    "HipSimulation": ("engine", "HipSimulation"),
    "FdtdDriver": ("driver", "FdtdDriver"),
    "GaussianSource": ("sources", "GaussianSource"),
    "FluxMonitor": ("dft", "FluxMonitor"),
    "HipStepPlan": ("fastpath", "HipStepPlan"),
    "plan_fast_path": ("fastpath", "plan_fast_path"),
    "DRIVER_SLOTS": ("fastpath", "DRIVER_SLOTS"),
    "gpu_compatibility": ("from_meep", "gpu_compatibility"),
    "lift_simulation": ("from_meep", "lift_simulation"),
    "run_on_gpu": ("from_meep", "run_on_gpu"),
    "GpuRunResult": ("from_meep", "GpuRunResult"),
    "BatchedRun": ("batch", "BatchedRun"),
    "SolveServer": ("serve", "SolveServer"),
}

__all__ = sorted(_LAZY)


def __getattr__(name):
    if name in _LAZY:
        module, attr = _LAZY[name]
        value = getattr(importlib.import_module(f".{module}", __name__), attr)
        globals()[name] = value
        return value
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
