"""Batched form: many independent simulations advanced by the same fused launches.

A ``BatchedRun`` takes device-resident simulations of any sizes (each with its own absorber,
poles, sources and flux planes) and steps them together: per half-step ONE fused launch walks
a work list of (simulation, tile) entries covering every member, and one sampling launch
serves every member's flux planes. Uneven grids share the machine instead of each leaving
most of the 256 CUs idle, which is where the aggregate speed-up over one-at-a-time comes
from for grids of 32^3..112^3 cells.

``synthetic_batch`` builds the benchmark's batch: edges log-uniform over [32, 112], sphere
radius uniform in [0.1, 0.35] x edge, eps uniform in [1.5, 12], a quarter of them with a
Lorentz pole (eps_inf 2.0, sigma 1.5, f0 1.1, gamma 0.05) where the sphere lies outside the
absorber, 8-cell UPML, one flux plane x 32 frequencies each, from one seed.
"""

from __future__ import annotations

import ctypes
import math
import weakref
from typing import List, Optional, Sequence

import numpy

from . import _native
from ._native import check
from .dispersion import LORENTZIAN, Susceptibility
from .driver import FdtdDriver
from .workloads import FluxPlane, Workload, meep_frequency_span, outside_absorber, sphere


class BatchedRun:
    """Step several runs through one batch: frozen ``FdtdDriver`` runs, or anything holding a
    ``HipSimulation`` as ``.sim`` and its grid as ``.grid`` (a ``capsule.LiftedRun``)."""

    def __init__(self, drivers: Sequence[FdtdDriver]):
        self.drivers = list(drivers)
        if not self.drivers:
            raise ValueError("a batch needs at least one simulation")
        libs = {id(d.sim.lib) for d in self.drivers if d.sim is not None}
        if len(libs) > 1:
            raise ValueError("a batch steps simulations of one subnormal policy")
        self.lib = self.drivers[0].sim.lib if self.drivers[0].sim is not None else _native.load()
        # a run on complex storage brings both of its parts (``sims``), stepped side by side
        self.members = []
        for drv in self.drivers:
            if drv.sim is None:
                raise ValueError("freeze every driver before batching it")
            self.members.extend(getattr(drv, "sims", None) or [drv.sim])
        handles = (ctypes.c_void_p * len(self.members))()
        for q, sim in enumerate(self.members):
            handles[q] = sim.handle.value
        out = ctypes.c_void_p()
        check(self.lib.fdtd_batch_create(handles, len(self.members), ctypes.byref(out)),
              "fdtd_batch_create", self.lib)
        self.handle = out
        for sim in self.members:
            sim._batch = weakref.ref(self)

    @property
    def cells(self) -> int:
        return sum(d.grid.total_cells for d in self.drivers)

    def run(self, steps: int):
        check(self.lib.fdtd_batch_run(self.handle, int(steps)), "fdtd_batch_run", self.lib)

    def synchronize(self):
        check(self.lib.fdtd_batch_synchronize(self.handle), "fdtd_batch_synchronize", self.lib)

    def step_h(self):
        check(self.lib.fdtd_batch_step_h(self.handle), "fdtd_batch_step_h", self.lib)

    def step_d(self):
        check(self.lib.fdtd_batch_step_d(self.handle), "fdtd_batch_step_d", self.lib)

    def set_config(self, config):
        """The batch's own launch configuration (``engine.LaunchConfig``); None: follow the
        process defaults."""
        words = config.words() if config is not None else None
        check(self.lib.fdtd_batch_set_config(
            self.handle, words.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
            if words is not None else None), "fdtd_batch_set_config", self.lib)

    def config(self):
        from .engine import LaunchConfig

        out = (ctypes.c_int32 * LaunchConfig.WORDS)()
        check(self.lib.fdtd_batch_get_config(self.handle, out), "fdtd_batch_get_config", self.lib)
        return LaunchConfig.from_words(list(out))

    def graph_stats(self):
        """(graphs captured, replays after their capture) over the batch's life."""
        out = (ctypes.c_int64 * 2)()
        check(self.lib.fdtd_batch_graph_stats(self.handle, out), "fdtd_batch_graph_stats",
              self.lib)
        return int(out[0]), int(out[1])

    def flush(self):
        check(self.lib.fdtd_batch_flush(self.handle), "fdtd_batch_flush", self.lib)

    def close(self):
        """End the batch; members keep their state and step count and can be read or stepped."""
        if self.handle is not None and self.handle.value:
            self.lib.fdtd_batch_destroy(self.handle)
        self.handle = None
        for sim in getattr(self, "members", ()):
            sim._batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synthetic_sim(rng: numpy.random.Generator, edge_range=(32, 112), steps: int = 3000,
                  nfreq: int = 32, lorentz_fraction: float = 0.25, pml_cells: int = 8) -> Workload:
    """One member of the synthetic batch (see the module docstring)."""
    lo, hi = edge_range
    edge = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
    edge = max(edge, 2 * pml_cells + 8)
    radius = float(rng.uniform(0.10, 0.35)) * edge
    eps = float(rng.uniform(1.5, 12.0))
    dispersive = bool(rng.uniform() < lorentz_fraction)
    resolution = 10.0
    wl = sphere(edge=edge, resolution=resolution, pml_cells=pml_cells, radius_cells=radius,
                eps=eps, fcen=0.15, df=0.1, nfreq=nfreq, planes="one", steps=steps,
                source_x=-0.2 * edge / resolution,
                plane_x=0.2 * edge / resolution)
    if dispersive:  # the pole inside the sphere and outside the absorber
        inside = (wl.inv["Ez"] < 1.0) & outside_absorber(wl.grid, pml_cells)
        sig = numpy.where(inside, numpy.float32(1.5), numpy.float32(0.0))
        wl.poles.append((Susceptibility(1.1, 0.05, LORENTZIAN),
                         {c: sig.astype(numpy.float32) for c in ("Ex", "Ey", "Ez")}))
        for c in ("Ex", "Ey", "Ez"):  # eps_inf 2.0 inside the sphere
            wl.inv[c] = numpy.where(wl.inv[c] < 1.0, numpy.float32(0.5), wl.inv[c]).astype(
                numpy.float32)
    wl.name = f"batch_member_{edge}{'_lorentz' if dispersive else ''}"
    return wl


def synthetic_batch(count: int = 64, seed: int = 20261203, **kw) -> List[Workload]:
    rng = numpy.random.default_rng(seed)
    return [synthetic_sim(rng, **kw) for _ in range(count)]
