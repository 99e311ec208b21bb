"""The user entry and the result surface: ``run_on_gpu`` / ``GpuRunResult`` (reference
``from_meep.py:454-990``, ``:9532``).

``GpuRunResult`` is the object a run hands back: the stepped driver plus what the run cost,
with MEEP's read-out names (``get_flux_spectrum`` is ``mp.get_fluxes``; ``get_flux_data`` the
opaque transforms for the normalization idiom; ``get_dft_array`` a DFT field monitor's reduced
block; ``get_array`` one field component; ``close``).
``run_driver`` steps an ``FdtdDriver`` built through the driver API (materials, poles,
absorber, sources, monitors) and returns one; its ``flux_data`` / ``minus_flux_data`` load an
earlier run's transforms first (MEEP's ``load_flux_data`` / ``load_minus_flux_data``).

Lifting an ``mp.Simulation`` (``lift_simulation``, ``gpu_compatibility``, ``run_on_gpu`` on a
MEEP object) needs MEEP itself: MEEP computes the sub-pixel chi1inv, the chunk division and
the source currents the lift reads. MEEP is not part of this package's dependencies; without
it these entry points raise ``ImportError`` naming it, and with it they refuse by name
(``MeepSimulationNotLiftable``) until the lift adapter is built.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import Any, Optional, Sequence

import numpy

from .dft import FluxDftData, FluxMonitor
from .driver import FdtdDriver

MEEP_VERSION = "1.35.0"  # the MEEP release the reference certifies its lift on (README.md:468)


def _import_meep():
    try:
        import meep  # noqa: F401
    except ImportError as exc:
        raise ImportError(
            f"lifting an mp.Simulation needs MEEP (pymeep {MEEP_VERSION}), which is not "
            f"installed; build the run through meep_gpu.FdtdDriver instead") from exc
    return meep


class MeepSimulationNotLiftable(ValueError):
    """An ``mp.Simulation`` this package will not run, with every reason."""

    def __init__(self, reasons: Sequence[str]):
        self.reasons = tuple(reasons)
        listed = "\n".join(f"  - {r}" for r in self.reasons)
        super().__init__(f"This mp.Simulation cannot be lifted onto the GPU stepper "
                         f"({len(self.reasons)} reason(s)):\n{listed}\nNothing was run.")


@dataclass(frozen=True)
class GpuCompatibility:
    """A verdict plus every blocking reason (reference ``from_meep.py:248``)."""
    supported: bool
    reasons: tuple

    def __bool__(self) -> bool:
        return self.supported

    def __str__(self) -> str:
        if self.supported:
            return "GpuCompatibility(supported)"
        listed = "\n".join(f"  - {r}" for r in self.reasons)
        return f"GpuCompatibility(unsupported), {len(self.reasons)} reason(s):\n{listed}"


_LIFT_PENDING = ("the MEEP lift adapter (sub-pixel chi1inv, chunk division and source currents "
                 "read from MEEP) is not part of this build; describe the run through "
                 "meep_gpu.FdtdDriver")


def gpu_compatibility(sim) -> GpuCompatibility:
    """Every reason ``sim`` cannot be lifted (needs MEEP)."""
    _import_meep()
    return GpuCompatibility(False, (_LIFT_PENDING,))


def lift_simulation(sim, **kwargs) -> FdtdDriver:
    _import_meep()
    raise MeepSimulationNotLiftable((_LIFT_PENDING,))


@dataclass
class GpuRunResult:
    """The outcome of a run: the stepped driver (not closed) and what the run cost."""
    driver: FdtdDriver
    steps: int
    meep_time: float
    wall_time_s: float

    @property
    def monitors(self) -> Sequence[FluxMonitor]:
        return tuple(self.driver.flux_monitors)

    def monitor_for(self, monitor) -> FluxMonitor:
        if isinstance(monitor, int):
            return self.driver.flux_monitors[monitor]
        if monitor in self.driver.flux_monitors:
            return monitor
        raise KeyError("not a flux monitor of this run")

    def get_flux_spectrum(self, monitor) -> numpy.ndarray:
        """The flux spectrum of a monitor (``mp.get_fluxes``), float64 per frequency."""
        return self.monitor_for(monitor).get_flux_spectrum()

    def get_flux_data(self, monitor) -> FluxDftData:
        """The monitor's accumulated transforms, MEEP's ``sim.get_flux_data(flux)`` (reference
        from_meep.py:575): opaque, for ``run_driver(..., minus_flux_data=[(monitor, data)])`` on
        a later run carrying the same monitor -- the two-run normalization idiom."""
        return self.monitor_for(monitor).get_dft_data()

    def get_dft_array(self, monitor, component: str, freq_index: int = 0) -> numpy.ndarray:
        """MEEP's ``sim.get_dft_array(dft_fields, component, freq_index)`` for a monitor of
        ``driver.add_dft_monitor`` (reference from_meep.py:886): the region's transform with
        its flat axes reduced by their interpolation weights and its invariant axes dropped."""
        if isinstance(monitor, int):
            monitor = self.driver.dft_monitors[monitor]
        if monitor not in self.driver.dft_monitors:
            raise KeyError("not a DFT field monitor of this run")
        return monitor.get_dft_array(component, freq_index)

    def get_array(self, component: str, center=None, size=None) -> numpy.ndarray:
        """One field component as MEEP's ``sim.get_array`` returns it, word for word (reference
        from_meep.py:956-980; ``readback.meep_array``): MEEP's centred lattice, its duplicate
        plane and its rank, over the whole cell or the ``center``/``size`` box (with MEEP's
        interpolation weights and collapse on an empty direction)."""
        return self.driver.get_meep_array(component, center=center, size=size)

    def get_field(self, component: str) -> numpy.ndarray:
        """One stored field component, float32 C-order (nx, ny, nz): the driver's own layout
        (reference from_meep.py:982)."""
        return self.driver.get_array(component)

    def close(self) -> None:
        self.driver.close()


def run_driver(driver: FdtdDriver, until: Optional[float] = None,
               steps: Optional[int] = None, flux_data: Sequence = (),
               minus_flux_data: Sequence = ()) -> GpuRunResult:
    """Step ``driver`` for ``steps`` steps or until simulation time ``until`` (MEEP's
    ``sim.run(until=...)``: ceil(until / dt) steps) and return the result object.

    ``flux_data`` / ``minus_flux_data``: ``(monitor, data)`` pairs loaded into the monitors
    before the first step, MEEP's ``sim.load_flux_data`` / ``sim.load_minus_flux_data`` (data
    from an earlier run's ``get_flux_data``; reference from_meep.py:575-620)."""
    if (until is None) == (steps is None):
        raise ValueError("give exactly one of until or steps")
    if until is not None:
        steps = int(math.ceil(until / driver.grid.dt - 1e-9))
    if flux_data or minus_flux_data:
        driver.freeze(max(int(steps) + driver.step_count, driver.max_steps))
        for monitor, data in flux_data:
            monitor.load_dft_data(data)
        for monitor, data in minus_flux_data:
            monitor.load_minus_dft_data(data)
    t0 = time.perf_counter()
    driver.run(steps=int(steps))
    driver.synchronize()
    wall = time.perf_counter() - t0
    return GpuRunResult(driver, int(steps), driver.time, wall)


def run_on_gpu(sim, until: Optional[float] = None, **kwargs) -> GpuRunResult:
    """MEEP's ``sim.run(until=...)`` with the loop on the GPU; an ``FdtdDriver`` runs directly,
    an ``mp.Simulation`` is lifted first (needs MEEP)."""
    if isinstance(sim, FdtdDriver):
        return run_driver(sim, until=until, steps=kwargs.get("steps"),
                          flux_data=kwargs.get("flux_data", ()),
                          minus_flux_data=kwargs.get("minus_flux_data", ()))
    driver = lift_simulation(sim, **kwargs)
    return run_driver(driver, until=until)
