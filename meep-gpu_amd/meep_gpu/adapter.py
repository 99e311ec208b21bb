"""The reference side of the seam: a live reference ``FdtdDriver`` stepped by this engine.

The reference builds one plan per configuration freeze with ``plan_fast_path(fields, pml,
grid, sources, chunk_layout)`` (reference fastpath.py:6986) and asks it once per sub-step,
in MEEP's order, ``fast.dispatch(slot, fields)`` (driver.py:3505-3596). A maintainer adds this
engine as the candidate of a ROCm host at that ladder's hardware rung with

    plan = adapter.hip_plan(driver)          # None, with adapter.last_refusal(), if refused

and the driver's consults then reach the GPU. What crosses the boundary, and how:

* **State in** (``upload_reference_fields``, at the plan): the live driver is read through
  the same object reader the lifted capsules go through (``capsule.lift_state``), by
  attribute access instead of a decoded file: the grid and its folds, the UPML vectors and
  the chunk division's membership lines, inverse epsilon, conductivity, poles (MEEP's ADE
  words), source point tables, every stored field, auxiliary and polarization volume, the
  step count and the one-time H start. ``capsule.LiftedRun`` pastes every word into a
  ``HipSimulation``. What the engine does not step is refused by name before anything is
  allocated, and the reference keeps its array path.
* **Currents**: each source's ``current(t, dt)`` (reference sources.py:1844-1936, MEEP's
  ``src_time::current``) evaluated on the host at the times the driver asks for them
  (magnetic at ``n*dt``, electric at ``n*dt + dt/2``; driver.py:3536, :3565), tabulated for
  the steps ahead and extended on the device when the run goes past them.
* **The consults** (``HipPlanAdapter.dispatch``): the reference's contract
  (fastpath.py:4631-4700) -- ``fields`` is an identity guard on the reference's own
  ``Fields``; ``True`` means the whole sub-step ran, ``False`` that nothing did; the engine's
  two fused launches lead at ``step_B`` and ``step_D`` and their trailing consults (the fills
  and their far passes, ``update_H``, ``update_E``, ``update_P``) answer with them for one
  step; a launch error raises.
* **State out** (``download_reference_fields``): after ``update_P`` the step's words are
  written back into the reference's arrays (``sync=True``, the compatibility mode): the
  driver's host-side monitors read them, and its host-side writes between consults (the
  source injection and the wall clear, which the fused launches already did on the device)
  land on a copy the next download replaces. A re-plan after ``invalidate_fast_path``
  (driver.py:3321) uploads what the last download left. With ``sync=False`` nothing comes
  back until ``download_reference_fields`` is called.
"""

from __future__ import annotations

import numbers
from typing import Any, Dict, List, Optional

import numpy

from . import capsule
from .fastpath import HipStepPlan

_LAST_REFUSAL: List[str] = []
_PRIMITIVE = (str, bytes, bool, numbers.Number, numpy.ndarray, numpy.generic, type(None))


def _wrap(value):
    if isinstance(value, _PRIMITIVE):
        return value
    if isinstance(value, dict):
        return {k: _wrap(v) for k, v in value.items()}
    if isinstance(value, (list, tuple)):
        return type(value)(_wrap(v) for v in value) if not hasattr(value, "_fields") \
            else type(value)(*(_wrap(v) for v in value))
    return LiveObject(value)


class LiveObject:
    """A live reference object read the way ``capsule.Node`` reads a decoded one: its class
    path ``module:qualname`` and its attributes by name (the capsule writer records exactly
    an object's attributes, reference capsule.py:1-60)."""

    def __init__(self, obj: Any):
        self.obj = obj
        kind = type(obj)
        self.cls = f"{kind.__module__}:{kind.__qualname__}"

    @property
    def attrs(self) -> Dict[str, Any]:
        return {k: _wrap(v) for k, v in vars(self.obj).items()}

    def __getitem__(self, name: str):
        try:
            return _wrap(getattr(self.obj, name))
        except AttributeError:
            raise KeyError(name) from None

    def get(self, name: str, default=None):
        return _wrap(getattr(self.obj, name, default))


def last_refusal() -> List[str]:
    """Why the last ``hip_plan`` returned None (empty when it returned a plan)."""
    return list(_LAST_REFUSAL)


def lift_live(driver: Any) -> capsule.LiftedState:
    """The engine's view of a live reference driver (refuses, by name, what it does not step)."""
    for src in getattr(driver, "_sources", ()) or ():
        if getattr(src, "is_integrated", False):
            raise capsule.CapsuleRefusal("a polarization (integrated) source crossing the "
                                         "seam is not built")
    return capsule.lift_state(LiveObject(driver))


def live_currents(driver: Any, state: capsule.LiftedState, steps: int) -> List[numpy.ndarray]:
    """Each source's current for steps ``0 .. steps-1`` after the lifted step, as (steps, 2)
    float64 (re, im), from the driver's own source objects."""
    tables = []
    for src, lifted in zip(driver._sources, state.sources):
        offset = 0.0 if lifted.magnetic else 0.5 * state.dt
        table = numpy.zeros((steps, 2), dtype=numpy.float64)
        for n in range(steps):
            value = complex(src.current((state.step + n) * state.dt + offset, state.dt))
            table[n] = (value.real, value.imag)
        tables.append(table)
    return tables


def hip_grid_from(grid: Any):
    """This engine's ``Grid`` for a reference ``Grid`` (its folds and metallic axes included)."""
    node = LiveObject(grid)
    mirrors = tuple((a, int(m["phase"])) for a, m in enumerate(node["mirrors_by_axis"])
                    if m is not None)
    metallic = tuple(pair[0] == "metallic" for pair in grid.boundaries)
    from .grid import Grid

    size = [0.0 if grid.is_invariant(a) else grid.shape_full[a] / grid.resolution
            for a in range(3)]
    return Grid(resolution=grid.resolution, cell_size=tuple(size), courant=grid.courant,
                dimensions=grid.dimensions, symmetry=mirrors, metallic=metallic)


#: the stored volumes that move back, by the reference's Fields attribute names
_FIELD_NAMES = tuple(p + a for p in ("B", "D", "E", "H", "fu_B", "fu_D", "f_w_E", "f_w_H",
                                     "f_cond_D", "f_cond_B") for a in "xyz")


def upload_reference_fields(run: capsule.LiftedRun, fields: Any) -> None:
    """Paste the reference ``Fields``' stored volumes into the engine (after a host-side
    write between plans): every field, auxiliary and polarization volume it holds."""
    sim = run.sim
    for name in _FIELD_NAMES:
        value = getattr(fields, name, None)
        if value is not None and run.engine_has(name):
            sim.upload(name, numpy.asarray(value, dtype=numpy.float32))
    for k, state in enumerate(getattr(fields, "polarizations", ()) or ()):
        for c, value in state.P.items():
            sim.upload(f"P[{k}].{c}", numpy.asarray(value, dtype=numpy.float32))
        for c, value in state.P_prev.items():
            sim.upload(f"P_prev[{k}].{c}", numpy.asarray(value, dtype=numpy.float32))


def download_reference_fields(run: capsule.LiftedRun, fields: Any) -> None:
    """Write the engine's current words into the reference ``Fields``' arrays, in place."""
    sim = run.sim
    for name in _FIELD_NAMES:
        target = getattr(fields, name, None)
        if target is None or not run.engine_has(name):
            continue
        target[...] = sim.download(name)
    for k, state in enumerate(getattr(fields, "polarizations", ()) or ()):
        for c in state.P:
            state.P[c][...] = sim.download(f"P[{k}].{c}")
        for c in state.P_prev:
            state.P_prev[c][...] = sim.download(f"P_prev[{k}].{c}")
    if hasattr(fields, "h_separated") and run.sim.step_count > run.state.step:
        fields.h_separated = True  # update_H has run (MEEP's one-time H start)


class HipPlanAdapter:
    """The engine's plan behind the reference's ``dispatch(slot, fields)`` contract."""

    def __init__(self, driver: Any, run: capsule.LiftedRun, planned: int, sync: bool = True):
        self.driver = driver
        self.fields = driver.fields
        self.run = run
        self.plan = HipStepPlan(run.sim)
        self.sync = bool(sync)
        self.planned = int(planned)  # source steps tabulated after the lifted step
        self._start = run.state.step

    @property
    def slots(self):
        return self.plan.slots

    def _extend(self, steps_after_start: int) -> None:
        if steps_after_start <= self.planned:
            return
        planned = max(steps_after_start, 2 * self.planned)
        index = 0
        for table, src in zip(live_currents(self.driver, self.run.state, planned),
                              self.run.state.sources):
            for _ in range(len(src.amps)):  # one engine source per lifted point
                self.run.sim.set_source_currents(index, table)
                index += 1
        self.planned = planned

    def dispatch(self, slot: str, fields: Any) -> bool:
        if fields is not self.fields:
            return False
        if slot == "step_B":
            self._extend(self.run.sim.step_count - self._start + 1)
        answered = self.plan.dispatch(slot, self.run.sim)
        if answered and slot == "update_P" and self.sync:
            download_reference_fields(self.run, self.fields)
        return answered

    def close(self) -> None:
        if self.run is not None:
            if self.sync:
                download_reference_fields(self.run, self.fields)
            self.run.close()
            self.run = None


def hip_plan(driver: Any, planned_steps: int = 1024, sync: bool = True,
             subnormals: str = "keep") -> Optional[HipPlanAdapter]:
    """The engine's plan for a live reference driver, or None (``last_refusal()`` says why)."""
    _LAST_REFUSAL.clear()
    try:
        state = lift_live(driver)
    except capsule.CapsuleRefusal as err:
        _LAST_REFUSAL.append(str(err))
        return None
    run = capsule.LiftedRun(state, live_currents(driver, state, int(planned_steps)),
                            subnormals=subnormals)
    return HipPlanAdapter(driver, run, int(planned_steps), sync=sync)
