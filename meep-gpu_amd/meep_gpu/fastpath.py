"""The dispatch seam: the driver's per-sub-step consult, answered by the HIP engine.

The reference driver asks one plan object, once per sub-step and in MEEP's order, whether it
will run that sub-step (``driver.py:3505-3596``): ``if fast is None or not
fast.dispatch(slot, fields): <array call>``. The plan is built once per configuration freeze
by ``plan_fast_path(fields, pml, grid, sources, chunk_layout)`` (``fastpath.py:6986``) and the
contract of ``dispatch`` (``fastpath.py:4631-4700``) is kept here:

* the slots are the driver's (``DRIVER_SLOTS``, ``fastpath.py:575``), and the two far-fill
  passes are consulted by their own pass names (``FAR_FILL_PASSES``);
* ``True`` means the whole sub-step ran, ``False`` means nothing was written, never half;
* ``fields`` is an identity guard: a plan answers only for the state it was built on;
* a fused pair's leading consult runs the pair, its trailing consults answer ``True`` only
  when the lead ran in the same step (a commitment lasts one step: a consult whose position
  does not advance opens a new step and clears it);
* an error inside a launch raises (``EngineError``); nothing falls back mid-step.

Here the fields are device-resident (a ``HipSimulation``) and the engine fuses each half-step
into one launch, so the plan has two fused units:

* ``step_B`` leads the magnetic pair: step_B, the magnetic deposit and update_H, in one
  launch; ``fill_B`` (and its far pass) and ``update_H`` are its trailing consults;
* ``step_D`` leads the electric pair: step_D, the electric deposit at t+dt/2, update_E and
  update_P, in one launch, then the counter advance and the flux sampling; ``fill_D`` (and its
  far pass), ``update_E`` and ``update_P`` are its trailing consults.

The symmetry fills of a mirror-folded grid run inside the pairs' launches (the engine's fold
kernel, after the sub-step's sources and before update_H / update_E, in the reference's
order), so the fill consults and their far passes answer with their pair.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Dict, List, Mapping, Optional, Tuple

DRIVER_SLOTS: Tuple[str, ...] = (
    "step_B", "fill_B", "update_H", "step_D", "fill_D", "update_E", "update_P",
)

FAR_FILL_PASSES: Mapping[str, str] = {
    "fill_B": "fill_folded_far_ghosts_B",
    "fill_D": "fill_folded_far_ghosts_D",
}
FAR_FILL_OWNERS: Mapping[str, str] = {v: k for k, v in FAR_FILL_PASSES.items()}

#: The fused units: leading slot first, then the slots its one launch stands in for.
FUSED_PAIRS: Tuple[Tuple[str, ...], ...] = (
    ("step_B", "fill_B", "update_H"),
    ("step_D", "fill_D", "update_E", "update_P"),
)

_PAIR_OF: Dict[str, Tuple[str, ...]] = {slot: pair for pair in FUSED_PAIRS for slot in pair}

_LAST_REPORT: Dict[str, Any] = {}


class FastPathError(RuntimeError):
    """A consult the plan cannot answer in a step it already began (no array path exists)."""


@dataclass
class HipStepPlan:
    """The engine's answer to the driver's consults for one frozen configuration."""
    fields: Any                                 # the HipSimulation the plan was built on
    slots: Tuple[str, ...] = DRIVER_SLOTS
    launches: List[int] = field(default_factory=lambda: [0])
    _committed: set = field(default_factory=set, repr=False)
    _last_position: List[int] = field(default_factory=lambda: [-1], repr=False)

    @property
    def replaces(self) -> Tuple[str, ...]:
        return self.slots

    @property
    def dispatches_a_kernel(self) -> bool:
        return bool(self.slots)

    def dispatch(self, slot: str, fields: Any) -> bool:
        """Run ``slot`` (or its fused pair) and answer True, or answer False having run nothing."""
        if fields is not self.fields:
            return False
        owner = FAR_FILL_OWNERS.get(slot)
        if owner is not None:
            # the far half of a fill: part of the pair's launch, and no new consult position
            return owner in self.slots and _PAIR_OF[owner][0] in self._committed
        if slot not in self.slots:
            return False
        position = self.slots.index(slot)
        if position <= self._last_position[0]:
            self._committed.clear()  # a position that does not advance opens a new step
        self._last_position[0] = position
        pair = _PAIR_OF[slot]
        if slot == pair[0]:
            if slot == "step_B":
                self.fields.step_h()
            else:
                self.fields.step_d()
            self.launches[0] += 1
            self._committed.add(slot)
            return True
        return pair[0] in self._committed


def plan_fast_path(fields: Any, pml: Any = None, grid: Any = None, sources: Any = None,
                   chunk_layout: Any = None) -> Optional[HipStepPlan]:
    """Build the plan for a frozen configuration, or return None with the named reason.

    The arguments are the reference's (``fastpath.py:6986``). ``fields`` is the device-resident
    state (``HipSimulation``, or ``ComplexSimulation``'s two linked parts); the engine already holds the absorber, materials, poles and
    sources uploaded at the freeze, so ``pml``, ``grid`` and ``sources`` are checked against
    it rather than re-read. ``chunk_layout`` must be the one-rank division (or None).
    """
    from .engine import ComplexSimulation, HipSimulation

    reasons = []
    if not isinstance(fields, (HipSimulation, ComplexSimulation)):
        reasons.append("fields are not device-resident (HipSimulation or ComplexSimulation)")
    elif fields.handle is None:
        reasons.append("the simulation is closed")
    else:
        if grid is not None and tuple(getattr(grid, "shape", fields.grid.shape)) != tuple(
                fields.grid.shape):
            reasons.append("grid shape differs from the resident fields")
        if pml is not None and (fields.pml is None) != (not getattr(pml, "is_active", True)):
            reasons.append("absorber differs from the resident fields")
    if chunk_layout is not None and getattr(chunk_layout, "num_chunks", 1) != 1:
        reasons.append("only the one-rank chunk division is stepped")
    _LAST_REPORT.clear()
    _LAST_REPORT.update({"table": "hip", "dispatched": not reasons, "reasons": reasons,
                         "fused": [list(p) for p in FUSED_PAIRS]})
    if reasons:
        return None
    return HipStepPlan(fields)


def last_dispatch_report() -> Dict[str, Any]:
    """The record of the last ``plan_fast_path`` call (``fastpath.last_dispatch_report``)."""
    return dict(_LAST_REPORT)
