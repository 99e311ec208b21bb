"""Run a reference fingerprint case on the HIP engine and digest it as the reference does.

The case's inputs (seeded state words, material volumes, pole sigmas) come from
``oracle.cases.build_state``; the layer table is the product's own (``meep_gpu.pml``), so the
digest pins the product's setup and kernels together.
"""

from __future__ import annotations

import math
from typing import Dict, Tuple

import numpy

from meep_gpu.chunks import IYEE, ChunkDivision
from meep_gpu.dispersion import DRUDE, LORENTZIAN, Susceptibility
from meep_gpu.engine import HipSimulation
from meep_gpu.grid import Grid
from meep_gpu.pml import PML
from meep_gpu.sources import meep_volume_source
from oracle import cases, fixture


def hip_simulation(state: cases.CaseState) -> HipSimulation:
    case = cases.CASE_BY_NAME[state.name]
    grid = Grid(resolution=fixture.RESOLUTION, cell_size=case.cell, courant=fixture.COURANT,
                dimensions=case.dimensions, symmetry=case.symmetry, metallic=case.metallic)
    assert grid.shape == state.shape, (grid.shape, state.shape)
    pml = (PML(grid, case.pml, ln_reflection=fixture.PML_LN_REFLECTION,
               wraps=tuple(not m for m in case.metallic)) if case.pml else None)
    sim = HipSimulation(grid, pml)
    for c, vol in state.inv.items():
        sim.set_inverse_epsilon(c, vol)
    for row, partners in state.offdiag.items():
        for partner, vol in partners.items():
            sim.set_offdiagonal(row, partner, vol)
    for pole in state.poles:
        term = Susceptibility(pole["frequency"], pole["gamma"],
                              DRUDE if pole["drude"] else LORENTZIAN)
        sim.add_pole(term, pole["sigma"])
    division = ChunkDivision.one_rank(grid, pml.faces if pml else ((0, 0),) * 3,
                                      tuple(not m for m in case.metallic))
    for comp, vol in state.conductivity.items():
        sim.set_conductivity(comp, vol, division.conductive_mask(vol, IYEE[comp]))
    for axis, metal in enumerate(case.metallic):
        if metal and grid.folds[axis] is None:  # a folded metallic axis is the fold's kind
            sim.set_boundary(axis, "metallic")
    for spec, placed in zip(case.sources, state.sources):
        # the product's placement; the stored word of step n serves a current's two slots, and
        # an integrated source's dipole at (n+1)*dt is the word of step n+1
        words = placed["table"][1:] if spec.integrated else placed["table"]
        current = numpy.array([[w.real, w.imag] for w in words], dtype=numpy.float64)
        for box in meep_volume_source(grid, spec.component, spec.center, spec.size,
                                      metallic=case.metallic, integrated=spec.integrated):
            sim.add_source(box.magnetic, "xyz".index(box.component[1]), box.lo, box.cnt,
                           box.amp_re, box.amp_im, current, grid.dt, box.integrated)
    for n, vol in state.volumes.items():
        sim.upload(n, vol)
    if not state.stores_E:
        sim.prime_e()
    return sim


def hip_volumes(sim: HipSimulation, state: cases.CaseState) -> Dict[str, numpy.ndarray]:
    return {n: sim.download(n) for n in state.volumes}


def hip_case_digest(state: cases.CaseState, steps: int = fixture.STEPS
                    ) -> Tuple[str, Dict[str, numpy.ndarray]]:
    """(hex digest over steps 0..steps, the final stored volumes) for one case on the GPU."""
    sim = hip_simulation(state)
    digest = cases.new_digest(state.name)

    def record():
        vols = hip_volumes(sim, state)
        get_e = None
        if not state.stores_E:
            electric = {c: sim.download(c) for c in ("Ex", "Ey", "Ez")}
            get_e = electric.__getitem__
        cases.digest_update(digest, vols, state.stores_E, get_e)
        return vols

    vols = record()
    for _ in range(steps):
        sim.step_h()
        sim.step_d()
        vols = record()
    sim.close()
    return digest.hexdigest(), vols
