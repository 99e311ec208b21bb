"""Small workloads stepped on the HIP engine and on the oracle from identical inputs."""

from __future__ import annotations

import numpy

from meep_gpu import workloads
from meep_gpu.dft import FluxMonitor
from meep_gpu.driver import FdtdDriver
from meep_gpu.pml import normalize_faces
from meep_gpu.sources import GaussianSource, current_table, meep_volume_source
from oracle import workload as ow


def small(name: str):
    if name == "sphere":
        return workloads.sphere(edge=36, resolution=4.0, pml_cells=6, radius_cells=7.0,
                                eps=2.25, fcen=0.15, df=0.1, nfreq=8, planes="one",
                                steps=160, source_x=-2.0)
    if name == "slab":
        return workloads.slab(edge=40, resolution=8.0, pml_cells=6, thickness=6, nfreq=9,
                              steps=160)
    if name == "sphere_sources":
        # the sphere with a magnetic point source off the lattice and a finite electric sheet
        wl = small("sphere")
        grid = wl.grid
        for box in meep_volume_source(grid, "Hz", (0.33, -0.71, 0.52), (0.0, 0.0, 0.0)):
            wl.sources.append((box, GaussianSource(0.12, 0.08)))
        for box in meep_volume_source(grid, "Ex", (-0.4, 0.9, 0.1), (0.0, 1.3, 0.7)):
            wl.sources.append((box, GaussianSource(0.18, 0.1)))
        return wl
    if name == "sphere_integrated":
        # the sphere driven by an electric polarization (integrated) source as well
        wl = small("sphere")
        for box in meep_volume_source(wl.grid, "Ey", (0.41, 0.23, -0.37), (0.0, 0.0, 0.0),
                                      integrated=True):
            wl.sources.append((box, GaussianSource(0.15, 0.1)))
        return wl
    if name == "sphere_box":
        return workloads.sphere(edge=40, resolution=5.0, pml_cells=6, radius_cells=8.0,
                                nfreq=6, planes="box8", steps=120)
    raise KeyError(name)


def oracle_run(wl, steps: int, reference_precision: bool = True, dft_monitors=()):
    """The oracle's run of ``wl``: its flux planes, then one oracle region per DFT field
    monitor in ``dft_monitors`` (after them in the returned list)."""
    grid = wl.grid
    faces = normalize_faces(wl.pml_faces, grid) if wl.pml_faces else ((0, 0),) * 3
    deposits = []
    for b, fn in wl.sources:
        if b.integrated:  # the dipole, evaluated at the oracle's own time
            table = fn.dipole
        else:
            table = current_table(fn, grid.dt, steps + 1, b.magnetic)
        deposits.append((b.component, b.lo, b.cnt, b.amp_re, b.amp_im, table, b.integrated))
    planes = []
    for p in wl.planes:
        mon = FluxMonitor(p.normal, p.lo, p.cnt, p.frequencies, grid.dx, grid.dt)
        planes.append((p.normal, p.lo, p.cnt, mon.omegas, mon.scale))
    for m in dft_monitors:
        planes.append((-1, m.lo, m.cnt, m.omegas, m.scale, m.decimation, m.components,
                       m.yee_grid))
    poles = [(t.frequency, t.gamma, t.is_drude, s) for t, s in wl.poles]
    sim, oplanes = ow.build(grid.shape, grid.dt, grid.dx, faces, wl.inv, poles, deposits, planes,
                            grid.invariant)
    ow.run(sim, oplanes, steps, reference_precision)
    return sim, oplanes


def hip_run(wl, steps: int):
    drv = FdtdDriver.from_workload(wl)
    drv.freeze(planned_steps=steps + 1)
    drv.run(steps=steps)
    drv.synchronize()
    return drv
