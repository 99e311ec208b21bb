"""Sphere and Lorentz-slab workloads with sources and flux planes: HIP engine vs the oracle (GPU).

Fields are compared word for word. Spectra: the engine transforms in float64 (blocked over 128
steps), the oracle in float64 per step (the sums differ only in association, so 1e-9 relative
to the spectrum's peak is the stated tolerance), and the reference's own complex64
accumulation lies within 1e-4 of both (the stated flux tolerance).
"""

import numpy
import pytest

from tests import workload_runs as wr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["sphere", "slab", "sphere_box", "sphere_sources",
                                  "sphere_integrated"])
def test_workload_fields_and_spectra_match_the_oracle(name):
    print(f"workload {name}", flush=True)
    wl = wr.small(name)
    steps = wl.steps
    drv = wr.hip_run(wl, steps)
    sim, planes = wr.oracle_run(wl, steps)
    for comp in ("Ex", "Ey", "Ez", "Dx", "Dy", "Dz", "Bx", "By", "Bz", "Hx", "Hy", "Hz"):
        got = drv.sim.download(comp)
        want = sim.stored_volumes()[comp]
        assert got.view(numpy.uint32).tobytes() == want.view(numpy.uint32).tobytes(), comp
    for mon, plane in zip(drv.flux_monitors, planes):
        got = mon.get_flux_spectrum()
        ref64 = plane.flux(wl.grid.dx, "c128")
        ref32 = plane.flux(wl.grid.dx, "c64")
        peak = numpy.max(numpy.abs(ref64))
        assert peak > 0
        assert numpy.max(numpy.abs(got - ref64)) <= 1e-9 * peak
        assert numpy.max(numpy.abs(ref32 - ref64)) <= 1e-4 * peak
    drv.close()


def test_compact_inverse_epsilon_and_pole_rows_are_in_use():
    """The workloads above run with inverse epsilon as 1-byte table indices (the sphere has 78
    distinct words per component) and, for the slab, pole work only over the slab's x range: their
    word-for-word agreement with the oracle covers both forms."""
    from meep_gpu.driver import FdtdDriver
    from tests.workload_runs import small

    drv = FdtdDriver.from_workload(small("sphere"))
    drv.freeze(planned_steps=2)
    assert drv.sim.inverse_epsilon_bytes() == (1, 1, 1)
    drv.close()
    drv = FdtdDriver.from_workload(small("slab"))
    drv.freeze(planned_steps=2)
    assert 0.0 < drv.sim.pole_row_fraction() < 0.5
    drv.close()


def test_matrix_core_and_vector_flux_transforms_agree():
    """The float64 transform on the matrix cores (MFMA 16x16x4 f64, with 1, 2 and 4 cell tiles
    per wave) and as vector FMAs: the same products summed in another order, so the spectra
    agree to float64 rounding. The sphere box's planes (96 and 160 cells at this size) end in
    partial tiles, and 6 frequencies leave most of a wave's frequency tiles empty."""
    from meep_gpu.engine import set_flux_transform
    from tests.workload_runs import hip_run, small

    spectra = []
    for mfma, tiles in ((True, 1), (True, 2), (True, 4), (False, 2)):
        set_flux_transform(mfma, tiles)
        try:
            print(f"flux transform mfma={mfma} cell tiles {tiles}", flush=True)
            drv = hip_run(small("sphere_box"), 150)
            spectra.append([m.get_flux_spectrum() for m in drv.flux_monitors])
            drv.close()
        finally:
            set_flux_transform(True)
    vector = spectra[-1]
    for run in spectra[:-1]:
        for a, b in zip(run, vector):
            scale = numpy.abs(b).max()
            assert scale > 0
            assert numpy.abs(a - b).max() <= 1e-12 * scale


def test_the_work_split_covers_the_grid():
    """The box is whole rows whose y and z lie on no layer line (so only x-layer forms hold in
    it), the shell every other row: together they are every run of real cells (the parity tests
    above check the words)."""
    from meep_gpu.driver import FdtdDriver
    from tests.workload_runs import small

    drv = FdtdDriver.from_workload(small("sphere"))
    drv.freeze(planned_steps=2)
    nx, ny, nz = drv.grid.shape
    pitch = -(-nx // 4) * 4
    pml = drv.sim.pml
    for vec in (1, 2, 4):
        split = drv.sim.work_split(vec)
        x0, x1, y0, y1, z0, z1 = split["box"]
        assert (x0, x1) == (0, pitch) and y0 < y1 and z0 < z1
        for axis, (lo, hi) in ((1, (y0, y1)), (2, (z0, z1))):
            for half in (0, 1):
                assert not numpy.any(pml.members[(half, axis)][lo:hi])
                # the box is the longest such range: the rows just past it are on a line
                assert pml.members[(half, axis)][lo - 1] or pml.members[(1 - half, axis)][lo - 1]
        inside_rows = (y1 - y0) * (z1 - z0)
        assert split["interior_runs"] == (pitch // vec) * inside_rows
        assert split["shell_runs"] == (ny * nz - inside_rows) * -(-nx // vec)
    drv.close()
