"""The five benchmark configurations at their own sizes: HIP engine vs the oracle (GPU).

* sphere_80: 80^3 cells (512,000), 8-cell UPML, 1 flux plane x 50 frequencies, the whole
  2000-step run.
* sphere_192: 192^3 cells (7,077,888), 8 flux planes x 100 frequencies, 256 steps: two ring
  periods, both replayed from a captured graph with their float64 transform passes inside.
* slab_160: 160^3 cells (4,096,000), 10-cell UPML, the Lorentz slab, 2 planes x 200
  frequencies, 256 steps.
* the batch: all 64 members of seed 20261203 (edges 32-112, a quarter Lorentz-dispersive) in
  one batch, 160 steps (one graph period and a partial ring).
* the served stream: the first 8 requests of the seed-20261203 Poisson stream (edges 32-160,
  30 % Lorentz, 1-8 planes x 16-100 frequencies) with their steps cut to 128-287, admitted
  while others run and batched by the continuous-batching server.

Every stored field word equals the oracle's (E, D, B and H, whose words the reference stores);
every spectrum is within 1e-9 of the float64 oracle's peak (the same float64 products summed
in another order), and for the record the reference's complex64 accumulation lies within 1e-4
of the float64 one where the oracle keeps it.
"""

import hashlib

import numpy
import pytest

from tests import workload_runs as wr

pytestmark = pytest.mark.gpu

COMPONENTS = ("Ex", "Ey", "Ez", "Dx", "Dy", "Dz", "Bx", "By", "Bz", "Hx", "Hy", "Hz")


def _same_fields(sim, osim, label):
    want = osim.stored_volumes()
    for comp in COMPONENTS:
        got = sim.download(comp)
        assert got.view(numpy.uint32).tobytes() == want[comp].view(numpy.uint32).tobytes(), \
            f"{label}: {comp}"


def _spectra_match(monitors, oplanes, dx, label, reference_precision=False):
    assert len(monitors) == len(oplanes)
    for q, (mon, plane) in enumerate(zip(monitors, oplanes)):
        got = mon.get_flux_spectrum() if hasattr(mon, "get_flux_spectrum") else mon
        ref64 = plane.flux(dx, "c128")
        peak = numpy.max(numpy.abs(ref64))
        assert peak > 0, f"{label}: plane {q} saw no field"
        err = numpy.max(numpy.abs(got - ref64)) / peak
        assert err <= 1e-9, f"{label}: plane {q} error {err:.3e}"
        if reference_precision:
            ref32 = plane.flux(dx, "c64")
            assert numpy.max(numpy.abs(ref32 - ref64)) <= 1e-4 * peak, f"{label}: plane {q} c64"


@pytest.mark.parametrize("name,steps", [("sphere80", 2000), ("sphere192", 256),
                                        ("slab160", 256)])
def test_config_at_full_size_matches_the_oracle(name, steps):
    from meep_gpu import workloads

    wl = workloads.config(name)
    print(f"config {name} {wl.grid.shape} {steps} steps: engine", flush=True)
    drv = wr.hip_run(wl, steps)
    print(f"config {name}: oracle", flush=True)
    osim, oplanes = wr.oracle_run(wl, steps, reference_precision=(name == "sphere80"))
    _same_fields(drv.sim, osim, name)
    _spectra_match(drv.flux_monitors, oplanes, wl.grid.dx, name,
                   reference_precision=(name == "sphere80"))
    drv.close()


def test_the_64_member_batch_matches_the_oracle_member_by_member():
    from meep_gpu.batch import BatchedRun, synthetic_batch
    from meep_gpu.driver import FdtdDriver

    steps = 160
    wls = synthetic_batch(count=64, seed=20261203, steps=3000)
    drivers = []
    for wl in wls:
        d = FdtdDriver.from_workload(wl)
        d.freeze(planned_steps=steps + 1)
        drivers.append(d)
    print(f"batch of {len(drivers)}: {sum(w.cells for w in wls)} cells, {steps} steps",
          flush=True)
    batch = BatchedRun(drivers)
    batch.run(steps)
    batch.synchronize()
    batch.close()
    lorentz = 0
    for q, (wl, d) in enumerate(zip(wls, drivers)):
        if q % 8 == 0:
            print(f"batch member {q} {wl.name}: oracle", flush=True)
        lorentz += bool(wl.poles)
        osim, oplanes = wr.oracle_run(wl, steps, reference_precision=False)
        _same_fields(d.sim, osim, wl.name)
        _spectra_match(d.flux_monitors, oplanes, wl.grid.dx, wl.name)
        d.close()
    assert lorentz > 0


def _oracle_checksum(osim):
    h = hashlib.sha256()
    vols = osim.stored_volumes()
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        h.update(numpy.ascontiguousarray(vols[comp], dtype=numpy.float32).tobytes())
    return h.hexdigest()


def test_served_stream_results_match_the_oracle():
    from meep_gpu.serve import SolveServer, poisson_requests, request_workload

    reqs = poisson_requests(count=8, rate=40.0, seed=20261203)
    for r in reqs:
        r.steps = 128 + r.steps % 160
    print("served: " + ", ".join(f"{r.edge}^3x{r.steps}{'L' if r.dispersive else ''}"
                                 f"/{r.nplanes}p{r.nfreq}f" for r in reqs), flush=True)
    rep = SolveServer(max_slots=8, quantum=64, checksums=True).run(reqs)
    assert rep.peak_resident > 1  # solves were on the device together
    for res, r in zip(rep.results, reqs):
        assert res.rid == r.rid
        wl = request_workload(r)
        print(f"served request {r.rid}: oracle", flush=True)
        osim, oplanes = wr.oracle_run(wl, r.steps, reference_precision=False)
        assert res.checksum == _oracle_checksum(osim), f"request {r.rid}"
        _spectra_match(res.spectra, oplanes, wl.grid.dx, f"request {r.rid}")
