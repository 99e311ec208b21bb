"""Served form on the GPU: every solve's result is the same alone and in the mix (bit for bit:
fields and flux spectra), and the admission footprint bounds what a solve allocates."""

import numpy
import pytest

pytestmark = pytest.mark.gpu


def _requests():
    from meep_gpu.serve import SolveRequest

    return [SolveRequest(rid=0, arrival=0.0, edge=32, steps=260, dispersive=True, nplanes=2,
                         nfreq=8),
            SolveRequest(rid=1, arrival=0.0, edge=40, steps=130, nplanes=8, nfreq=6),
            SolveRequest(rid=2, arrival=0.0, edge=36, steps=200, nplanes=1, nfreq=16,
                         dispersive=True),
            SolveRequest(rid=3, arrival=0.0, edge=44, steps=90, nplanes=3, nfreq=5),
            SolveRequest(rid=4, arrival=0.0, edge=32, steps=170, nplanes=4, nfreq=7)]


def test_each_solve_is_the_same_alone_and_in_the_mix():
    from meep_gpu.serve import SolveServer, run_alone

    reqs = _requests()
    # five at once into three slots: two wait for completions, the batch regroups
    srv = SolveServer(max_slots=3, quantum=64, checksums=True)
    rep = srv.run(reqs)
    assert rep.peak_resident == 3
    assert rep.regroupings >= 3
    for req, res in zip(reqs, rep.results):
        print(f"request {req.rid} edge {req.edge} steps {req.steps}", flush=True)
        spectra, checksum = run_alone(req)
        assert res.checksum == checksum, req.rid
        assert len(res.spectra) == req.nplanes
        for a, b in zip(res.spectra, spectra):
            assert numpy.array_equal(a, b), req.rid
            assert numpy.all(numpy.isfinite(a))
    s = rep.summary()
    assert s["p99_latency_s"] >= s["p50_latency_s"] > 0


def test_served_results_match_the_oracle():
    """Four solves (one with a Lorentz pole in its sphere) arriving together into two slots, so
    the batch regroups as they complete: each result read out of the mix against the oracle's
    run of the same request. Fields: the final E and H words (their sha256, as the server
    records them) equal the oracle's. Spectra: within 1e-9 of the oracle's float64 peak (the
    transform sums the same float64 products blocked over 128 steps)."""
    import hashlib

    from meep_gpu.serve import SolveRequest, SolveServer, request_workload
    from tests import workload_runs as wr

    reqs = [SolveRequest(rid=0, arrival=0.0, edge=30, steps=150, dispersive=True, nplanes=3,
                         nfreq=6),
            SolveRequest(rid=1, arrival=0.0, edge=34, steps=110, nplanes=2, nfreq=5),
            SolveRequest(rid=2, arrival=0.0, edge=28, steps=190, nplanes=1, nfreq=7),
            SolveRequest(rid=3, arrival=0.0, edge=32, steps=70, nplanes=4, nfreq=4,
                         dispersive=True)]
    srv = SolveServer(max_slots=2, quantum=48, checksums=True)
    rep = srv.run(reqs)
    assert rep.peak_resident == 2
    for req, res in zip(reqs, rep.results):
        print(f"request {req.rid} edge {req.edge} steps {req.steps} against the oracle",
              flush=True)
        sim, planes = wr.oracle_run(request_workload(req), req.steps,
                                    reference_precision=False)
        stored = sim.stored_volumes()
        h = hashlib.sha256()
        for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
            h.update(numpy.ascontiguousarray(stored[comp], dtype=numpy.float32).tobytes())
        assert res.checksum == h.hexdigest(), req.rid
        assert len(res.spectra) == len(planes) == req.nplanes
        for got, plane in zip(res.spectra, planes):
            want = plane.flux(request_workload(req).grid.dx, "c128")
            peak = numpy.max(numpy.abs(want))
            assert peak > 0, req.rid
            assert numpy.max(numpy.abs(got - want)) <= 1e-9 * peak, req.rid


def test_the_admission_footprint_bounds_the_allocation():
    from meep_gpu.serve import HipBackend, request_footprint, request_workload

    be = HipBackend()
    for req in _requests():
        drv = be.admit(req, request_workload(req))
        try:
            actual = drv.sim.device_bytes
            bound = request_footprint(req)
            assert actual <= bound, (req.rid, actual, bound)
            assert bound < 1.15 * actual + (1 << 20), (req.rid, actual, bound)
        finally:
            drv.close()
