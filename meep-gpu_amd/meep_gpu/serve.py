"""Served form: a request loop that runs FDTD solves of different sizes side by side on one GPU.

Requests arrive over time (``poisson_requests``: a Poisson stream of sphere solves, edge
log-uniform over 32..160 cells, 30 % Lorentz-dispersive, 1..8 flux planes x 16..100
frequencies, 1000..8000 steps). ``SolveServer.run`` is continuous batching, not a queue:

* a request that has arrived is ADMITTED as soon as a slot is free and its HBM footprint
  (``footprint_bytes``: every field, absorber and ADE volume, ring, phase table and float64
  accumulator it will allocate) fits under the budget beside the solves already resident --
  first come, first admitted, so a large request is never starved by smaller ones behind it;
* every resident solve keeps its fields, absorber and ADE state and DFT accumulators on the
  device from admission to completion, and all of them advance together through the same
  fused launches (one batch, ``fdtd_batch_run``), in chunks of at most ``quantum`` steps that
  end where the next resident solve completes;
* while the device runs a chunk, the host admits what has arrived and fits -- allocates its
  device state and uploads its materials and source tables -- and builds the inputs of the
  requests still waiting, so admission costs the device nothing: the new solves join the
  shared launches at the end of the chunk;
* a solve that completes leaves the batch at once, and is read out (spectra) and freed while
  the device runs the next chunk (or at once when nothing else is resident); its slot and
  bytes go to the queue when it is freed.

Each solve steps exactly the words it would step alone (a batch member's launches are the
lone simulation's launches), which the GPU tests check request by request. The report gives
p50/p99 latency (arrival to result) beside aggregate Mcell-steps/s over the stream.

The scheduler is written against a small backend interface (``prepare``, ``admit``,
``group``, ``advance``, ``wait``, ``ungroup``, ``finish``) so its admission and ordering
logic is testable without a GPU; ``HipBackend`` is the engine.
"""

from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence

import numpy

from .dispersion import LORENTZIAN, Susceptibility
from .workloads import FluxPlane, Workload, meep_frequency_span, outside_absorber, sphere

# --- requests --------------------------------------------------------------------------------


@dataclass
class SolveRequest:
    """One FDTD solve: a dielectric sphere under a UPML with flux planes around it."""
    rid: int
    arrival: float        # seconds after the stream starts
    edge: int
    steps: int
    radius_frac: float = 0.25
    eps: float = 2.25
    dispersive: bool = False
    nplanes: int = 1
    nfreq: int = 32
    pml_cells: int = 8
    resolution: float = 10.0

    @property
    def cells(self) -> int:
        return self.edge ** 3

    def workload(self) -> Workload:
        return request_workload(self)


def _request_planes(edge: int, pml: int, radius: float, nplanes: int,
                    freqs: numpy.ndarray) -> List[FluxPlane]:
    """Up to 8 planes: the six faces of a box around the sphere, then two transmission
    planes near the absorber; all inside the non-absorbing interior."""
    c = edge // 2
    lo_i, hi_i = pml + 1, edge - pml - 2
    b = max(2, min(int(round(radius * 1.25)), c - lo_i, hi_i - c))
    planes = []
    for axis in range(3):
        for pos in (c - b, c + b):
            lo = [c - b, c - b, c - b]
            cnt = [2 * b, 2 * b, 2 * b]
            lo[axis], cnt[axis] = pos, 1
            planes.append(FluxPlane(axis, tuple(lo), tuple(cnt), freqs))
    inner = edge - 2 * pml
    for pos in (lo_i, hi_i):
        planes.append(FluxPlane(0, (pos, pml, pml), (1, inner, inner), freqs))
    return planes[:nplanes]


def request_workload(r: SolveRequest) -> Workload:
    radius = r.radius_frac * r.edge
    wl = sphere(edge=r.edge, resolution=r.resolution, pml_cells=r.pml_cells,
                radius_cells=radius, eps=r.eps, fcen=0.15, df=0.1, nfreq=r.nfreq,
                planes="none", steps=r.steps, source_x=-0.2 * r.edge / r.resolution)
    wl.planes = _request_planes(r.edge, r.pml_cells, radius, r.nplanes,
                                meep_frequency_span(0.15, 0.1, r.nfreq))
    if r.dispersive:  # the Lorentz pole inside the sphere (out of the absorber), eps_inf 2.0
        inside = (wl.inv["Ez"] < 1.0) & outside_absorber(wl.grid, r.pml_cells)
        sig = numpy.where(inside, numpy.float32(1.5), numpy.float32(0.0)).astype(numpy.float32)
        wl.poles.append((Susceptibility(1.1, 0.05, LORENTZIAN), {c: sig for c in ("Ex", "Ey", "Ez")}))
        for c in ("Ex", "Ey", "Ez"):
            wl.inv[c] = numpy.where(wl.inv[c] < 1.0, numpy.float32(0.5), wl.inv[c]).astype(
                numpy.float32)
    wl.name = f"request_{r.rid}_{r.edge}{'_lorentz' if r.dispersive else ''}"
    return wl


def request_footprint(r: SolveRequest) -> int:
    """Device bytes the solve will hold while resident."""
    from .engine import footprint_bytes

    radius = r.radius_frac * r.edge
    planes = _request_planes(r.edge, r.pml_cells, radius, r.nplanes, numpy.zeros(r.nfreq))
    return footprint_bytes((r.edge,) * 3, r.pml_cells > 0,
                           driven_poles=[(3, 3)] if r.dispersive else [],
                           planes=[(p.cells, r.nfreq) for p in planes],
                           source_steps=r.steps + 1, source_points=2)


def poisson_requests(count: int = 32, rate: float = 4.0, seed: int = 20261203,
                     edge_range=(32, 160), dispersive_fraction: float = 0.3,
                     plane_range=(1, 8), nfreq_range=(16, 100), steps_range=(1000, 8000)
                     ) -> List[SolveRequest]:
    """A Poisson arrival stream (exponential gaps of mean 1/rate seconds) from one seed."""
    rng = numpy.random.default_rng(seed)
    t = 0.0
    out = []
    lo, hi = edge_range
    for rid in range(count):
        t += float(rng.exponential(1.0 / rate))
        edge = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        edge = max(edge, 32)
        out.append(SolveRequest(
            rid=rid, arrival=t, edge=edge,
            steps=int(rng.integers(steps_range[0], steps_range[1] + 1)),
            radius_frac=float(rng.uniform(0.10, 0.35)), eps=float(rng.uniform(1.5, 12.0)),
            dispersive=bool(rng.uniform() < dispersive_fraction),
            nplanes=int(rng.integers(plane_range[0], plane_range[1] + 1)),
            nfreq=int(rng.integers(nfreq_range[0], nfreq_range[1] + 1))))
    return out


# --- results ---------------------------------------------------------------------------------


@dataclass
class SolveResult:
    rid: int
    cells: int
    steps: int
    arrival: float
    admitted: float
    finished: float
    spectra: List[numpy.ndarray] = field(default_factory=list)
    checksum: Optional[str] = None   # sha256 of the final E and H words, when asked for

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def queued(self) -> float:
        return self.admitted - self.arrival


@dataclass
class ServeReport:
    results: List[SolveResult]
    makespan: float
    peak_resident: int
    peak_bytes: int
    regroupings: int

    @property
    def cell_steps(self) -> int:
        return sum(r.cells * r.steps for r in self.results)

    @property
    def mcell_steps_per_s(self) -> float:
        return self.cell_steps / self.makespan / 1e6 if self.makespan > 0 else 0.0

    def latency_percentile(self, q: float) -> float:
        lat = numpy.array([r.latency for r in self.results])
        return float(numpy.percentile(lat, q)) if lat.size else 0.0

    def summary(self) -> Dict[str, object]:
        return {"requests": len(self.results),
                "p50_latency_s": round(self.latency_percentile(50), 4),
                "p99_latency_s": round(self.latency_percentile(99), 4),
                "mean_queue_s": round(float(numpy.mean([r.queued for r in self.results]))
                                      if self.results else 0.0, 4),
                "aggregate_mcell_steps_per_s": round(self.mcell_steps_per_s, 1),
                "makespan_s": round(self.makespan, 3),
                "peak_resident": self.peak_resident,
                "peak_resident_bytes": self.peak_bytes,
                "regroupings": self.regroupings}


# --- the scheduler ---------------------------------------------------------------------------


@dataclass
class _Slot:
    request: SolveRequest
    handle: object
    remaining: int
    admitted: float
    nbytes: int


class SolveServer:
    """Continuous batching of FDTD solves under an HBM budget (see the module docstring)."""

    def __init__(self, budget_bytes: Optional[int] = None, max_slots: int = 64,
                 quantum: int = 256, max_batch_cells: Optional[int] = 4 << 20, backend=None,
                 footprint: Callable = request_footprint,
                 clock: Callable[[], float] = time.perf_counter,
                 sleep: Callable[[float], None] = time.sleep, checksums: bool = False):
        if max_slots < 1 or quantum < 1:
            raise ValueError("max_slots and quantum must be >= 1")
        self.backend = backend if backend is not None else HipBackend(checksums=checksums)
        if budget_bytes is None:
            free, _ = self.backend.device_memory()
            budget_bytes = int(0.85 * free)
        self.budget = int(budget_bytes)
        self.max_slots = int(max_slots)
        # resident cells past which a batch already saturates the device: admitting more only
        # shares it thinner (every resident solve's latency grows), so further requests wait
        self.max_batch_cells = int(max_batch_cells) if max_batch_cells else None
        self.quantum = int(quantum)
        self.footprint = footprint
        self.clock = clock
        self.sleep = sleep

    def run(self, requests: Sequence[SolveRequest]) -> ServeReport:
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in reqs:
            if self.footprint(r) > self.budget:
                raise ValueError(f"request {r.rid} ({r.edge}^3) needs {self.footprint(r)} bytes, "
                                 f"more than the whole budget of {self.budget}")
        be = self.backend
        t0 = self.clock()
        nxt = 0
        queue: deque = deque()
        prepared: Dict[int, object] = {}
        running: List[_Slot] = []
        results: List[SolveResult] = []
        used = 0
        batch = None
        regroup = True
        peak_n = peak_b = regroupings = 0

        def arrivals(now: float):
            nonlocal nxt
            while nxt < len(reqs) and reqs[nxt].arrival <= now:
                queue.append(reqs[nxt])
                nxt += 1

        staged: List[_Slot] = []  # admitted while the device ran; they join at the next regroup
        finishing: List[_Slot] = []  # completed: read out while the next chunk runs

        def finish_done():
            nonlocal used
            for s in finishing:
                spectra, checksum = be.finish(s.handle)
                results.append(SolveResult(s.request.rid, s.request.cells, s.request.steps,
                                           s.request.arrival, s.admitted, self.clock() - t0,
                                           spectra, checksum))
                used -= s.nbytes  # its bytes stay counted until they are freed
            finishing.clear()

        def admit(into: List[_Slot]):
            """Admit from the head of the queue while a slot and the bytes are there."""
            nonlocal used, peak_n, peak_b
            while queue and len(running) + len(staged) < self.max_slots:
                r = queue[0]
                need = self.footprint(r)
                if used + need > self.budget:
                    break
                resident = sum(x.request.cells for x in running) + sum(
                    x.request.cells for x in staged)
                if (self.max_batch_cells is not None and resident > 0
                        and resident + r.cells > self.max_batch_cells):
                    break
                queue.popleft()
                inputs = prepared.pop(r.rid, None)
                if inputs is None:
                    inputs = be.prepare(r)
                handle = be.admit(r, inputs)
                into.append(_Slot(r, handle, r.steps, self.clock() - t0, need))
                used += need
            peak_n = max(peak_n, len(running) + len(staged))
            peak_b = max(peak_b, used)

        while nxt < len(reqs) or queue or running or staged or finishing:
            arrivals(self.clock() - t0)
            if staged:
                running.extend(staged)
                staged.clear()
                regroup = True
            if not running:
                finish_done()  # nothing on the device to overlap the read-out with
                admit(running)
                regroup = True
                if not running:
                    if nxt < len(reqs):
                        self.sleep(max(0.0, reqs[nxt].arrival - (self.clock() - t0)))
                    continue
            if regroup:
                if batch is not None:
                    be.ungroup(batch)
                batch = be.group([s.handle for s in running])
                regroup = False
                regroupings += 1
            chunk = min(self.quantum, min(s.remaining for s in running))
            be.advance(batch, chunk)
            # overlapped with the device: read out and free the solves that completed, admit
            # (allocate and upload) what has arrived and fits, and build the host inputs of the
            # requests still waiting
            finish_done()
            arrivals(self.clock() - t0)
            admit(staged)
            for r in list(queue)[: self.max_slots]:
                if r.rid not in prepared:
                    prepared[r.rid] = be.prepare(r)
            be.wait(batch)
            for s in running:
                s.remaining -= chunk
            done = [s for s in running if s.remaining == 0]
            if done:
                be.ungroup(batch)
                batch = None
                regroup = True
                finishing.extend(done)
                running = [s for s in running if s.remaining > 0]
        if batch is not None:
            be.ungroup(batch)
        makespan = (max(r.finished for r in results) - min(r.arrival for r in results)
                    if results else 0.0)
        results.sort(key=lambda r: r.rid)
        return ServeReport(results, makespan, peak_n, peak_b, regroupings)


# --- the HIP backend -------------------------------------------------------------------------


def final_checksum(sim) -> str:
    import hashlib

    h = hashlib.sha256()
    for comp in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
        h.update(sim.download(comp).tobytes())
    return h.hexdigest()


class HipBackend:
    """Solves resident on the GPU, advanced together by ``BatchedRun``."""

    def __init__(self, checksums: bool = False):
        self.checksums = checksums

    def device_memory(self):
        from .engine import device_memory

        return device_memory()

    def prepare(self, r: SolveRequest):
        """Host-side inputs (materials, planes, source tables are built at admission)."""
        return request_workload(r)

    def admit(self, r: SolveRequest, wl: Workload):
        from .driver import FdtdDriver

        drv = FdtdDriver.from_workload(wl)
        drv.freeze(planned_steps=r.steps + 1)
        return drv

    def group(self, drivers):
        from .batch import BatchedRun

        return BatchedRun(drivers)

    def advance(self, batch, steps: int):
        batch.run(steps)

    def wait(self, batch):
        batch.synchronize()

    def ungroup(self, batch):
        batch.close()

    def finish(self, drv):
        spectra = [m.get_flux_spectrum() for m in drv.flux_monitors]
        checksum = final_checksum(drv.sim) if self.checksums else None
        drv.close()
        return spectra, checksum


def run_alone(r: SolveRequest, checksum: bool = True):
    """The same solve stepped by itself (the reference for the in-mix results)."""
    from .driver import FdtdDriver

    drv = FdtdDriver.from_workload(request_workload(r))
    drv.freeze(planned_steps=r.steps + 1)
    drv.run(steps=r.steps)
    drv.synchronize()
    spectra = [m.get_flux_spectrum() for m in drv.flux_monitors]
    cs = final_checksum(drv.sim) if checksum else None
    drv.close()
    return spectra, cs
