"""The served form's scheduler on the CPU: admission by footprint, first-come ordering, chunking
that ends where a resident solve completes, and the report's figures. A fake backend stands in
for the engine and advances a simulated clock by the cell-steps each chunk carries."""

import pytest

from meep_gpu.serve import ServeReport, SolveRequest, SolveServer, poisson_requests


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(0.0, dt)


class FakeBackend:
    """Solves as step counters; a chunk costs cells * steps / rate seconds of the clock."""

    def __init__(self, clock, rate=1e9):
        self.clock = clock
        self.rate = rate
        self.advanced = {}
        self.groups = []
        self.live = set()
        self.prepared = []

    def device_memory(self):
        return (1 << 30, 1 << 31)

    def prepare(self, r):
        self.prepared.append(r.rid)
        return ("inputs", r.rid)

    def admit(self, r, inputs):
        assert inputs == ("inputs", r.rid)
        self.live.add(r.rid)
        self.advanced[r.rid] = 0
        return r

    def group(self, handles):
        self.groups.append(tuple(h.rid for h in handles))
        return list(handles)

    def advance(self, batch, steps):
        assert steps > 0
        for r in batch:
            self.advanced[r.rid] += steps
            assert self.advanced[r.rid] <= r.steps, "a solve was stepped past its length"
        self.clock.t += sum(r.cells for r in batch) * steps / self.rate

    def wait(self, batch):
        pass

    def ungroup(self, batch):
        pass

    def finish(self, r):
        assert self.advanced[r.rid] == r.steps
        self.live.remove(r.rid)
        return [], None


def _server(budget, slots=64, quantum=256, rate=1e9):
    clock = FakeClock()
    be = FakeBackend(clock, rate)
    srv = SolveServer(budget_bytes=budget, max_slots=slots, quantum=quantum, max_batch_cells=None,
                      backend=be, footprint=lambda r: r.cells * 100, clock=clock,
                      sleep=clock.sleep)
    return srv, be


def test_every_request_completes_with_its_own_step_count():
    reqs = poisson_requests(count=20, rate=50.0, seed=3, edge_range=(32, 64),
                            steps_range=(100, 900))
    srv, be = _server(budget=10 ** 12)
    rep = srv.run(reqs)
    assert [r.rid for r in rep.results] == list(range(20))
    assert all(be.advanced[r.rid] == r.steps for r in reqs)
    assert not be.live
    for res in rep.results:
        assert res.arrival <= res.admitted <= res.finished


def test_admission_never_exceeds_the_byte_budget_or_the_slots():
    reqs = [SolveRequest(rid=i, arrival=0.0, edge=40, steps=300 + 50 * i) for i in range(10)]
    budget = 3 * 40 ** 3 * 100  # three resident at a time
    srv, be = _server(budget=budget)
    rep = srv.run(reqs)
    assert rep.peak_bytes <= budget
    assert rep.peak_resident == 3
    assert max(len(g) for g in be.groups) == 3
    srv, be = _server(budget=10 ** 12, slots=2)
    rep = srv.run(reqs)
    assert rep.peak_resident == 2


def test_first_come_first_admitted_even_when_a_smaller_request_would_fit():
    big = SolveRequest(rid=1, arrival=0.1, edge=60, steps=200)
    reqs = [SolveRequest(rid=0, arrival=0.0, edge=50, steps=2000), big,
            SolveRequest(rid=2, arrival=0.2, edge=32, steps=200)]
    budget = (50 ** 3 + 60 ** 3) * 100 - 1  # the big one must wait for rid 0
    srv, _ = _server(budget=budget, rate=1e8)
    rep = srv.run(reqs)
    adm = {r.rid: r.admitted for r in rep.results}
    assert adm[1] <= adm[2], "a later, smaller request overtook the head of the queue"
    assert rep.results[1].queued > 0


def test_chunks_end_where_a_resident_solve_completes_and_regroup():
    reqs = [SolveRequest(rid=0, arrival=0.0, edge=32, steps=300),
            SolveRequest(rid=1, arrival=0.0, edge=32, steps=1000)]
    srv, be = _server(budget=10 ** 12, quantum=256)
    rep = srv.run(reqs)
    assert be.groups == [(0, 1), (1,)]
    assert rep.regroupings == 2
    assert rep.results[0].finished < rep.results[1].finished


def test_a_completed_solve_is_read_out_while_the_next_chunk_runs():
    """The read-out of a completed solve follows the launch of the next chunk of the solves
    still resident; with nothing else resident it is read out at once. Its bytes stay counted
    until it is freed, so a request needing them is admitted only after the read-out."""
    log = []

    class Logged(FakeBackend):
        def admit(self, r, inputs):
            log.append(("admit", r.rid))
            return super().admit(r, inputs)

        def advance(self, batch, steps):
            log.append(("advance", tuple(r.rid for r in batch), steps))
            super().advance(batch, steps)

        def finish(self, r):
            log.append(("finish", r.rid))
            return super().finish(r)

    clock = FakeClock()
    be = Logged(clock)
    fp = lambda r: r.cells * 100  # noqa: E731
    reqs = [SolveRequest(rid=0, arrival=0.0, edge=32, steps=300),
            SolveRequest(rid=1, arrival=0.0, edge=32, steps=1000),
            SolveRequest(rid=2, arrival=0.0, edge=32, steps=10)]
    srv = SolveServer(budget_bytes=2 * 32 ** 3 * 100, max_slots=64, quantum=256,
                      max_batch_cells=None, backend=be, footprint=fp, clock=clock,
                      sleep=clock.sleep)
    srv.run(reqs)
    assert log[:4] == [("admit", 0), ("admit", 1), ("advance", (0, 1), 256),
                       ("advance", (0, 1), 44)]
    # rid 0 completed: rid 1 goes on, then rid 0 is read out, then its bytes admit rid 2
    assert log[4:7] == [("advance", (1,), 256), ("finish", 0), ("admit", 2)]
    assert log.index(("finish", 1)) == len(log) - 1  # the last solve: read out at once
    assert [e for e in log if e[0] == "finish"] == [("finish", 0), ("finish", 2), ("finish", 1)]


def test_requests_are_prepared_while_the_device_runs():
    reqs = [SolveRequest(rid=0, arrival=0.0, edge=64, steps=4000),
            SolveRequest(rid=1, arrival=0.05, edge=32, steps=100)]
    srv, be = _server(budget=64 ** 3 * 100 + 10, rate=1e8)  # rid 1 must wait for a slot
    srv.run(reqs)
    assert be.prepared == [0, 1]


def test_a_request_larger_than_the_budget_is_refused_by_name():
    srv, _ = _server(budget=1000)
    with pytest.raises(ValueError, match="request 0"):
        srv.run([SolveRequest(rid=0, arrival=0.0, edge=32, steps=10)])


def test_report_percentiles_and_aggregate():
    reqs = poisson_requests(count=30, rate=20.0, seed=5, edge_range=(32, 48),
                            steps_range=(100, 400))
    srv, _ = _server(budget=10 ** 12, rate=5e8)
    rep = srv.run(reqs)
    s = rep.summary()
    assert s["requests"] == 30
    assert 0 < s["p50_latency_s"] <= s["p99_latency_s"]
    assert s["aggregate_mcell_steps_per_s"] == pytest.approx(
        rep.cell_steps / rep.makespan / 1e6, rel=1e-3)
    assert isinstance(rep, ServeReport)


def test_the_served_stream_matches_its_stated_distribution():
    reqs = poisson_requests(count=400, rate=4.0, seed=20261203)
    edges = [r.edge for r in reqs]
    assert min(edges) >= 32 and max(edges) <= 160
    assert 0.2 < sum(r.dispersive for r in reqs) / len(reqs) < 0.4
    assert all(1 <= r.nplanes <= 8 and 16 <= r.nfreq <= 100 for r in reqs)
    assert all(1000 <= r.steps <= 8000 for r in reqs)
    gaps = [b.arrival - a.arrival for a, b in zip(reqs, reqs[1:])]
    assert 0.2 < sum(gaps) / len(gaps) < 0.3


def test_resident_cells_cap_holds_requests_back_until_completions():
    reqs = [SolveRequest(rid=i, arrival=0.0, edge=40, steps=200 + 100 * i) for i in range(6)]
    clock = FakeClock()
    be = FakeBackend(clock)
    srv = SolveServer(budget_bytes=10 ** 12, max_batch_cells=2 * 40 ** 3, backend=be,
                      footprint=lambda r: r.cells * 100, clock=clock, sleep=clock.sleep)
    rep = srv.run(reqs)
    assert rep.peak_resident == 2
    assert max(len(g) for g in be.groups) == 2
    assert [r.rid for r in sorted(rep.results, key=lambda r: r.admitted)] == list(range(6))
