"""MEEP-identity capsules (tests/golden/meep_identity): the oracle built from a lifted state,
and the replay loop both tiers judge MEEP's stored words with.

The capsules, currents and digests are the reference's Tier A keep and flush fixtures, copied as data
(tests/golden/meep_identity/provenance.json). The product reads them with
``meep_gpu.capsule``; the oracle here steps the same pasted state on the CPU.
"""

from __future__ import annotations

import os
from typing import Dict, List

import numpy

from meep_gpu import capsule

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.join(HERE, "golden", "meep_identity")
FIXTURES = os.path.join(ROOT, "keep")
#: The reference's two subnormal policies: MEEP's words with subnormals kept, and flushed.
MODES = ("keep", "flush")
#: The holey waveguide of MEEP's test_holey_wvg_bands (complex storage, a Bloch phase of -1
#: along x, an odd Y fold over periodic boundaries, an off-diagonal chi1inv, a PML on high y).
HOLEY = "t.test_holey_wvg_bands.TestHoleyWvgBands.test_fields_at_kx"
#: The cases the engine steps (the ring needs cylindrical storage, refused by name).
CASES = ("a", "b", "ex.material-dispersion", "ex.absorber-1d", "c", "c_prime", HOLEY)
#: Those on a mirror-folded grid (Y and Z odd mirrors over metallic boundaries, an Hx point
#: source on both planes; the holey waveguide's odd Y mirror over periodic boundaries).
FOLDED = ("c", "c_prime", HOLEY)


def case_dir(case: str, mode: str = "keep") -> str:
    return os.path.join(ROOT, mode, case)


def runtime(case: str, mode: str = "keep") -> str:
    """The C++ runtime MEEP ran with where the case was recorded."""
    return capsule.recording_runtime(os.path.join(case_dir(case, mode), "currents.json"))


def load(case: str, max_poles: int = 8, mode: str = "keep"):
    """(state, digests, currents) of one case in one subnormal policy."""
    folder = case_dir(case, mode)
    state = capsule.read_capsule(os.path.join(folder, "capsule.capsule"), max_poles=max_poles)
    digests = capsule.read_digests(os.path.join(folder, "digests.json"))
    currents = capsule.read_currents(os.path.join(folder, "currents.json"), state,
                                     digests.n_steps)
    return state, digests, currents


class OracleRun:
    """A lifted state stepped by the CPU restatement, every word pasted as the engine pastes it."""

    def __init__(self, state: capsule.LiftedState, currents, subnormals: str = "keep"):
        from oracle.engine import ComplexOracle, OracleSim, PointDeposit

        self.flush = subnormals == "flush"
        self.state = state
        if state.complex:
            parts = [self._part(state, currents, state.volumes),
                     self._part(state, currents, state.volumes_im)]
            self.complex = ComplexOracle(parts[0], parts[1], state.bloch_kinds,
                                         state.bloch_phase)
            # the imaginary part's deposits are the imaginary line of the same sources
            deposits, parts[0].deposits, parts[1].deposits = parts[0].deposits, [], []
            for dep in deposits:
                self.complex.add_deposit(dep)
            self.sim, self.sims = parts[0], parts
            return
        self.complex = None
        self.sim = self._part(state, currents, state.volumes)
        self.sims = [self.sim]

    @staticmethod
    def _part(state: capsule.LiftedState, currents, volumes):
        from oracle.engine import OracleSim, PointDeposit

        faces = state.pml.faces if state.pml is not None else ((0, 0), (0, 0), (0, 0))
        sim = OracleSim(state.shape, state.dt, state.dx, invariant=state.invariant,
                        pml_faces=faces)
        folds = state.grid().folds
        sim.set_metallic(tuple(m and f is None for m, f in zip(state.metallic, folds)))
        sim.set_folding(folds)
        if state.pml is not None:
            sim.set_layer_tables(state.pml.vectors, state.pml.members)
        sim.set_inverse_epsilon(dict(state.inv_eps), dict(state.offdiag) or None)
        for pole in state.poles:
            sigma = {c: pole.sigma.get(c, 0.0) for c in ("Ex", "Ey", "Ez")}
            added = sim.add_pole(pole.frequency, pole.gamma, pole.kind == "drude", sigma)
            added.consts = numpy.array(pole.constants, dtype=numpy.float32)
        if state.conductivity:
            sim.set_conductivity(dict(state.conductivity))
        for src, table in zip(state.sources, currents):
            flat = numpy.ravel_multi_index(tuple(src.index.T), state.shape).astype(numpy.int64)
            start = state.step
            times = {}
            for n in range(table.shape[0]):
                t = (start + n) * state.dt + (0.0 if src.magnetic else 0.5 * state.dt)
                times[t] = complex(table[n, 0], table[n, 1])
            sim.add_deposit(PointDeposit(src.component, flat, src.amps.real.copy(),
                                         src.amps.imag.copy(), times.__getitem__))
        if state.offdiag or state.complex:
            sim.enable_field_storage()
        for name, vol in volumes.items():
            if not numpy.any(vol):
                continue
            if name.startswith("P"):
                raise NotImplementedError("pasting a nonzero polarization")
            sim.v[name][...] = vol
        sim.step_count = state.step
        sim.time = state.step * state.dt
        sim.h_separated = state.h_separated
        return sim

    @property
    def step_count(self) -> int:
        return self.sim.step_count

    @property
    def grid(self):
        return self.state.grid()

    def advance(self, steps: int):
        from oracle.engine import flushing

        with flushing(self.flush):
            for _ in range(steps):
                (self.complex or self.sim).step()

    def words(self, component: str, part: int = 0) -> numpy.ndarray:
        sim = self.sims[part]
        if component[0] == "E":
            return sim.get_E(component)
        if component[0] == "H" and not sim.pml:
            return sim.v["B" + component[1]]
        return sim.v[component]

    runtime = None  # the C++ runtime of MEEP's complex powers (recording_runtime)

    def read(self, name: str) -> numpy.ndarray:
        """A stored volume: float32, or complex64 from the two parts."""
        if len(self.sims) == 1:
            return self.words(name)
        out = numpy.empty(self.state.shape, dtype=numpy.complex64)
        out.real, out.imag = self.words(name, 0), self.words(name, 1)
        return out

    def meep_array(self, component: str) -> numpy.ndarray:
        """The product's readback over the oracle's words, under the run's subnormal policy."""
        from meep_gpu.readback import meep_array
        from oracle.engine import flushing

        state = self.state
        with flushing(self.flush):
            return meep_array(self.read, component, state.grid(),
                              tuple(not m for m in state.metallic), state.division,
                              state.pml is not None, self.step_count,
                              phases=state.read_phases(), runtime=self.runtime)

    def close(self):
        pass


def replay_digests(run, digests: capsule.DigestSet, steps=None) -> Dict[int, List[str]]:
    """Advance ``run`` through the compared steps; the components differing from MEEP's words
    at each (all compared steps when ``steps`` is None)."""
    out = {}
    for step in (digests.steps if steps is None else steps):
        run.advance(step - run.step_count)
        print(f"compared step {step}", flush=True)
        out[step] = capsule.compare_step(run, digests, step)
    return out


def replay_reader(run, digests: capsule.DigestSet, steps=None) -> Dict[int, List[str]]:
    """As ``replay_digests``, judging the ``get_array`` words (MEEP's reader digests)."""
    out = {}
    for step in (digests.steps if steps is None else steps):
        run.advance(step - run.step_count)
        out[step] = capsule.compare_reader(run, digests, step)
    return out


def live_driver(case: str, mode: str = "keep"):
    """The capsule's object graph rebuilt as live Python objects (classes named by the
    graph's class paths, attributes as decoded), each source given ``current(t, dt)`` from the
    recorded currents table: a stand-in for a live reference driver at the seam."""
    import json

    folder = case_dir(case, mode)
    root = capsule.read_capsule_document(os.path.join(folder, "capsule.capsule"))[1]
    memo = {}
    classes = {}

    def build(value):
        if isinstance(value, capsule.Node):
            if id(value) in memo:
                return memo[id(value)]
            module, qual = value.cls.split(":")
            kind = classes.get(value.cls)
            if kind is None:
                kind = type(qual.split(".")[-1], (), {})
                kind.__module__, kind.__qualname__ = module, qual
                classes[value.cls] = kind
            obj = kind.__new__(kind)
            memo[id(value)] = obj
            for k, v in value.attrs.items():
                setattr(obj, k, build(v))
            return obj
        if isinstance(value, dict):
            return {k: build(v) for k, v in value.items()}
        if isinstance(value, (list, tuple)) and not hasattr(value, "_fields"):
            return type(value)(build(v) for v in value)
        return value

    driver = build(root)
    with open(os.path.join(folder, "currents.json")) as fh:
        doc = json.load(fh)
    for src, rec in zip(driver._sources, doc["sources"]):
        table = {}
        for key, value in rec["calls"]["current"]:
            t_text, dt_text = key.split("|")
            table[(capsule._float_bits(t_text), capsule._float_bits(dt_text))] = complex(
                capsule._float_bits(value["c"][0]), capsule._float_bits(value["c"][1]))
        src.current = lambda t, dt, table=table: table[(t, dt)]
    return driver
