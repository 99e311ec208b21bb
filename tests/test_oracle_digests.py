"""The oracle reproduces the reference's recorded fingerprint digests (CPU)."""

import pytest

from oracle import cases


@pytest.mark.parametrize("name", [case.name for case in cases.CASES])
def test_oracle_reproduces_recorded_digest(name, golden_digests):
    assert cases.run_case(name) == golden_digests["case_sha256"][name]


def test_a_planted_mistake_in_the_oracle_moves_the_digest(golden_digests, monkeypatch):
    """An off-by-one in a single PML coefficient must change the pml_3d digest."""
    from oracle import layer

    original = layer.axis_coefficients

    def skewed(n, faces, offset, dt, dx, order=2, ln_reflection=None, wraps=True):
        kms, sinv, kps = original(n, faces, offset, dt, dx, order, ln_reflection, wraps)
        if offset == 0.5 and faces[0] > 0:
            kms = kms.copy()
            kms[1] = kms[2]
        return kms, sinv, kps

    monkeypatch.setattr(layer, "axis_coefficients", skewed)
    assert cases.run_case("pml_3d") != golden_digests["case_sha256"]["pml_3d"]


@pytest.mark.parametrize("mistake", ["condinv_in_float32", "condfac_without_the_half"])
def test_a_planted_conductivity_mistake_moves_the_digests(golden_digests, monkeypatch, mistake):
    """condinv formed in float32 (rounded twice), or condfac missing the 1/2 of sigma*dt/2,
    moves both conductive digests. (condfac formed in double instead of float32 would not:
    with dt = 2**-4 the fixture's products are exact either way.)"""
    import numpy

    from oracle import engine

    original = engine.OracleSim.set_conductivity

    def planted(self, sigma, cells=None):
        original(self, sigma, cells)
        for name, vol in sigma.items():
            if name not in self.condinv:
                continue
            v = numpy.ascontiguousarray(vol, dtype=numpy.float32)
            if mistake == "condinv_in_float32":
                dt = numpy.float32(self.dt)
                self.condinv[name] = (numpy.float32(1.0)
                                      / (numpy.float32(1.0) + v * dt * numpy.float32(0.5)))
            else:
                self.condfac[name] = (numpy.float32(1.0) - v * numpy.float32(self.dt)
                                      ).astype(numpy.float32)

    monkeypatch.setattr(engine.OracleSim, "set_conductivity", planted)
    for name in ("conductive_3d", "conductive_pml_3d"):
        assert cases.run_case(name) != golden_digests["case_sha256"][name], (mistake, name)


def test_keeping_the_conductivity_in_every_chunk_moves_the_block_digest(golden_digests,
                                                                        monkeypatch):
    """conductive_chunks_pml_2d: stepping the conductive forms in the five chunks that keep no
    conductivity (sigma 0 there, so only the PML forms' rounding differs) is caught."""
    import numpy

    from oracle import chunks

    name = "conductive_chunks_pml_2d"
    assert cases.run_case(name) == golden_digests["case_sha256"][name]
    monkeypatch.setattr(chunks, "conductive_cells",
                        lambda shape, *a, **k: numpy.ones(tuple(shape), dtype=bool))
    assert cases.run_case(name) != golden_digests["case_sha256"][name]


@pytest.mark.parametrize("mistake", ["magnetic_after_update_H", "magnetic_a_step_late"])
def test_a_planted_deposit_mistake_moves_the_repair_digest(golden_digests, monkeypatch, mistake):
    """repair_pml_2d pins the magnetic deposit: into B before H is formed from it, with the
    stored current word of its own step. (Its points are pinned by the reference's recorded
    point counts: a zero-weight neighbour would not move this digest.)"""
    from oracle import engine

    name = "repair_pml_2d"
    if mistake == "magnetic_after_update_H":
        def step(self):
            self.step_B()
            self.zero_metal("B")
            self.update_H()
            self.deposit(True, self.time)
            self.step_D()
            self.deposit(False, self.time + 0.5 * self.dt)
            self.zero_metal("D")
            self.update_E()
            self.update_P()
            self.step_count += 1
            self.time = self.step_count * self.dt
        monkeypatch.setattr(engine.OracleSim, "step", step)
    elif mistake == "magnetic_a_step_late":
        original = engine.OracleSim.deposit

        def late(self, magnetic, time_value):
            original(self, magnetic, time_value + (self.dt if magnetic else 0.0))
        monkeypatch.setattr(engine.OracleSim, "deposit", late)
    assert cases.run_case(name) != golden_digests["case_sha256"][name]


@pytest.mark.parametrize("mistake", ["dipole_deposited_as_a_current", "dipole_times_dt",
                                     "dipole_a_half_step_early"])
def test_a_planted_integrated_source_mistake_moves_the_digests(golden_digests, monkeypatch,
                                                               mistake):
    """sources_pml_2d and conductive_source_2d pin MEEP's polarization source: never deposited
    into D, its dipole at (n+1)*dt subtracted from the copy of D the constitutive update reads,
    with no dt (and, in the conductive case, no condinv)."""
    from oracle import engine

    if mistake == "dipole_deposited_as_a_current":
        original = engine.OracleSim.deposit

        def both(self, magnetic, time_value, integrated=False):
            if integrated:
                return
            for dep in self.deposits:
                dep.integrated = False
            original(self, magnetic, time_value)
        monkeypatch.setattr(engine.OracleSim, "deposit", both)
    elif mistake == "dipole_times_dt":
        original = engine.OracleSim.deposit

        def scaled(self, magnetic, time_value, integrated=False):
            if not integrated:
                return original(self, magnetic, time_value)
            for dep in self.deposits:
                if dep.integrated:
                    dep.amp_re = dep.amp_re * self.dt
            original(self, magnetic, time_value, integrated=True)
            for dep in self.deposits:
                if dep.integrated:
                    dep.amp_re = dep.amp_re / self.dt
        monkeypatch.setattr(engine.OracleSim, "deposit", scaled)
    else:
        original = engine.OracleSim.f_minus_p_begin

        def early(self, magnetic, time_value):
            return original(self, magnetic, time_value - 0.5 * self.dt)
        monkeypatch.setattr(engine.OracleSim, "f_minus_p_begin", early)
    for name in ("sources_pml_2d", "conductive_source_2d"):
        assert cases.run_case(name) != golden_digests["case_sha256"][name], (mistake, name)
