"""MEEP identity on the CPU: the lifted capsules decode, their tables are this package's own
word for word, and the oracle steps them to MEEP 1.35.0's stored words at every compared step.

Fixtures: the reference's Tier A keep set (tests/golden/meep_identity, copied as data; the
reference's replay is meep_identity.py ``replay``). Each digest is the SHA-256 of MEEP's own
float32 words of one component at one step (steps 0-20, then every tenth of the run).
"""

import hashlib
import json
import math
import os

import numpy
import pytest

from meep_gpu import capsule
from meep_gpu.absorber import AbsorbingFace, graded_sigma
from meep_gpu.dispersion import Susceptibility
from meep_gpu.pml import PML
from tests import capsule_cases as cc



def test_fixture_files_are_the_reference_ledger_s():
    with open(os.path.join(cc.ROOT, "provenance.json")) as fh:
        prov = json.load(fh)
    assert set(prov["sets"]) == set(cc.MODES)
    for mode, cases in prov["sets"].items():
        assert set(cases) == set(cc.CASES)
        for case, files in cases.items():
            for name, digest in files.items():
                with open(os.path.join(cc.case_dir(case, mode), name), "rb") as fh:
                    assert hashlib.sha256(fh.read()).hexdigest() == digest, (mode, case, name)


@pytest.mark.parametrize("case", cc.CASES)
def test_capsules_decode(case):
    state, digests, currents = cc.load(case)
    assert digests.mode == "keep" and digests.case == case
    assert state.step == 0 and not state.h_separated
    assert len(currents) == len(state.sources) == 1
    assert currents[0].shape == (digests.n_steps, 2)
    # the source is live: MEEP's words move off zero
    last = digests.stored[digests.steps[-1]]
    comp = state.sources[0].component
    primary = ("B" if comp[0] == "H" else "D") + comp[1]
    assert last[primary] != digests.stored[0][primary]


def _decoded(case):
    return capsule.read_capsule_document(os.path.join(cc.case_dir(case), "capsule.capsule"))[1]


@pytest.mark.parametrize("change, reason", [
    (lambda d: d["fields"]["grid"].attrs.__setitem__("cylindrical", True), "cylindrical"),
    (lambda d: d["fields"].attrs.__setitem__("force_complex_fields", True), "complex"),
    (lambda d: d.attrs.__setitem__("_chi3", 1.0), "chi2/chi3"),
    (lambda d: d["fields"]["grid"].attrs.__setitem__(
        "boundaries", (("metallic", "periodic"), ("metallic", "metallic"),
                       ("periodic", "periodic"))), "boundaries"),
    (lambda d: d["_sources"][0].attrs.__setitem__("_point_ix", None), "point tables"),
])
def test_what_the_engine_does_not_step_is_refused_by_name(change, reason):
    driver = _decoded("a")
    change(driver)
    with pytest.raises(capsule.CapsuleRefusal, match=reason):
        capsule.lift_state(driver)


def test_more_poles_than_the_engine_holds_are_refused():
    driver = _decoded("ex.absorber-1d")
    assert len(capsule.lift_state(driver).poles) == 5
    with pytest.raises(capsule.CapsuleRefusal, match="5 poles"):
        capsule.lift_state(driver, max_poles=4)


def test_a_currents_table_at_other_times_is_refused(tmp_path):
    state, digests, _ = cc.load("a")
    with open(os.path.join(cc.case_dir("a"), "currents.json")) as fh:
        doc = json.load(fh)
    calls = doc["sources"][0]["calls"]["current"]
    calls[0], calls[1] = calls[1], calls[0]
    path = tmp_path / "currents.json"
    path.write_text(json.dumps(doc))
    with pytest.raises(capsule.CapsuleRefusal, match="deposits at t="):
        capsule.read_currents(path, state, digests.n_steps)


@pytest.mark.parametrize("case", ("a", "b"))
def test_the_lifted_pml_is_this_package_s_inside_metallic_walls(case):
    """MEEP's UPML words on a metallic (non-wrapping) axis, and the chunk division's lines,
    equal the table this package grades itself (pml.py with wraps=False)."""
    state, _, _ = cc.load(case)
    assert state.metallic == (True, True, False)
    grid = state.grid()
    faces = {"x": state.pml.faces[0], "y": state.pml.faces[1]}
    own = PML(grid, faces, order=state.pml.order, R_asymptotic=state.pml.R_asymptotic,
              wraps=(False, False, True))
    for key, vectors in state.pml.vectors.items():
        for mine, lifted in zip(own.vectors[key], vectors):
            assert numpy.array_equal(mine.view(numpy.uint32), lifted.view(numpy.uint32)), key
        assert numpy.array_equal(own.members[key], state.pml.members[key]), key
    # the wrapping table owns the wall row (the image of the top plane); with equal faces its
    # sigma there is the same maximum either way
    wrapped = PML(grid, faces, order=state.pml.order, R_asymptotic=state.pml.R_asymptotic)
    assert wrapped.members[(0, 0)][0] and not state.pml.members[(0, 0)][0]


def test_a_metallic_axis_grades_row_zero_from_the_low_wall():
    """Unequal faces tell the two gradings apart: on a wrapping axis integer row 0 is the top
    wall's image (the high face's sigma), on a metallic one it is the low wall (pml.py:578-582)."""
    from meep_gpu.grid import Grid

    grid = Grid(resolution=10, cell_size=(4.0, 4.0, 0.0), dimensions=2)
    faces = {"x": (6, 3)}
    walls = PML(grid, faces, wraps=(False, True, True))
    periodic = PML(grid, faces)
    kms_w, kms_p = walls.vectors[(0, 0)][0], periodic.vectors[(0, 0)][0]
    assert (kms_w != kms_p).nonzero()[0].tolist() == [0]
    low_face = PML(grid, {"x": (6, 0)}, wraps=(False, True, True)).vectors[(0, 0)][0]
    high_face = PML(grid, {"x": (0, 3)}).vectors[(0, 0)][0]
    assert kms_w[0] == low_face[0] and kms_p[0] == high_face[0]


def test_the_lifted_absorber_is_this_package_s_on_a_metallic_axis():
    """mp.Absorber on a metallic z: the float64 sigma MEEP's lift stored equals
    absorber.graded_sigma with no seam wrap, bit for bit, and so does its float32."""
    state, _, _ = cc.load("ex.absorber-1d")
    assert state.metallic == (False, False, True) and state.dimensions == 1
    layers = [AbsorbingFace(t, axis, side, r) for t, axis, side, r in state.absorber_layers]
    mine = graded_sigma(state.grid(), layers, wraps=(True, True, False))
    assert set(mine) == set(state.absorber_sigma)
    for name, lifted in state.absorber_sigma.items():
        assert numpy.array_equal(mine[name].view(numpy.uint64), lifted.view(numpy.uint64)), name
        assert numpy.array_equal(mine[name].astype(numpy.float32), state.conductivity[name])
    # a one-sided layer tells the seam rule apart: integer row 0 is the far wall's image on a
    # wrapping axis (inside a high layer), the near wall on a metallic one (outside it)
    high = [AbsorbingFace(1.0, 2, "high")]
    walls = graded_sigma(state.grid(), high, wraps=(True, True, False))["Dx"]
    periodic = graded_sigma(state.grid(), high)["Dx"]
    assert walls[0, 0, 0] == 0.0 and periodic[0, 0, 0] > 0.0
    assert numpy.array_equal(walls[0, 0, 1:], periodic[0, 0, 1:])


def test_the_lifted_conductivity_words_are_the_oracle_s():
    """condfac = 1 - sigma*(float32(dt)*0.5) in float32 and condinv = 1/(1 + (sigma*dt)*0.5)
    in double, rounded once (fields.py:775-870): MEEP's words, as the oracle forms them."""
    from oracle.engine import OracleSim

    state, _, _ = cc.load("ex.absorber-1d")
    sim = OracleSim(state.shape, state.dt, state.dx, invariant=state.invariant)
    sim.set_conductivity(dict(state.conductivity))
    for name in state.conductivity:
        assert numpy.array_equal(sim.condfac[name], state.condfac[name]), name
        assert numpy.array_equal(sim.condinv[name], state.condinv[name]), name


@pytest.mark.parametrize("case", ("ex.material-dispersion", "ex.absorber-1d"))
def test_the_lifted_ade_words_are_this_package_s(case):
    state, _, _ = cc.load(case)
    assert state.poles
    for pole in state.poles:
        words = Susceptibility(pole.frequency, pole.gamma, pole.kind).realnum_constants(state.dt)
        assert numpy.array_equal(numpy.array(words, dtype=numpy.float32), pole.constants)


@pytest.mark.parametrize("mode", cc.MODES)
@pytest.mark.parametrize("case", cc.CASES)
def test_the_oracle_steps_meep_s_words(case, mode):
    """Every compared step's stored words of all twelve components are MEEP's, with subnormals
    kept, and flushed under x86 FTZ|DAZ (the policy each set was minted under)."""
    state, digests, currents = cc.load(case, mode=mode)
    assert digests.mode == mode
    got = cc.replay_digests(cc.OracleRun(state, currents, subnormals=mode), digests)
    assert set(got) == set(digests.steps)
    assert {step: d for step, d in got.items() if d} == {}


def test_the_flush_set_needs_the_flush():
    """The flush set's words leave the keep policy's once subnormals appear (step 60 of ``a``),
    so the oracle's FTZ|DAZ is what reproduces them."""
    state, digests, currents = cc.load("a", mode="flush")
    got = cc.replay_digests(cc.OracleRun(state, currents, subnormals="keep"), digests,
                            steps=[200])
    assert got[200]


def test_the_replay_catches_planted_mistakes():
    """Two mistakes a restatement could make, each caught by MEEP's words at the last step:
    the walls dropped (the axes wrap), and one UPML word inside the layer one ulp off."""
    state, digests, currents = cc.load("a")
    last = digests.steps[-1]
    state.metallic = (False, False, False)
    got = cc.replay_digests(cc.OracleRun(state, currents), digests, steps=[last])
    assert got[last]

    state, digests, currents = cc.load("a")
    kms, sinv, kps = state.pml.vectors[(1, 0)]
    kms = kms.copy()
    kms[185] = numpy.nextafter(kms[185], numpy.float32(0))
    state.pml.vectors[(1, 0)] = (kms, sinv, kps)
    got = cc.replay_digests(cc.OracleRun(state, currents), digests, steps=[last])
    assert got[last]


def test_a_planted_pole_word_is_caught():
    state, digests, currents = cc.load("ex.material-dispersion")
    words = state.poles[0].constants.copy()
    words[1] = numpy.nextafter(words[1], numpy.float32(2))
    state.poles[0].constants = words
    got = cc.replay_digests(cc.OracleRun(state, currents), digests, steps=[40])
    assert "Ez" in got[40]


def test_a_folded_capsule_s_conductivity_is_refused_by_name():
    driver = capsule.read_capsule_document(os.path.join(cc.case_dir("c"), "capsule.capsule"))[1]
    sigma = numpy.ones(tuple(driver["fields"]["grid"]["shape"]), dtype=numpy.float32)
    driver["fields"].attrs["_conductivity"] = {"Dx": sigma}
    with pytest.raises(capsule.CapsuleRefusal, match="mirror-folded"):
        capsule.lift_state(driver)


@pytest.mark.parametrize("planted", ["even_parity", "no_fill", "no_zero_top"])
def test_the_folded_replay_catches_planted_mistakes(planted, monkeypatch):
    """MEEP's words of ``c`` (odd Y and Z mirrors over metallic boundaries) catch a fold taken
    even, the ghost planes left unfilled, and the zero past a folded axis's top read as a
    periodic wrap (the oracle's C grid told the axes are not folded)."""
    from oracle import engine

    state, digests, currents = cc.load("c")
    run = cc.OracleRun(state, currents)
    sim = run.sim
    if planted == "even_parity":
        sim.folds = tuple(None if f is None else (1, f[1], f[2]) for f in sim.folds)
    elif planted == "no_fill":
        monkeypatch.setattr(engine.OracleSim, "fill_folds", lambda self, field, far: None)
    else:
        real = engine.OracleSim._grid

        def grid(self):
            g = real(self)
            for a in range(3):
                g.fold[a] = 0
            return g

        monkeypatch.setattr(engine.OracleSim, "_grid", grid)
    got = cc.replay_digests(run, digests, steps=[20])
    assert got[20]


@pytest.mark.parametrize("mode", cc.MODES)
@pytest.mark.parametrize("case", cc.CASES)
def test_get_array_is_meep_s_on_the_oracle_s_words(case, mode):
    """readback.meep_array over the oracle's words gives MEEP's own sim.get_array words (the
    identity sets' reader digests) at every compared step: the centred lattice, MEEP's
    collapse of a one-cell axis, under a PML the chunks' H/B alias rule, on a folded grid the
    mirror images with their parities, and on complex storage with a Bloch phase the window's
    connected words and the duplicate plane's phase."""
    state, digests, currents = cc.load(case, mode=mode)
    assert digests.reader
    run = cc.OracleRun(state, currents, subnormals=mode)
    run.runtime = cc.runtime(case, mode)
    got = cc.replay_reader(run, digests)
    assert {step: d for step, d in got.items() if d} == {}


@pytest.mark.parametrize("planted", ["no_alias_rule", "no_duplicate_plane"])
def test_get_array_planted_mistakes_are_caught(planted, monkeypatch):
    """Reading B without the chunks' H/B alias rule (case b under its PML, where the interior
    chunk's boundary words differ from step 200), or dropping the periodic duplicate plane
    (the one-cell Lorentz case), leaves MEEP's get_array words."""
    from meep_gpu import readback

    case = "b" if planted == "no_alias_rule" else "ex.material-dispersion"
    state, digests, currents = cc.load(case)
    run = cc.OracleRun(state, currents)
    if planted == "no_alias_rule":
        original = readback.centred_sums
        monkeypatch.setattr(readback, "centred_sums",
                            lambda read, comp, inv, metal, div, pml, first_step, frame=None:
                            original(read, comp, inv, metal, div, False, first_step, frame))
    else:
        run.state.metallic = (True, True, False)  # no duplicate plane on x and y
    got = cc.replay_reader(run, digests, steps=[1, 200])
    assert any(got.values())


def test_get_array_over_a_box_is_a_slice_of_the_whole_cell():
    """MEEP's get_array over a center/size box. With its faces on integer-lattice coordinates,
    midway between centred points, it visits the centred points from the one below the low
    face to the one above the high face, each with weight 1: a slice of the whole cell. A line
    there (an empty direction) is MEEP's collapse of the two centred points beside it, weights
    0.5 and 0.5, summed by float32 += from zero. The whole-cell box is the whole cell."""
    import numpy

    from meep_gpu.readback import meep_array

    state, digests, currents = cc.load("a")
    run = cc.OracleRun(state, currents)
    run.advance(40)
    whole = run.meep_array("Ez")
    args = (run.words, "Ez", state.grid(), (False, False, True), state.division, True, 40)
    dx = state.dx
    assert numpy.any(whole[90:110, 95:105])

    def at(k):  # integer-lattice point k (stored index) as a coordinate: (k - 100) dx
        return (k - 100) * dx

    box = meep_array(*args, center=(0.5 * (at(92) + at(108)), 0.5 * (at(96) + at(104)), 0.0),
                     size=(at(108) - at(92), at(104) - at(96), 0.0))
    assert numpy.array_equal(box, whole[91:109, 95:105])
    line = meep_array(*args, center=(at(101), 0.5 * (at(96) + at(104)), 0.0),
                      size=(0.0, at(104) - at(96), 0.0))
    half = lambda v: (0.5 * v.astype(numpy.float64)).astype(numpy.float32)  # noqa: E731
    want = (numpy.float32(0.0) + half(whole[100, 95:105])) + half(whole[101, 95:105])
    assert numpy.array_equal(line, want)
    assert numpy.array_equal(
        meep_array(*args, center=(-0.025, -0.025, 0.0), size=(9.95, 9.95, 0.0)), whole)


@pytest.mark.parametrize("planted", ["even_mirror", "no_reflection_sign", "no_window",
                                     "unphased_duplicate"])
def test_folded_and_complex_read_out_catches_planted_mistakes(planted, monkeypatch):
    """MEEP's get_array words of ``c`` (odd Y and Z folds) catch a mirror ghost read with the
    even sign and images unfolded without their parity; those of the holey waveguide (complex,
    a Bloch factor of -1 on x) catch the window's word below read as stored row 0 and the
    duplicate plane read without pow(eikna, -1)."""
    from meep_gpu import readback

    case = "c" if planted in ("even_mirror", "no_reflection_sign") else cc.HOLEY
    if planted == "even_mirror":
        monkeypatch.setattr(readback.ReadFrame, "parity", lambda self, comp, axis: 1)
    elif planted == "no_reflection_sign":
        monkeypatch.setattr(readback, "phase_shift", lambda frame, comp, mask: (1.0, 0.0))
    elif planted == "no_window":
        monkeypatch.setattr(readback.ReadFrame, "window_axes", lambda self, comp: ())
    else:
        monkeypatch.setattr(readback, "shift_phase", lambda frame, ishift: (1.0, 0.0))
    state, digests, currents = cc.load(case)
    run = cc.OracleRun(state, currents)
    run.runtime = cc.runtime(case)
    got = cc.replay_reader(run, digests, steps=[digests.steps[-1]])
    assert any(got.values())


def test_complex_powers_of_both_runtimes():
    """pow(complex<double>, int) as libstdc++ forms it (repeated squaring; 1 over the power by
    libgcc's division) and as libc++ does (exp(n log x)), against closed forms: a unit phase's
    inverse is its conjugate, its square its product with itself, and the Bloch factor MEEP
    recorded for the holey waveguide inverts to (-1, -residue) under libgcc's division."""
    import cmath
    import random

    from meep_gpu.readback import _divide, cmul, cpow_int

    rng = random.Random(7)
    for _ in range(200):
        z = cmath.exp(1j * rng.uniform(-math.pi, math.pi))
        x = (z.real, z.imag)
        for runtime in ("libstdc++", "libc++"):
            inv = cpow_int(x, -1, runtime)
            assert abs(complex(*inv) - z.conjugate()) < 4e-16
            sq = cpow_int(x, 2, runtime)
            assert abs(complex(*sq) - complex(*cmul(x, x))) < 4e-16
            assert cpow_int(x, 0, runtime) == (1.0, 0.0)
        a, b, c, d = (rng.uniform(-3, 3) for _ in range(4))
        q = complex(a, b) / complex(c, d)
        got = complex(*_divide(a, b, c, d))
        assert abs(got - q) <= 4e-16 * abs(q)
    eikna = (-1.0, 8.572527594031472e-16)
    assert cpow_int(eikna, -1, "libstdc++") == (-1.0, -8.572527594031472e-16)
    assert cpow_int((1.0, 0.0), -1, "libstdc++") == (1.0, 0.0)
    assert math.copysign(1.0, cpow_int((1.0, 0.0), -1, "libc++")[1]) > 0


@pytest.mark.parametrize("case", cc.FOLDED)
def test_the_folded_division_rule_answers_as_meep_s_boxes(case):
    """The one-rank division of a folded grid (MEEP's effort cut over the halved axis's owned
    window, a periodic fold's far ghost row in the last piece) answers the read-out's question
    -- which chunk owns each word of a B component's centred sum -- as the lifted chunk boxes
    of MEEP's own run do, for every B component, at the first step and after. (A periodic
    fold's far ghost row lies past every box and the read-out never takes its sums.)"""
    from meep_gpu.chunks import ChunkDivision, IYEE

    state, _, _ = cc.load(case)
    grid = state.grid()
    ours = ChunkDivision.one_rank(grid, state.pml.faces, tuple(not m for m in state.metallic))
    window = tuple(slice(0, grid.owned(a)) for a in range(3))
    for comp in ("Bx", "By", "Bz"):
        axes = [a for a in range(3) if not state.invariant[a] and IYEE[comp][a] == 0]
        for first in (True, False):
            got = ours.reader_masks(comp, axes, state.shape, first)
            want = state.division.reader_masks(comp, axes, state.shape, first)
            for g, w in zip(got[0] + got[1], want[0] + want[1]):
                assert numpy.array_equal(numpy.broadcast_to(g, state.shape)[window],
                                         numpy.broadcast_to(w, state.shape)[window]), (comp, first)
