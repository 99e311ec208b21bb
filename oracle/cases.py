"""The reference's fingerprint fixture cases, built and digested on the oracle (test infrastructure).

Restates, for the real-storage periodic Cartesian cases, ``array_path_oracle.py``'s
``CASES`` :427-475, ``_build`` :636-706, ``_stored_volumes`` :709-722, ``_digest_state``
:1048-1058 and ``_run_case`` :1118-1140. A case's digest is

    sha256("array_path_oracle/real-v6|<case>\\n"
           + for each step 0..6, for each stored volume in sorted name order:
                 "<name>|<f4|(nx, ny, nz)\\n" + the float32 bytes, C order
             and, when E is not stored, "get_E:<Ex|Ey|Ez>|<f4|(..)\\n" + (inv_eps*D) bytes)

The recorded values are in tests/golden/array_path_digests.json (reference
meep_gpu/cuda_kernels/fingerprints.json, driver_dispatch.array_path_oracle.case_sha256).

``build_state`` returns a plain description (shape, dt, layer faces, material volumes, poles
and the seeded state words) that both the oracle and the HIP engine's tests load, so the two
are compared on identical inputs.
"""

from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass, field
from typing import Any, Callable, Dict, List, Optional, Sequence, Tuple

import numpy

from . import chunks, fixture, layer
from . import sources as placement
from .engine import OracleSim, PointDeposit

_E = ("Ex", "Ey", "Ez")
#: Yee shifts of the conductive components (D like E, B like H; MEEP's iyee_shift)
_IYEE = {"Dx": (1, 0, 0), "Dy": (0, 1, 0), "Dz": (0, 0, 1),
         "Bx": (0, 1, 1), "By": (1, 0, 1), "Bz": (1, 1, 0)}


@dataclass(frozen=True)
class Source:
    """One stored-current source of a case (array_path_oracle.py:393-399): a custom waveform
    that returns the case's stored current word for each step."""
    component: str
    center: Tuple[float, float, float]
    size: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    #: MEEP's ``is_integrated``: the stored word is the dipole, which the constitutive update
    #: subtracts from D (sources.py:1697-1760) instead of a current deposited into it
    integrated: bool = False


@dataclass(frozen=True)
class Case:
    name: str
    cell: Tuple[float, float, float]
    dimensions: int = 3
    pml: Any = None
    offdiagonal: bool = False
    susceptibilities: Tuple[str, ...] = ()
    d_conductivity: bool = False
    b_conductivity: bool = False
    metallic: Tuple[bool, bool, bool] = (False, False, False)
    #: A conductivity kept only inside this block of fractions of each axis
    #: (array_path_oracle.py:626-633).
    conductivity_block: Any = None
    sources: Tuple[Source, ...] = ()
    #: The fused route's step order (array_path_oracle.py:1061-1109): the curl and the
    #: constitutive update run before the deposit and the deposit points' constitutive words
    #: are recomputed afterwards, which leaves the words of the plain order.
    repair: bool = False
    #: mirror planes as (axis, phase) (array_path_oracle.py:409, grid.Mirror)
    symmetry: Tuple[Tuple[str, int], ...] = ()


#: The fixture cases this engine steps: real storage, no symmetry; periodic or metallic axes;
#: a conductivity everywhere or in a block (kept per one-rank chunk).
CASES: Tuple[Case, ...] = (
    Case("pml_3d", (1.5, 1.25, 1.0), pml={"x": (2, 3), "y": (2.5, 2), "z": (0, 2)}),
    Case("pml_2d", (2.0, 1.5, 0.0), dimensions=2, pml={"x": (2.5, 1.5), "y": (1.75, 3)}),
    Case("no_pml_3d", (1.0, 0.75, 0.5)),
    Case("offdiag_3d", (1.0, 0.75, 0.75), offdiagonal=True),
    Case("offdiag_pml_3d", (1.5, 1.25, 1.25), pml=2, offdiagonal=True),
    Case("dispersive_pml_2d", (1.5, 1.25, 0.0), dimensions=2, pml={"x": 2, "y": 2},
         susceptibilities=("lorentz", "drude")),
    Case("dispersive_3d", (1.0, 0.75, 0.75), susceptibilities=("lorentz",)),
    Case("conductive_3d", (1.0, 0.75, 0.75), d_conductivity=True, b_conductivity=True),
    Case("conductive_pml_3d", (1.5, 1.25, 1.25), pml=2, d_conductivity=True,
         b_conductivity=True),
    Case("metallic_2d", (1.5, 1.25, 0.0), dimensions=2, pml={"y": 2},
         metallic=(True, False, False)),
    # an even mirror on x and an odd one on y over periodic boundaries: both axes store the
    # far ghost slot; a layer on the high faces of both and on z
    Case("folded_3d", (1.5, 1.25, 1.0), pml={"x": (0, 2), "y": (0, 2), "z": 2},
         symmetry=(("X", 1), ("Y", -1))),
    # the block conductivity reaches the interior chunk, the x-high and y-high edge chunks
    # and their corner; the other five chunks keep none and step the lossless PML forms
    Case("conductive_chunks_pml_2d", (2.0, 2.0, 0.0), dimensions=2, pml={"x": 2, "y": 2},
         d_conductivity=True, b_conductivity=True, metallic=(True, True, False),
         conductivity_block=((0.5, 1.0), (0.5, 1.0), (0.0, 1.0))),
    # an electric point between four samples and a magnetic one on a sample row in x
    Case("repair_pml_2d", (1.5, 1.5, 0.0), dimensions=2, pml={"x": 2, "y": 2},
         sources=(Source("Ez", (0.0625, 0.0625, 0.0)), Source("Hz", (-0.1875, 0.125, 0.0))),
         repair=True),
    # points, a finite sheet and an integrated (polarization) source under a 3-cell layer
    Case("sources_pml_2d", (2.0, 2.0, 0.0), dimensions=2, pml={"x": 3, "y": 3},
         sources=(Source("Ez", (-0.8125, 0.0625, 0.0)),
                  Source("Ex", (0.0, 0.25, 0.0), size=(0.75, 0.0, 0.0)),
                  Source("Hz", (0.1875, -0.1875, 0.0)),
                  Source("Ey", (0.125, 0.0, 0.0), integrated=True))),
    # the same kinds into conductive D components: the current scaled by condinv, the
    # integrated dipole not
    Case("conductive_source_2d", (1.5, 1.25, 0.0), dimensions=2, d_conductivity=True,
         sources=(Source("Ez", (0.0625, -0.0625, 0.0)),
                  Source("Ex", (0.0, 0.125, 0.0), size=(0.5, 0.0, 0.0)),
                  Source("Ey", (-0.125, 0.0, 0.0), integrated=True))),
)

#: Cases of this build's own, in the fixture cases' form (seeded the same way, by name) but with
#: no recorded digest: the engine is checked against the oracle word for word on them.
#: ``offdiag_pml_metallic_3d`` puts MEEP's default metallic walls around an off-diagonal
#: chi1inv and a PML on every face, two of them metallic.
VARIANTS: Tuple[Case, ...] = (
    Case("offdiag_pml_metallic_3d", (1.5, 1.25, 1.25), pml=2, offdiagonal=True,
         metallic=(True, False, True)),
    # folds over metallic boundaries (no stored far wall), an odd full count on z (9 -> 6 and
    # 11 -> 7 stored), walls on x, a layer on every boundary face
    Case("folded_metallic_3d", (1.5, 1.125, 1.375), pml={"x": 2, "y": (0, 2), "z": (0, 2)},
         symmetry=(("Y", -1), ("Z", 1)), metallic=(True, True, True)),
    # an odd periodic fold on x (11 -> 8 stored, far image row 5) and a metallic one on z,
    # with D and B conductivities, a Lorentz pole and a layer on y and the high faces
    Case("folded_conductive_pole_3d", (1.375, 1.0, 1.25), pml={"x": (0, 2), "y": 2, "z": (0, 2)},
         symmetry=(("X", -1), ("Z", 1)), metallic=(False, False, True),
         d_conductivity=True, b_conductivity=True, susceptibilities=("lorentz",)),
)

CASE_BY_NAME = {case.name: case for case in CASES + VARIANTS}


@dataclass
class CaseState:
    """Everything a stepper needs for one case, as plain arrays."""
    name: str
    shape: Tuple[int, int, int]
    dt: float
    dx: float
    invariant: Tuple[bool, bool, bool]
    faces: Tuple[Tuple[float, float], ...]
    inv: Dict[str, numpy.ndarray]
    offdiag: Dict[str, Dict[str, numpy.ndarray]]
    poles: List[Dict[str, Any]]          # {frequency, gamma, drude, sigma: {Ex: vol}}
    conductivity: Dict[str, numpy.ndarray]  # sigma by component, "Dx".."Bz" (absent: none)
    volumes: Dict[str, numpy.ndarray]    # the seeded stored state, by digest name
    stores_E: bool
    pml: bool
    #: {component, index (C-order flat), amp_re, amp_im, integrated, table (the stored word per
    #: step: the current, or an integrated source's dipole)}
    sources: List[Dict[str, Any]] = field(default_factory=list)
    #: per axis None or a mirror fold (phase, kind, far row): kind 1 over metallic boundaries,
    #: 2 over periodic ones (OracleSim.set_folding)
    folds: Tuple[Any, Any, Any] = (None, None, None)


def folded_axes(case: Case, full: Sequence[int]) -> Tuple[Tuple[int, ...], Tuple[Any, ...]]:
    """Stored counts and folds of the case's mirror planes (grid.py:568-575): a folded axis
    keeps MEEP's halved ``N - N//2 + 1`` cells, one more over periodic boundaries (the far
    ghost slot), whose image row is ``N - stored + 2`` (stepping._far_reflect_rows :2361)."""
    counts = list(full)
    folds: List[Any] = [None, None, None]
    for axis_name, phase in case.symmetry:
        a = "XYZ".index(axis_name)
        periodic = not case.metallic[a]
        stored = full[a] - full[a] // 2 + 1 + (1 if periodic else 0)
        counts[a] = stored
        folds[a] = (int(phase), 2 if periodic else 1,
                    full[a] - stored + 2 if periodic else None)
    return tuple(counts), tuple(folds)


def _grid_shape(case: Case) -> Tuple[Tuple[int, int, int], Tuple[bool, bool, bool]]:
    invariant = tuple(case.dimensions == 2 and axis == 2 for axis in range(3))
    shape = tuple(1 if invariant[a] else layer.cell_count(case.cell[a], fixture.RESOLUTION)
                  for a in range(3))
    return folded_axes(case, shape)[0], invariant


def _within(case: Case, values: numpy.ndarray) -> numpy.ndarray:
    """``values`` inside the case's conductivity block, exact zero outside
    (array_path_oracle.py:626-633: the block's rows are ``round(fraction * n)``)."""
    if case.conductivity_block is None:
        return values
    kept = numpy.zeros(values.shape, dtype=bool)
    kept[tuple(slice(int(round(lo * n)), int(round(hi * n)))
               for (lo, hi), n in zip(case.conductivity_block, values.shape))] = True
    return numpy.where(kept, values, numpy.float32(0.0)).astype(numpy.float32)


def build_state(name: str) -> CaseState:
    case = CASE_BY_NAME[name]
    shape, invariant = _grid_shape(case)
    dx = 1.0 / fixture.RESOLUTION
    dt = fixture.COURANT / fixture.RESOLUTION
    epsilon = {c: fixture.volume(fixture.seed(name, "eps", c), shape, 1.0, 3.0) for c in _E}
    inv = {c: (1.0 / epsilon[c].astype(numpy.float64)).astype(numpy.float32) for c in _E}
    offdiag: Dict[str, Dict[str, numpy.ndarray]] = {}
    if case.offdiagonal:
        for row in _E:
            for partner in _E:
                if partner == row:
                    continue
                vol = fixture.volume(fixture.seed(name, "chi1inv", row, partner), shape,
                                     -0.0625, 0.0625, zero_fraction=0.4)
                if numpy.any(vol):
                    offdiag.setdefault(row, {})[partner] = vol
    poles = []
    for index, kind in enumerate(case.susceptibilities):
        angular = fixture.LORENTZ_ANGULAR if kind == "lorentz" else fixture.DRUDE_ANGULAR
        poles.append({
            "frequency": angular / (2.0 * math.pi), "gamma": 0.0625 * (index + 1),
            "drude": kind == "drude",
            "sigma": {c: fixture.volume(fixture.seed(name, "pol", str(index), c), shape,
                                        0.0, 1.0, zero_fraction=0.3) for c in _E}})
    conductivity: Dict[str, numpy.ndarray] = {}
    for side, flag in (("D", case.d_conductivity), ("B", case.b_conductivity)):
        if flag:  # array_path_oracle.py:662-670
            for c in "xyz":
                vol = _within(case, fixture.volume(
                    fixture.seed(name, "sigma_" + side.lower(), side + c), shape, 0.0, 2.0,
                    zero_fraction=0.3))
                if numpy.any(vol):
                    conductivity[side + c] = vol
    faces = layer.pml_faces(case.pml, case.dimensions)
    pml = any(lo > 0 or hi > 0 for lo, hi in faces)
    # an electric integrated source makes E a stored array (driver.py:2718-2733)
    stores_E = (pml or bool(offdiag) or bool(poles)
                or any(src.integrated and src.component[0] == "E" for src in case.sources))
    names = ["D" + c for c in "xyz"] + ["B" + c for c in "xyz"]
    if stores_E:
        names += ["E" + c for c in "xyz"]
    if pml:
        names += ["H" + c for c in "xyz"]
        names += [p + f + c for p in ("fu_",) for f in "BD" for c in "xyz"]
        names += ["f_w_" + f + c for f in "EH" for c in "xyz"]
        names += ["f_cond_" + n for n in conductivity]
    for k, pole in enumerate(poles):
        for c in _E:
            if numpy.any(pole["sigma"][c] != 0):
                names += [f"P[{k}].{c}", f"P_prev[{k}].{c}"]
    volumes = {n: fixture.state_words(fixture.seed(name, "state", n), shape) for n in names}
    placed = []
    for index, src in enumerate(case.sources):  # array_path_oracle.py:692-703
        cells, amp_re, amp_im = placement.deposit_points(
            shape, invariant, fixture.RESOLUTION, src.component, src.center, src.size,
            wraps=tuple(not m for m in case.metallic))
        placed.append({"component": src.component, "index": cells, "amp_re": amp_re,
                       "amp_im": amp_im, "integrated": src.integrated,
                       "table": fixture.stored_currents(fixture.seed(name, "current", str(index)))})
    full = tuple(1 if invariant[a] else layer.cell_count(case.cell[a], fixture.RESOLUTION)
                 for a in range(3))
    return CaseState(name, shape, dt, dx, invariant, faces, inv, offdiag, poles, conductivity,
                     volumes, stores_E, pml, placed, folded_axes(case, full)[1])


def oracle_sim(state: CaseState) -> OracleSim:
    sim = OracleSim(state.shape, state.dt, state.dx, state.invariant, state.faces,
                    ln_reflection=fixture.PML_LN_REFLECTION)
    sim.set_inverse_epsilon(state.inv, state.offdiag)
    for pole in state.poles:
        sim.add_pole(pole["frequency"], pole["gamma"], pole["drude"], pole["sigma"])
    case = CASE_BY_NAME[state.name]
    cells = {n: chunks.conductive_cells(state.shape, state.invariant, case.metallic,
                                        state.faces, vol, _IYEE[n])
             for n, vol in state.conductivity.items()}
    sim.set_conductivity(state.conductivity, cells)
    # a folded metallic axis stores no wall (stepping._zero_metal :3145-3160): it is a fold
    sim.set_metallic(tuple(m and f is None for m, f in zip(case.metallic, state.folds)))
    sim.set_folding(state.folds)
    if state.stores_E:
        sim.enable_field_storage()
    for n, vol in state.volumes.items():
        target = sim.stored_volumes()[n]
        target[...] = vol
    for src in state.sources:
        table = src["table"]
        sim.add_deposit(PointDeposit(
            src["component"], src["index"], src["amp_re"], src["amp_im"],
            lambda t, table=table: fixture.stored_current_at(table, t, state.dt),
            integrated=src["integrated"]))
    return sim


def digest_update(digest, volumes: Dict[str, numpy.ndarray], stores_E: bool,
                  get_E: Optional[Callable[[str], numpy.ndarray]] = None) -> None:
    for n in sorted(volumes):
        values = numpy.ascontiguousarray(volumes[n])
        digest.update(f"{n}|{values.dtype.str}|{values.shape}\n".encode("utf-8"))
        digest.update(values.tobytes())
    if not stores_E:
        for c in _E:
            values = numpy.ascontiguousarray(get_E(c))
            digest.update(f"get_E:{c}|{values.dtype.str}|{values.shape}\n".encode())
            digest.update(values.tobytes())


def new_digest(name: str):
    return hashlib.sha256(f"{fixture.FIXTURE_SET}|{name}\n".encode("utf-8"))


def run_case(name: str, steps: int = fixture.STEPS) -> str:
    """The case's digest when the oracle steps it ``steps`` times."""
    state = build_state(name)
    sim = oracle_sim(state)
    digest = new_digest(name)
    digest_update(digest, sim.stored_volumes(), sim.stores_E, sim.get_E)
    for _ in range(steps):
        sim.step()
        digest_update(digest, sim.stored_volumes(), sim.stores_E, sim.get_E)
    return digest.hexdigest()
