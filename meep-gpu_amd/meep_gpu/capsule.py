"""Lifted MEEP simulations from capsules: load one with no MEEP and step it on the GPU.

A capsule (reference capsule.py:1-60) is the reference's lifted driver saved as data: a zip
holding ``capsule.json`` (the construction log and the whole object graph below the driver,
every float scalar as the hex of its IEEE-754 bytes) and one ``.npy`` member per array. The
MEEP-identity fixtures (reference meep_identity.py:1-60) pair each capsule with the currents
its sources returned (``currents.json``) and the SHA-256 of MEEP's own float32 words per
compared step and component (``digests.json``). Stepping the capsule here and hashing the
engine's words against those digests is the reference's Tier A replay, with the HIP engine in
place of its array path.

What is read, and how:

* only JSON and ``numpy.load(..., allow_pickle=False)``: nothing in a capsule executes;
* every array member is checked against the sha256 of its bytes the document records;
* the object graph is decoded by tag: ``a`` array, ``f`` float64 bits, ``c`` complex,
  ``t`` tuple, ``l`` list, ``d`` dict, ``i``/``o``/``s`` an object (id, class, attributes),
  ``r`` a reference to an object (written before or after it), ``b``/``g`` a NumPy scalar, ``nt``/``v`` a
  named tuple;
* the lifted state is pasted, not recomputed: the UPML vectors, inverse epsilon, ADE words,
  conductivity and the sources' point amplitudes go to the engine as the capsule holds them,
  and the PML membership lines come from the lifted chunk division (chunk_map.py:497-570).

Complex storage (a Bloch phase, ``force_complex_fields``) is taken as two real parts,
``volumes`` and ``volumes_im``, with each axis's Bloch connection (``bloch_kinds``,
``bloch_phase``: MEEP's copy, negation or product by the phase, boundaries.cpp:359-366).

What is refused, by name (``CapsuleRefusal``): cylindrical storage, chi2/chi3, monitors, a chunk division no membership line describes,
more poles than the engine holds, a source without point tables, a currents table whose
times are not the driver's (electric sources at ``n*dt + 0.5*dt``, magnetic at ``n*dt``;
driver.py:3536, :3565), and on a mirror-folded grid a conductivity and MEEP's ``get_array``
read-out (which unfolds the images). A folded grid steps with the engine's folds; its digests
hash every stored cell but the mirror ghosts (meep_identity.py:18-24, ``cell_classes``
:146-174), which MEEP does not store.
"""

from __future__ import annotations

import hashlib
import io
import json
import struct
import zipfile
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Tuple

import numpy

from . import _native
from .chunks import IYEE, BoxDivision, ChunkDivision, from_boxes
from .grid import Grid

CAPSULE_SCHEMA = ("meep_gpu.capsule", 1)
CURRENTS_SCHEMA = ("meep_gpu.capsule_currents", 1)
DIGESTS_SCHEMA = ("meep_gpu.meep_identity_digests", 1)

#: The twelve components a digest set covers (meep_identity.py COMPONENTS).
COMPONENTS = ("Ex", "Ey", "Ez", "Dx", "Dy", "Dz", "Hx", "Hy", "Hz", "Bx", "By", "Bz")
_AXES = ("x", "y", "z")


class CapsuleRefusal(ValueError):
    """A capsule this engine cannot load, with the reason."""


# --- decoding -------------------------------------------------------------------------------
@dataclass
class Node:
    """One decoded object of the graph: its class path and attributes."""
    cls: str
    attrs: Dict[str, object] = field(default_factory=dict)

    def __getitem__(self, name: str):
        return self.attrs[name]

    def get(self, name: str, default=None):
        return self.attrs.get(name, default)


def _float_bits(text: str) -> float:
    return struct.unpack("<d", bytes.fromhex(text))[0]


def _index_objects(node, raw: Dict[int, dict]):
    """Every object node of the document by id (attributes are written in sorted order, so a
    reference can come before the object it names)."""
    if isinstance(node, list):
        for v in node:
            _index_objects(v, raw)
    elif isinstance(node, dict):
        if set(node) == {"i", "o", "s"}:
            raw[node["i"]] = node
        for v in node.values():
            _index_objects(v, raw)


def _decode(node, arrays: Sequence[numpy.ndarray], seen: Dict[int, Node],
            raw: Dict[int, dict]):
    if node is None or isinstance(node, (bool, int, float, str)):
        return node
    if isinstance(node, list):
        return [_decode(v, arrays, seen, raw) for v in node]
    if not isinstance(node, dict):
        raise CapsuleRefusal(f"unknown capsule node {node!r}")
    keys = set(node)
    if keys == {"a"}:
        return arrays[node["a"]]
    if keys == {"f"}:
        return _float_bits(node["f"])
    if keys == {"c"}:
        return complex(_float_bits(node["c"][0]), _float_bits(node["c"][1]))
    if keys == {"t"}:
        return tuple(_decode(v, arrays, seen, raw) for v in node["t"])
    if keys == {"l"}:
        return [_decode(v, arrays, seen, raw) for v in node["l"]]
    if keys == {"d"}:
        out = {}
        for key, value in node["d"]:
            key = _decode(key, arrays, seen, raw)
            out[tuple(key) if isinstance(key, list) else key] = _decode(value, arrays, seen, raw)
        return out
    if keys == {"r"}:
        ref = node["r"]
        if ref not in seen:
            if ref not in raw:
                raise CapsuleRefusal(f"a reference to object {ref}, which the capsule lacks")
            return _decode(raw[ref], arrays, seen, raw)
        return seen[ref]
    if keys == {"b", "g"}:
        return numpy.frombuffer(bytes.fromhex(node["b"]), dtype=numpy.dtype(node["g"]))[0]
    if keys == {"nt", "v"}:
        return Node(node["nt"], {"values": tuple(_decode(v, arrays, seen, raw) for v in node["v"])})
    if keys == {"i", "o", "s"}:
        if node["i"] in seen:
            return seen[node["i"]]
        obj = Node(node["o"])
        seen[node["i"]] = obj  # registered first: attributes may refer back to it
        obj.attrs = {k: _decode(v, arrays, seen, raw) for k, v in node["s"].items()}
        return obj
    raise CapsuleRefusal(f"unknown capsule node tags {sorted(keys)}")


def read_capsule_document(path) -> Tuple[dict, Node]:
    """The capsule's JSON document and its decoded driver object, every array checked."""
    with zipfile.ZipFile(path) as archive:
        doc = json.loads(archive.read("capsule.json"))
        if (doc.get("schema"), doc.get("version")) != CAPSULE_SCHEMA:
            raise CapsuleRefusal(f"capsule schema {doc.get('schema')!r} version "
                                 f"{doc.get('version')!r}; this reader takes {CAPSULE_SCHEMA}")
        arrays = []
        for entry in doc["arrays"]:
            array = numpy.load(io.BytesIO(archive.read(entry["member"])), allow_pickle=False)
            if (array.dtype != numpy.dtype(entry["dtype"])
                    or list(array.shape) != list(entry["shape"])):
                raise CapsuleRefusal(f"{entry['path']}: member is {array.dtype} {array.shape}, "
                                     f"the document says {entry['dtype']} {entry['shape']}")
            if hashlib.sha256(numpy.ascontiguousarray(array).tobytes()).hexdigest() != entry["sha256"]:
                raise CapsuleRefusal(f"{entry['path']}: bytes differ from the recorded sha256")
            arrays.append(array)
    raw: Dict[int, dict] = {}
    _index_objects(doc["state"], raw)
    return doc, _decode(doc["state"], arrays, {}, raw)


# --- the lifted state -----------------------------------------------------------------------
@dataclass
class LayerTables:
    """A lifted UPML: the coefficient vectors as stored, and the division's membership lines,
    in the shape ``HipSimulation`` takes a ``PML`` (keys ``(half, axis)``)."""
    vectors: Dict[Tuple[int, int], Tuple[numpy.ndarray, numpy.ndarray, numpy.ndarray]]
    members: Dict[Tuple[int, int], numpy.ndarray]
    faces: Tuple[Tuple[float, float], ...]
    order: int
    R_asymptotic: float

    @property
    def is_active(self) -> bool:
        return True


@dataclass
class LiftedPole:
    kind: str
    frequency: float
    gamma: float
    constants: numpy.ndarray            # the four float32 words, as lifted
    sigma: Dict[str, object]            # "Ex".. -> float32 volume or float (driven only)


@dataclass
class LiftedSource:
    component: str                      # "Ex".."Ez" (into D) or "Hx".."Hz" (into B)
    index: numpy.ndarray                # (npoints, 3) int64 stored-cell indices
    amps: numpy.ndarray                 # (npoints,) complex128

    @property
    def magnetic(self) -> bool:
        return self.component[0] in "HB"


@dataclass
class LiftedState:
    shape: Tuple[int, int, int]
    dimensions: int
    resolution: float
    courant: float
    dx: float
    dt: float
    invariant: Tuple[bool, bool, bool]
    metallic: Tuple[bool, bool, bool]
    inv_eps: Dict[str, numpy.ndarray]
    pml: Optional[LayerTables]
    conductivity: Dict[str, numpy.ndarray]      # "Dx".."Bz" float32 sigma, nonzero only
    condfac: Dict[str, numpy.ndarray]
    condinv: Dict[str, numpy.ndarray]
    absorber_sigma: Dict[str, numpy.ndarray]    # float64, before rounding (provenance)
    absorber_layers: List[Tuple[float, int, str, float]]
    poles: List[LiftedPole]
    sources: List[LiftedSource]
    volumes: Dict[str, numpy.ndarray]           # every stored float32 field volume, engine names
    step: int
    h_separated: bool
    absent: Tuple[str, ...]
    division: ChunkDivision = None  # the lifted chunk division (per-chunk conductivity)
    #: mirror planes ((axis 0..2, phase +-1), ...) and the full (unfolded) cell counts
    symmetry: Tuple[Tuple[int, int], ...] = ()
    shape_full: Tuple[int, int, int] = None
    #: complex storage: the imaginary parts of the stored volumes (``volumes`` holds the real
    #: ones), and per axis the Bloch connection (0 none, 1 copy, 2 negate, 3 a product) and
    #: the phase's two parts (float64; the engine rounds them to float32)
    complex: bool = False
    volumes_im: Dict[str, numpy.ndarray] = field(default_factory=dict)
    bloch_kinds: Tuple[int, int, int] = (0, 0, 0)
    bloch_phase: Tuple[Tuple[float, float], ...] = ((1.0, 0.0),) * 3

    def read_phases(self) -> Tuple[Optional[complex], ...]:
        """Per axis MEEP's Bloch factor in double (``eikna``), None where the axis has none."""
        return tuple(None if k == 0 else complex(*p)
                     for k, p in zip(self.bloch_kinds, self.bloch_phase))
    #: off-diagonal chi1inv rows, {"Ex": {"Ey": float32 volume}}
    offdiag: Dict[str, Dict[str, numpy.ndarray]] = field(default_factory=dict)

    def grid(self) -> Grid:
        """The engine grid; refuses when its dx, dt or cell counts are not the lifted words."""
        full = self.shape_full or self.shape
        size = [0.0 if self.invariant[a] else full[a] / self.resolution for a in range(3)]
        grid = Grid(resolution=self.resolution, cell_size=tuple(size), courant=self.courant,
                    dimensions=self.dimensions, symmetry=self.symmetry, metallic=self.metallic)
        if tuple(grid.shape) != tuple(self.shape) or grid.invariant != self.invariant:
            raise CapsuleRefusal(f"grid {grid.shape} {grid.invariant} is not the lifted "
                                 f"{self.shape} {self.invariant}")
        if grid.dx != self.dx or grid.dt != self.dt:
            raise CapsuleRefusal(f"dx/dt {grid.dx!r}/{grid.dt!r} differ from the lifted "
                                 f"{self.dx!r}/{self.dt!r}")
        return grid


def _axis_frame(frame: Node) -> Tuple[int, int, int, bool, bool]:
    return (int(frame["origin"]), int(frame["stored"]), int(frame["owned"]),
            bool(frame["wraps"]), bool(frame["invariant"]))


def _folded_frame(frame: Node) -> bool:
    return frame.get("near_ghost_row") is not None or frame.get("far_ghost_row") is not None


def layout_membership(layout: Node, axis: int, half: bool) -> numpy.ndarray:
    """Per stored row: does the chunk owning it carry this axis's layer? A restatement of
    ``ChunkLayout.membership`` (chunk_map.py:497-570): every stored row's doubled coordinate
    ``origin + 2 i + half``, moved into the window on a wrapping axis; owned when it lies in
    ``(origin, origin + 2 owned]``; a member when every box covering it carries the axis and
    one does. Refuses a division no line describes (boxes that disagree on one row)."""
    lines = layout.get("_lines") or {}
    if (axis, bool(half)) in lines:
        return numpy.asarray(lines[(axis, bool(half))], dtype=bool)
    origin, stored, owned, wraps, invariant = _axis_frame(layout["frames"][axis])
    if invariant:
        return numpy.zeros(stored, dtype=bool)
    points = origin + 2 * numpy.arange(stored, dtype=numpy.int64) + (1 if half else 0)
    lo, hi = origin, origin + 2 * owned
    if wraps:
        points = numpy.where(points <= lo, points + (hi - lo), points)
        points = numpy.where(points > hi, points - (hi - lo), points)
    is_owned = (points > lo) & (points <= hi)
    any_true = numpy.zeros(stored, dtype=bool)
    any_false = numpy.zeros(stored, dtype=bool)
    for (box_lo, box_hi), flags in zip(layout["boxes"], layout["pml_flags"]):
        covered = (points > box_lo[axis]) & (points <= box_hi[axis])
        if flags[axis]:
            any_true |= covered
        else:
            any_false |= covered
    if numpy.any(any_true & any_false & is_owned):
        raise CapsuleRefusal(f"the chunk division is not separable along axis {axis}")
    return any_true & is_owned


def _flat(vector) -> numpy.ndarray:
    return numpy.ascontiguousarray(numpy.asarray(vector, dtype=numpy.float32).reshape(-1))


def _lift_pml(pml: Node, layout: Node, shape) -> LayerTables:
    vectors, members = {}, {}
    for half in (0, 1):
        suffix = "_h" if half else ""
        for axis, name in enumerate(_AXES):
            kms = _flat(pml[f"kms_{name}{suffix}"])
            sinv = _flat(pml[f"sinv_{name}{suffix}"])
            kps = _flat(pml[f"kps_{name}{suffix}"])
            if not kms.size == sinv.size == kps.size == shape[axis]:
                raise CapsuleRefusal(f"PML vectors of axis {name} have {kms.size} words, the "
                                     f"grid {shape[axis]} rows")
            vectors[(half, axis)] = (kms, sinv, kps)
            members[(half, axis)] = layout_membership(layout, axis, bool(half))
    faces = tuple(tuple(float(v) for v in f) for f in pml["thickness_by_face"])
    return LayerTables(vectors, members, faces, int(pml["order"]), float(pml["R_asymptotic"]))


def _volume(value) -> numpy.ndarray:
    array = numpy.asarray(value)
    if array.dtype != numpy.float32:
        raise CapsuleRefusal(f"a field volume is {array.dtype}, not float32")
    return numpy.ascontiguousarray(array)


def lift_state(driver: Node, max_poles: int = _native.FDTD_MAX_POLES) -> LiftedState:
    """The engine's view of a decoded lifted driver, refusing what the engine does not step."""
    if driver.cls != "meep_gpu.driver:FdtdDriver":
        raise CapsuleRefusal(f"the capsule holds a {driver.cls}, not an FdtdDriver")
    fields = driver["fields"]
    grid = fields["grid"]
    symmetry = tuple((axis, int(m["phase"])) for axis, m in enumerate(grid["mirrors_by_axis"])
                     if m is not None)
    if grid["cylindrical"]:
        raise CapsuleRefusal("cylindrical coordinates are not built")
    complex_fields = bool(fields["force_complex_fields"] or driver.get("force_complex_fields")
                          or grid["has_bloch"])
    kinds, phases = [0, 0, 0], [(1.0, 0.0)] * 3
    for axis, phase in enumerate(grid["bloch_phases"] or (None, None, None)):
        if phase is None:
            continue
        phase = complex(phase)
        kinds[axis], phases[axis] = bloch_connection(phase), (phase.real, phase.imag)
        if kinds[axis] == 3 and not complex_fields:
            raise CapsuleRefusal("a Bloch phase on real storage")
    if float(driver["_chi2"]) != 0.0 or float(driver["_chi3"]) != 0.0 \
            or fields["_chi2_components"] or fields["_chi3_components"]:
        raise CapsuleRefusal("chi2/chi3 nonlinearity is not built")
    offdiag = {row: {partner: _volume(vol) for partner, vol in partners.items()}
               for row, partners in (fields["_chi1inv_offdiagonal"] or {}).items()}
    if driver["_flux_monitors"] or driver["_dft_monitors"]:
        raise CapsuleRefusal("a capsule carries no monitors")
    if any(fields.get(f"f_bfast_{c}") is not None for c in ("Bx", "By", "Bz", "Dx", "Dy", "Dz")):
        raise CapsuleRefusal("the special-kz (bfast) form is not built")

    shape = tuple(int(n) for n in grid["shape"])
    invariant = tuple(bool(v) for v in grid["invariant_axes"])
    bounds = grid["boundaries"]
    metallic = []
    for axis in range(3):
        low, high = bounds[axis]
        if low != high or low not in ("periodic", "metallic"):
            raise CapsuleRefusal(f"boundaries {low!r}/{high!r} on axis {axis} are not built")
        metallic.append(low == "metallic")

    inv_eps = {}
    for c in ("Ex", "Ey", "Ez"):
        vol = fields["_inv_eps_components"].get(c, fields["inv_eps"])
        inv_eps[c] = _volume(vol)

    layout = driver["_chunk_layout"]
    pml = None
    if driver["pml"] is not None and fields["_pml_active"]:
        pml = _lift_pml(driver["pml"], layout, shape)

    def side(table) -> Dict[str, numpy.ndarray]:
        return {k: _volume(v) for k, v in (table or {}).items()}

    conductivity = {k: v for k, v in side(fields["_conductivity"]).items() if numpy.any(v)}
    absorber_sigma = {k: numpy.asarray(v, dtype=numpy.float64)
                      for k, v in (driver["_absorber_sigma"] or {}).items()}
    layers = [(float(l["thickness"]), int(l["axis"]), str(l["side"]), float(l["R_asymptotic"]))
              for l in driver["_absorber_layers"]]

    poles = []
    for state in fields["polarizations"]:
        term = state["susceptibility"]
        words = numpy.array(state["constants"]["values"], dtype=numpy.float32)
        driven = tuple(state["_driven"])
        sigma = {}
        for c, value in state["sigma"].items():
            if c not in driven:
                continue
            sigma[c] = _volume(value) if numpy.ndim(value) else float(value)
        poles.append(LiftedPole(str(term["kind"]), float(term["frequency"]),
                                float(term["gamma"]), words, sigma))
    if len(poles) > max_poles:
        raise CapsuleRefusal(f"{len(poles)} poles; the engine holds {max_poles}")

    sources = []
    for src in driver["_sources"]:
        ix = src.get("_point_ix")
        if ix is None or src.get("_point_amps") is None:
            raise CapsuleRefusal(f"source {src.cls} has no point tables")
        index = numpy.stack([numpy.asarray(src[f"_point_i{a}"], dtype=numpy.int64).reshape(-1)
                             for a in _AXES], axis=1)
        amps = numpy.asarray(src["_point_amps"], dtype=numpy.complex128).reshape(-1)
        component = str(src["component"])
        if component not in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"):
            raise CapsuleRefusal(f"source component {component!r} is not built")
        if numpy.any(index < 0) or numpy.any(index >= numpy.array(shape)):
            raise CapsuleRefusal(f"source {src.cls} has points outside the grid")
        sources.append(LiftedSource(component, index, amps))

    volumes, volumes_im = {}, {}

    def put(name, value):
        array = numpy.asarray(value)
        if array.dtype == numpy.complex64:
            if not complex_fields:
                raise CapsuleRefusal(f"{name} is complex on a real-storage run")
            volumes[name] = numpy.ascontiguousarray(array.real)
            volumes_im[name] = numpy.ascontiguousarray(array.imag)
            return
        volumes[name] = _volume(array)
        if complex_fields:
            raise CapsuleRefusal(f"{name} is {array.dtype} on a complex-storage run")

    for prefix in ("B", "D", "E", "H", "fu_B", "fu_D", "f_w_E", "f_w_H", "f_cond_D", "f_cond_B"):
        for a in _AXES:
            value = fields.get(prefix + a)
            if value is not None:
                put(prefix + a, value)
    for k, state in enumerate(fields["polarizations"]):
        for c, value in state["P"].items():
            put(f"P[{k}].{c}", value)
        for c, value in state["P_prev"].items():
            put(f"P_prev[{k}].{c}", value)

    frames = [_axis_frame(frame) for frame in layout["frames"]]
    boxes = [(tuple(int(v) for v in lo), tuple(int(v) for v in hi)) for lo, hi in layout["boxes"]]
    flags = [tuple(bool(f) for f in fl) for fl in layout["pml_flags"]]
    if not any(_folded_frame(frame) for frame in layout["frames"]):
        try:
            division = from_boxes(frames, boxes, flags)
        except ValueError as err:
            raise CapsuleRefusal(f"the lifted chunk division: {err}") from None
    elif conductivity:  # the per-chunk conductivity of a folded division is not built
        raise CapsuleRefusal("a conductivity on a mirror-folded grid is not built")
    else:  # the boxes answer the read-out's question (which chunk owns a word)
        division = BoxDivision(tuple(frames), tuple(boxes), tuple(flags))
    return LiftedState(
        division=division, shape=shape, dimensions=int(grid["dimensions"]), resolution=float(grid["resolution"]),
        courant=float(grid["courant"]), dx=float(grid["dx"]), dt=float(grid["dt"]),
        invariant=invariant, metallic=tuple(metallic), inv_eps=inv_eps, pml=pml,
        conductivity=conductivity, condfac=side(fields["_condfac"]),
        condinv=side(fields["_condinv"]), absorber_sigma=absorber_sigma, absorber_layers=layers,
        poles=poles, sources=sources, volumes=volumes, step=int(driver["step_count"]),
        h_separated=bool(fields["h_separated"]), absent=tuple(fields.get("absent_components") or ()),
        symmetry=symmetry, shape_full=tuple(int(n) for n in grid["shape_full"]),
        complex=complex_fields, volumes_im=volumes_im, bloch_kinds=tuple(kinds),
        bloch_phase=tuple(phases), offdiag=offdiag)


#: ``phase_isclose`` (reference part_products.py CONNECT_TOLERANCE, boundaries.cpp:359-360)
CONNECT_TOLERANCE = 1e-13


def bloch_connection(phase: complex) -> int:
    """MEEP's connection kind for a Bloch phase: 1 a copy (within 1e-13 of +1), 2 a negation
    (of -1), 3 a product by the phase."""
    if abs(phase.imag) < CONNECT_TOLERANCE and abs(phase.real - 1.0) < CONNECT_TOLERANCE:
        return 1
    if abs(phase.imag) < CONNECT_TOLERANCE and abs(phase.real + 1.0) < CONNECT_TOLERANCE:
        return 2
    return 3


def read_capsule(path, max_poles: int = _native.FDTD_MAX_POLES) -> LiftedState:
    return lift_state(read_capsule_document(path)[1], max_poles=max_poles)


# --- currents and digests -------------------------------------------------------------------
def recording_runtime(path) -> str:
    """The C++ runtime MEEP ran with where a currents table was recorded (its provenance's
    ``system``): libc++ on Darwin, libstdc++ elsewhere."""
    doc = json.loads(Path(path).read_text())
    system = str(doc.get("provenance", {}).get("system", ""))
    return "libc++" if system == "Darwin" else "libstdc++"


def read_currents(path, state: LiftedState, steps: int) -> List[numpy.ndarray]:
    """Each source's current per step, ``(steps, 2)`` float64 (re, im), from a recorded
    currents table; its times must be the ones the driver asks for at each step."""
    doc = json.loads(Path(path).read_text())
    if (doc.get("schema"), doc.get("version")) != CURRENTS_SCHEMA:
        raise CapsuleRefusal(f"currents schema {doc.get('schema')!r}; this reader takes "
                             f"{CURRENTS_SCHEMA}")
    if len(doc["sources"]) != len(state.sources):
        raise CapsuleRefusal(f"{len(doc['sources'])} current tables for "
                             f"{len(state.sources)} sources")
    start = int(doc["provenance"].get("start_step", 0))
    tables = []
    for rec, src in zip(doc["sources"], state.sources):
        if rec["component"] != src.component:
            raise CapsuleRefusal(f"currents of {rec['component']} for a {src.component} source")
        calls = rec["calls"]["current"]
        if len(calls) < steps:
            raise CapsuleRefusal(f"{len(calls)} recorded currents for {steps} steps")
        table = numpy.zeros((steps, 2), dtype=numpy.float64)
        for n in range(steps):
            key, value = calls[n]
            t_text, dt_text = key.split("|")
            t, dt = _float_bits(t_text), _float_bits(dt_text)
            step = start + n
            want = step * state.dt + (0.0 if src.magnetic else 0.5 * state.dt)
            if dt != state.dt or t != want:
                raise CapsuleRefusal(f"current call {n} asks t={t!r}, dt={dt!r}; step {step} "
                                     f"deposits at t={want!r}")
            table[n] = (_float_bits(value["c"][0]), _float_bits(value["c"][1]))
        tables.append(table)
    return tables


@dataclass
class DigestSet:
    """MEEP's words per compared step and component (meep_identity.py:17-30)."""
    case: str
    mode: str
    n_steps: int
    steps: List[int]
    stored: Dict[int, Dict[str, str]]
    ghost_cells: Dict[str, int]
    #: the same steps' digests of MEEP's sim.get_array words (the whole cell)
    reader: Dict[int, Dict[str, str]] = field(default_factory=dict)


def read_digests(path) -> DigestSet:
    doc = json.loads(Path(path).read_text())
    if (doc.get("schema"), doc.get("version")) != DIGESTS_SCHEMA:
        raise CapsuleRefusal(f"digests schema {doc.get('schema')!r}; this reader takes "
                             f"{DIGESTS_SCHEMA}")
    stored = {int(step): dict(words) for step, words in doc["stored"].items()}
    reader = {int(step): dict(words) for step, words in doc.get("reader", {}).items()}
    return DigestSet(doc["case"], doc["mode"], int(doc["n_steps"]),
                     [int(s) for s in doc["steps"]], stored, dict(doc["ghost_cells"]), reader)


def word_digest(words: numpy.ndarray) -> str:
    """SHA-256 of one component's words: little-endian float32 in C order (x, y, z); complex
    words as interleaved (real, imaginary) float32 pairs."""
    kind = "<c8" if numpy.asarray(words).dtype.kind == "c" else "<f4"
    return hashlib.sha256(numpy.ascontiguousarray(words, dtype=kind).tobytes()).hexdigest()


# --- stepping a lifted state on the GPU -----------------------------------------------------
def _source_boxes(src: LiftedSource):
    """One deposit box per point, in the table's order (MEEP's src_vol order)."""
    for (i, j, k), amp in zip(src.index, src.amps):
        yield (int(i), int(j), int(k)), amp


class LiftedRun:
    """A lifted state resident on the GPU: every word pasted, stepped by the engine built with
    the given subnormal policy ("keep" or "flush", the identity set's ``mode``). Complex
    storage is two linked simulations, the real and the imaginary parts (``fdtd_link_parts``),
    stepped together as a batch of two."""

    def __init__(self, state: LiftedState, currents: Sequence[numpy.ndarray],
                 subnormals: str = "keep"):
        from .engine import HipSimulation, link_parts

        self.state = state
        grid = state.grid()
        self.sims = []
        try:
            for part in range(2 if state.complex else 1):
                self.sims.append(HipSimulation(grid, state.pml, subnormals=subnormals))
                self._configure(self.sims[-1], state, currents, part)
            if state.complex:
                link_parts(self.sims[0], self.sims[1], state.bloch_kinds, state.bloch_phase)
        except Exception:
            for sim in self.sims:
                sim.close()
            raise
        self.sim = self.sims[0]

    @staticmethod
    def _configure(sim, state: LiftedState, currents, part: int = 0):
        for axis in range(3):
            if state.metallic[axis] and sim.grid.folds[axis] is None:  # a fold carries its own
                sim.set_boundary(axis, "metallic")
        for c, vol in state.inv_eps.items():
            sim.set_inverse_epsilon(c, vol)
        for row, partners in state.offdiag.items():
            for partner, vol in partners.items():
                sim.set_offdiagonal(row, partner, vol)
        for pole in state.poles:
            sim.add_pole(None, pole.sigma, constants=pole.constants)
        for name, vol in state.conductivity.items():
            sim.set_conductivity(name, vol, state.division.conductive_mask(vol, IYEE[name]))
        for src, table in zip(state.sources, currents):
            axis = "xyz".index(src.component[1])
            for lo, amp in _source_boxes(src):
                # the imaginary part takes MEEP's second line, (amp_re*c_im) + (amp_im*c_re):
                # the real line of the amplitude (amp_im, -amp_re) (step.cpp:309-314)
                re, im = (amp.real, amp.imag) if part == 0 else (amp.imag, -amp.real)
                sim.add_source(src.magnetic, axis, lo, (1, 1, 1), numpy.array([re]),
                               numpy.array([im]), table, state.dt)
        stores_e = state.pml is not None or bool(state.poles) or bool(state.offdiag)
        volumes = state.volumes if part == 0 else state.volumes_im
        for name, vol in volumes.items():
            if not numpy.any(vol):
                continue  # the engine's volumes start at +0
            if name[0] == "H" and state.pml is None:
                raise CapsuleRefusal("a stored H apart from B without an absorber")
            if name[0] == "E" and not stores_e:
                continue  # E is formed from D below
            sim.upload(name, vol)
        sim.set_step(state.step, state.h_separated)
        if not stores_e:
            sim.prime_e()  # E = chi1inv*(D - sum P), as the reference's get_E forms it

    @property
    def grid(self) -> Grid:
        return self.sim.grid

    def engine_has(self, name: str) -> bool:
        """Whether the engine holds a stored volume of this name: D, B and E always; H and
        the layer auxiliaries under an absorber; f_cond where that component has a
        conductivity under one."""
        if name[:-1] in ("B", "D", "E"):
            return True
        if self.state.pml is None:
            return False
        if name.startswith("f_cond_"):
            return name[len("f_cond_"):] in self.state.conductivity
        return name[:-1] in ("H", "fu_B", "fu_D", "f_w_E", "f_w_H")

    @property
    def step_count(self) -> int:
        return self.sim.step_count

    def advance(self, steps: int):
        if steps <= 0:
            return
        if len(self.sims) == 1:
            self.sim.run(int(steps))
            return
        import types

        from .batch import BatchedRun

        batch = BatchedRun([types.SimpleNamespace(sim=s, grid=s.grid) for s in self.sims])
        try:
            batch.run(int(steps))
            batch.synchronize()
        finally:
            batch.close()

    def words(self, component: str, part: int = 0) -> numpy.ndarray:
        """One component's stored words (of one part) in the reference layout (H is B's word
        off the layer's H rows, and everywhere without an absorber)."""
        if component[0] == "H" and self.state.pml is None:
            component = "B" + component[1]
        return self.sims[part].download(component)

    #: the C++ runtime of MEEP's complex powers in the read-out (None: this host's); a
    #: recorded run's is its recording host's (``recording_runtime``)
    runtime: Optional[str] = None

    def meep_array(self, component: str, runtime: Optional[str] = None) -> numpy.ndarray:
        """MEEP's whole-cell ``get_array(component)`` of the current words, with the lifted
        chunk division's H/B alias rule (``readback.meep_array``): float32, or complex64 from
        complex storage."""
        from .readback import meep_array

        state = self.state
        words = {}  # downloaded before the host arithmetic changes mode

        def read(name):
            if name not in words:
                parts = [self.words(name, p) for p in range(len(self.sims))]
                if len(parts) == 1:
                    words[name] = parts[0]
                else:
                    words[name] = numpy.empty(parts[0].shape, dtype=numpy.complex64)
                    words[name].real, words[name].imag = parts
            return words[name]

        read(component)
        if component[0] == "B" and state.pml is not None:
            read("H" + component[1])
        with _native.host_flush(self.sim.subnormals == "flush"):
            return meep_array(read, component, state.grid(),
                              tuple(not m for m in state.metallic), state.division,
                              state.pml is not None, self.step_count,
                              phases=state.read_phases(), runtime=runtime or self.runtime)

    def close(self):
        for sim in self.sims:
            sim.close()


def compare_reader(run, digests: DigestSet, step: int) -> List[str]:
    """The components whose ``get_array`` words differ from MEEP's at one compared step."""
    want = digests.reader.get(step, {})
    return [c for c in COMPONENTS if c in want and word_digest(run.meep_array(c)) != want[c]]


def ghost_cells(grid: Grid, component: str) -> numpy.ndarray:
    """The mirror ghosts of one component's stored array (reference meep_identity.py:146-174):
    on a folded axis stored cell 0 where the component's Yee shift is 0, and over periodic
    boundaries the last slot where it is 1. MEEP stores neither."""
    ghost = numpy.zeros(grid.shape, dtype=bool)
    for axis, fold in enumerate(grid.folds):
        if fold is None:
            continue
        face = [slice(None)] * 3
        if IYEE[component][axis] == 0:
            face[axis] = 0
        elif fold[1] == 2:
            face[axis] = -1
        else:
            continue
        ghost[tuple(face)] = True
    return ghost


def compared_words(run, component: str) -> numpy.ndarray:
    """One component's words on the cells MEEP stores: every stored cell but the mirror
    ghosts, in C order (x, y, z); a complex component's two parts interleaved per cell (as
    complex64 lays them out, reference meep_identity.py:18-24)."""
    grid = run.grid
    folded = any(f is not None for f in grid.folds)
    keep = ~ghost_cells(grid, component) if folded else None
    parts = [run.words(component)]
    if getattr(run.state, "complex", False):
        parts.append(run.words(component, part=1))
    if keep is not None:
        parts = [p[keep] for p in parts]
    if len(parts) == 1:
        return parts[0]
    return numpy.stack([p.reshape(-1) for p in parts], axis=-1).reshape(-1)


def compare_step(run, digests: DigestSet, step: int) -> List[str]:
    """The components whose words differ from MEEP's at one compared step."""
    want = digests.stored[step]
    return [c for c in COMPONENTS
            if c in want and word_digest(compared_words(run, c)) != want[c]]


def replay(capsule_dir, max_poles: int = _native.FDTD_MAX_POLES) -> Dict[int, List[str]]:
    """Step a MEEP-identity case (``capsule.capsule``, ``currents.json``, ``digests.json`` in
    one folder) on the GPU, with the subnormal policy its digests were minted under, and judge
    MEEP's stored words at every compared step. Returns the differing components per compared
    step (empty lists when every word is MEEP's)."""
    folder = Path(capsule_dir)
    state = read_capsule(folder / "capsule.capsule", max_poles=max_poles)
    digests = read_digests(folder / "digests.json")
    currents = read_currents(folder / "currents.json", state, digests.n_steps)
    if digests.mode not in ("keep", "flush"):
        raise CapsuleRefusal(f"digests minted under subnormal mode {digests.mode!r}")
    run = LiftedRun(state, currents, subnormals=digests.mode)
    try:
        out = {}
        for step in digests.steps:
            run.advance(step - run.step_count)
            out[step] = compare_step(run, digests, step)
        return out
    finally:
        run.close()
