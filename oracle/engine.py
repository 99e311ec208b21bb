"""ctypes driver for the C restatement (ORACLE / test infrastructure only).

``OracleSim`` holds the state the reference's ``Fields`` holds for a real-storage periodic
Cartesian run (reference fields.py:605-760): D and B always; E when stored (an absorber, a
susceptibility or an off-diagonal row); H, fu_B, fu_D, f_w_H, f_w_E under an absorber; P and
P_prev per pole. ``step()`` runs the reference driver's order (driver.py:3419-3596):
step_B, magnetic deposits, update_H, step_D, electric deposits, update_E, update_P.

Volumes are NumPy float32 C-order (nx, ny, nz) arrays, the reference's own layout.
"""

from __future__ import annotations

import contextlib
import ctypes
import os
import subprocess
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy

from . import layer

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_FP = ctypes.POINTER(ctypes.c_float)
_U8P = ctypes.POINTER(ctypes.c_uint8)


class _OGrid(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3), ("present", ctypes.c_int * 3), ("dtdx", ctypes.c_float),
                ("pml", ctypes.c_int),
                ("kms", (_FP * 3) * 2), ("sinv", (_FP * 3) * 2), ("kps", (_FP * 3) * 2),
                ("mem", (_U8P * 3) * 2), ("metal", ctypes.c_int * 3),
                ("fold", ctypes.c_int * 3), ("fphase", ctypes.c_int * 3),
                ("pkind", ctypes.c_int * 3), ("ph", (ctypes.c_float * 2) * 3),
                ("part", ctypes.c_int)]


MAX_POLES = 8  # ORACLE_MAX_POLES of the C restatement


class _OElectric(ctypes.Structure):
    _fields_ = [("npol", ctypes.c_int), ("p", (_FP * 3) * MAX_POLES), ("u", (_FP * 3) * 3),
                ("offdiag", ctypes.c_int)]


def build() -> str:
    """Compile the restatement (make in this directory); returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.oracle_deposit.argtypes = [_FP, ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, _FP]
        _lib.oracle_update_p.argtypes = [ctypes.c_void_p, ctypes.c_int, _FP, _FP, _FP, _FP,
                                         ctypes.c_float, _FP, _FP, _FP]
        _lib.oracle_set_flush.argtypes = [ctypes.c_int]
        _lib.oracle_flush_state.restype = ctypes.c_int
    return _lib


@contextlib.contextmanager
def flushing(on: bool = True):
    """Run the restatement (and this thread's NumPy) with x86 FTZ|DAZ set, then clear it."""
    L = lib()
    L.oracle_set_flush(1 if on else 0)
    try:
        yield
    finally:
        L.oracle_set_flush(0)


def _fp(array: Optional[numpy.ndarray]):
    if array is None:
        return ctypes.cast(None, _FP)
    assert array.dtype == numpy.float32 and array.flags.c_contiguous
    return array.ctypes.data_as(_FP)


COMPONENTS = ("x", "y", "z")


@dataclass
class Pole:
    """One Lorentz/Drude term: frequency and gamma in MEEP units, sigma per component."""
    frequency: float
    gamma: float
    drude: bool
    sigma: Dict[str, object]          # 'Ex'.. -> float32 volume or scalar
    P: Dict[str, numpy.ndarray] = field(default_factory=dict)
    P_prev: Dict[str, numpy.ndarray] = field(default_factory=dict)
    consts: Optional[numpy.ndarray] = None


@dataclass
class PointDeposit:
    """A source's deposit table: component, flat C-order cell indices, amplitudes, current(t)."""
    component: str                    # 'Ex'.. for electric (deposits into D), 'Hx'.. magnetic
    index: numpy.ndarray              # int64 flat indices
    amp_re: numpy.ndarray
    amp_im: numpy.ndarray
    current: object                   # callable t -> complex (the dipole, when integrated)
    #: MEEP's integrated (polarization) source: never deposited; its dipole is subtracted
    #: from the copy of D (B) the constitutive update reads (reference sources.py:1697-1760)
    integrated: bool = False

    @property
    def magnetic(self) -> bool:
        return self.component[0] in "HB"


class OracleSim:
    """A real-storage periodic Cartesian run of the reference array path, on the CPU."""

    def __init__(self, shape: Sequence[int], dt: float, dx: float,
                 invariant: Sequence[bool] = (False, False, False),
                 pml_faces=((0, 0), (0, 0), (0, 0)), pml_order: int = 2,
                 ln_reflection: Optional[float] = None):
        self.shape = tuple(int(n) for n in shape)
        self._layer_args = (pml_order, ln_reflection)
        self.dt = float(dt)
        self.dx = float(dx)
        self.invariant = tuple(bool(v) for v in invariant)
        self.faces = tuple(tuple(f) for f in pml_faces)
        self.pml = any(lo > 0 or hi > 0 for lo, hi in self.faces)
        self.time = 0.0
        self.step_count = 0
        self.h_separated = False
        zeros = lambda: numpy.zeros(self.shape, dtype=numpy.float32)
        self.v: Dict[str, numpy.ndarray] = {}
        for c in COMPONENTS:
            self.v["D" + c] = zeros()
            self.v["B" + c] = zeros()
        self.inv = {("E" + c): numpy.ones(self.shape, dtype=numpy.float32) for c in COMPONENTS}
        self.offdiag: Dict[str, Dict[str, numpy.ndarray]] = {}
        self.poles: List[Pole] = []
        self.deposits: List[PointDeposit] = []
        self.stores_E = False
        self.metal = (False, False, False)  # metallic (PEC) axes
        # per axis: None, or a mirror fold (phase, kind, far row): kind 1 over metallic
        # boundaries, 2 over periodic ones with the far ghost imaging stored row `far row`
        self.folds: Tuple[Optional[Tuple[int, int, Optional[int]]], ...] = (None, None, None)
        # Bloch connections per axis (0 none, 1 copy, 2 negate, 3 a product by ph[axis]) and,
        # for complex storage, which part this is and the other part's run
        self.bloch_kinds = (0, 0, 0)
        self.bloch_ph = ((1.0, 0.0),) * 3
        self.part = 0
        self.other: Optional["OracleSim"] = None
        self.condfac: Dict[str, numpy.ndarray] = {}  # "Dx".."Bz": MEEP's conductivity words
        self.condinv: Dict[str, numpy.ndarray] = {}
        # coefficient vectors and membership lines [half][axis]
        self._vec = {}
        for half in (0, 1):
            for axis in range(3):
                n = self.shape[axis]
                faces = self.faces[axis] if not self.invariant[axis] else (0, 0)
                kms, sinv, kps = layer.axis_coefficients(n, faces, 0.5 * half, self.dt, self.dx,
                                                         pml_order, ln_reflection)
                mem = (numpy.zeros(n, dtype=bool) if self.invariant[axis]
                       else layer.membership_line(n, faces, bool(half)))
                self._vec[(half, axis)] = (kms, sinv, kps, mem.astype(numpy.uint8))
        if self.pml:
            self.enable_field_storage()
            for c in COMPONENTS:
                for name in ("H" + c, "fu_B" + c, "fu_D" + c, "f_w_E" + c, "f_w_H" + c):
                    self.v[name] = zeros()

    # --- configuration -------------------------------------------------------------------
    def enable_field_storage(self):
        if not self.stores_E:
            for c in COMPONENTS:
                self.v["E" + c] = numpy.zeros(self.shape, dtype=numpy.float32)
            self.stores_E = True

    def set_inverse_epsilon(self, inv: Dict[str, numpy.ndarray],
                            offdiag: Optional[Dict[str, Dict[str, numpy.ndarray]]] = None):
        self.inv = {k: numpy.ascontiguousarray(v, dtype=numpy.float32) for k, v in inv.items()}
        self.offdiag = {}
        for row, partners in (offdiag or {}).items():
            for partner, vol in partners.items():
                vol = numpy.ascontiguousarray(vol, dtype=numpy.float32)
                if numpy.any(vol):
                    self.offdiag.setdefault(row, {})[partner] = vol
        if self.offdiag:
            self.enable_field_storage()

    def add_pole(self, frequency: float, gamma: float, drude: bool, sigma: Dict[str, object]):
        pole = Pole(frequency, gamma, drude, dict(sigma))
        pole.consts = numpy.array(layer.ade_constants(frequency, gamma, drude, self.dt),
                                  dtype=numpy.float32)
        for c in COMPONENTS:
            s = pole.sigma["E" + c]
            trivial = (numpy.ndim(s) == 0 and float(s) == 0.0) or \
                (numpy.ndim(s) > 0 and not numpy.any(numpy.asarray(s) != 0.0))
            if trivial:
                continue
            if numpy.ndim(s) > 0:
                pole.sigma["E" + c] = numpy.ascontiguousarray(s, dtype=numpy.float32)
            pole.P["E" + c] = numpy.zeros(self.shape, dtype=numpy.float32)
            pole.P_prev["E" + c] = numpy.zeros(self.shape, dtype=numpy.float32)
        self.poles.append(pole)
        self.enable_field_storage()
        return pole

    def set_conductivity(self, sigma: Dict[str, numpy.ndarray],
                         cells: Optional[Dict[str, numpy.ndarray]] = None):
        """D and B conductivities by component ("Dx".."Bz"), as fields.set_d_conductivity /
        set_b_conductivity form their words (reference fields.py:775-870): condfac =
        1 - sigma*dt2 in float32 with dt2 = float32(dt)*0.5, condinv = 1/(1 + (sigma*dt)*0.5)
        in double from the float32 word, rounded once. An all-zero volume is no conductivity.
        Under an absorber each conductive component gets an f_cond history (fields.py:754).
        ``cells[name]``, a bool volume, marks the cells whose chunk keeps the array; the others
        step the lossless forms (structure.cpp:883-900), held as a condinv word of 0."""
        dt2 = numpy.float32(numpy.float32(self.dt) * numpy.float32(0.5))
        for name, vol in sigma.items():
            vol = numpy.ascontiguousarray(vol, dtype=numpy.float32)
            if not numpy.any(vol):
                continue
            self.condfac[name] = (numpy.float32(1.0) - vol * dt2).astype(numpy.float32)
            wide = vol.astype(numpy.float64)
            self.condinv[name] = (1.0 / (1.0 + (wide * self.dt) * 0.5)).astype(numpy.float32)
            if cells is not None and name in cells:
                self.condinv[name][~numpy.asarray(cells[name], dtype=bool)] = 0.0
            if self.pml:
                self.v["f_cond_" + name] = numpy.zeros(self.shape, dtype=numpy.float32)

    def set_metallic(self, axes: Sequence[bool]):
        """Metallic (PEC) walls on the given axes (stepping.py:140-146): zero ghosts past the
        wall, the wall's own samples unstepped and held at zero. A metallic axis does not wrap
        (grid.py:1267-1289), so its layer is graded from the walls without a wrap of the
        integer row 0 and that row (the wall) is owned by no chunk (pml.py:578-582)."""
        self.metal = tuple(bool(a) for a in axes)
        order, ln_r = self._layer_args
        for axis in range(3):
            if not self.metal[axis] or self.invariant[axis]:
                continue
            for half in (0, 1):
                n = self.shape[axis]
                kms, sinv, kps = layer.axis_coefficients(n, self.faces[axis], 0.5 * half,
                                                         self.dt, self.dx, order, ln_r,
                                                         wraps=False)
                mem = layer.membership_line(n, self.faces[axis], bool(half), wraps=False)
                self._vec[(half, axis)] = (kms, sinv, kps, mem.astype(numpy.uint8))

    def set_folding(self, folds):
        """Mirror folds (grid.py:404-575): per axis None or ``(phase, kind, far_row)``, kind 1
        over metallic boundaries and 2 over periodic ones. A folded axis does not wrap, so its
        layer is graded from the window top (``2 * owned`` half cells: pml.py:600-611) without
        a wrap of row 0, and the membership line is ``one_rank_line`` on the stored count
        (chunk_map.py:180-233)."""
        self.folds = tuple(folds)
        order, ln_r = self._layer_args
        for axis, fold in enumerate(self.folds):
            if fold is None or self.invariant[axis]:
                continue
            n = self.shape[axis]
            owned = n - 1 if fold[1] == 2 else n
            for half in (0, 1):
                kms, sinv, kps = layer.axis_coefficients(n, self.faces[axis], 0.5 * half,
                                                         self.dt, self.dx, order, ln_r,
                                                         wraps=False, wall_high=2 * owned)
                mem = layer.membership_line(n, self.faces[axis], bool(half), wraps=False)
                self._vec[(half, axis)] = (kms, sinv, kps, mem.astype(numpy.uint8))

    def fill_folds(self, field: str, far: bool):
        """The ghost planes of the folded axes, after the sub-step's sources.

        Near (``far`` False; fill_symmetry_bc_D/_B, stepping.py:2004-2063): stored cell 0 of
        every component with Yee shift 0 on a folded axis takes ``parity * cell 2``, the axes
        in x, y, z order. Far (fill_folded_far_ghosts_*, :2083-2145): on a folded periodic
        axis the last slot of every shift-1 component takes ``parity * cell far_row``.
        The parity is MEEP's phase_shift (fields.mirror_parity, fields.py:121): a vector
        (E, D) is odd about the plane its own axis is normal to and even otherwise, a
        pseudovector (H, B) the opposite, both times the plane's declared phase."""
        for axis, fold in enumerate(self.folds):
            if fold is None or self.invariant[axis]:
                continue
            phase, kind, far_row = fold
            if far and kind != 2:
                continue
            for c, name in enumerate(COMPONENTS):
                shift0 = (axis == c) if field == "B" else (axis != c)
                if shift0 == far:
                    continue
                own = axis == c
                parity = phase * ((1 if own else -1) if field == "B" else (-1 if own else 1))
                vol = self.v[field + name]
                dst = [slice(None)] * 3
                src = [slice(None)] * 3
                dst[axis] = -1 if far else 0
                src[axis] = far_row if far else 2
                vol[tuple(dst)] = vol[tuple(src)] * numpy.float32(parity)

    def set_layer_tables(self, vectors, members):
        """Paste a layer table as lifted (``{(half, axis): (kms, sinv, kps)}`` float32 and
        ``{(half, axis): bool line}``) in place of the one this run graded itself."""
        if not self.pml:
            raise ValueError("this run has no absorber to paste a table into")
        for key, (kms, sinv, kps) in vectors.items():
            mem = numpy.asarray(members[key], dtype=numpy.uint8)
            self._vec[key] = tuple(numpy.ascontiguousarray(v, dtype=numpy.float32)
                                   for v in (kms, sinv, kps)) + (numpy.ascontiguousarray(mem),)

    def zero_metal(self, field: str):
        """zero_metal_D / zero_metal_B (stepping.py:3094-3183): stored cell 0 of every
        component whose Yee shift on a metallic axis is 0 (B_c along c, D_c along the other two)
        is written zero, after the sub-step's sources."""
        for a in range(3):
            if not self.metal[a]:
                continue
            for c, name in enumerate(COMPONENTS):
                on_wall = (a == c) if field == "B" else (a != c)
                if on_wall:
                    index = [slice(None)] * 3
                    index[a] = 0
                    self.v[field + name][tuple(index)] = 0.0

    def _conductive(self, field: str):
        cf = (_FP * 3)(*[_fp(self.condfac.get(field + c)) for c in COMPONENTS])
        ci = (_FP * 3)(*[_fp(self.condinv.get(field + c)) for c in COMPONENTS])
        fc = self._triple("f_cond_" + field)
        return cf, ci, fc

    def add_deposit(self, deposit: PointDeposit):
        self.deposits.append(deposit)

    # --- views ----------------------------------------------------------------------------
    def stored_volumes(self) -> Dict[str, numpy.ndarray]:
        out = dict(self.v)
        for k, pole in enumerate(self.poles):
            for c, a in pole.P.items():
                out[f"P[{k}].{c}"] = a
            for c, a in pole.P_prev.items():
                out[f"P_prev[{k}].{c}"] = a
        return out

    def get_E(self, component: str) -> numpy.ndarray:
        if self.stores_E:
            return self.v[component]
        return (self.inv[component] * self.v["D" + component[1]]).astype(numpy.float32)

    # --- the C structures -----------------------------------------------------------------
    def _grid(self) -> _OGrid:
        g = _OGrid()
        for a in range(3):
            g.n[a] = self.shape[a]
            g.present[a] = 0 if self.invariant[a] else 1
        g.dtdx = numpy.float32(self.dt / self.dx)
        g.pml = 1 if self.pml else 0
        for a in range(3):
            g.metal[a] = 1 if self.metal[a] else 0
            g.fold[a] = 0 if self.folds[a] is None else self.folds[a][1]
            g.fphase[a] = 0 if self.folds[a] is None else int(self.folds[a][0])
            g.pkind[a] = int(self.bloch_kinds[a])
            g.ph[a][0] = float(numpy.float32(self.bloch_ph[a][0]))
            g.ph[a][1] = float(numpy.float32(self.bloch_ph[a][1]))
        g.part = int(self.part)
        for half in (0, 1):
            for a in range(3):
                kms, sinv, kps, mem = self._vec[(half, a)]
                g.kms[half][a] = _fp(kms)
                g.sinv[half][a] = _fp(sinv)
                g.kps[half][a] = _fp(kps)
                g.mem[half][a] = mem.ctypes.data_as(_U8P)
        return g

    def _triple(self, prefix: str):
        arr = (_FP * 3)()
        for a, c in enumerate(COMPONENTS):
            arr[a] = _fp(self.v.get(prefix + c))
        return arr

    def _electric(self) -> _OElectric:
        if len(self.poles) > MAX_POLES:
            raise ValueError(f"{len(self.poles)} poles; the restatement holds {MAX_POLES}")
        x = _OElectric()
        x.npol = len(self.poles)
        for k, pole in enumerate(self.poles):
            for a, c in enumerate(COMPONENTS):
                x.p[k][a] = _fp(pole.P.get("E" + c))
        x.offdiag = 1 if self.offdiag else 0
        for r, rc in enumerate(COMPONENTS):
            for p, pc in enumerate(COMPONENTS):
                x.u[r][p] = _fp(self.offdiag.get("E" + rc, {}).get("E" + pc))
        return x

    # --- sub-steps ------------------------------------------------------------------------
    def _e_sources(self):
        if self.stores_E:
            return self._triple("E"), []
        arrays = [self.get_E("E" + c) for c in COMPONENTS]
        return (_FP * 3)(*[_fp(a) for a in arrays]), arrays

    def step_B(self):
        L = lib()
        g = self._grid()
        fu = self._triple("fu_B") if self.pml else None
        src, keep = self._e_sources()
        other, keep_o = self.other._e_sources() if self.other is not None else (None, [])
        L.oracle_step_curl(ctypes.byref(g), 0, self._triple("B"), src, fu, *self._conductive("B"),
                           other)

    def step_D(self):
        L = lib()
        g = self._grid()
        fu = self._triple("fu_D") if self.pml else None
        other = None
        if self.other is not None:
            other = self.other._triple("H") if self.other.pml else self.other._triple("B")
        L.oracle_step_curl(ctypes.byref(g), 1, self._triple("D"),
                           self._triple("H") if self.pml else self._triple("B"), fu,
                           *self._conductive("D"), other)

    def update_H(self):
        if not self.pml:
            return
        g = self._grid()
        lib().oracle_update_h(ctypes.byref(g), self._triple("H"), self._triple("B"),
                              self._triple("f_w_H"), 1 if self.h_separated else 0)
        self.h_separated = True

    def update_E(self):
        if not self.pml and not self.stores_E:
            return
        g = self._grid()
        x = self._electric()
        inv = (_FP * 3)(*[_fp(self.inv["E" + c]) for c in COMPONENTS])
        fw = self._triple("f_w_E") if self.pml else None
        xo = ctypes.byref(self.other._electric()) if self.other is not None else None
        do = self.other._triple("D") if self.other is not None else None
        lib().oracle_update_e(ctypes.byref(g), ctypes.byref(x), self._triple("E"),
                              self._triple("D"), inv, fw, xo, do)

    def update_P(self):
        g = self._grid()
        for pole in self.poles:
            for c in COMPONENTS:
                name = "E" + c
                if name not in pole.P:
                    continue
                s = pole.sigma[name]
                vol = s if numpy.ndim(s) > 0 else None
                scalar = 0.0 if vol is not None else float(numpy.float32(s))
                p, pp = pole.P[name], pole.P_prev[name]
                lib().oracle_update_p(ctypes.byref(g), COMPONENTS.index(c),
                                      pole.consts.ctypes.data_as(_FP), _fp(p), _fp(pp), _fp(vol),
                                      ctypes.c_float(scalar), _fp(self.v[name]),
                                      _fp(self.v.get("f_w_" + name)), _fp(pp))
                pole.P[name], pole.P_prev[name] = pp, p

    def deposit(self, magnetic: bool, time_value: float, integrated: bool = False):
        """The sources' deposit into D (B); with ``integrated``, the integrated sources' dipole
        subtracted instead, with no dt and no conductivity (``f_minus_p``)."""
        for dep in self.deposits:
            if dep.magnetic != magnetic or dep.integrated != integrated:
                continue
            cur = complex(dep.current(time_value))
            if (numpy.copysign(1.0, cur.real) > 0 and cur.real == 0.0
                    and numpy.copysign(1.0, cur.imag) > 0 and cur.imag == 0.0):
                continue
            name = ("B" if magnetic else "D") + dep.component[1]
            target = self.v[name]
            idx = numpy.ascontiguousarray(dep.index, dtype=numpy.int64)
            condinv = self.condinv.get(name)
            if condinv is not None:
                condinv = numpy.ascontiguousarray(condinv).reshape(-1)
            lib().oracle_deposit(_fp(target.reshape(-1)),
                                 idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                 numpy.ascontiguousarray(dep.amp_re, dtype=numpy.float64)
                                 .ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 numpy.ascontiguousarray(dep.amp_im, dtype=numpy.float64)
                                 .ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                 int(idx.size), cur.real, cur.imag,
                                 1.0 if integrated else self.dt,
                                 None if integrated else _fp(condinv))

    def f_minus_p_begin(self, magnetic: bool, time_value: float):
        """MEEP's ``f_minus_p`` in place (reference sources.py:1697-1747, update_eh.cpp:66-140):
        the words at the integrated sources' cells are saved, the dipole subtracted from them
        and the wall clear run again; returns what ``f_minus_p_end`` restores, or None."""
        deps = [d for d in self.deposits if d.integrated and d.magnetic == magnetic]
        if not deps:
            return None
        name = "B" if magnetic else "D"
        saved = []
        for c in COMPONENTS:
            cells = [numpy.asarray(d.index, dtype=numpy.int64) for d in deps
                     if d.component[1] == c]
            if cells:
                flat = numpy.unique(numpy.concatenate(cells))
                saved.append((name + c, flat, self.v[name + c].reshape(-1)[flat].copy()))
        self.deposit(magnetic, time_value, integrated=True)
        self._boundary_passes(name)
        return saved

    def _boundary_passes(self, family: str):
        """The family's boundary passes run again over ``f_minus_p`` (reference
        sources.py:1749 ``_boundary_passes``): the mirror ghosts from their images, the walls
        cleared, the far ghosts of periodic folds (each a no-op without folds or walls)."""
        self.fill_folds(family, far=False)
        self.zero_metal(family)
        self.fill_folds(family, far=True)

    def f_minus_p_end(self, saved):
        if saved is None:
            return
        for name, flat, words in saved:
            self.v[name].reshape(-1)[flat] = words
        self._boundary_passes(name[0])

    def step(self):
        """The reference driver's step (driver.py:3524-3590)."""
        self.step_B()
        self.deposit(True, self.time)
        self.fill_folds("B", far=False)
        self.zero_metal("B")
        self.fill_folds("B", far=True)
        held = self.f_minus_p_begin(True, self.time + 0.5 * self.dt)
        self.update_H()
        if self.pml:  # a stored H: B takes its words back at once
            self.f_minus_p_end(held)
            held = None
        self.step_D()
        self.f_minus_p_end(held)  # without a stored H the D curl read f_minus_p as H
        self.deposit(False, self.time + 0.5 * self.dt)
        self.fill_folds("D", far=False)
        self.zero_metal("D")
        self.fill_folds("D", far=True)
        held = self.f_minus_p_begin(False, self.time + self.dt)
        self.update_E()
        self.f_minus_p_end(held)
        self.update_P()
        self.step_count += 1
        self.time = self.step_count * self.dt


class ComplexOracle:
    """Complex field storage as MEEP steps it (part_products.py:1-40): two real parts, each
    stepped with the same real coefficients (MEEP's DOCMP loops), coupled only where a Bloch
    connection forms a product by a phase that is neither +1 nor -1. ``re`` and ``im`` are
    OracleSims set up alike (materials, layer, folds, poles); a source's imaginary line
    ``(amp_re*c_im) + (amp_im*c_re)`` is the real line of amplitudes ``(amp_im, -amp_re)``
    (sources.py:1357-1482, step.cpp:309-314), so ``add_deposit`` gives the imaginary part
    those. Every sub-step runs for both parts before the next starts."""

    def __init__(self, re: OracleSim, im: OracleSim, kinds, phases):
        self.re, self.im = re, im
        for part, (sim, other) in enumerate(((re, im), (im, re))):
            sim.part = part
            sim.other = other
            sim.bloch_kinds = tuple(int(k) for k in kinds)
            sim.bloch_ph = tuple((float(a), float(b)) for a, b in phases)

    @property
    def parts(self):
        return (self.re, self.im)

    @property
    def step_count(self) -> int:
        return self.re.step_count

    def add_deposit(self, deposit: PointDeposit):
        self.re.add_deposit(deposit)
        self.im.add_deposit(PointDeposit(
            deposit.component, deposit.index, numpy.asarray(deposit.amp_im, dtype=numpy.float64),
            -numpy.asarray(deposit.amp_re, dtype=numpy.float64), deposit.current,
            integrated=deposit.integrated))

    def step(self):
        """One step of both parts. An electric polarization (integrated) source writes each
        part's ``f_minus_p`` around update_E (reference sources.py:1697-1747: the real line
        from the real part, the imaginary line from the imaginary part, as a deposit splits)."""
        re, im = self.re, self.im
        if any(d.integrated and d.magnetic for d in re.deposits):
            raise NotImplementedError("a magnetic polarization source on complex storage")
        for s in (re, im):
            s.step_B()
        for s in (re, im):
            s.deposit(True, s.time)
            s.fill_folds("B", far=False)
            s.zero_metal("B")
            s.fill_folds("B", far=True)
        for s in (re, im):
            s.update_H()
        for s in (re, im):
            s.step_D()
        for s in (re, im):
            s.deposit(False, s.time + 0.5 * s.dt)
            s.fill_folds("D", far=False)
            s.zero_metal("D")
            s.fill_folds("D", far=True)
        held = [s.f_minus_p_begin(False, s.time + s.dt) for s in (re, im)]
        for s in (re, im):
            s.update_E()
        for s, h in zip((re, im), held):
            s.f_minus_p_end(h)
        for s in (re, im):
            s.update_P()
        for s in (re, im):
            s.step_count += 1
            s.time = s.step_count * s.dt
