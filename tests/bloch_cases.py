"""Complex-storage Bloch runs for the tests: one seeded description stepped by the oracle (two
real parts, ``oracle.engine.ComplexOracle``) and by the engine (two linked simulations).

A run is a small 3-D cell with a Bloch phase exp(i*theta) on x (and optionally y), a PML on
z, random inverse epsilon, an optional off-diagonal chi1inv row pair (Ex-Ey) and a Lorentz
pole, and seeded complex state. ``supercell`` lays the same run out as ``copies`` cells along
x with plain periodic wraps and Bloch-periodic data, the check that pins the connection.
"""

from __future__ import annotations

import cmath
import math
from dataclasses import dataclass, field
from typing import Dict, Optional, Tuple

import numpy

from oracle import engine, layer

COMPONENTS = "xyz"
DT, DX = 0.0625, 0.125


@dataclass
class BlochRun:
    shape: Tuple[int, int, int]
    thetas: Tuple[float, float, float]
    inv: Dict[str, numpy.ndarray]
    offdiag: Dict[str, Dict[str, numpy.ndarray]]
    pole_sigma: Optional[Dict[str, numpy.ndarray]]
    re: Dict[str, numpy.ndarray]
    im: Dict[str, numpy.ndarray]
    faces: Tuple = ((0, 0), (0, 0), (2, 2))
    kinds: Tuple[int, int, int] = field(init=False)
    phases: Tuple = field(init=False)

    def __post_init__(self):
        kinds, phases = [], []
        for theta in self.thetas:
            if theta is None:
                kinds.append(0)
                phases.append((1.0, 0.0))
                continue
            z = cmath.exp(1j * theta)
            kinds.append(3)
            phases.append((z.real, z.imag))
        self.kinds, self.phases = tuple(kinds), tuple(phases)


NAMES = [p + c for p in ("D", "B", "E", "H", "fu_B", "fu_D", "f_w_E", "f_w_H")
         for c in COMPONENTS]


def make(seed: int, shape=(8, 6, 6), thetas=(math.pi / 2, None, None), offdiag: bool = True,
         pole: bool = False) -> BlochRun:
    rng = numpy.random.default_rng(seed)

    def vol(lo, hi):
        return rng.uniform(lo, hi, shape).astype(numpy.float32)

    inv = {"E" + c: (1.0 / vol(1.0, 3.0).astype(numpy.float64)).astype(numpy.float32)
           for c in COMPONENTS}
    od = {}
    if offdiag:
        od = {"Ex": {"Ey": vol(-0.06, 0.06)}, "Ey": {"Ex": vol(-0.06, 0.06)}}
    sig = {"E" + c: vol(0.0, 1.0) for c in COMPONENTS} if pole else None
    re = {n: vol(-1.0, 1.0) for n in NAMES}
    im = {n: vol(-1.0, 1.0) for n in NAMES}
    return BlochRun(shape, tuple(thetas), inv, od, sig, re, im)


def _oracle_part(run: BlochRun, volumes, shape=None, inv=None, od=None, sig=None):
    shape = shape or run.shape
    sim = engine.OracleSim(shape, DT, DX, pml_faces=run.faces, ln_reflection=34.5)
    sim.set_inverse_epsilon(inv or run.inv, od if od is not None else run.offdiag)
    sim.enable_field_storage()
    if (sig if sig is not None else run.pole_sigma) is not None:
        sim.add_pole(0.5, 0.0625, False, sig if sig is not None else run.pole_sigma)
    for n, v in volumes.items():
        sim.v[n][...] = v
    return sim


def oracle(run: BlochRun) -> engine.ComplexOracle:
    return engine.ComplexOracle(_oracle_part(run, run.re), _oracle_part(run, run.im),
                                run.kinds, run.phases)


def hip(run: BlochRun):
    """The engine's two linked parts (HipSimulations), every word pasted as the oracle's."""
    from meep_gpu.engine import HipSimulation, link_parts
    from meep_gpu.grid import Grid
    from meep_gpu.pml import PML

    grid = Grid(resolution=1.0 / DX, cell_size=tuple(n * DX for n in run.shape), courant=DT / DX)
    assert grid.shape == run.shape
    pml = PML(grid, run.faces, ln_reflection=34.5)
    parts = []
    for volumes in (run.re, run.im):
        sim = HipSimulation(grid, pml)
        for c, v in run.inv.items():
            sim.set_inverse_epsilon(c, v)
        for row, partners in run.offdiag.items():
            for partner, v in partners.items():
                sim.set_offdiagonal(row, partner, v)
        if run.pole_sigma is not None:
            from meep_gpu.dispersion import LORENTZIAN, Susceptibility

            sim.add_pole(Susceptibility(0.5, 0.0625, LORENTZIAN), run.pole_sigma)
        for n, v in volumes.items():
            sim.upload(n, v)
        parts.append(sim)
    link_parts(parts[0], parts[1], run.kinds, run.phases)
    return parts


def _yee0(name: str, axis: int) -> bool:
    """Yee shift 0 along `axis` (E/D off their own axis, H/B and their auxiliaries on it)."""
    comp = "xyz".index(name[-1])
    electric = name[-2] in "ED"
    return (axis != comp) if electric else (axis == comp)


def supercell(run: BlochRun, copies: int):
    """The run as ``copies`` cells along x with plain periodic wraps (real and imaginary parts
    independent): data f(x + m*L) = exp(i*m*theta) f(x) in MEEP's stored window, where stored
    shift-0 row 0 of the Bloch cell is the point at the window top (= the next copy's row 0).
    Returns the two oracle parts."""
    theta = run.thetas[0]
    nx = run.shape[0]
    big = (nx * copies,) + run.shape[1:]

    def tile(v):
        return numpy.concatenate([v] * copies, axis=0)

    inv = {k: tile(v) for k, v in run.inv.items()}
    od = {r: {p: tile(v) for p, v in d.items()} for r, d in run.offdiag.items()}
    sig = {k: tile(v) for k, v in run.pole_sigma.items()} if run.pole_sigma else None
    re_big, im_big = {}, {}
    for name in NAMES:
        f = run.re[name].astype(numpy.complex128) + 1j * run.im[name].astype(numpy.complex128)
        out = numpy.empty(big, dtype=numpy.complex128)
        for m in range(copies):
            blk = f * cmath.exp(1j * m * theta)
            if _yee0(name, 0):
                # row 0 of this copy is the previous copy's window top
                blk[0] = f[0] * cmath.exp(1j * (m - 1) * theta)
            out[m * nx:(m + 1) * nx] = blk
        re_big[name] = out.real.astype(numpy.float32)
        im_big[name] = out.imag.astype(numpy.float32)
    a = _oracle_part(run, re_big, big, inv, od, sig)
    b = _oracle_part(run, im_big, big, inv, od, sig)
    return a, b


def step_pair(a, b, steps: int):
    for _ in range(steps):
        a.step()
        b.step()


def bloch_window(sim_re, sim_im, name: str, run: BlochRun) -> numpy.ndarray:
    """The supercell's first copy in the Bloch cell's stored window (complex)."""
    nx = run.shape[0]
    f = sim_re.v[name].astype(numpy.complex128) + 1j * sim_im.v[name].astype(numpy.complex128)
    out = f[:nx].copy()
    if _yee0(name, 0):
        out[0] = f[nx]  # the window top: the second copy's row 0
    return out
