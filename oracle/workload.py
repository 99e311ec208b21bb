"""Step a synthetic workload on the oracle, with sources and flux planes (test infrastructure).

Takes the plain description a workload carries (grid shape, dt/dx, absorber faces, inverse
permittivity volumes, poles, deposit boxes with their current tables, flux planes) and
runs the reference driver's order (driver.py:3419-3596) on the C restatement, accumulating
each flux plane after every step as ``FluxMonitor.update`` does (dft.py:3031-3063): the four
tangential components sampled at cell centres (dft.py:352), E at t and H at t - dt/2 with
``t = step_count*dt`` after the increment, phase ``exp(i*omega*t)*dt/sqrt(2*pi)``.

Two accumulations are kept per plane: the reference's own precision (complex64 phase times
float32 sample added in float32, per part) and a float64 one, the yardstick for the engine's
float64 transform. The flux is ``sum Re(E1*conj(H2) - E2*conj(H1)) * dx**2`` over the plane
(dft.py:3107-3130 with unit interior weights).
"""

from __future__ import annotations

import ctypes
import math
from typing import Dict, List, Sequence

import numpy

from . import engine as oe

_IYEE_MASK = {"Ex": 1, "Ey": 2, "Ez": 4, "Hx": 6, "Hy": 5, "Hz": 3}  # bit a: Yee shift 1 on axis a
_AX = "xyz"


class OraclePlane:
    """A flux plane (the four tangential components, cell-centred), or, with ``components``, a
    DFT field region (reference DFTMonitor, dft.py:1960-2300): those components over the box,
    cell-centred or (``yee``) the stored word at each cell's own Yee site."""

    def __init__(self, normal: int, lo, cnt, omegas: numpy.ndarray, scale: float,
                 decimation: int = 1, components=None, yee: bool = False):
        self.normal = int(normal)
        self.decimation = int(decimation)  # accumulate when step_count % decimation == 0
        self.lo = tuple(int(v) for v in lo)
        self.cnt = tuple(int(v) for v in cnt)
        self.omegas = numpy.asarray(omegas, dtype=numpy.float64)
        self.scale = float(scale)
        self.yee = bool(yee)
        if components is None:
            t1, t2 = (self.normal + 1) % 3, (self.normal + 2) % 3
            self.components = ("E" + _AX[t1], "E" + _AX[t2], "H" + _AX[t1], "H" + _AX[t2])
        else:
            self.components = tuple(components)
        cells = int(numpy.prod(self.cnt))
        nf = self.omegas.size
        self.acc64 = numpy.zeros((len(self.components), nf, cells), dtype=numpy.complex128)
        self.acc32 = numpy.zeros((len(self.components), nf, cells), dtype=numpy.complex64)

    def sample(self, sim: oe.OracleSim, name: str, field: numpy.ndarray) -> numpy.ndarray:
        if self.yee:  # the stored words of the box, C order (x slowest), periodic indices
            idx = [numpy.arange(l, l + c) % n for l, c, n in zip(self.lo, self.cnt, field.shape)]
            return numpy.ascontiguousarray(field[numpy.ix_(*idx)], dtype=numpy.float32).reshape(-1)
        g = sim._grid()
        out = numpy.empty(self.cnt, dtype=numpy.float32)
        shift = (ctypes.c_int * 3)(*[(_IYEE_MASK[name] >> a) & 1 for a in range(3)])
        lo = (ctypes.c_int * 3)(*self.lo)
        cnt = (ctypes.c_int * 3)(*self.cnt)
        oe.lib().oracle_centre_sample(ctypes.byref(g), oe._fp(field), shift, lo, cnt, oe._fp(out))
        return out.reshape(-1)

    def update(self, sim: oe.OracleSim, time_value: float, reference_precision: bool = True):
        dt = sim.dt
        cells = int(numpy.prod(self.cnt))
        for q, name in enumerate(self.components):
            field = sim.get_E(name) if name[0] == "E" else sim.v["H" + name[1]] if sim.pml \
                else sim.v["B" + name[1]]
            s = self.sample(sim, name, numpy.ascontiguousarray(field))
            t = time_value if name[0] == "E" else time_value - 0.5 * dt
            phase = numpy.exp(1j * self.omegas * t) * self.scale
            ph64 = numpy.ascontiguousarray(phase.astype(numpy.complex128)).view(numpy.float64)
            acc = self.acc64[q]
            oe.lib().oracle_dft_accumulate_c128(
                acc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                ph64.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), oe._fp(s),
                ctypes.c_int(self.omegas.size), ctypes.c_int64(cells))
            if reference_precision:
                p32 = numpy.ascontiguousarray(phase.astype(numpy.complex64)).view(numpy.float32)
                a32 = self.acc32[q]
                oe.lib().oracle_dft_accumulate_c64(
                    a32.view(numpy.float32).ctypes.data_as(oe._FP), p32.ctypes.data_as(oe._FP),
                    oe._fp(s), ctypes.c_int(self.omegas.size), ctypes.c_int64(cells))

    def flux(self, dx: float, precision: str = "c128") -> numpy.ndarray:
        acc = self.acc64 if precision == "c128" else self.acc32.astype(numpy.complex128)
        e1, e2, h1, h2 = acc
        return numpy.sum(numpy.real(e1 * numpy.conj(h2) - e2 * numpy.conj(h1)), axis=1) * dx * dx


def meep_ladder(where_min: float, where_max: float, a: float):
    """The oracle's restatement of MEEP's monitor ladder on the centred grid (reference
    dft.py:604-690: vec2diel_floor/ceil, then compute_boundary_weights' case ladder; MEEP
    loop_in_chunks.cpp:233-297). Returns (first doubled site, per-site weights)."""
    lo_site = 1 + 2 * math.floor(where_min * a - 0.5)
    hi_site = 1 + 2 * math.ceil(where_max * a - 0.5)
    w0 = 1.0 - where_min * a + 0.5 * lo_site
    w1 = 1.0 + where_max * a - 0.5 * hi_site
    n = (hi_site - lo_site) // 2 + 1
    if hi_site - lo_site >= 6:
        ends = (w0 * w0 / 2, 1 - (1 - w0) ** 2 / 2, w1 * w1 / 2, 1 - (1 - w1) ** 2 / 2)
    elif hi_site - lo_site == 4:
        mid = 1 - (1 - w0) ** 2 / 2 - (1 - w1) ** 2 / 2
        ends = (w0 * w0 / 2, mid, w1 * w1 / 2, mid)
    elif where_min == where_max:
        ends = (w0, w1, w1, w0)
    elif hi_site - lo_site == 2:
        first = w0 * w0 / 2 - (1 - w1) ** 2 / 2
        last = w1 * w1 / 2 - (1 - w0) ** 2 / 2
        ends = (first, last, last, first)
    else:
        ends = (1.0, 1.0, 1.0, 1.0)
    # IVEC_LOOP_WEIGHT1x: site 0 -> s0, site 1 -> s1, last -> e0, second last -> e1, in that
    # order of precedence; interior sites 1
    w = numpy.ones(n)
    for q in range(n):
        for site, value in ((0, ends[0]), (1, ends[1]), (n - 1, ends[2]), (n - 2, ends[3])):
            if q == site and not 1 < q < n - 2:
                w[q] = value
                break
    return lo_site, w


def region_planes(shape, dx, center, size, invariant=(False, False, False)):
    """Oracle planes for a MEEP FluxRegion: [(lo, cnt, normal weight)], the transverse weights
    (C-order over cnt) and the normal axis. Cell i of an axis of n cells is centred at
    (i + 1/2 - n/2) * dx, MEEP's doubled site 2i + 1 - n."""
    a = 1.0 / dx
    normal = [d for d in range(3) if not invariant[d] and size[d] == 0][0]
    lo, cnt, tw, nplanes = [0, 0, 0], [1, 1, 1], [], []
    for d in range(3):
        if invariant[d]:
            tw.append(numpy.ones(1))
            continue
        site, w = meep_ladder(center[d] - size[d] / 2, center[d] + size[d] / 2, a)
        first = (site - 1 + shape[d]) // 2
        if d == normal:
            nplanes = [(first + q, w[q]) for q in range(w.size)]
            tw.append(numpy.ones(1))
            lo[d] = first
        else:
            lo[d], cnt[d] = first, w.size
            tw.append(w)
    weights = tw[0][:, None, None] * tw[1][None, :, None] * tw[2][None, None, :]
    out = []
    for index, w in nplanes:
        l = list(lo)
        l[normal] = index
        out.append((tuple(l), tuple(cnt), w))
    return out, weights, normal


def region_flux(oracle_planes, weights, normal_weights, area):
    """sum_planes w_n * sum_cells w * Re(E1 conj(H2) - E2 conj(H1)) * dA from the float64
    accumulations (dft.py:3107-3130)."""
    total = 0.0
    for p, wn in zip(oracle_planes, normal_weights):
        e1, e2, h1, h2 = p.acc64
        s = numpy.real(e1 * numpy.conj(h2) - e2 * numpy.conj(h1)) * weights.reshape(1, -1)
        total = total + wn * s.sum(axis=1) * area
    return total


def build(shape, dt, dx, faces, inv: Dict[str, numpy.ndarray], poles: Sequence, deposits,
          planes, invariant=(False, False, False), ln_reflection=None):
    """An ``OracleSim`` plus its planes. ``poles``: (frequency, gamma, drude, sigma dict);
    ``deposits``: (component, lo, cnt, amp_re, amp_im, current[, integrated]) with ``current``
    a (steps, 2) table whose row n is the word at n*dt or (n+1/2)*dt (the row
    ``floor(t/dt + 1/4)``), or a callable of the time; ``planes``: (normal, lo, cnt, omegas,
    scale)."""
    sim = oe.OracleSim(shape, dt, dx, invariant, faces, ln_reflection=ln_reflection)
    sim.set_inverse_epsilon(inv)
    for frequency, gamma, drude, sigma in poles:
        sim.add_pole(frequency, gamma, drude, sigma)
    nx, ny, nz = shape
    for component, lo, cnt, amp_re, amp_im, current, *rest in deposits:
        integrated = bool(rest[0]) if rest else False
        ii, jj, kk = numpy.meshgrid(*[numpy.arange(l, l + c) for l, c in zip(lo, cnt)],
                                    indexing="ij")
        flat = ((ii * ny + jj) * nz + kk).reshape(-1).astype(numpy.int64)
        if callable(current):
            cur = current
        else:
            table = numpy.asarray(current, dtype=numpy.float64).reshape(-1, 2)

            def cur(t, table=table):
                n = int(math.floor(t / dt + 0.25))
                return complex(table[n, 0], table[n, 1]) if 0 <= n < len(table) else 0j

        sim.add_deposit(oe.PointDeposit(component, flat,
                                        numpy.asarray(amp_re, dtype=numpy.float64).reshape(-1),
                                        numpy.asarray(amp_im, dtype=numpy.float64).reshape(-1),
                                        cur, integrated))
    out_planes = [OraclePlane(*p) for p in planes]
    return sim, out_planes


def run(sim: oe.OracleSim, planes: List[OraclePlane], steps: int,
        reference_precision: bool = True):
    for _ in range(steps):
        sim.step()
        for p in planes:
            if sim.step_count % p.decimation == 0:  # dft.py:3057
                p.update(sim, sim.time, reference_precision)
