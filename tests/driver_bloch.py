"""Driver-level complex storage: one seeded description built through ``FdtdDriver``
(``force_complex_fields``, ``k_point``, ``symmetry``, ``boundaries``) and the same description
stepped by the oracle's two real parts (``oracle.engine.ComplexOracle``)."""

from __future__ import annotations

import numpy

from meep_gpu.dispersion import LORENTZIAN, Susceptibility
from meep_gpu.driver import FdtdDriver
from meep_gpu.pml import normalize_faces
from meep_gpu.sources import GaussianSource, current_table
from oracle import engine as oe

RUNS = {
    # Bloch products on x and y, a z PML, an off-diagonal Ex-Ey row, a Lorentz pole, complex
    # electric and magnetic point sources and a finite electric sheet
    "xy_products": dict(cell=(1.6, 1.2, 2.0), k=(0.3, -0.45, 0.0), pml={"z": 2},
                        offdiag=True, pole=True, symmetry=(), boundaries=None),
    # the zone edge on x (the connection is a negation) beside a metallic y axis
    "x_edge_metallic_y": dict(cell=(1.6, 1.2, 1.6), k=(0.5 / 1.6, 0.0, 0.0), pml={"z": 2},
                              offdiag=False, pole=True, symmetry=(),
                              boundaries={"y": "metallic"}),
    # products on x and y with an electric polarization (integrated) source as well: each
    # part subtracts its own line of the dipole before update_E
    "xy_products_integrated": dict(cell=(1.6, 1.2, 2.0), k=(0.3, -0.45, 0.0), pml={"z": 2},
                                   offdiag=False, pole=False, symmetry=(), boundaries=None,
                                   integrated=True),
    # products on x and y, an off-diagonal Ex-Ey row and a polarization source: the row's
    # neighbour reads take the dipole-less words too
    "xy_products_offdiag_integrated": dict(cell=(1.6, 1.2, 2.0), k=(0.3, -0.45, 0.0),
                                           pml={"z": 2}, offdiag=True, pole=True, symmetry=(),
                                           boundaries=None,
                                           integrated=(("Ey", (0.13, -0.07, 0.21)),
                                                       ("Ex", (-0.27, 0.31, -0.13)))),
    # polarization sources on a folded axis: Ez one cell above the odd Y mirror (its image
    # row feeds the near ghost row) and Ey below the periodic fold's second mirror (the far
    # ghost row)
    "odd_y_mirror_integrated": dict(cell=(1.6, 2.4, 1.6), k=(0.7, 0.0, 0.0), pml={"z": 2},
                                    offdiag=False, pole=False, symmetry=(("Y", -1),),
                                    boundaries=None,
                                    integrated=(("Ez", (0.13, 0.2, 0.21)),
                                                ("Ey", (-0.31, 1.1, -0.17)))),
    # a product on x beside an odd mirror on y (the fold and the connection together)
    "x_product_odd_y_mirror": dict(cell=(1.6, 2.4, 1.6), k=(0.7, 0.0, 0.0), pml={"z": 2},
                                   offdiag=False, pole=False, symmetry=(("Y", -1),),
                                   boundaries=None),
}

SOURCES = (  # component, center, size, frequency, fwidth, amplitude
    ("Ez", (0.05, 0.1, -0.1), (0.0, 0.0, 0.0), 1.1, 0.5, 0.7 + 0.4j),
    ("Hx", (-0.2, 0.3, 0.3), (0.0, 0.0, 0.0), 0.9, 0.4, -0.3 + 1.1j),
    ("Ey", (0.1, -0.1, 0.2), (0.4, 0.0, 0.4), 1.3, 0.6, 1.0 + 0.0j),
)


def driver(name: str, seed: int = 11, complex_fields: bool = True, **overrides) -> FdtdDriver:
    """The run ``name`` (its entries replaced by ``overrides``) on complex storage, or on real
    storage with ``complex_fields=False``."""
    spec = dict(RUNS[name], **overrides)
    rng = numpy.random.default_rng(seed)
    drv = FdtdDriver(cell_size=spec["cell"], resolution=5.0, force_complex_fields=complex_fields,
                     symmetry=spec["symmetry"], k_point=spec["k"],
                     boundaries=spec["boundaries"], max_steps=400)
    shape = drv.grid.shape
    eps = {c: rng.uniform(1.0, 3.0, shape) for c in ("Ex", "Ey", "Ez")}
    od = None
    if spec["offdiag"]:
        od = {"Ex": {"Ey": rng.uniform(-0.06, 0.06, shape)},
              "Ey": {"Ex": rng.uniform(-0.06, 0.06, shape)}}
    drv.set_epsilon_components(eps, od)
    if spec["pole"]:
        drv.add_susceptibility(Susceptibility(1.2, 0.1, LORENTZIAN),
                               {c: rng.uniform(0.0, 0.8, shape).astype(numpy.float32)
                                for c in ("Ex", "Ey", "Ez")})
    drv.setup_pml(spec["pml"])
    for comp, center, size, f, fw, amp in SOURCES:
        if spec["symmetry"]:
            continue  # placement under a mirror is refused; boxes go in below
        drv.add_volume_source(comp, center, size, GaussianSource(f, fw), amplitude=amp)
    integrated = spec.get("integrated")
    if integrated is True:  # an electric polarization source off every cell face
        integrated = (("Ey", (0.13, -0.07, 0.21)),)
    for comp, center in integrated or ():
        drv.add_volume_source(comp, center, (0.0, 0.0, 0.0), GaussianSource(1.2, 0.5),
                              amplitude=0.6 - 0.8j, is_integrated=True)
    if spec.get("flux"):  # a z-normal flux plane between the sources and the absorber
        drv.add_flux_monitor(2, (0, 0, shape[2] - 3), (shape[0], shape[1], 1),
                             frequencies=[0.9, 1.1, 1.3])
    if spec["symmetry"]:
        from meep_gpu.sources import DepositBox

        for comp, lo, fn, amp in (("Ez", (3, 4, 5), GaussianSource(1.1, 0.5), 0.7 + 0.4j),
                                  ("Hy", (5, 2, 3), GaussianSource(0.9, 0.4), -0.3 + 1.1j)):
            drv.add_source(DepositBox(comp, lo, (1, 1, 1), numpy.array([amp.real * 125.0]),
                                      numpy.array([amp.imag * 125.0])), fn)
    return drv


def oracle(drv: FdtdDriver, steps: int) -> oe.ComplexOracle:
    """The driver's configuration as the oracle's two parts (read from the driver's own
    tables: inverse epsilon, off-diagonal rows, poles, deposit boxes and folds)."""
    grid = drv.grid
    faces = normalize_faces(drv._pml_spec, grid) if drv._pml_spec else ((0, 0),) * 3
    parts = []
    for part in range(2):
        sim = oe.OracleSim(grid.shape, grid.dt, grid.dx, grid.invariant, faces)
        # a folded metallic axis stores no wall: it is a fold
        sim.set_metallic(tuple(not w and f is None for w, f in zip(drv.wraps, grid.folds)))
        sim.set_folding(grid.folds)
        sim.set_inverse_epsilon(drv._inv, drv._offdiag)
        for term, sigma in drv._poles:
            sim.add_pole(term.frequency, term.gamma, term.kind != LORENTZIAN, sigma)
        nx, ny, nz = grid.shape
        for box, fn in drv._sources:
            ii, jj, kk = numpy.meshgrid(*[numpy.arange(l, l + c) for l, c in
                                          zip(box.lo, box.cnt)], indexing="ij")
            flat = ((ii * ny + jj) * nz + kk).reshape(-1).astype(numpy.int64)
            table = current_table(fn, grid.dt, steps + 1, box.magnetic)

            def cur(t, table=table, dt=grid.dt):
                n = int(numpy.floor(t / dt + 0.25))
                return complex(table[n, 0], table[n, 1]) if 0 <= n < len(table) else 0j

            if box.integrated:  # the dipole, evaluated at the oracle's own time
                cur = fn.dipole

            re = numpy.asarray(box.amp_re, dtype=numpy.float64).reshape(-1)
            im = numpy.asarray(box.amp_im, dtype=numpy.float64).reshape(-1)
            if part == 1:
                re, im = im, -re
            sim.add_deposit(oe.PointDeposit(box.component, flat, re, im, cur, box.integrated))
        parts.append(sim)
    from meep_gpu.capsule import bloch_connection

    phases = drv.bloch_phases
    kinds = [0 if z is None else bloch_connection(z) for z in phases]
    return oe.ComplexOracle(parts[0], parts[1], kinds,
                            [(1.0, 0.0) if z is None else (z.real, z.imag) for z in phases])


def oracle_planes(drv: FdtdDriver):
    """Per flux monitor of the driver, the oracle's plane on each part (real, imaginary)."""
    from oracle import workload as ow

    return [tuple(ow.OraclePlane(m.normal, m.lo, m.cnt, m.omegas, m.scale) for _ in range(2))
            for m in drv.flux_monitors]


def step_oracle(ref: oe.ComplexOracle, planes, steps: int) -> None:
    for _ in range(steps):
        ref.step()
        for pair in planes:
            for plane, part in zip(pair, ref.parts):
                plane.update(part, part.time, reference_precision=False)


def complex_flux(pair, dx: float) -> numpy.ndarray:
    """The flux of the complex field from the two parts' transforms (each complex itself)."""
    e1, e2, h1, h2 = pair[0].acc64 + 1j * pair[1].acc64
    return numpy.sum(numpy.real(e1 * numpy.conj(h2) - e2 * numpy.conj(h1)), axis=1) * dx * dx
