"""The driver's complex storage, Bloch phases, mirrors and boundaries on the CPU: MEEP's
``eikna`` rule against the phase MEEP itself recorded in the holey-waveguide capsule, the
zone-edge branch, and every pairing the reference refuses."""

import math
import os

import numpy
import pytest

from meep_gpu.driver import FdtdDriver, bloch_phase
from meep_gpu.sources import DepositBox, GaussianSource

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOLEY = os.path.join(ROOT, "tests", "golden", "meep_identity", "keep",
                     "t.test_holey_wvg_bands.TestHoleyWvgBands.test_fields_at_kx",
                     "capsule.capsule")


def test_phase_equals_meep_s_recorded_eikna():
    """MEEP's own Bloch factor read from its memory (the capsule's grid) for k = 3.5 on a
    20-cell axis at resolution 20: -1 plus the rounding residue of sin(7*pi) in the imaginary
    part, which the zone-edge branch must not absorb (k * n = 70, not resolution / 2)."""
    from meep_gpu import capsule

    grid = capsule.read_capsule_document(HOLEY)[1]["fields"]["grid"]
    recorded = complex(grid["bloch_phases"][0])
    k = float(grid["k_point"][0])
    drv = FdtdDriver(cell_size=(grid["nx_full"] / grid["resolution"],
                                grid["ny_full"] / grid["resolution"], 0.0),
                     resolution=grid["resolution"], dimensions=2, force_complex_fields=True,
                     symmetry=(("Y", -1),), k_point=(k, 0.0, 0.0))
    assert drv.bloch_phases[0] == recorded and recorded.imag != 0.0
    assert drv.bloch_phases[1] is None and drv.bloch_phases[2] is None
    assert tuple(drv.grid.shape) == tuple(grid["shape"])


def test_zone_edge_is_exactly_minus_one_and_zero_k_has_no_phase():
    assert bloch_phase(0.3125, 8, 5.0) == complex(-1.0, 0.0)
    angle = 0.3125 * ((2 * math.pi / 5.0) * 8)
    assert math.sin(angle) != 0.0  # the branch is what makes it exact
    assert bloch_phase(0.0, 8, 5.0) is None
    z = bloch_phase(0.3, 8, 5.0)
    assert abs(z - complex(math.cos(2 * math.pi * 0.3 * 1.6), math.sin(2 * math.pi * 0.3 * 1.6))) \
        < 1e-14


def test_refused_pairings():
    cell = (1.6, 1.2, 1.6)
    with pytest.raises(ValueError, match="complex fields"):
        FdtdDriver(cell, 5.0, k_point=(0.3, 0, 0))
    with pytest.raises(ValueError, match="mirror"):
        FdtdDriver(cell, 5.0, force_complex_fields=True, symmetry=("X",), k_point=(0.3, 0, 0))
    with pytest.raises(ValueError, match="invariant"):
        FdtdDriver((1.6, 1.2, 0.0), 5.0, dimensions=2, force_complex_fields=True,
                   k_point=(0, 0, 0.3))
    with pytest.raises(ValueError, match="lattice vector"):
        FdtdDriver(cell, 5.0, force_complex_fields=True, k_point=(0.3, 0, 0),
                   boundaries={"x": "metallic"})
    drv = FdtdDriver(cell, 5.0, force_complex_fields=True, k_point=(0.3, 0, 0))
    with pytest.raises(ValueError, match="PML"):
        drv.setup_pml(2)
    drv.setup_pml({"z": 2})  # the grating arrangement: absorber off the phased axis
    drv.add_flux_region((0, 0, 0.3), (1.2, 0.8, 0), fcen=1.0, df=0.5, nfreq=3)
    with pytest.raises(ValueError, match="Bloch-phased boundary"):
        drv.add_volume_source("Ez", (0.0, 0.0, 0.0), (1.6, 0.0, 0.0), GaussianSource(1.0, 0.5))
    with pytest.raises(ValueError, match="stored row 0"):  # Ez is Yee-shift 0 on x
        drv.add_source(DepositBox("Ez", (0, 2, 2), (1, 1, 1), numpy.ones(1), numpy.zeros(1)),
                       GaussianSource(1.0, 0.5))
    # Ex is Yee-shift 1 on x: row 0 is an owned cell there
    drv.add_source(DepositBox("Ex", (0, 2, 2), (1, 1, 1), numpy.ones(1), numpy.zeros(1)),
                   GaussianSource(1.0, 0.5))
    # a polarization (integrated) source is taken on complex storage, as each part's line
    drv.add_source(DepositBox("Ey", (3, 2, 2), (1, 1, 1), numpy.ones(1), numpy.zeros(1),
                              integrated=True), GaussianSource(1.0, 0.5))


def test_boundaries_set_a_fold_s_stored_count():
    drv = FdtdDriver((1.6, 2.4, 1.6), 5.0, symmetry=(("Y", -1),))
    assert drv.grid.shape == (8, 8, 8)  # 12 - 6 + 1, plus the periodic fold's far slot
    drv.set_boundaries({"y": "metallic"})
    assert drv.grid.shape == (8, 7, 8) and drv._inv["Ex"].shape == (8, 7, 8)
    drv.set_epsilon_components({c: 2.0 for c in ("Ex", "Ey", "Ez")})
    with pytest.raises(ValueError, match="before any material"):
        drv.set_boundaries({"y": "periodic"})
    assert drv.grid.shape == (8, 7, 8)
    assert FdtdDriver((1.6, 2.4, 1.6), 5.0, symmetry=("Y",),
                      boundaries="metallic").grid.shape == (8, 7, 8)
    boxes = drv.add_volume_source("Ez", (0, 0.2, 0), (0, 0, 0), GaussianSource(1.0, 0.5))
    assert boxes and all(0 < b.lo[1] < drv.grid.shape[1] for b in boxes)
    with pytest.raises(ValueError, match="mirror-folded"):
        drv.add_dft_monitor([1.0], ("Ez",))


def test_boundary_spellings_agree():
    from meep_gpu.driver import _boundary_kinds

    assert _boundary_kinds(None) == {}
    assert _boundary_kinds("metallic") == {0: "metallic", 1: "metallic", 2: "metallic"}
    assert _boundary_kinds({"Y": "metallic"}) == {1: "metallic"}
    assert _boundary_kinds(("periodic", ("metallic", "metallic"), "periodic")) == \
        {0: "periodic", 1: "metallic", 2: "periodic"}
    with pytest.raises(ValueError, match="agree"):
        _boundary_kinds({"x": ("metallic", "periodic")})


def test_a_workload_driver_takes_boundaries():
    """``from_workload`` keeps what ``set_boundaries`` rebuilds the grid from."""
    from meep_gpu import workloads

    wl = workloads.sphere(edge=24, resolution=4.0, pml_cells=4, radius_cells=4.0, eps=2.25,
                          fcen=0.15, df=0.1, nfreq=2, planes="none", steps=10)
    drv = FdtdDriver.from_workload(wl)
    drv.set_boundaries({"x": "metallic"})
    assert drv.wraps == (False, True, True) and drv.grid.shape == wl.grid.shape


def test_complex_oracle_polarization_source_is_two_real_runs():
    """With no Bloch phase the two parts of the complex oracle never meet: each part is the
    real oracle run of its own line of the sources (the real line; the imaginary line as the
    amplitudes (amp_im, -amp_re)), the polarization source's dipole included. A dropped or
    misplaced ``f_minus_p`` around the complex step's update_E moves one of them."""
    from tests import driver_bloch as db

    steps = 30
    drv = db.driver("xy_products_integrated", k=(0.0, 0.0, 0.0))
    assert drv.bloch_phases == [None, None, None] or all(z is None for z in drv.bloch_phases)
    assert any(box.integrated for box, _ in drv._sources)
    linked = db.oracle(drv, steps)
    alone = db.oracle(drv, steps)
    for _ in range(steps):
        linked.step()
        for part in alone.parts:
            part.step()
    for a, b in zip(linked.parts, alone.parts):
        for key, want in b.stored_volumes().items():
            got = a.stored_volumes()[key]
            assert got.view(numpy.uint32).tobytes() == want.view(numpy.uint32).tobytes(), key
    assert numpy.any(linked.im.v["Ey"] != 0)
    drv.close()
