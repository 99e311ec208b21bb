"""Sources and flux regions on a mirror-folded grid, on the CPU: MEEP's placement of the three
folded capsules' point sources (their lifted source tables: MEEP's own indices and
amplitudes) reproduced from the request, a request below the plane served at its image with
the component's parity, and a flux region's rows and signed weights."""

import os

import numpy
import pytest

from meep_gpu import capsule
from meep_gpu.grid import Grid
from meep_gpu.sources import meep_volume_source
from tests import capsule_cases as cc


def _table(boxes):
    got = {}
    for b in boxes:
        for idx in numpy.ndindex(*b.cnt):
            key = tuple(int(l + i) for l, i in zip(b.lo, idx))
            got[key] = got.get(key, 0) + complex(b.amp_re[idx], b.amp_im[idx])
    return got


@pytest.mark.parametrize("case", cc.FOLDED)
def test_folded_point_sources_are_placed_as_meep_placed_them(case):
    doc = capsule.read_capsule_document(os.path.join(cc.case_dir(case), "capsule.capsule"))[1]
    state, _, _ = cc.load(case)
    for src, lifted in zip(doc["_sources"], state.sources):
        boxes = meep_volume_source(state.grid(), src["component"], src["center"], src["size"],
                                   complex(src["amplitude"]), metallic=tuple(state.metallic))
        want = {tuple(int(v) for v in ix): complex(a) for ix, a in zip(lifted.index, lifted.amps)}
        assert _table(boxes) == want


@pytest.mark.parametrize("phase", [1, -1])
@pytest.mark.parametrize("comp", ["Ez", "Ey", "Hx", "Hy"])
def test_a_request_below_the_plane_lands_at_its_image(comp, phase):
    """A point below the Y mirror deposits where the point above it does, times the
    component's parity (a vector odd about the plane normal to its own axis, a pseudovector
    the other way round, both times the declared phase)."""
    g = Grid(resolution=10, cell_size=(1.6, 3.2, 1.2), symmetry=(("Y", phase),))
    above = _table(meep_volume_source(g, comp, (0.1, 0.37, 0.05), (0, 0, 0)))
    below = _table(meep_volume_source(g, comp, (0.1, -0.37, 0.05), (0, 0, 0)))
    own = comp[1] == "y"
    parity = -phase if own != (comp[0] == "H") else phase
    assert set(above) == set(below) and above
    for key in above:
        assert below[key] == parity * above[key]


def test_flux_region_rows_and_signed_weights():
    from meep_gpu.dft import FluxMonitor

    g = Grid(resolution=10, cell_size=(3.2, 3.2, 2.4), symmetry=(("Y", 1),))
    f = numpy.array([1.0])
    straddle = FluxMonitor.from_region(g, (0.5, 0.0, 0.0), (0.0, 2.0, 1.6), f)
    # y from -1 to 1: the ladder's 22 centres at doubled -21..21 (end weights 1/8 and 7/8);
    # stored row j holds the centre at doubled 2j - 1: the owned pixel at -1 (row 0) and +1
    # (row 1) once each, every other row twice (a centre and the image of its mirror), the
    # two 1/8 ends meeting on row 11 -- 20 cells in all, as unfolded
    wy = straddle.cell_weights[0, :, straddle.cnt[2] // 2]
    assert straddle.lo[1] == 0
    assert list(wy) == [1.0, 1.0] + [2.0] * 8 + [1.75, 0.25] and wy.sum() == 20.0
    above = FluxMonitor.from_region(g, (0.0, 0.6, 0.0), (2.0, 0.0, 1.6), f)
    below = FluxMonitor.from_region(g, (0.0, -0.6, 0.0), (2.0, 0.0, 1.6), f)
    assert [r for r, _ in above.normal_planes] == [r for r, _ in below.normal_planes]
    assert [w for _, w in below.normal_planes] == [-w for _, w in above.normal_planes]


def test_a_folded_dft_region_reads_its_images_with_the_parity():
    """A DFT region over a folded axis registers the stored rows its centres land on, and its
    array reads every centre of the full region from its row, the images times the parity."""
    from meep_gpu.dft import DftFieldsMonitor

    g = Grid(resolution=10, cell_size=(1.6, 3.2, 1.2), symmetry=(("Y", -1),))
    m = DftFieldsMonitor.from_region(g, ("Ez", "Ey"), (0.0, -0.1, 0.0), (0.4, 1.0, 0.0),
                                     numpy.array([1.0]), 1.0)
    rows, flips, declared = m.folded[1]
    # y in [-0.6, 0.4]: the ladder's centres at doubled -13 .. 9; -13 .. -3 are the images
    # of 13 .. 3 (rows 7 .. 2), -1 is the owned pixel (row 0), 1 .. 9 rows 1 .. 5
    assert list(flips) == [1] * 6 + [0] * 6 and declared == -1
    assert list(m.lo[1] + rows) == [7, 6, 5, 4, 3, 2, 0, 1, 2, 3, 4, 5]

    class Fake:
        def flux_read(self, rid):
            nz = m.cnt
            data = numpy.zeros((2, 1) + nz, dtype=numpy.complex128)
            data[0, 0] = numpy.arange(nz[1])[None, :, None] + 1.0  # Ez: the row number + 1
            data[1, 0] = numpy.arange(nz[1])[None, :, None] + 1.0
            return data

    m._sim, m._slots = Fake(), {"Ez": (0, 0), "Ey": (0, 1)}
    ez = m.get_dft_array("Ez")[0]   # Ez under an odd Y mirror: parity -1 (even x declared -1)
    ey = m.get_dft_array("Ey")[0]   # Ey: its own axis, parity +1
    assert list(ez.real) == [-(r + 1.0) if f else r + 1.0 for r, f in zip(rows, flips)]
    assert list(ey.real) == [r + 1.0 for r in rows]
