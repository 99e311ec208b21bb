"""Host transfers on the GPU: ``fdtd_upload``/``fdtd_download`` of a whole volume return the
words sent, bit for bit, through the pinned staging path (fixed 16 MiB halves, so a large
volume streams through them in several pieces), for a padded row pitch and an unpadded one."""

import numpy
import pytest

from meep_gpu.engine import HipSimulation
from meep_gpu.grid import Grid

pytestmark = pytest.mark.gpu


def _words(shape, seed):
    """Random float32 bit patterns: normals, subnormals, signed zeros, infinities, NaNs."""
    rng = numpy.random.default_rng(seed)
    bits = rng.integers(0, 1 << 32, size=shape, dtype=numpy.uint64).astype(numpy.uint32)
    return bits.view(numpy.float32)


@pytest.mark.parametrize("cells", [
    (233, 229, 236),  # 12.6 M cells: 50 MB a volume, four pieces; pitch 236 pads every row
    (64, 48, 40),     # pitch == nx, one piece
    (7, 1, 1),        # 1-D: the grid (1, 1, 7), nx 1 padded to a pitch of 4
])
def test_upload_download_round_trip(cells):
    nx, ny, nz = cells
    dims = 3 if ny > 1 else 1
    size = (nx / 4.0, ny / 4.0, nz / 4.0) if dims == 3 else (0.0, 0.0, nx / 4.0)
    grid = Grid(resolution=4, cell_size=size, dimensions=dims)
    print(f"transfer round trip {grid.shape}", flush=True)
    sim = HipSimulation(grid)
    try:
        for seed, name in enumerate(("Ex", "Bz", "Dy")):
            sent = _words(grid.shape, 100 + seed)
            sim.upload(name, sent)
            got = sim.download(name)
            assert got.shape == sent.shape
            assert got.view(numpy.uint32).tobytes() == sent.view(numpy.uint32).tobytes(), name
        # the volumes uploaded first are still intact after the later transfers
        assert sim.download("Ex").view(numpy.uint32).tobytes() == \
            _words(grid.shape, 100).view(numpy.uint32).tobytes()
    finally:
        sim.close()
