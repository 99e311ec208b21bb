"""Flux-data save/load and minus-flux on the host (CPU): the opaque data's layout, the checks a
load makes before it writes anything, and the minus convention (load, then scale by -1) --
reference dft.py:2618-2733, :3261-3316 (MEEP's get_flux_data / load_flux_data /
load_minus_flux_data). A recording stand-in takes the place of the device simulation."""

import numpy
import pytest

from meep_gpu.dft import FluxDftData, FluxMonitor, FluxPlaneDft, tangential_names


class _Recorder:
    """flux_read / flux_write / flux_scale of a device simulation, on host arrays."""

    def __init__(self):
        self.boxes, self.acc, self.calls = [], [], []

    def add_flux_plane(self, normal, lo, cnt, omegas, scale, decimation):
        self.boxes.append(tuple(cnt))
        rng = numpy.random.default_rng(len(self.boxes))
        shape = (4, len(omegas)) + tuple(cnt)
        self.acc.append(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
        return len(self.boxes) - 1

    def flux_read(self, pid):
        return self.acc[pid].copy()

    def flux_write(self, pid, block):
        self.calls.append(("write", pid))
        assert block.shape == self.acc[pid].shape
        self.acc[pid] = numpy.array(block, dtype=numpy.complex128)

    def flux_scale(self, pid, scale):
        self.calls.append(("scale", pid, complex(scale)))
        self.acc[pid] = self.acc[pid] * complex(scale)


def _monitor(normal=0, nplanes=2, freqs=(0.1, 0.15, 0.2)):
    lo, cnt = [3, 4, 5], [6, 7, 8]
    lo[normal], cnt[normal] = 9, 1
    planes = [(9 + q, 0.5) for q in range(nplanes)]
    mon = FluxMonitor(normal, tuple(lo), tuple(cnt), numpy.array(freqs), 0.1, 0.05,
                      normal_planes=planes)
    mon.attach(_Recorder())
    return mon


@pytest.mark.parametrize("normal", (0, 1, 2))
def test_the_data_names_each_planes_tangential_blocks(normal):
    mon = _monitor(normal)
    data = mon.get_dft_data()
    assert isinstance(data, FluxDftData) and len(data.planes) == 2
    e_names, h_names = tangential_names(normal)
    assert e_names == ("E" + "xyz"[(normal + 1) % 3], "E" + "xyz"[(normal + 2) % 3])
    for pid, rec in zip(mon.plane_ids, data.planes):
        acc = mon._sim.acc[pid]
        assert set(rec.E) == set(e_names) and set(rec.H) == set(h_names)
        for q, name in enumerate(e_names + h_names):
            block = (rec.E if q < 2 else rec.H)[name]
            assert numpy.array_equal(block, acc[q])
            assert block.shape == (3,) + mon.plane_boxes().__next__()[1]
        assert rec.frequencies == (0.1, 0.15, 0.2)


def test_load_replaces_every_plane_and_minus_negates_after():
    a, b = _monitor(), _monitor()
    for pid in b.plane_ids:  # b's own accumulators differ from a's
        b._sim.acc[pid] = b._sim.acc[pid] * 3.0
    data = a.get_dft_data()
    b._sim.calls.clear()
    b.load_minus_dft_data(data)
    assert b._sim.calls == [("write", 0), ("write", 1), ("scale", 0, -1), ("scale", 1, -1)]
    for pa, pb in zip(a.plane_ids, b.plane_ids):
        assert numpy.array_equal(b._sim.acc[pb], -a._sim.acc[pa])
    b.load_dft_data(data)
    for pa, pb in zip(a.plane_ids, b.plane_ids):
        assert numpy.array_equal(b._sim.acc[pb], a._sim.acc[pa])


def _altered(data, plane=0, **changes):
    recs = list(data.planes)
    recs[plane] = FluxPlaneDft(**{**dict(E=recs[plane].E, H=recs[plane].H,
                                          frequencies=recs[plane].frequencies), **changes})
    return FluxDftData(planes=tuple(recs))


def test_a_refused_load_writes_nothing():
    a, b = _monitor(), _monitor()
    data = a.get_dft_data()
    before = [x.copy() for x in b._sim.acc]
    e_names, h_names = tangential_names(0)
    bad = [
        (TypeError, {"E": 1}),                                               # not the data object
        (ValueError, FluxDftData(planes=data.planes[:1])),                   # plane count
        (ValueError, _altered(data, 1, frequencies=(0.1, 0.15, 0.2000001))),  # frequencies
        (ValueError, _altered(data, 1, frequencies=(0.1, 0.15))),            # frequency count
        (ValueError, _altered(data, 1, E={e_names[0]: data.planes[1].E[e_names[0]]})),  # set
        (ValueError, _altered(data, 1, H={h_names[0]: data.planes[1].H[h_names[0]],
                                          h_names[1]: data.planes[1].H[h_names[1]][:, :, :2]})),
    ]
    for exc, payload in bad:
        with pytest.raises(exc):
            b.load_dft_data(payload)
        assert not [c for c in b._sim.calls if c[0] == "write"]
        for x, y in zip(before, b._sim.acc):
            assert numpy.array_equal(x, y)


def test_an_unattached_monitor_refuses():
    mon = FluxMonitor(0, (1, 1, 1), (1, 2, 2), numpy.array([0.1]), 0.1, 0.05)
    with pytest.raises(RuntimeError):
        mon.get_dft_data()
    with pytest.raises(RuntimeError):
        mon.scale_dfts(-1)


def test_automatic_decimation_is_meep_s_rule():
    """decimation=0 resolves at the freeze to floor(1/(dt*(f_max + src_max))) with a Gaussian
    source's |f| + fwidth/2 (reference driver.py:3101-3163): the two recorded examples' nominal
    values 49 and 36, and 1 without a Gaussian source or a nonzero monitor frequency."""
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.sources import ContinuousSource, GaussianSource

    def rule(cell, fcen, df, src_f, src_w, kind=GaussianSource):
        drv = FdtdDriver(cell_size=cell, resolution=20.0, dimensions=2)
        drv._sources.append((None, kind(src_f, src_w) if kind is GaussianSource else kind(src_f)))
        freqs = numpy.linspace(fcen - df / 2, fcen + df / 2, 5)
        return drv.automatic_decimation(freqs)

    assert rule((10.0, 6.0, 0.0), 0.25, 0.3, 0.25, 0.3) == 49
    assert rule((8.0, 6.0, 0.0), 0.4, 0.3, 0.4, 0.3) == 36
    assert rule((8.0, 6.0, 0.0), 0.4, 0.3, 0.4, 0.3, ContinuousSource) == 1
    assert rule((8.0, 6.0, 0.0), 0.0, 0.0, 0.4, 0.3) == 1
