"""The allocation-free flux spectrum (fdtd_flux_spectrum_noalloc) against the allocating one
(fdtd_flux_spectrum), through the C ABI (GPU).

Two calls in a row on one GPU, the second on a smaller plane (fewer cells, fewer frequencies),
share the library's static scratch: each must equal what fdtd_flux_spectrum returns for the
same plane word for word (the same two kernels, the same order of summation), with cell
weights and without. The caller's device buffers are allocated by the test through the HIP
runtime the engine itself links.
"""

import ctypes

import numpy
import pytest

from tests import workload_runs as wr

pytestmark = pytest.mark.gpu


def _hip():
    lib = ctypes.CDLL("libamdhip64.so.7")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return lib


H2D, D2H = 1, 2


def _dev_copy(hip, host):
    ptr = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(ptr), max(host.nbytes, 8)) == 0
    assert hip.hipMemcpy(ptr, host.ctypes.data, host.nbytes, H2D) == 0
    return ptr


def _noalloc(sim, plane, nf, weights_dev, hip):
    from meep_gpu._native import check

    out = numpy.full(nf, numpy.nan)
    out_dev = _dev_copy(hip, out)
    try:
        check(sim.lib.fdtd_flux_spectrum_noalloc(sim.handle, int(plane), weights_dev, out_dev,
                                                 nf), "fdtd_flux_spectrum_noalloc", sim.lib)
        sim.synchronize()
        assert hip.hipMemcpy(out.ctypes.data, out_dev, out.nbytes, D2H) == 0
    finally:
        hip.hipFree(out_dev)
    return out


def test_no_alloc_spectrum_equals_the_allocating_one_twice_in_a_row():
    from meep_gpu._native import EngineError

    hip = _hip()
    big = wr.hip_run(wr.small("sphere_box"), 150)    # 8 planes x 6 frequencies
    small = wr.hip_run(wr.small("sphere"), 140)      # 1 plane x 8 frequencies
    cases = []
    # the largest plane of the box first (the interior plane, 28 x 28 cells), then the small run's
    pid = max(range(len(big.sim.planes)), key=lambda q: numpy.prod(big.sim.planes[q][2]))
    cases.append((big.sim, pid, True))
    cases.append((small.sim, 0, False))
    rng = numpy.random.default_rng(7)
    sizes = []
    for sim, plane, weighted in cases:
        _, _, cnt, nf, _ = sim.planes[plane]
        sizes.append((int(numpy.prod(cnt)), nf))
        w = rng.uniform(0.25, 1.0, size=tuple(cnt)) if weighted else None
        print(f"no-alloc spectrum: plane {plane}, {sizes[-1][0]} cells x {nf} frequencies, "
              f"weights {weighted}", flush=True)
        want = sim.flux_spectrum(plane, w)  # the allocating entry point
        w_dev = None
        if w is not None:
            w_dev = _dev_copy(hip, numpy.ascontiguousarray(w.transpose(2, 1, 0)))
        try:
            got = _noalloc(sim, plane, nf, w_dev, hip)
        finally:
            if w_dev is not None:
                hip.hipFree(w_dev)
        assert numpy.all(numpy.isfinite(got))
        assert numpy.max(numpy.abs(want)) > 0
        assert got.tobytes() == want.tobytes(), (plane, got, want)
    assert sizes[1][0] < sizes[0][0]  # the second input is the smaller one
    # more frequencies than the scratch holds are refused before any launch
    with pytest.raises(EngineError, match="at most 1024"):
        drv = _many_frequencies()
        try:
            _noalloc(drv.sim, 0, 1100, None, hip)
        finally:
            drv.close()
    big.close()
    small.close()


def _many_frequencies():
    from meep_gpu import workloads
    from meep_gpu.driver import FdtdDriver

    wl = workloads.sphere(edge=24, resolution=4.0, pml_cells=4, radius_cells=4.0, nfreq=1100,
                          planes="one", steps=10, source_x=-1.0, plane_x=1.0)
    drv = FdtdDriver.from_workload(wl)
    drv.freeze(planned_steps=11)
    drv.run(steps=10)
    drv.synchronize()
    return drv
