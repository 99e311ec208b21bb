"""The oracle against MEEP's own recorded output for the reference's two 2-D examples (CPU).

Pins, against MEEP itself rather than against a restatement: the 2-D curl and UPML, MEEP's
zero-size source deposit (interpolation weights, amplitude times a^2), the Gaussian source's
current, the Lorentz ADE, the cell-centred DFT with E at t and H at t - dt/2, MEEP's flux
region ladder and its decimation. Tolerances: 2e-5 relative on the final Ez norms, 1e-4 of
the spectrum's peak on every flux value (measured: <= 1.1e-5 and <= 2.7e-6).
"""

import pytest

from tests.meep_examples import build, centred_ez_norms, recorded, spectrum_error

FIELD_TOL = 2e-5
SPECTRUM_TOL = 1e-4


def run_oracle(name, decimation=None):
    from meep_gpu.dft import FluxMonitor, meep_frequency_span
    from meep_gpu.sources import current_table
    from oracle import workload as ow

    ex = build(name)
    spec, grid = ex.spec, ex.grid
    dec = spec["decimation"] if decimation is None else decimation
    steps = spec["steps"]
    cells = int(round(spec["pml"] * spec["resolution"]))
    faces = ((cells, cells), (cells, cells), (0, 0))
    fl = spec["flux"]
    freqs = meep_frequency_span(fl["fcen"], fl["df"], fl["nfreq"])
    mon = FluxMonitor(0, (0, 0, 0), (1, 1, 1), freqs, grid.dx, grid.dt, dec)
    planes, layout = [], []
    for region in fl["regions"]:
        boxes, weights, _ = ow.region_planes(grid.shape, grid.dx, region["center"],
                                             region["size"], grid.invariant)
        layout.append((len(planes), [w for _, _, w in boxes], weights))
        for lo, cnt, _ in boxes:
            planes.append((0, lo, cnt, mon.omegas, mon.scale, dec))
    poles = [(t.frequency, t.gamma, t.is_drude, s) for t, s in ex.poles]
    table = current_table(ex.source, grid.dt, steps + 1, False)
    deposits = [(b.component, b.lo, b.cnt, b.amp_re, b.amp_im, table) for b in ex.boxes]
    sim, oplanes = ow.build(grid.shape, grid.dt, grid.dx, faces, ex.inv, poles, deposits,
                            planes, grid.invariant)
    ow.run(sim, oplanes, steps, reference_precision=False)
    spectra = []
    for first, nw, weights in layout:
        spectra.append(ow.region_flux(oplanes[first:first + len(nw)], weights, nw, grid.dx))
    return centred_ez_norms(sim.get_E("Ez")), spectra


@pytest.mark.parametrize("name", ["slab_flux_2d", "dispersive_slab_2d"])
def test_oracle_reproduces_meeps_recorded_fields_and_spectra(name):
    want = recorded()[name]["meep"]
    (mx, l2), spectra = run_oracle(name)
    assert abs(mx / want["ez_max_abs"] - 1) < FIELD_TOL
    assert abs(l2 / want["ez_l2_norm"] - 1) < FIELD_TOL
    for got, ref in zip(spectra, want["flux"]):
        assert spectrum_error(got, ref) < SPECTRUM_TOL, (got, ref)


def test_the_nominal_decimation_rule_would_miss_meep():
    """Sensitivity: MEEP's spectra discriminate the accumulation stride (nominal rule: 49)."""
    want = recorded()["slab_flux_2d"]["meep"]
    _, spectra = run_oracle("slab_flux_2d", decimation=49)
    assert max(spectrum_error(g, r) for g, r in zip(spectra, want["flux"])) > 10 * SPECTRUM_TOL


def test_ladder_weights_integrate_the_requested_extent():
    """Closed form of MEEP's ladder: an extended direction's weights sum to extent * a, a flat
    one's to 1 (a linear interpolation onto the requested coordinate)."""
    import random

    from oracle.workload import meep_ladder

    rng = random.Random(3)
    for _ in range(500):
        a = rng.choice([10.0, 12.0, 20.0, 24.0])
        lo = rng.uniform(-3, 3)
        ext = rng.choice([0.0, rng.uniform(1.0 / a, 3.0)])
        _, w = meep_ladder(lo, lo + ext, a)
        assert abs(w.sum() - (ext * a if ext else 1.0)) < 1e-9
        if ext == 0.0:  # the two bracketing centres, weights by distance
            assert w.size == 2 and min(w) >= 0.0
