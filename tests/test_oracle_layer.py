"""Setup-time words: ADE constants and one update against MEEP's words, PML sigma closed form,
and the product's layer table against the oracle's (CPU)."""

import json
import math
import os

import numpy
import pytest

from oracle import layer

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ade_words.json")


def _word(x):
    return "%08x" % numpy.float32(x).view(numpy.uint32)


def _float(word):
    return numpy.array([int(word, 16)], dtype=numpy.uint32).view(numpy.float32)[0]


def _keys():
    return list(json.load(open(GOLD))["constant_words"])


@pytest.mark.parametrize("key", _keys())
def test_ade_constants_are_meeps_words(key):
    gold = json.load(open(GOLD))["constant_words"][key]
    f, g, kind, dt = key.split("|")
    words = layer.ade_constants(float(f), float(g), kind == "drude", float(dt))
    assert [_word(w) for w in words] == list(gold)


@pytest.mark.parametrize("key", ["1.1|0.05|lorentz|0.05", "0.8|0.12|drude|0.05"])
def test_one_ade_update_matches_meeps_words(key):
    """The C restatement's update_P on MEEP's operand words (signed zeros, subnormals)."""
    import ctypes
    from oracle import engine

    data = json.load(open(GOLD))
    f, g, kind, dt = key.split("|")
    consts = numpy.array(layer.ade_constants(float(f), float(g), kind == "drude", float(dt)),
                         dtype=numpy.float32)
    ops = data["update_operands_p_pp_s_w"]
    p = numpy.array([_float(o[0]) for o in ops], dtype=numpy.float32).reshape(len(ops), 1, 1)
    pp = numpy.array([_float(o[1]) for o in ops], dtype=numpy.float32).reshape(p.shape)
    s = numpy.array([_float(o[2]) for o in ops], dtype=numpy.float32).reshape(p.shape)
    w = numpy.array([_float(o[3]) for o in ops], dtype=numpy.float32).reshape(p.shape)
    sim = engine.OracleSim(p.shape, float(dt), 0.1)
    out = numpy.empty_like(p)
    g_ = sim._grid()
    engine.lib().oracle_update_p(ctypes.byref(g_), 0, consts.ctypes.data_as(engine._FP),
                                 engine._fp(p), engine._fp(pp), engine._fp(s), ctypes.c_float(0),
                                 engine._fp(w), engine._fp(None), engine._fp(out))
    assert [_word(v) for v in out.reshape(-1)] == list(data["update_results"][key])


def test_sigma_at_the_wall_is_the_closed_form():
    """sig at the outermost half-cell sample: (dt/2)*prefac*u^2 with u = (2n-1)/(2n)."""
    n_pml, dt, dx = 8, 0.05, 0.1
    sig = layer.graded_sigma(80, (n_pml, n_pml), 0.5, dt, dx)
    prefac = -math.log(1e-15) / (4.0 * n_pml * dx / 3.0)
    u = (2 * n_pml - 1) / (2 * n_pml)
    assert sig[0] == pytest.approx(0.5 * dt * prefac * u * u, rel=1e-6)
    assert sig[79] == pytest.approx(0.5 * dt * prefac * u * u, rel=1e-6)
    assert numpy.all(sig[n_pml:80 - n_pml] == 0)
    # integer lattice: index 0 is the wrapped wall plane, graded from the upper wall at u = 1
    sig0 = layer.graded_sigma(80, (n_pml, n_pml), 0.0, dt, dx)
    assert sig0[0] == pytest.approx(0.5 * dt * prefac, rel=1e-6)


def test_membership_lines_match_meeps_one_rank_chunks():
    line = layer.membership_line(20, (3, 2), half_integer=False)
    # low face: span int(3+1.5)=4 owns integer rows 1..4; high: span 3 owns 18, 19 and row 0
    assert numpy.flatnonzero(line).tolist() == [0, 1, 2, 3, 4, 18, 19]
    half = layer.membership_line(20, (3, 2), half_integer=True)
    assert numpy.flatnonzero(half).tolist() == [0, 1, 2, 3, 17, 18, 19]


@pytest.mark.parametrize("spec", [{"x": (2, 3), "y": (2.5, 2), "z": (0, 2)}, 8, {"z": 5.5}])
def test_product_layer_table_equals_the_oracles(spec):
    from meep_gpu.grid import Grid
    from meep_gpu.pml import PML

    grid = Grid(resolution=10, cell_size=(3.2, 2.4, 2.0))
    pml = PML(grid, spec)
    faces = layer.pml_faces(spec)
    for half in (0, 1):
        for axis in range(3):
            ref = layer.axis_coefficients(grid.shape[axis], faces[axis], 0.5 * half, grid.dt,
                                          grid.dx)
            for a, b in zip(pml.vectors[(half, axis)], ref):
                assert a.tobytes() == b.tobytes()
            assert numpy.array_equal(pml.members[(half, axis)],
                                     layer.membership_line(grid.shape[axis], faces[axis],
                                                           bool(half)))
