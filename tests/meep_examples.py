"""The reference's two 2-D MEEP examples with recorded MEEP output, as inputs both steppers take.

A slab normal to one axis on a grid whose samples of the tangential E components fall exactly
on the slab faces: MEEP's sub-pixel averaging of such a sample is the parallel average,
chi1inv = 1/((eps_in + eps_out)/2) (half the pixel on each side), every other sample is pure,
and a susceptibility takes the closed block's sigma on the face (no averaging of sigma).
The source is MEEP's zero-size Ez source (interpolation weights times a per zero-size axis),
the monitors MEEP's FluxRegions with MEEP's resolved decimation.
"""

from __future__ import annotations

import json
import os
from dataclasses import dataclass
from typing import Dict, List

import numpy

from meep_gpu.dispersion import LORENTZIAN, Susceptibility
from meep_gpu.grid import Grid
from meep_gpu.sources import DepositBox, GaussianSource, meep_volume_source

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "meep_examples.json")


def recorded() -> Dict[str, dict]:
    with open(GOLDEN) as fh:
        return json.load(fh)


@dataclass
class Example:
    name: str
    spec: dict
    grid: Grid
    inv: Dict[str, numpy.ndarray]
    poles: list
    boxes: List[DepositBox]
    source: GaussianSource


def build(name: str) -> Example:
    spec = recorded()[name]
    grid = Grid(resolution=spec["resolution"], cell_size=tuple(spec["cell"]), dimensions=2)
    slab = spec["slab"]
    axis, half = slab["axis"], slab["width"] / 2
    n = grid.shape[axis]
    inv, sig = {}, {}
    for c in ("Ex", "Ey", "Ez"):
        shift = 1 if "xyz".index(c[1]) == axis else 0
        # doubled coordinate of each sample along the slab normal, origin at the cell centre
        site = 2 * numpy.arange(n) + shift - n
        width2 = int(round(2 * half * spec["resolution"]))
        inside = numpy.abs(site) < width2
        face = numpy.abs(site) == width2
        v = numpy.ones(n, dtype=numpy.float32)
        v[inside] = numpy.float32(1.0 / slab["epsilon"])
        v[face] = numpy.float32(1.0 / (0.5 * (slab["epsilon"] + 1.0)))
        s = numpy.where(inside | face, numpy.float32(slab.get("lorentz", {}).get("sigma", 0.0)),
                        numpy.float32(0.0)).astype(numpy.float32)
        shape = [1, 1, 1]
        shape[axis] = n
        inv[c] = numpy.ascontiguousarray(numpy.broadcast_to(v.reshape(shape), grid.shape))
        sig[c] = numpy.ascontiguousarray(numpy.broadcast_to(s.reshape(shape), grid.shape))
    poles = []
    if "lorentz" in slab:
        lz = slab["lorentz"]
        poles.append((Susceptibility(lz["frequency"], lz["gamma"], LORENTZIAN), sig))
    src = spec["source"]
    boxes = meep_volume_source(grid, src["component"], tuple(src["center"]), (0.0, 0.0, 0.0))
    return Example(name, spec, grid, inv, poles, boxes,
                   GaussianSource(src["frequency"], src["fwidth"]))


def centred_ez_norms(ez: numpy.ndarray):
    """MEEP's get_array(Ez) on a 2-D grid: each cell centre the average of its four Ez samples
    (periodic neighbours), summed as 0.25*(((a+b)+c)+d); returns (max abs, l2 norm)."""
    e = numpy.asarray(ez, dtype=numpy.float64).reshape(ez.shape[0], ez.shape[1])
    b = numpy.roll(e, -1, 0)
    c = numpy.roll(e, -1, 1)
    d = numpy.roll(b, -1, 1)
    cent = 0.25 * (((e + b) + c) + d)
    return float(numpy.abs(cent).max()), float(numpy.sqrt((cent * cent).sum()))


def spectrum_error(got, want) -> float:
    """max |got - want| over the spectrum relative to the spectrum's largest |want|."""
    got, want = numpy.asarray(got), numpy.asarray(want)
    return float(numpy.abs(got - want).max() / numpy.abs(want).max())
