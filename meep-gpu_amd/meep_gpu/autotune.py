"""Launch-shape autotuner: pick the tile shape the fused pairs run fastest with on this GPU.

The candidates are the engine's launch shapes (``engine.LaunchConfig``): cells per thread along x
(1, 2 or 4 -- wider runs mean fewer index computations and wider loads, narrower ones fewer
registers and more resident waves), one workgroup per tile or a persistent grid of 4/8
workgroups per CU, the XCD-contiguous tile order on or off, and plain or non-temporal output
stores; then the shell kernels' widths, and for the D pair at one cell a thread the halo tile:
runs of cells, or LDS-staged tiles of 128 x 2, 64 x 4 or 32 x 8 cells whose row below and
column left are staged once per tile (``TILE_ROWS``). Each is timed on a scratch
simulation with the same grid, absorber and poles as the run (no sources, no monitors), with
HIP events around every launch; the H pair and the D pair each get the shape they ran fastest
with (the D pair streams more arrays than the H pair and their best shapes differ). Results
are cached per (grid shape, absorber, poles, device) in a JSON file, so a run tunes once.
Candidates are timed under the scratch simulation's (or the batch's) own configuration, so
tuning changes no other simulation; the result is installed on ``target`` (a simulation or a
batch) when one is given, else as the process defaults.
"""

from __future__ import annotations

import json
import os
import time
from typing import Dict, Iterable, List, Optional, Tuple

import numpy

from .engine import (HipSimulation, LaunchConfig, device_info, set_launch_pair,
                     set_launch_shell, set_launch_tile)
from .grid import Grid
from .pml import PML

Shape = Tuple[int, int, int, bool]
PairShapes = Tuple[Shape, Shape]  # (H pair, D pair)

#: the D pair's box over runs of cells (0) or over LDS-staged halo tiles of 2, 4, 8 rows
TILE_ROWS = (0, 2, 4, 8)

CANDIDATES: Tuple[Shape, ...] = tuple(
    (vec, bpc, remap, nt) for vec in (1, 2, 4) for bpc in (0, 4, 8) for remap in (0, 1)
    for nt in (False, True))


def plane_superblock(sim, vec: int) -> int:
    """Super-block size (work items) of about one eighth of a plane of the box, for the tile
    order that deals super-blocks round robin to the 8 XCDs (0 when under 2 items)."""
    x0, x1, y0, y1, z0, z1 = sim.work_split(vec)["box"]
    items = -(-((x1 - x0) // vec) * (y1 - y0) // 256)
    sb = -(-items // 8)
    return sb if sb >= 2 else 0


def candidates_for(sim) -> Tuple[Shape, ...]:
    """The generic candidates plus, per vector width, the plane-eighth super-block order."""
    extra = []
    for vec in (1, 2, 4):
        sb = plane_superblock(sim, vec)
        if sb:
            extra += [(vec, bpc, sb, nt) for bpc in (0, 4, 8) for nt in (False, True)]
    return CANDIDATES + tuple(extra)

_CACHE = os.environ.get("MEEP_GPU_TUNE_CACHE",
                        os.path.join(os.path.expanduser("~"), ".cache", "meep_gpu_tune.json"))


def _key(grid: Grid, pml: Optional[PML], npoles: int) -> str:
    faces = pml.faces if pml is not None else None
    return json.dumps([list(grid.shape), faces, npoles, device_info()])


def _load_cache() -> Dict[str, list]:
    try:
        with open(_CACHE) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def _store_cache(cache: Dict[str, list]) -> None:
    try:
        os.makedirs(os.path.dirname(_CACHE), exist_ok=True)
        with open(_CACHE, "w") as fh:
            json.dump(cache, fh)
    except OSError:
        pass


def measure(grid: Grid, pml: Optional[PML], poles: Iterable = (),
            candidates: Optional[Iterable[Shape]] = None, repeats: int = 10,
            inv: Optional[dict] = None) -> List[Tuple[Shape, float, float]]:
    """(shape, mean H-pair ms, mean D-pair ms) for every candidate, on a scratch simulation
    (with the run's inverse epsilon when given: its storage form changes the D pair)."""
    sim = HipSimulation(grid, pml)
    for c, vol in (inv or {}).items():
        sim.set_inverse_epsilon(c, vol)
    for term, sigma in poles:
        sim.add_pole(term, sigma)
    if candidates is None:
        candidates = candidates_for(sim)
    out = []
    try:
        for shape in candidates:
            sim.set_config(LaunchConfig.of_shapes((shape, shape)))
            sim.time_pairs(2)  # first launch of a shape builds its tile walk
            h, d = sim.time_pairs(repeats)
            out.append((shape, h, d))
    finally:
        sim.close()
    return out


#: widths tried for the shell kernels (the rows on a y or z layer line)
SHELL_WIDTHS = (1, 2)


def tune_shell(grid: Grid, pml: Optional[PML], poles: Iterable = (), repeats: int = 10,
               inv: Optional[dict] = None, shapes: Optional[PairShapes] = None
               ) -> Tuple[int, int]:
    """With the pairs' ``shapes``, the fastest shell width of each pair."""
    if pml is None:
        return (1, 1)
    shapes = shapes or ((1, 8, 0, False), (1, 8, 0, False))
    sim = HipSimulation(grid, pml)
    for c, vol in (inv or {}).items():
        sim.set_inverse_epsilon(c, vol)
    for term, sigma in poles:
        sim.add_pole(term, sigma)
    times = []
    try:
        for sv in SHELL_WIDTHS:
            sim.set_config(LaunchConfig.of_shapes(shapes, shell=(sv, sv)))
            sim.time_pairs(2)
            h, d = sim.time_pairs(repeats)
            times.append((sv, h, d))
    finally:
        sim.close()
    return (min(times, key=lambda r: r[1])[0], min(times, key=lambda r: r[2])[0])


def tune_tile(grid: Grid, pml: Optional[PML], poles: Iterable = (), repeats: int = 10,
              inv: Optional[dict] = None, shapes: Optional[PairShapes] = None,
              shell: Tuple[int, int] = (1, 1)) -> int:
    """With the pairs' ``shapes`` and shell widths, the D pair's fastest halo tile (0: runs);
    0 where the D pair is not at one cell a thread or there is no absorber (no box split)."""
    shapes = shapes or ((1, 8, 0, False), (1, 8, 0, False))
    if pml is None or shapes[1][0] != 1:
        return 0
    sim = HipSimulation(grid, pml)
    for c, vol in (inv or {}).items():
        sim.set_inverse_epsilon(c, vol)
    for term, sigma in poles:
        sim.add_pole(term, sigma)
    times = []
    try:
        for rows in TILE_ROWS:
            sim.set_config(LaunchConfig.of_shapes(shapes, shell=shell, tile=rows))
            sim.time_pairs(2)
            times.append((rows, sim.time_pairs(repeats)[1]))
    finally:
        sim.close()
    return min(times, key=lambda r: r[1])[0]


def install(shapes: PairShapes, shell: Optional[Tuple[int, int]] = None, target=None,
            tile: Optional[int] = None) -> None:
    """The pairs' shapes (and shell widths, and the D pair's halo tile) as ``target``'s own
    configuration (a simulation or a batch; its graph and transform modes kept), or as the
    process defaults."""
    if target is not None:
        now = target.config()
        sh = shell or (now.h[4], now.d[4])
        rows = now.d[5] if tile is None else tile
        target.set_config(LaunchConfig.of_shapes(shapes, shell=sh, graphs=now.graphs,
                                                 flux_mfma=now.flux_mfma,
                                                 tile=rows if shapes[1][0] == 1 else 0))
        return
    set_launch_pair(0, *shapes[0])
    set_launch_pair(1, *shapes[1])
    if shell is not None:
        set_launch_shell(0, shell[0])
        set_launch_shell(1, shell[1])
    if tile is not None:
        set_launch_tile(tile if shapes[1][0] == 1 else 0)


def best_pair_shapes(results) -> PairShapes:
    """The fastest shape for each fused pair among ``measure``'s results."""
    h = min(results, key=lambda r: r[1])[0]
    d = min(results, key=lambda r: r[2])[0]
    return tuple(h), tuple(d)


def tune(grid: Grid, pml: Optional[PML], poles: Iterable = (), repeats: int = 10,
         use_cache: bool = True, inv: Optional[dict] = None, target=None) -> PairShapes:
    """Choose the fastest launch shape of each pair (and shell widths) for this grid, install
    them on ``target`` (a simulation or batch) or as the process defaults; returns the shapes."""
    poles = list(poles)
    key = _key(grid, pml, len(poles))
    cache = _load_cache() if use_cache else {}
    entry = cache.get(key)
    if isinstance(entry, list) and len(entry) == 4 and all(len(x) == 4 for x in entry[:2]):
        shapes = (tuple(entry[0]), tuple(entry[1]))
        install(shapes, tuple(entry[2]), target, int(entry[3]))
        return shapes  # type: ignore[return-value]
    shapes = best_pair_shapes(measure(grid, pml, poles, repeats=repeats, inv=inv))
    shell = tune_shell(grid, pml, poles, repeats=repeats, inv=inv, shapes=shapes)
    tile = tune_tile(grid, pml, poles, repeats=repeats, inv=inv, shapes=shapes, shell=shell)
    install(shapes, shell, target, tile)
    if use_cache:
        cache[key] = [list(shapes[0]), list(shapes[1]), list(shell), tile]
        _store_cache(cache)
    return shapes


def tune_batch(batch, candidates: Iterable[Shape] = CANDIDATES, steps: int = 16
               ) -> PairShapes:
    """Choose launch shapes for a ``BatchedRun`` by stepping it; installed as the batch's own
    configuration (no other simulation's launches change).

    Each candidate runs ``steps`` warm steps (its tile list is built on first use), then
    ``steps`` timed ones, for both pairs at once; then the best shape's D pair is held while
    the H pair tries every candidate. The batch really advances, so its source tables must
    cover 2 * steps * (2 * len(candidates) + len(TILE_ROWS)) steps.
    """
    candidates = list(candidates)
    base = batch.config()

    def timed(shapes, tile: int = 0) -> float:
        batch.set_config(LaunchConfig.of_shapes(shapes, shell=(base.h[4], base.d[4]),
                                                graphs=base.graphs, flux_mfma=base.flux_mfma,
                                                tile=tile))
        batch.run(steps)
        batch.synchronize()
        t0 = time.perf_counter()
        batch.run(steps)
        batch.synchronize()
        return time.perf_counter() - t0

    best = min(candidates, key=lambda sh: timed((sh, sh)))
    h_best = min(candidates, key=lambda sh: timed((sh, best)))
    shapes = (tuple(h_best), tuple(best))
    tile = 0
    if best[0] == 1:  # the D pair's halo tiles at one cell a thread
        tile = min(TILE_ROWS, key=lambda rows: timed(shapes, rows))
    install(shapes, target=batch, tile=tile)
    return shapes
