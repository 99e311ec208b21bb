"""Benchmark: the plain single-simulation FDTD step on one MI355X.

Workload (default): the 192^3 dielectric sphere (7,077,888 cells, 24 px/unit, 8-cell UPML on
all faces, sphere radius 38.4 cells, eps 2.25 with a sub-pixel diagonal chi1inv, Ez Gaussian
point source fcen 0.15 / df 0.1), 8 flux planes x 100 frequencies transformed in float64.
A step is one whole time step: the fused H pair, the fused D pair and the flux sampling; every
128 samples a plane's ring is transformed in float64, and the timed window ends with the pass
over the samples still buffered, so every sample taken in the window is transformed inside it.
Inputs are resident in HBM before timing starts.

Prints ONE JSON line. ``value`` is Mcell-steps/s = cells * steps / elapsed; ``roofline`` is
for the D pair (the heavier fused launch): the stencil's compulsory bytes per launch (every
array once, neighbours free, auxiliaries on their layer rows, inverse epsilon as float32) over its mean launch time from HIP events around every launch in the timed region,
with the engine's compacted layout's own byte count beside it; ``cpu_baseline`` times the C
restatement (OpenMP on the host cores) on a bounded sample of the same workload, and the same
leg states each spectrum's error against that float64 oracle run (``precision``).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "meep-gpu_amd"))
sys.path.insert(0, ROOT)

import numpy  # noqa: E402

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(sim) -> dict:
    """Compulsory HBM bytes per launch of each fused pair (every array touched once, neighbour
    reads free, auxiliaries only on their membership rows, inverse epsilon at the width the
    engine stores it: 1 or 2 byte table indices where it has few distinct words)."""
    grid = sim.grid
    n = grid.total_cells
    pml = sim.pml
    frac = (lambda axis, half: pml.member_fraction(axis, half)) if pml else (lambda a, h: 0.0)
    # H pair: E r 12, B r+w 24; per component c: fu_B r/w 8 on its dsigu line (half lattice),
    # H r/w + f_w_H r/w 16 on its own-axis line (integer lattice; H equals B elsewhere and is
    # not stored there).
    h = n * (12 + 24)
    d = n * (12 + 24 + 12 + sum(sim.inverse_epsilon_bytes()))  # H-or-B r, D r/w, E w, inv r
    for c in range(3):
        dsigu = (c + 2) % 3
        h += n * (8 * frac(dsigu, 1) + 16 * frac(c, 0))
        d += n * (8 * frac(dsigu, 0) + 12 * frac(c, 1))
    rows = sim.pole_row_fraction()
    for driven in sim.poles:  # P r, P_prev r, P' w, sigma r on the rows that carry pole work
        d += n * 16 * sum(driven) * rows
    return {"h": h, "d": d}


def _oracle_inputs(wl, steps: int):
    from meep_gpu.dft import FluxMonitor
    from meep_gpu.pml import normalize_faces
    from meep_gpu.sources import current_table

    grid = wl.grid
    faces = normalize_faces(wl.pml_faces, grid) if wl.pml_faces else ((0, 0),) * 3
    deposits = []
    for box, fn in wl.sources:
        deposits.append((box.component, box.lo, box.cnt, box.amp_re, box.amp_im,
                         current_table(fn, grid.dt, steps, box.magnetic)))
    planes = []
    for p in wl.planes:
        mon = FluxMonitor(p.normal, p.lo, p.cnt, p.frequencies, grid.dx, grid.dt)
        planes.append((p.normal, p.lo, p.cnt, mon.omegas, mon.scale))
    poles = [(t.frequency, t.gamma, t.is_drude, s) for t, s in wl.poles]
    return faces, deposits, planes, poles


def spectrum_errors(wl, steps: int, oplanes) -> dict:
    """The engine's spectra after ``steps`` steps of ``wl`` against the float64 oracle run of
    the same steps (``oplanes``): per spectrum, max |engine - oracle| over the oracle's peak."""
    from meep_gpu.driver import FdtdDriver

    drv = FdtdDriver.from_workload(wl)
    drv.freeze(planned_steps=steps + 1)
    drv.run(steps=steps)
    drv.synchronize()
    errs = []
    for mon, plane in zip(drv.flux_monitors, oplanes):
        got = mon.get_flux_spectrum()
        ref = plane.flux(wl.grid.dx, "c128")
        peak = float(numpy.max(numpy.abs(ref)))
        errs.append(float(numpy.max(numpy.abs(got - ref)) / peak) if peak > 0 else None)
    drv.close()
    return {"steps": steps, "spectra": len(errs), "frequencies": int(wl.planes[0].frequencies.size),
            "max_rel_error_vs_fp64_oracle": [None if e is None else float(f"{e:.3e}") for e in errs],
            "tolerance": 1e-9}


def cpu_baseline(wl, seconds: float):
    """Time the C restatement (oracle) on the same workload for about ``seconds``; then step
    the engine for the same steps and state each spectrum's error against that float64 run."""
    from oracle import workload as ow

    grid = wl.grid
    steps_cap = 2000
    faces, deposits, planes, poles = _oracle_inputs(wl, steps_cap)
    sim, oplanes = ow.build(grid.shape, grid.dt, grid.dx, faces, wl.inv, poles, deposits, planes,
                            grid.invariant)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ow.run(sim, oplanes, 1, reference_precision=False)  # warm (page in, thread pool)
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds and steps < steps_cap - 2:
        ow.run(sim, oplanes, 1, reference_precision=False)
        steps += 1
    elapsed = time.perf_counter() - t0
    base = {"value": round(grid.total_cells * steps / elapsed / 1e6, 2), "unit": "Mcell-steps/s",
            "cores": threads, "kind": "port",
            "sample": f"{wl.name}: {steps} steps of the C restatement (OpenMP, {threads} threads) "
                      f"incl. the {len(wl.planes)} flux planes in float64, {elapsed:.1f} s"}
    return base, spectrum_errors(wl, steps + 1, oplanes)


def oracle_precision(name: str, steps: int) -> dict:
    """Each spectrum's error against a float64 oracle run of ``steps`` steps of config ``name``."""
    from meep_gpu import workloads
    from oracle import workload as ow

    wl = workloads.config(name)
    grid = wl.grid
    faces, deposits, planes, poles = _oracle_inputs(wl, steps + 1)
    sim, oplanes = ow.build(grid.shape, grid.dt, grid.dx, faces, wl.inv, poles, deposits, planes,
                            grid.invariant)
    ow.run(sim, oplanes, steps, reference_precision=False)
    return spectrum_errors(wl, steps, oplanes)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_pairs_sphere192.txt")


def measured_traffic(pair: str, kernels: str):
    """HBM bytes per step of one fused pair (``pair`` "h_pair" or "d_pair", its launches named by
    ``kernels``) from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (scaled by the copy
    calibration it states), or None when it has no line for that pair and those kernels."""
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 4 and parts[0] == "traffic" and parts[1] == pair \
                        and parts[2] == kernels:
                    return int(parts[3])
    except OSError:
        pass
    return None


def pair_kernels(sim, shape) -> dict:
    """The kernels each fused pair launches for this workload under ``shape`` ((H, D) of
    (cells per thread, workgroups per CU, tile order, non-temporal stores, shell cells per
    thread, the D pair's halo tile rows)): the box's rows and the shell's rows as two lean
    kernels where the simulation runs the lossless forms, which the benchmark workloads do."""
    (hv, _, _, hnt, hsv, _), (dv, _, _, dnt, dsv, tile) = shape
    widths = set(sim.inverse_epsilon_bytes())
    inv = 1 if widths == {1} else 0
    pole = "true" if len(sim.poles) == 1 else "false"
    b = lambda x: "true" if x else "false"
    box = (f"dtile_kernel<{tile},{b(dnt)},{inv},{pole}>" if tile and dv == 1 else
           f"dshell_kernel<{dv},{b(dnt)},{inv},{pole},true>")
    return {"h_pair": f"hshell_kernel<{hv},{b(hnt)},true>+hshell_kernel<{hsv},{b(hnt)},false>",
            "d_pair": f"{box}+dshell_kernel<{dsv},{b(dnt)},{inv},{pole},false>"}


def compulsory_bytes(sim) -> dict:
    """Algorithmic HBM bytes per launch of each fused pair, the compulsory traffic of the
    stencil with every array touched once per launch and neighbour reads free, independent of
    this engine's storage choices: H pair E r 12, B r/w 24, H w 12 and, per component on its
    layer rows, fu_B r/w 8,
    H r 4, f_w_H r/w 8; D pair H r 12, D r/w 24, inverse epsilon r 12 (float32), E w 12 and,
    per component on its layer rows, fu_D r/w 8, E r 4, f_w_E r/w 8; plus P r, P_prev r, P' w
    and sigma r per driven pole component on the rows where a pole acts."""
    grid = sim.grid
    n = grid.total_cells
    pml = sim.pml
    frac = (lambda axis, half: pml.member_fraction(axis, half)) if pml else (lambda a, h: 0.0)
    h = n * 48.0
    d = n * 60.0
    for c in range(3):
        dsigu = (c + 2) % 3
        h += n * (8 * frac(dsigu, 1) + 12 * frac(c, 0))
        d += n * (8 * frac(dsigu, 0) + 12 * frac(c, 1))
    rows = sim.pole_row_fraction()
    for driven in sim.poles:
        d += n * 16 * sum(driven) * rows
    return {"h": h, "d": d}


def step_bytes(sim) -> float:
    b = algorithmic_bytes(sim)
    return float(b["h"] + b["d"])


def stream_ceiling(cells: int) -> dict:
    """The HBM rate the D pair's stream pattern reaches with no arithmetic: 9 read and 6 write
    float32 streams of ``cells`` (H-or-B, D, inv in; D, E out), one touch per thread."""
    from meep_gpu.engine import stream_probe

    ms = stream_probe(9, 6, cells, 0, 20)
    gbs = 15 * cells * 4 / (ms * 1e-3) / 1e9
    return {"GB/s": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4),
            "probe": f"fdtd_stream_probe 9 reads + 6 writes x {cells} floats, no arithmetic"}


def config_line(name: str, steps: int) -> dict:
    """Mcell-steps/s of another benchmark configuration (autotuned, graphs, flux every step)."""
    from meep_gpu import autotune, workloads
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.pml import PML

    wl = workloads.config(name)
    drv = FdtdDriver.from_workload(wl)
    sim = drv.freeze(planned_steps=steps + WARM + 1)
    shapes = autotune.tune(wl.grid, PML(wl.grid, wl.pml_faces) if wl.pml_faces else None,
                           wl.poles, use_cache=False, inv=wl.inv, target=sim)
    sim.run(WARM)  # the first step graph is captured (and replayed once) here
    sim.synchronize()
    t0 = time.perf_counter()
    sim.run(steps)
    sim.flush()  # the samples the window took and has not transformed yet
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    finite = all(numpy.all(numpy.isfinite(m.get_flux_spectrum())) for m in drv.flux_monitors)
    tile = sim.config().d[5]
    drv.close()
    return {"workload": f"{wl.name}: {wl.grid.total_cells} cells, {len(wl.poles)} Lorentz pole(s), "
                        f"{len(wl.planes)} flux planes x {wl.planes[0].frequencies.size} freqs",
            "steps_timed": steps, "value": round(wl.grid.total_cells * steps / elapsed / 1e6, 1),
            "unit": "Mcell-steps/s", "ms_per_step": round(elapsed * 1e3 / steps, 4),
            "launch_shape": {"h_pair": list(shapes[0]), "d_pair": list(shapes[1]),
                             "d_halo_tile_rows": tile},
            "spectra_finite": bool(finite)}


#: warm-up steps of the configuration, batched and one-at-a-time legs: one ring period, so the
#: first 128-step graph is captured before any timed window and the rings start it empty
WARM = 128

#: False under --no-graphs: every leg, the one-at-a-time baselines' "graphed" runs included,
#: launches eagerly (kernel tracing then sees no graph replay)
GRAPHS = True


def batched_line(count: int, steps: int, seed: int) -> dict:
    """The batched form: ``count`` uneven synthetic simulations stepped by shared launches,
    against the same simulations stepped one at a time."""
    from meep_gpu import autotune
    from meep_gpu.batch import BatchedRun, synthetic_batch
    from meep_gpu.driver import FdtdDriver

    t_build = time.perf_counter()
    wls = synthetic_batch(count=count, seed=seed, steps=steps)
    drivers = []
    for wl in wls:
        drv = FdtdDriver.from_workload(wl)
        drv.freeze(planned_steps=3 * steps + 3 * WARM
                    + 2 * 16 * (2 * len(autotune.CANDIDATES) + len(autotune.TILE_ROWS)) + 400)
        drivers.append(drv)
    build_s = time.perf_counter() - t_build
    cells = sum(d.grid.total_cells for d in drivers)
    batch = BatchedRun(drivers)
    shape = autotune.tune_batch(batch, steps=16)
    tile = batch.config().d[5]
    batch.flush()
    batch.run(WARM)  # warm: the first capture of the step graph, rings empty after it
    batch.synchronize()
    t0 = time.perf_counter()
    batch.run(steps)
    batch.flush()
    batch.synchronize()
    batched_s = time.perf_counter() - t0
    batch.close()
    nbytes = sum(step_bytes(d.sim) for d in drivers)
    # one at a time: each simulation alone, in turn, with its own autotuned shapes (its own
    # launch configuration); first as the plain form launches a step (eagerly), then with the
    # step graphs of the batched form
    import dataclasses

    alone = {}
    for graphs in (False, True):
        alone_s = 0.0
        for q, d in enumerate(drivers):
            if q % 16 == 0:
                print(f"bench: one at a time (graphs {graphs}), simulation {q}/{len(drivers)}",
                      file=sys.stderr, flush=True)
            autotune.tune(d.grid, d.sim.pml, [], repeats=3, target=d.sim)
            d.sim.set_config(dataclasses.replace(d.sim.config(), graphs=graphs and GRAPHS))
            d.sim.flush()
            d.sim.run(WARM)
            d.sim.synchronize()
            t0 = time.perf_counter()
            d.sim.run(steps)
            d.sim.flush()
            d.sim.synchronize()
            alone_s += time.perf_counter() - t0
        alone[graphs] = cells * steps / alone_s / 1e6
    finite = all(numpy.all(numpy.isfinite(m.get_flux_spectrum()))
                 for d in drivers for m in d.flux_monitors)
    for d in drivers:
        d.close()
    value = cells * steps / batched_s / 1e6
    achieved = nbytes * steps / batched_s / 1e9
    return {"workload": f"{count} synthetic sphere simulations (seed {seed}), edges 32-112 "
                        f"log-uniform, 25% Lorentz, UPML 8, 1 flux plane x 32 freqs",
            "cells": cells, "steps": steps, "value": round(value, 1), "unit": "Mcell-steps/s",
            "ms_per_step": round(batched_s * 1e3 / steps, 4),
            "one_at_a_time": round(alone[False], 1),
            "speedup_vs_one_at_a_time": round(value / alone[False], 2),
            "one_at_a_time_graphed": round(alone[True], 1),
            "speedup_vs_one_at_a_time_graphed": round(value / alone[True], 2),
            "one_at_a_time_note": ("each simulation alone with its own autotuned shapes: plain "
                                   "form (one launch sequence per step) and with 128-step graphs"),
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                         "note": "algorithmic bytes of both fused pairs per step, all members"},
            "launch_shape": {"h_pair": list(shape[0]), "d_pair": list(shape[1]),
                             "d_halo_tile_rows": tile},
            "graphs": GRAPHS, "spectra_finite": bool(finite),
            "host_build_s": round(build_s, 2)}


def served_line(count: int, rate: float, seed: int) -> dict:
    """The served form: a Poisson stream of solves under continuous batching, beside the same
    stream served one solve at a time (one slot, each solve with its step graphs)."""
    from meep_gpu.serve import SolveServer, poisson_requests

    reqs = poisson_requests(count=count, rate=rate, seed=seed)
    rep = SolveServer(max_slots=64, quantum=256).run(reqs)
    print("bench: served form, one solve at a time", file=sys.stderr, flush=True)
    one = SolveServer(max_slots=1, quantum=256).run(reqs).summary()
    s = rep.summary()
    return {"workload": f"{count} Poisson arrivals at {rate}/s (seed {seed}): sphere solves, "
                        f"edge 32-160 log-uniform, 30% Lorentz, 1-8 planes x 16-100 freqs, "
                        f"1000-8000 steps",
            "cell_steps": rep.cell_steps, **s,
            "one_at_a_time": {k: one[k] for k in ("p50_latency_s", "p99_latency_s",
                                                   "aggregate_mcell_steps_per_s", "makespan_s")},
            "one_at_a_time_note": ("the same stream through one slot: each solve alone, with "
                                   "its 128-step graphs" if GRAPHS else
                                   "the same stream through one slot: each solve alone"),
            "speedup_vs_one_at_a_time": round(s["aggregate_mcell_steps_per_s"]
                                              / max(one["aggregate_mcell_steps_per_s"], 1e-9),
                                              2)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="sphere192")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true",
                    help="skip the CPU restatement and the float64 oracle comparison")
    ap.add_argument("--precision-steps", type=int, default=160,
                    help="steps of the Lorentz slab's float64 oracle comparison")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the sphere80 and slab160 lines")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=3000)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--serve-count", type=int, default=256)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-saturated-rate", type=float, default=64.0,
                    help="a second stream whose arrivals outpace the device (capacity); 0: none")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every step eagerly, in every leg (for kernel tracing)")
    ap.add_argument("--launch", default="",
                    metavar="V,BPC,REMAP,NT[,SV][/V,BPC,REMAP,NT[,SV[,TILE]]]",
                    help="fixed launch shapes (H pair[/D pair], SV the shell's cells per "
                         "thread) instead of the autotuned ones (profiling runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="time a plain device copy of BYTES (profile counter calibration)")
    args = ap.parse_args(argv)
    if args.calibrate_copy:
        from meep_gpu.engine import copy_probe

        ms = copy_probe(args.calibrate_copy, 10)
        print(json.dumps({"copy_bytes": args.calibrate_copy, "mean_ms": ms,
                          "GB/s": 2 * args.calibrate_copy / (ms * 1e-3) / 1e9}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not run here (the engine is a "
              f"single-GPU engine)",
              file=sys.stderr)
        return 2

    from meep_gpu import workloads
    from meep_gpu.driver import FdtdDriver
    from meep_gpu.engine import device_info

    from meep_gpu import autotune
    from meep_gpu.pml import PML

    if args.no_graphs:
        from meep_gpu.engine import set_graphs

        global GRAPHS
        GRAPHS = False
        set_graphs(False)
    wl = workloads.config(args.workload)
    from meep_gpu.engine import get_launch_shell, get_launch_tile, set_launch_shell

    if args.launch:
        parts = args.launch.split("/")
        shapes = []
        tile = 0
        for pair, part in enumerate((parts[0], parts[-1])):
            fields = [int(x) for x in part.split(",")]
            v, bpc, remap, nt = fields[:4]
            set_launch_shell(pair, fields[4] if len(fields) > 4 else 1)
            if pair == 1 and len(fields) > 5:
                tile = fields[5]
            shapes.append((v, bpc, int(remap), bool(nt)))
        shape = tuple(shapes)
        autotune.install(shape, tile=tile)
    else:
        shape = autotune.tune(wl.grid, PML(wl.grid, wl.pml_faces) if wl.pml_faces else None,
                              wl.poles, use_cache=False, inv=wl.inv)
    shape = tuple(tuple(sh) + (get_launch_shell(pair), get_launch_tile() if pair else 0)
                  for pair, sh in enumerate(shape))
    drv = FdtdDriver.from_workload(wl)
    sim = drv.freeze(planned_steps=args.warmup + 2 * args.steps + 1)
    print(f"bench: {wl.name} {wl.grid.shape} on {device_info()}", file=sys.stderr, flush=True)
    sim.run(args.warmup)
    sim.flush()  # the window starts with empty rings: it transforms exactly its own samples
    sim.synchronize()
    t0 = time.perf_counter()
    total_ms, h_ms, d_ms = sim.run_timed(args.steps)  # ends with the pass over what is buffered
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    # the production path beside it: the same number of steps through fdtd_run (replayed
    # 128-step graphs, no per-pair events), rings transformed in the window as above
    production = None
    if GRAPHS:
        sim.flush()
        sim.synchronize()
        t1 = time.perf_counter()
        sim.run(args.steps)
        sim.flush()
        sim.synchronize()
        production = time.perf_counter() - t1
    cells = wl.grid.total_cells
    value = cells * args.steps / elapsed / 1e6
    nbytes = compulsory_bytes(sim)
    compact = algorithmic_bytes(sim)
    achieved = nbytes["d"] / (d_ms * 1e-3) / 1e9
    compact_achieved = compact["d"] / (d_ms * 1e-3) / 1e9
    kernels = pair_kernels(sim, shape)
    dkernel = kernels["d_pair"]
    traffic = measured_traffic("d_pair", dkernel)
    spectra_ok = all(numpy.all(numpy.isfinite(m.get_flux_spectrum())) for m in drv.flux_monitors)
    line = {
        "metric": "Mcell-steps/s (plain single-simulation FDTD step, fused HIP kernels)",
        "value": round(value, 1),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (fields), f64 (source deposit, flux transform)",
        "data": "synthetic",
        "config": {"workload": f"{wl.name}: {cells} cells {wl.grid.shape}, UPML 8, sphere r "
                               f"{0.2 * wl.grid.shape[0]:.1f} cells eps 2.25, Ez Gaussian point "
                               f"source, {len(wl.planes)} flux planes x "
                               f"{wl.planes[0].frequencies.size if wl.planes else 0} freqs",
                   "cells": cells, "steps_timed": args.steps},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_source": ("profiles/traffic_pairs_sphere192.txt (FETCH_SIZE and "
                                        "WRITE_SIZE per step over both launches, "
                                        "copy-calibrated)" if traffic else
                                        f"no profile line for {dkernel}"),
                     "kernel": (f"{dkernel} (step_D+deposit+update_E+update_P: the box's rows "
                                f"and the shell's rows)"),
                     "bytes_per_launch": int(nbytes["d"]), "mean_launch_ms": round(d_ms, 5),
                     "bytes_model": ("every array once per launch, neighbours free, layer "
                                     "auxiliaries on their rows, inverse epsilon float32"),
                     "compact_layout": {"bytes_per_launch": int(compact["d"]),
                                        "achieved": round(compact_achieved, 1),
                                        "frac": round(compact_achieved / HBM_PEAK_GBS, 4),
                                        "note": "the bytes this engine's layout must move: "
                                                "inverse epsilon as 1-byte table indices, H "
                                                "stored only on its layer rows"},
                     "h_pair": {"kernel": kernels["h_pair"],
                                "bytes_per_launch": int(nbytes["h"]),
                                "mean_launch_ms": round(h_ms, 5),
                                "achieved": round(nbytes["h"] / (h_ms * 1e-3) / 1e9, 1),
                                "frac": round(nbytes["h"] / (h_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                              4),
                                "compact_bytes_per_launch": int(compact["h"]),
                                "traffic": measured_traffic("h_pair", kernels["h_pair"])},
                     "step_bytes_per_cell": round((nbytes["h"] + nbytes["d"]) / cells, 2),
                     "compact_step_bytes_per_cell": round((compact["h"] + compact["d"]) / cells,
                                                          2)},
        "spectra_finite": bool(spectra_ok),
        "production_path": (None if production is None else {
            "value": round(cells * args.steps / production / 1e6, 1),
            "ms_per_step": round(production * 1e3 / args.steps, 4),
            "note": ("the same number of steps through fdtd_run after the timed window: "
                     "replayed 128-step graphs without the per-pair events `value` is timed "
                     "with")}),
        "launch_shape": {name: {"cells_per_thread": sh[0], "blocks_per_cu": sh[1],
                                "xcd_remap": int(sh[2]), "nt_stores": bool(sh[3]),
                                "shell_cells_per_thread": sh[4], "halo_tile_rows": sh[5]}
                         for name, sh in (("h_pair", shape[0]), ("d_pair", shape[1]))},
    }
    drv.close()
    line["roofline"]["stream_ceiling"] = stream_ceiling(2 * cells)
    if not args.no_configs:
        print("bench: other configurations", file=sys.stderr, flush=True)
        line["configs"] = {"sphere80": config_line("sphere80", 2000),
                           "slab160": config_line("slab160", 2000)}
    if not args.no_batched:
        print("bench: batched form", file=sys.stderr, flush=True)
        line["batched"] = batched_line(args.batch_count, args.batch_steps, 20261203)
    if not args.no_served:
        print("bench: served form", file=sys.stderr, flush=True)
        line["served"] = served_line(args.serve_count, args.serve_rate, 20261203)
        if args.serve_saturated_rate > 0:
            print("bench: served form, saturated", file=sys.stderr, flush=True)
            line["served_saturated"] = served_line(args.serve_count, args.serve_saturated_rate,
                                                   20261203)
    if not args.no_cpu_baseline:
        print("bench: CPU restatement and the float64 oracle comparison", file=sys.stderr,
              flush=True)
        line["cpu_baseline"], own = cpu_baseline(wl, args.cpu_seconds)
        line["precision"] = {wl.name: own}
        if not args.no_configs:
            line["precision"]["slab_160"] = oracle_precision("slab160", args.precision_steps)
        line["precision"]["note"] = ("fields float32 word for word with the oracle; phase and "
                                     "spectral sums float64 on both sides")
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
