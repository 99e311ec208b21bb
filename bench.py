"""Benchmark: the plain single-simulation FDTD step on one MI355X.

Workload (default): the 192^3 dielectric sphere (7,077,888 cells, 24 px/unit, 8-cell UPML on
all faces, sphere radius 38.4 cells, eps 2.25 with a sub-pixel diagonal chi1inv, Ez Gaussian
point source fcen 0.15 / df 0.1), 8 flux planes x 100 frequencies transformed in float64.
A step is one whole time step: the fused H pair, the fused D pair and the flux sampling (and,
every 32 steps, the float64 transform pass). Inputs are resident in HBM before timing starts.

Prints ONE JSON line. ``value`` is Mcell-steps/s = cells * steps / elapsed; ``roofline`` is
for the D pair (the heavier fused launch): algorithmic bytes per launch over its mean launch
time from HIP events around every launch in the timed region; ``cpu_baseline`` times the C
restatement (OpenMP on the host cores) on a bounded sample of the same workload.
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


def cpu_baseline(wl, seconds: float) -> dict:
    """Time the C restatement (oracle) on the same workload for about ``seconds``."""
    from meep_gpu.dft import FluxMonitor
    from meep_gpu.pml import PML, normalize_faces
    from meep_gpu.sources import current_table
    from oracle import workload as ow

    grid = wl.grid
    faces = normalize_faces(wl.pml_faces, grid) if wl.pml_faces else ((0, 0),) * 3
    steps_cap = 2000
    deposits = []
    for box, fn in wl.sources:
        deposits.append((box.component, box.lo, box.cnt, box.amp_re, box.amp_im,
                         current_table(fn, grid.dt, steps_cap, box.magnetic)))
    planes = []
    for p in wl.planes:
        mon = FluxMonitor(p.normal, p.lo, p.cnt, p.frequencies, grid.dx, grid.dt)
        planes.append((p.normal, p.lo, p.cnt, mon.omegas, mon.scale))
    poles = [(t.frequency, t.gamma, t.is_drude, s) for t, s in wl.poles]
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
    return {"value": round(grid.total_cells * steps / elapsed / 1e6, 2), "unit": "Mcell-steps/s",
            "cores": threads, "kind": "port",
            "sample": f"{wl.name}: {steps} steps of the C restatement (OpenMP, {threads} threads) "
                      f"incl. the {len(wl.planes)} flux planes in float64, {elapsed:.1f} s"}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_pairs_sphere192.txt")


def measured_traffic(kernel: str):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (scaled by the copy calibration it states), or None when it has no line for it."""
    try:
        with open(TRAFFIC_FILE) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kernel:
                    return int(parts[2])
    except OSError:
        pass
    return None


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
    shapes = autotune.tune(wl.grid, PML(wl.grid, wl.pml_faces) if wl.pml_faces else None,
                           wl.poles, use_cache=False, inv=wl.inv)
    drv = FdtdDriver.from_workload(wl)
    sim = drv.freeze(planned_steps=steps + 101)
    sim.run(100)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.run(steps)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    finite = all(numpy.all(numpy.isfinite(m.get_flux_spectrum())) for m in drv.flux_monitors)
    drv.close()
    return {"workload": f"{wl.name}: {wl.grid.total_cells} cells, {len(wl.poles)} Lorentz pole(s), "
                        f"{len(wl.planes)} flux planes x {wl.planes[0].frequencies.size} freqs",
            "steps_timed": steps, "value": round(wl.grid.total_cells * steps / elapsed / 1e6, 1),
            "unit": "Mcell-steps/s", "ms_per_step": round(elapsed * 1e3 / steps, 4),
            "launch_shape": {"h_pair": list(shapes[0]), "d_pair": list(shapes[1])},
            "spectra_finite": bool(finite)}


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
        drv.freeze(planned_steps=3 * steps + 3 * 64 + 4 * 16 * len(autotune.CANDIDATES) + 400)
        drivers.append(drv)
    build_s = time.perf_counter() - t_build
    cells = sum(d.grid.total_cells for d in drivers)
    batch = BatchedRun(drivers)
    shape = autotune.tune_batch(batch, steps=16)
    batch.run(64)  # warm: first capture of the step graph
    batch.synchronize()
    t0 = time.perf_counter()
    batch.run(steps)
    batch.synchronize()
    batched_s = time.perf_counter() - t0
    batch.close()
    nbytes = sum(step_bytes(d.sim) for d in drivers)
    # one at a time: each simulation alone, in turn, with its own autotuned shapes; first as
    # the plain form launches a step (eagerly), then with the step graphs of the batched form
    from meep_gpu.engine import set_graphs

    alone = {}
    for graphs in (False, True):
        set_graphs(graphs and GRAPHS)
        alone_s = 0.0
        for q, d in enumerate(drivers):
            if q % 16 == 0:
                print(f"bench: one at a time (graphs {graphs}), simulation {q}/{len(drivers)}",
                      file=sys.stderr, flush=True)
            autotune.tune(d.grid, d.sim.pml, [], repeats=3)
            d.sim.run(64)
            d.sim.synchronize()
            t0 = time.perf_counter()
            d.sim.run(steps)
            d.sim.synchronize()
            alone_s += time.perf_counter() - t0
        alone[graphs] = cells * steps / alone_s / 1e6
    set_graphs(GRAPHS)
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
            "launch_shape": {"h_pair": list(shape[0]), "d_pair": list(shape[1])}, "graphs": GRAPHS, "spectra_finite": bool(finite),
            "host_build_s": round(build_s, 2)}


def served_line(count: int, rate: float, seed: int) -> dict:
    """The served form: a Poisson stream of solves under continuous batching, beside the same
    stream served one solve at a time (one slot)."""
    from meep_gpu.serve import SolveServer, poisson_requests

    from meep_gpu.engine import set_graphs

    reqs = poisson_requests(count=count, rate=rate, seed=seed)
    rep = SolveServer(max_slots=64, quantum=256).run(reqs)
    one = {}
    for graphs in (False, True):
        set_graphs(graphs and GRAPHS)
        one[graphs] = SolveServer(max_slots=1, quantum=256).run(reqs).summary()
    set_graphs(GRAPHS)
    s, o, og = rep.summary(), one[False], one[True]
    return {"workload": f"{count} Poisson arrivals at {rate}/s (seed {seed}): sphere solves, "
                        f"edge 32-160 log-uniform, 30% Lorentz, 1-8 planes x 16-100 freqs, "
                        f"1000-8000 steps",
            "cell_steps": rep.cell_steps, **s,
            "one_at_a_time": {k: o[k] for k in ("p50_latency_s", "p99_latency_s",
                                                 "aggregate_mcell_steps_per_s", "makespan_s")},
            "speedup_vs_one_at_a_time": round(s["aggregate_mcell_steps_per_s"]
                                              / max(o["aggregate_mcell_steps_per_s"], 1e-9), 2),
            "one_at_a_time_graphed": {k: og[k] for k in ("p50_latency_s", "p99_latency_s",
                                                         "aggregate_mcell_steps_per_s",
                                                         "makespan_s")},
            "speedup_vs_one_at_a_time_graphed": round(
                s["aggregate_mcell_steps_per_s"] / max(og["aggregate_mcell_steps_per_s"], 1e-9), 2)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--workload", default="sphere192")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the sphere80 and slab160 lines")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=3000)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--serve-count", type=int, default=32)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-saturated-rate", type=float, default=32.0,
                    help="a second stream whose arrivals outpace the device (capacity)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every step eagerly, in every leg (for kernel tracing)")
    ap.add_argument("--launch", default="", metavar="V,BPC,REMAP,NT[/V,BPC,REMAP,NT]",
                    help="fixed launch shapes (H pair[/D pair]) instead of the autotuned ones "
                         "(profiling runs)")
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
              f"single-GPU engine; how a run would shard is described in DESIGN.md)",
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
    if args.launch:
        parts = args.launch.split("/")
        shapes = []
        for part in (parts[0], parts[-1]):
            v, bpc, remap, nt = (int(x) for x in part.split(","))
            shapes.append((v, bpc, int(remap), bool(nt)))
        shape = tuple(shapes)
        autotune.install(shape)
    else:
        shape = autotune.tune(wl.grid, PML(wl.grid, wl.pml_faces) if wl.pml_faces else None,
                              wl.poles, use_cache=False, inv=wl.inv)
    drv = FdtdDriver.from_workload(wl)
    sim = drv.freeze(planned_steps=args.warmup + args.steps + 1)
    print(f"bench: {wl.name} {wl.grid.shape} on {device_info()}", file=sys.stderr, flush=True)
    sim.run(args.warmup)
    sim.synchronize()
    t0 = time.perf_counter()
    total_ms, h_ms, d_ms = sim.run_timed(args.steps)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    cells = wl.grid.total_cells
    value = cells * args.steps / elapsed / 1e6
    nbytes = algorithmic_bytes(sim)
    achieved = nbytes["d"] / (d_ms * 1e-3) / 1e9
    dshape = shape[1]
    widths = set(sim.inverse_epsilon_bytes())
    inv_mode = {1: 1, 2: 2, 4: 0}[widths.pop()] if len(widths) == 1 else 3
    # the lossless periodic instantiation (the workload has no conductivity and no metallic wall)
    dkernel = f"dpair_kernel<{dshape[0]},{'true' if dshape[3] else 'false'},{inv_mode},false>"
    traffic = measured_traffic(dkernel)
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
                     "traffic_source": (f"profiles/traffic_pairs_sphere192.txt ({dkernel}: "
                                        "FETCH_SIZE x2 + WRITE_SIZE x1, copy-calibrated)"
                                        if traffic else f"no profile line for {dkernel}"),
                     "kernel": f"{dkernel} (step_D+deposit+update_E+update_P)",
                     "bytes_per_launch": int(nbytes["d"]), "mean_launch_ms": round(d_ms, 5),
                     "h_pair": {"bytes_per_launch": int(nbytes["h"]),
                                "mean_launch_ms": round(h_ms, 5),
                                "achieved": round(nbytes["h"] / (h_ms * 1e-3) / 1e9, 1)},
                     "step_model_bytes_per_cell": round((nbytes["h"] + nbytes["d"]) / cells, 2)},
        "spectra_finite": bool(spectra_ok),
        "launch_shape": {name: {"cells_per_thread": sh[0], "blocks_per_cu": sh[1],
                                "xcd_remap": int(sh[2]), "nt_stores": bool(sh[3])}
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
        line["cpu_baseline"] = cpu_baseline(wl, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
