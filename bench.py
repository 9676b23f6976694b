#!/usr/bin/env python3
"""Throughput of the coupled per-timestep lattice path on one MI355X.

A step is one SimulationRunner iteration (SimulationRunner.cpp:72-100) of the
2048x2048 tissue workload: D2Q9 BGK + one D2Q5 species, 400 cells (radius
40 +- 4, 150 markers each, remeshed to ~600) with membrane tension and an area
regulator, box outlet.  Prints one JSON line on stdout; progress goes to
stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def tissue_case(nx=2048, ny=2048, cols=20, rows=20, seed=1):
    from lbibcell_amd import workloads
    return workloads.hex_tissue(nx, ny, cols, rows, radius=40.0, rjit=4.0, n=150, pos_jit=2.0, seed=seed)


SIGNAL = dict(production=1e-3, decay=1e-5, switchoff=5000, init_inside=0.0, prod_domain=-1)
TAU_F, TAU_C = 1.0, 0.8
TENSION, AREA = 0.01, (200.0, 1e-5, 0.004, 1.2)


def build_device(nx, ny, nodes, conns, precision):
    from lbibcell_amd import capi, runner
    sim = runner.Simulation(nx, ny, TAU_F, precision=precision)
    sim.add_species("tutorial_01_CDESolverD2Q5_SIGNAL", capi.species_spec(capi.SPECIES_SIGNAL, TAU_C, **SIGNAL))
    sim.add_geometry(nodes, conns)
    for d in sorted({c[2] for c in conns}):
        sim.set_celltype(d, 2)
    sim.setup()
    sim.init_solvers()
    sim.add_mass_solver("MassSolverBoxOutlet")
    sim.add_area_regulator(*AREA, freq=1)
    sim.add_tension(TENSION, 10)
    return sim


def cpu_baseline(nx, ny, nodes, conns, budget_s=15.0, max_steps=50):
    """The oracle (a CPU restatement of the reference path, OpenMP on the
    race-free lattice loops) on the same workload, timed for a bounded sample."""
    from oracle import orc
    threads = min(16, os.cpu_count() or 1)
    o = orc.Oracle(nx, ny, TAU_F)
    o.set_threads(threads)
    o.add_species(orc.SP_SIGNAL, TAU_C, [SIGNAL["production"], SIGNAL["decay"], SIGNAL["switchoff"],
                                         SIGNAL["init_inside"], SIGNAL["prod_domain"]])
    for n in nodes:
        o.add_node(*n)
    for c in conns:
        o.add_connection(*c)
    for d in sorted({c[2] for c in conns}):
        o.set_celltype(d, 2)
    o.cure()
    o.init_solvers()
    o.box_outlet(True)
    o.add_bio(1, list(AREA), 1)
    o.add_bio(0, [TENSION], 10)
    o.run(2)            # first steps remesh the initial polygons
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps and time.perf_counter() - t0 < budget_s:
        o.run(1)
        steps += 1
    dt = time.perf_counter() - t0
    return {"value": nx * ny * steps / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"{steps} coupled steps (after 2 warm-up steps) of the same {nx}x{ny} 400-cell workload, "
                      f"oracle with {threads} OpenMP threads on the lattice loops, serial IB/geometry"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="fp64", choices=["fp64"])
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--run-ahead", type=int, default=16,
                    help="iterations enqueued between device status reads")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the lattice path runs on a single MI355X per process; "
              "multi-GPU slab decomposition is a design note, not built", file=sys.stderr)
        sys.exit(2)
    from lbibcell_amd import capi
    prec = {"fp64": capi.FP64}[args.precision]
    nx = ny = args.size
    cols = rows = max(1, round(20 * args.size / 2048))
    nodes, conns = tissue_case(nx, ny, cols, rows)
    log(f"workload {nx}x{ny}, {len(set(c[2] for c in conns))} cells, {len(nodes)} initial markers")
    sim = build_device(nx, ny, nodes, conns, prec)
    sim.set_run_ahead(args.run_ahead)
    log("device setup done; warming up")
    sim.run(args.warmup)
    sim.sync()
    ctx = sim.context()
    ctx.profile_enable(True)
    ctx.profile_read()
    sim.reset_times()
    log(f"timing {args.steps} steps")
    sim.sync()
    t0 = time.perf_counter()
    sim.run(args.steps)
    sim.sync()
    t1 = time.perf_counter()
    ms, launches, bpl = ctx.profile_read()
    ctx.profile_enable(False)
    dt = t1 - t0
    n = nx * ny
    mlups = n * args.steps / dt / 1e6
    kern_s = ms / 1e3 / max(launches, 1)
    achieved = bpl / kern_s / 1e9
    copy_bytes = 1 << 30
    copy_ms = ctx.profile_copy(copy_bytes, 10)
    copy_gbs = 2 * copy_bytes / (copy_ms / 1e3) / 1e9
    nmark = len(sim.markers()[0])
    times = sim.times()
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_stream_collide_{args.precision}_{nx}.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("bytes_per_launch")
    out = {
        "metric": "MLUPS (coupled D2Q9+D2Q5+IB lattice step)",
        "value": mlups,
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64"}[args.precision],
        "data": "synthetic (seeded hex tissue geometry)",
        "config": {"workload": f"tissue {nx}x{ny}: D2Q9+D2Q5, {len(set(c[2] for c in conns))} cells r=40+-4, "
                               f"150 markers/cell (remeshed), tension 0.01, area regulator 1.2xA0, box outlet",
                   "precision": args.precision, "markers_after_remesh": nmark, "run_ahead": args.run_ahead},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_stream_collide_f64<1,0>", "kernel_ms_avg": kern_s * 1e3,
                     "algorithmic_bytes_per_launch": bpl,
                     "copy_probe_gbs": copy_gbs, "frac_of_copy": achieved / copy_gbs},
        "host_seconds_per_step": {k: v / max(times["steps"], 1) for k, v in times.items()
                                  if k not in ("steps", "stops")},
        "resolved_stops": times["stops"],
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(nx, ny, nodes, conns)
    sim.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
