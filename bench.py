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
    # one GPU's share of the box's host: 16 threads (OMP_NUM_THREADS on the box)
    box_cpus = os.cpu_count() or 1
    threads = min(int(os.environ.get("OMP_NUM_THREADS", "16") or 16), box_cpus)
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
    before = o.stage_times()
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps and time.perf_counter() - t0 < budget_s:
        o.run(1)
        steps += 1
    dt = time.perf_counter() - t0
    after = o.stage_times()
    return {"value": nx * ny * steps / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "host_cpus_visible": box_cpus,
            "seconds_per_step_by_stage": {k: (after[k] - before[k]) / max(steps, 1) for k in after},
            "sample": f"{steps} coupled steps (after 2 warm-up steps) of the same {nx}x{ny} 400-cell workload, "
                      f"oracle with {threads} OpenMP threads (one GPU's share of the box's host) on the "
                      f"lattice loops, serial IB/geometry"}


PRECISIONS = {"fp64": 0, "fp32": 1, "fp16": 2, "bf16": 3, "fp16r": 4}
DTYPES = {"fp64": "f64", "fp32": "f32", "fp16": "f32 arithmetic, f16 storage (f - w)",
          "bf16": "f32 arithmetic, bf16 storage (f - w)",
          "fp16r": "f32 arithmetic, f32 moments + f16 remainders (f - feq) storage"}
# the lattice kernel each precision uses on one lattice with the default
# shape (one site per lane in 256-wide groups)
# the lattice kernel's HIP-event timing brackets every PROFILE_EVERY-th launch
# of the timed region (each event pair idles the stream for a few microseconds)
PROFILE_EVERY = 4
KERNELS = {"fp64": "k_stream_collide_f64<1,0>", "fp32": "k_stream_collide_mx<1,0,FP32,1>",
           "fp16": "k_stream_collide_mx<1,0,FP16_DEV,1>", "bf16": "k_stream_collide_mx<1,0,BF16_DEV,1>",
           "fp16r": "k_stream_collide_rem<1,0>"}


def measure(nx, ny, nodes, conns, prec_name, steps, warmup, run_ahead):
    """Build the workload at one storage precision, warm up, time `steps`
    coupled steps.  Returns timings, the lattice-kernel roofline figures and
    the final fields (for the error of mixed precision against FP64)."""
    sim = build_device(nx, ny, nodes, conns, PRECISIONS[prec_name])
    sim.set_run_ahead(run_ahead)
    log(f"[{prec_name}] device setup done; warming up")
    sim.run(warmup)
    sim.sync()
    ctx = sim.context()
    ctx.profile_enable(True, every=PROFILE_EVERY)
    ctx.profile_read()
    sim.reset_times()
    log(f"[{prec_name}] timing {steps} steps")
    sim.sync()
    t0 = time.perf_counter()
    sim.run(steps)
    sim.sync()
    t1 = time.perf_counter()
    ms, launches, bpl = ctx.profile_read()
    ctx.profile_enable(False)
    ctx.ns = 1
    u, v = ctx.get_velocity()
    fields = dict(speed=np.hypot(u, v), conc=ctx.get_species(0).sum(axis=0), area=sim.areas())
    out = dict(dt=t1 - t0, kernel_s=ms / 1e3 / max(launches, 1), bytes_per_launch=bpl,
               markers=len(sim.markers()[0]), times=sim.times(), fields=fields, ctx=ctx, sim=sim)
    return out


def build_batched(ens, precision):
    """The batched ensemble as one atlas context: closed periodic domains
    (per-domain tau_f and tau_c), the SIGNAL species produced inside cells and
    the membrane tension.  No area regulator or box outlet: in domains of
    64-512 sites the outlet's drift carries growing cells to the domain edge
    within ~150 steps, and a sweep runs thousands."""
    from lbibcell_amd import capi, runner
    sim = runner.Simulation(ens["nx"], ens["ny"], TAU_F, precision=precision)
    sim.add_species("tutorial_01_CDESolverD2Q5_SIGNAL", capi.species_spec(capi.SPECIES_SIGNAL, TAU_C, **SIGNAL))
    sim.set_batch_layout(ens["rects"])
    sim.add_geometry(ens["nodes"], ens["conns"])
    sim.setup()
    sim.init_solvers()
    sim.add_tension(TENSION, 10)
    return sim


def measure_batched(ens, prec_name, steps, warmup, run_ahead, graph_steps=10):
    """Batched ensemble at one storage precision: autotune the lattice launch
    shape after the warm-up, then time `steps` coupled steps of all domains,
    replayed as recorded step graphs of graph_steps iterations (blocks start
    after iterations that are multiples of graph_steps)."""
    sim = build_batched(ens, PRECISIONS[prec_name])
    sim.set_run_ahead(run_ahead)
    log(f"[batched {prec_name}] setup done; warming up")
    sim.run(warmup)
    sim.sync()
    ctx = sim.context()
    chosen, rows = ctx.autotune(reps=5)
    log(f"[batched {prec_name}] tuned lattice shape {chosen[:3]} ({chosen[3] * 1e3:.1f} us)")
    if graph_steps:
        sim.set_graph_steps(graph_steps)
        sim.run(graph_steps)          # records the block once, outside the timed region
        sim.sync()
    g0 = sim.graph_stats()
    ctx.profile_enable(True, every=PROFILE_EVERY)
    ctx.profile_read()
    sim.reset_times()
    sim.sync()
    t0 = time.perf_counter()
    sim.run(steps)
    sim.sync()
    t1 = time.perf_counter()
    ms, launches, bpl = ctx.profile_read()
    ctx.profile_enable(False)
    ctx.ns = 1
    u, v = ctx.get_velocity()
    live = np.zeros((ens["ny"], ens["nx"]), dtype=bool)
    for x0, y0, w, h, *_ in ens["rects"]:
        live[y0:y0 + h, x0:x0 + w] = True
    fields = dict(speed=np.hypot(u, v)[live], conc=ctx.get_species(0).sum(axis=0)[live], area=sim.areas())
    g1 = sim.graph_stats()
    # inside step graphs the per-launch events see no lattice launch: the
    # kernel time is then the tuner's average over launches on the live state
    kernel_s = ms / 1e3 / launches if launches else chosen[3] / 1e3
    out = dict(dt=t1 - t0, kernel_s=kernel_s, kernel_time_source="events" if launches else "autotuner",
               bytes_per_launch=bpl, tuned=chosen,
               tried=rows, markers=len(sim.markers()[0]), times=sim.times(), fields=fields,
               graph={"steps_per_graph": g1[0], "launches_in_timed_region": g1[1] - g0[1],
                      "recordings": g1[2], "note": sim.graph_note()})
    sim.close()
    return out


MS_TAUS = (0.8, 0.9, 1.0, 1.1)
MS_SOURCE = 1e-3
JUNCTION = dict(radius=2.0, k=0.02, l0=0.5, freq=10)


def multispecies_case(size=4096, ncell=1600, seed=1):
    """1600 cells r = 30 +- 5 (120 markers each) grown as one touching
    cluster, and a 4x4 linear reaction matrix with entries N(0, (1e-4)^2)."""
    from lbibcell_amd import workloads
    nodes, conns = workloads.packed_tissue(size, size, ncell, 30.0, 5.0, 120, seed=seed)
    K = np.random.default_rng(seed + 100).normal(0.0, 1e-4, size=(4, 4))
    return nodes, conns, K


def measure_multispecies(size, nodes, conns, K, prec_name, steps, warmup, run_ahead):
    """Four coupled D2Q5 species (LINEAR kind: R_s = sum_j K[s][j] C_j, plus
    a source inside cells for the first), membrane tension and cell
    junctions within 2 lattice units, box outlet; times `steps` coupled
    steps after `warmup`."""
    from lbibcell_amd import capi, runner
    sim = runner.Simulation(size, size, TAU_F, precision=PRECISIONS[prec_name])
    for s in range(4):
        sim.add_species(f"LINEAR_{s}", capi.species_spec(capi.SPECIES_LINEAR, MS_TAUS[s], init_inside=0.1,
                                                          source_inside=MS_SOURCE if s == 0 else 0.0,
                                                          krow=tuple(K[s])))
    sim.add_geometry(nodes, conns)
    sim.setup()
    sim.init_solvers()
    sim.add_mass_solver("MassSolverBoxOutlet")
    sim.add_tension(TENSION, 10)
    sim.add_junctions(**JUNCTION)
    sim.set_run_ahead(run_ahead)
    log(f"[multispecies {prec_name}] setup done; warming up")
    sim.run(warmup)
    sim.sync()
    ctx = sim.context()
    chosen, rows = ctx.autotune(reps=5)
    log(f"[multispecies {prec_name}] tuned lattice shape {chosen[:3]} ({chosen[3] * 1e3:.1f} us)")
    ctx.profile_enable(True, every=PROFILE_EVERY)
    ctx.profile_read()
    sim.reset_times()
    sim.sync()
    t0 = time.perf_counter()
    sim.run(steps)
    sim.sync()
    t1 = time.perf_counter()
    ms, launches, bpl = ctx.profile_read()
    ctx.profile_enable(False)
    ids, _, _ = sim.markers()
    off, t, _, _ = ctx.get_forces(len(ids))
    ctx.ns = 4
    u, v = ctx.get_velocity()
    fields = dict(speed=np.hypot(u, v), conc=[ctx.get_species(s).sum(axis=0) for s in range(4)], area=sim.areas())
    out = dict(dt=t1 - t0, kernel_s=ms / 1e3 / max(launches, 1), bytes_per_launch=bpl, markers=len(ids),
               junction_structs=int(np.sum(t == 7)), times=sim.times(), fields=fields, tuned=chosen, tried=rows)
    sim.close()
    return out


def measure_adjoint(steps, K, size=512):
    """Adjoint workload: size^2 lattice, 3x3 cells (radius 40, 150 markers),
    the target is the final concentration under hidden parameters (tau_c 0.8,
    kappa 0.01, per-cell rates); loss and gradient are evaluated at a guess
    (tau_c 0.9, kappa 0.006, 0.7 x rates).  FP32 is the measured path, FP64
    its gradient reference; central differences check the FP64 gradient."""
    from lbibcell_amd import adjoint, capi, workloads
    rings, dom, rates = workloads.adjoint_case(size, size, 3, 3, 40.0, 150, seed=4)
    guess = (0.9, 0.006, rates * 0.7)
    out = {}
    for name, prec in (("fp32", capi.FP32), ("fp64", capi.FP64)):
        p = adjoint.AdjointProblem(size, size, 1.0, rings, dom, decay=1e-5, precision=prec, checkpoint_every=K)
        _, tgt = p.forward(steps, 0.8, 0.01, rates, want_field=True)
        p.set_target(tgt)
        p.forward(steps, *guess)          # warm-up of the timed pair
        p.gradient()
        loss = p.forward(steps, *guess)
        g = p.gradient()
        fwd_ms, grad_ms = p.timing()
        r = {"loss": loss, "d_tau_c": g[0], "d_kappa": g[1], "d_rates": list(g[2]), "forward_ms": fwd_ms,
             "gradient_ms": grad_ms, "adjoint_to_forward": grad_ms / fwd_ms}
        if name == "fp64":
            def fd(idx, h):
                vals = []
                for sgn in (1, -1):
                    tc, kp, rt = guess[0], guess[1], guess[2].copy()
                    if idx == 0:
                        tc += sgn * h
                    elif idx == 1:
                        kp += sgn * h
                    else:
                        rt[0] += sgn * h
                    vals.append(p.forward(steps, tc, kp, rt))
                return (vals[0] - vals[1]) / (2 * h)
            fds = [fd(0, 1e-5), fd(1, 1e-6), fd(2, 1e-7)]
            r["central_difference_rel_error"] = {
                "tau_c": abs(g[0] - fds[0]) / abs(fds[0]), "kappa": abs(g[1] - fds[1]) / abs(fds[1]),
                "rate_cell1": abs(g[2][0] - fds[2]) / abs(fds[2])}
        out[name] = r
        p.close()
    g32, g64 = out["fp32"], out["fp64"]
    out["fp32_vs_fp64_rel_error"] = {
        "tau_c": abs(g32["d_tau_c"] - g64["d_tau_c"]) / abs(g64["d_tau_c"]),
        "kappa": abs(g32["d_kappa"] - g64["d_kappa"]) / abs(g64["d_kappa"]),
        "rates": float(np.max(np.abs(np.array(g32["d_rates"]) - np.array(g64["d_rates"]))
                              / np.abs(np.array(g64["d_rates"]))))}
    return {"workload": f"adjoint {size}x{size}, 9 cells r=40 (150 markers), {steps} steps, checkpoints every {K}, "
                        "loss = L2 mismatch of the final concentration against a target from hidden parameters",
            "steps": steps, "checkpoint_every": K, **out}


def adjoint_ms_case(size, ncell, batch):
    """The four-species junction workload as `batch` x `batch` independent
    problems of (size / batch)^2 in one size^2 lattice (ncell cells in all):
    marker rings, cell map, per-cell rates, junction pairs (within each
    problem) and the rectangles."""
    from lbibcell_amd import workloads
    side = size // batch
    rings, rates, pairs, rects = [], [], [], []
    dom = np.zeros((size, size), dtype=np.uint32)
    base_m = base_c = 0
    for by in range(batch):
        for bx in range(batch):
            k = by * batch + bx
            r, d, rt = workloads.tissue_adjoint_case(side, side, ncell // (batch * batch), 30.0, 5.0, 120, seed=1 + k)
            pr = workloads.junction_pairs(r, JUNCTION["radius"])
            off = np.array([bx * side, by * side], dtype=np.float64)
            rings += [q + off for q in r]
            dd = d.astype(np.uint32)
            dom[by * side:(by + 1) * side, bx * side:(bx + 1) * side] = np.where(dd > 0, dd + base_c, 0)
            rates.append(rt)
            pairs.append(pr + base_m)
            rects.append((bx * side, by * side, side, side))
            base_m += sum(len(q) for q in r)
            base_c += len(r)
    return dict(rings=rings, dom=dom, rates=np.concatenate(rates), pairs=np.concatenate(pairs),
                rects=rects if batch > 1 else None)


def measure_adjoint_ms(size, steps, K, ncell=1600, batch=2, snapshots=20):
    """Adjoint of the four-species junction workload: size^2 lattice holding
    batch x batch independent problems (one summed loss, shared parameters),
    ncell touching cells in all (r = 30 +- 5, 120 markers), junction springs
    between the closest markers of different cells within 2 lu, four species
    with a linear reaction matrix N(0, (1e-4)^2) and per-cell sources of
    species 0.  The target is the final concentration of every species under
    hidden parameters (FP64); loss and gradient are evaluated at a guess in
    three set-ups: FP32 with checkpoints every K steps, the mixed-precision
    one (FP16 deviation storage, FP32 arithmetic, revolve with `snapshots`
    slots), and the FP64 reference gradient (revolve) that both are compared
    with."""
    from lbibcell_amd import adjoint, capi
    ns = 4
    case = adjoint_ms_case(size, ncell, batch)
    Km = np.random.default_rng(101).normal(0.0, 1e-4, size=(ns, ns))
    g0 = np.zeros((ns, 5, size, size))
    decay = [1e-5] * ns
    hidden = dict(tau_c=list(MS_TAUS), kappa=0.01, k_junction=JUNCTION["k"], K=Km, rates=case["rates"])
    guess = dict(tau_c=[t + 0.05 for t in MS_TAUS], kappa=0.006, k_junction=0.6 * JUNCTION["k"], K=0.7 * Km,
                 rates=0.7 * case["rates"])

    def problem(precision, target, storage=None, revolve=None):
        p = adjoint.AdjointProblem(size, size, 1.0, case["rings"], case["dom"], decay=decay[0], precision=precision,
                                   checkpoint_every=K, storage=storage, revolve_snapshots=revolve)
        p.set_species(g0, target, decay)
        p.set_junctions(case["pairs"], JUNCTION["l0"])
        if case["rects"]:
            p.set_batch_layout(case["rects"])
        return p

    p = problem(capi.FP64, np.zeros((ns, size, size)), revolve=snapshots)
    _, target = p.forward_ms(steps, want_field=True, **hidden)
    p.close()
    runs = {}
    for name, prec, storage, rev in (("fp64_revolve", capi.FP64, None, snapshots), ("fp32", capi.FP32, None, None),
                                     ("fp16_revolve", capi.FP32, capi.FP16_DEV, snapshots)):
        log(f"  adjoint four species: {name}")
        p = problem(prec, target, storage, rev)
        p.forward_ms(steps, **guess)          # warm-up of the timed pair
        p.gradient_ms()
        loss = p.forward_ms(steps, **guess)
        g = p.gradient_ms()
        fwd_ms, grad_ms = p.timing()
        st = p.stats()
        p.close()
        runs[name] = dict(loss=loss, g=g, forward_ms=fwd_ms, gradient_ms=grad_ms, stats=st)
    ref = runs["fp64_revolve"]["g"]

    def rel(g):
        return {k: float(np.max(np.abs(np.asarray(g[k]) - np.asarray(ref[k]))) /
                         max(float(np.max(np.abs(np.asarray(ref[k])))), 1e-300)) for k in ref}

    out = {"workload": f"adjoint {size}x{size} holding {batch * batch} independent {size // batch}^2 problems "
                       f"(one batched lattice, summed loss), {ncell} touching cells r=30+-5 (120 markers), "
                       f"{len(case['pairs'])} junction springs within {JUNCTION['radius']} lu, 4 species with a linear "
                       f"reaction matrix, {steps} steps; loss = L2 mismatch of the final concentrations against a "
                       "target from hidden parameters",
           "steps": steps, "markers": int(sum(len(r) for r in case["rings"])),
           "junction_springs": int(len(case["pairs"])), "problems": batch * batch}
    for name, r in runs.items():
        e = {"loss": r["loss"], "forward_ms": r["forward_ms"], "gradient_ms": r["gradient_ms"],
             "adjoint_to_forward": r["gradient_ms"] / r["forward_ms"],
             "checkpointing": (f"revolve, {snapshots} snapshot slots" if name.endswith("revolve")
                               else f"every {K} steps + {K}-state segment store"),
             "storage": {"fp64_revolve": "fp64", "fp32": "fp32", "fp16_revolve": "fp16 deviations, fp32 arithmetic"}[name],
             "checkpoint_store_gb": r["stats"]["store_bytes"] / 1e9,
             "gradient_recomputed_steps": r["stats"]["gradient_steps"]}
        if name != "fp64_revolve":
            e["gradient_rel_error_vs_fp64"] = rel(r["g"])
        out[name] = e
    g = ref
    out["fp64_gradient"] = {"d_tau_c": list(g["tau_c"]), "d_kappa": g["kappa"], "d_k_junction": g["k_junction"],
                            "d_K": g["K"].tolist(), "d_rates_norm": float(np.linalg.norm(g["rates"]))}
    return out


LONG_CHECKS = (100, 1000, 5000)


# The sensitivity of the coupled model itself: FP64 runs whose initial marker
# positions differ from the reference run's by at most JITTER_LU lattice units
# (1e-12: far below any rounding of the path; 1e-8: about what FP32
# arithmetic moves a marker by in a hundred steps).  Their distance from the
# FP64 run shows how far the model carries a small difference in the input.
JITTER_LU = (1e-12, 1e-8)


def _jittered(nodes, amp, seed=7):
    rng = np.random.default_rng(seed)
    d = rng.uniform(-amp, amp, size=(len(nodes), 2))
    return [(nid, x + d[k, 0], y + d[k, 1]) for k, (nid, x, y) in enumerate(nodes)]


def long_window_errors(kind, precisions, steps):
    """L2 errors of every storage precision against the FP64 run over a long
    window, sampled at LONG_CHECKS (<= steps).  kind "tissue": the 2048^2 cells
    of the headline laid out 100 sites from the edges, area regulator KP 1e-6
    and clamp 4e-4 (the headline's own regulator carries the grown tissue into
    the lattice edge in iteration 88, as it does in the reference), box outlet;
    kind "ensemble": the 256-domain batched ensemble.  "fp64_jitter<amp>" are
    the FP64 runs from initial positions moved by at most amp.  Returns
    {prec: {step: {speed, concentration, cell_area}}} against FP64, the same
    against FP32 for the 16-bit stores (their storage error alone), plus the
    seconds each run took."""
    from lbibcell_amd import capi, runner, workloads
    checks = [c for c in LONG_CHECKS if c <= steps]
    fields, secs = {}, {}
    jit = {f"fp64_jitter{a:.0e}": a for a in JITTER_LU}
    for p in ["fp64"] + list(jit) + [q for q in precisions if q != "fp64"]:
        prec = "fp64" if p in jit else p
        if kind == "tissue":
            nodes, conns = workloads.hex_tissue(2048, 2048, 20, 20, radius=40.0, rjit=4.0, n=150, pos_jit=2.0,
                                                seed=1, margin=100.0)
            if p in jit:
                nodes = _jittered(nodes, jit[p])
            sim = runner.Simulation(2048, 2048, TAU_F, precision=PRECISIONS[prec])
            sim.add_species("tutorial_01_CDESolverD2Q5_SIGNAL",
                            capi.species_spec(capi.SPECIES_SIGNAL, TAU_C, **SIGNAL))
            sim.add_geometry(nodes, conns)
            for d in sorted({c[2] for c in conns}):
                sim.set_celltype(d, 2)
            sim.setup()
            sim.init_solvers()
            sim.add_mass_solver("MassSolverBoxOutlet")
            sim.add_area_regulator(200.0, 1e-6, 4e-4, 1.2, freq=1)
            sim.add_tension(TENSION, 10)
            live = None
        else:
            ens = workloads.batched_ensemble(ndom=256, seed=1)
            if p in jit:
                ens = dict(ens, nodes=_jittered(ens["nodes"], jit[p]))
            sim = build_batched(ens, PRECISIONS[prec])
            live = np.zeros((ens["ny"], ens["nx"]), dtype=bool)
            for x0, y0, w, h, *_ in ens["rects"]:
                live[y0:y0 + h, x0:x0 + w] = True
        log(f"[long window {kind} {p}] {steps} steps")
        t0 = time.perf_counter()
        fields[p], done = {}, 0
        for c in checks:
            sim.run(c - done)
            done = c
            ctx = sim.context()
            ctx.ns = 1
            u, v = ctx.get_velocity()
            sp, C = np.hypot(u, v), ctx.get_species(0).sum(axis=0)
            if live is not None:
                sp, C = sp[live], C[live]
            fields[p][c] = (sp, C, sim.areas())
        secs[p] = time.perf_counter() - t0
        sim.close()
    def errors(ref, skip):
        out = {}
        for p in fields:
            if p in skip:
                continue
            out[p] = {}
            for c in checks:
                sp, C, a = fields[p][c]
                sr, Cr, ar = fields[ref][c]
                common = sorted(set(a) & set(ar))
                out[p][str(c)] = {"speed": l2_rel(sp, sr), "concentration": l2_rel(C, Cr),
                                  "cell_area": l2_rel([a[d] for d in common], [ar[d] for d in common])}
        return out
    res = {"errors_vs_fp64": errors("fp64", {"fp64"}), "seconds": secs, "steps": steps}
    if "fp32" in fields:
        res["errors_vs_fp32"] = errors("fp32", {"fp64", "fp32"} | set(jit))
    return res


def l2_rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", default="fp64", choices=list(PRECISIONS),
                    help="storage precision of the headline line (fp64 is the parity path)")
    ap.add_argument("--mixed", default="fp32,fp16,bf16,fp16r",
                    help="comma-separated mixed precisions also measured and compared with fp64 ('' = none)")
    ap.add_argument("--size", type=int, default=2048)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--run-ahead", type=int, default=16,
                    help="iterations enqueued between device status reads")
    ap.add_argument("--batched", default="fp16,fp16r,fp32,fp64",
                    help="storage precisions of the batched-ensemble measurement ('' = skip); "
                         "fp64 also serves as the reference of the mixed errors")
    ap.add_argument("--batched-domains", type=int, default=256)
    ap.add_argument("--batched-steps", type=int, default=20)
    ap.add_argument("--adjoint-steps", type=int, default=1000, help="0 = skip the adjoint measurement")
    ap.add_argument("--adjoint-ms-steps", type=int, default=200,
                    help="steps of the four-species junction adjoint window (0 = skip)")
    ap.add_argument("--adjoint-ms-size", type=int, default=4096)
    ap.add_argument("--adjoint-ms-batch", type=int, default=2,
                    help="the four-species adjoint lattice holds batch x batch independent problems")
    ap.add_argument("--adjoint-ms-snapshots", type=int, default=20, help="revolve snapshot slots")
    ap.add_argument("--multispecies", default="fp16,fp16r,fp64",
                    help="storage precisions of the 4096^2 four-species junction workload ('' = skip); "
                         "fp64 also serves as the reference of the mixed errors")
    ap.add_argument("--multispecies-size", type=int, default=4096)
    ap.add_argument("--long-window", default="tissue,ensemble",
                    help="workloads whose mixed-precision errors are also measured over a long window ('' = skip)")
    ap.add_argument("--long-steps", type=int, default=5000)
    ap.add_argument("--long-precisions", default="fp32,fp16r,fp16")
    ap.add_argument("--multispecies-steps", type=int, default=20)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the lattice path runs on a single MI355X per process; "
              "multi-GPU slab decomposition is a design note, not built", file=sys.stderr)
        sys.exit(2)
    nx = ny = args.size
    cols = rows = max(1, round(20 * args.size / 2048))
    nodes, conns = tissue_case(nx, ny, cols, rows)
    ncell = len(set(c[2] for c in conns))
    log(f"workload {nx}x{ny}, {ncell} cells, {len(nodes)} initial markers")
    n = nx * ny
    variants = [args.precision] + [p for p in args.mixed.split(",") if p and p != args.precision]
    if "fp64" not in variants and variants[1:]:
        variants.insert(1, "fp64")   # mixed errors are taken against the FP64 path
    res = {}
    copy_gbs = None
    for pname in variants:
        r = measure(nx, ny, nodes, conns, pname, args.steps, args.warmup, args.run_ahead)
        if copy_gbs is None:
            copy_bytes = 1 << 30
            copy_gbs = 2 * copy_bytes / (r["ctx"].profile_copy(copy_bytes, 10) / 1e3) / 1e9
        r["sim"].close()
        del r["sim"], r["ctx"]
        res[pname] = r

    def roofline(pname):
        r = res[pname]
        achieved = r["bytes_per_launch"] / r["kernel_s"] / 1e9
        traffic = None
        tfile = os.path.join(ROOT, "profiles", f"traffic_stream_collide_{pname}_{nx}.json")
        if os.path.exists(tfile):
            traffic = json.load(open(tfile)).get("bytes_per_launch")
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": KERNELS[pname],
                "kernel_ms_avg": r["kernel_s"] * 1e3, "algorithmic_bytes_per_launch": r["bytes_per_launch"],
                "copy_probe_gbs": copy_gbs, "frac_of_copy": achieved / copy_gbs}

    head = res[args.precision]
    times = head["times"]
    out = {
        "metric": "MLUPS + %HBM roofline (plain vs mixed/batched); adjoint-to-forward time ratio",
        "value": n * args.steps / head["dt"] / 1e6,
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["dt"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPES[args.precision],
        "data": "synthetic (seeded hex tissue geometry)",
        "config": {"workload": f"tissue {nx}x{ny}: D2Q9+D2Q5, {ncell} cells r=40+-4, "
                               f"150 markers/cell (remeshed), tension 0.01, area regulator 1.2xA0, box outlet",
                   "precision": args.precision, "markers_after_remesh": head["markers"],
                   "run_ahead": args.run_ahead},
        "roofline": roofline(args.precision),
        "host_seconds_per_step": {k: v / max(times["steps"], 1) for k, v in times.items()
                                  if k not in ("steps", "stops")},
        "resolved_stops": times["stops"],
    }
    mixed = {}
    for pname in variants:
        if pname == "fp64":
            continue
        r, ref = res[pname], res.get("fp64")
        mixed[pname] = {
            "value": n * args.steps / r["dt"] / 1e6, "unit": "MLUPS", "ms_per_step": r["dt"] / args.steps * 1e3,
            "dtype": DTYPES[pname], "roofline": roofline(pname)}
        if ref is not None:
            common = sorted(set(r["fields"]["area"]) & set(ref["fields"]["area"]))
            a_m = np.array([r["fields"]["area"][d] for d in common])
            a_r = np.array([ref["fields"]["area"][d] for d in common])
            mixed[pname]["l2_rel_error_vs_fp64"] = {
                "speed": l2_rel(r["fields"]["speed"], ref["fields"]["speed"]),
                "concentration": l2_rel(r["fields"]["conc"], ref["fields"]["conc"]),
                "cell_area": l2_rel(a_m, a_r), "after_steps": args.warmup + args.steps}
    if mixed:
        out["mixed_precision"] = mixed
    bvars = [p for p in args.batched.split(",") if p]
    if bvars:
        from lbibcell_amd import workloads
        ens = workloads.batched_ensemble(ndom=args.batched_domains, seed=1)
        nlive = sum(r[2] * r[3] for r in ens["rects"])
        log(f"batched ensemble: {len(ens['rects'])} domains, {nlive} sites, atlas {ens['nx']}x{ens['ny']}, "
            f"{sum(len(d) for d in ens['domains'])} cells, {len(ens['nodes'])} initial markers")
        # warm-up to iteration 11, one recorded block, then timed blocks from 21
        bres = {p: measure_batched(ens, p, args.batched_steps, 11, args.run_ahead) for p in bvars}
        batched = {"workload": f"{len(ens['rects'])} periodic domains, sides log-uniform in [64, 512], "
                               f"1-16 cells r 8-30 each, tau_f U(0.6,1.5), tau_c U(0.55,1.2), one atlas "
                               f"{ens['nx']}x{ens['ny']} ({nlive} live sites)",
                   "sites": nlive, "cells": sum(len(d) for d in ens["domains"]), "steps": args.batched_steps}
        for p, r in bres.items():
            achieved = r["bytes_per_launch"] / r["kernel_s"] / 1e9
            e = {"value": nlive * args.batched_steps / r["dt"] / 1e6, "unit": "MLUPS",
                 "ms_per_step": r["dt"] / args.batched_steps * 1e3, "dtype": DTYPES[p],
                 "markers_after_remesh": r["markers"], "step_graphs": r["graph"],
                 "lattice_kernel_time_from": r["kernel_time_source"],
                 "host_seconds_per_step": {k: v / max(r["times"]["steps"], 1) for k, v in r["times"].items()
                                           if k not in ("steps", "stops")},
                 "resolved_stops": r["times"]["stops"],
                 "lattice_shape": {"variant": r["tuned"][0], "workgroup": r["tuned"][1],
                                   "sites_per_lane": r["tuned"][2]},
                 "autotune_us": {f"{t[0]}/{t[1]}/{t[2]}": round(t[3] * 1e3, 2) for t in r["tried"]},
                 "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": achieved / HBM_PEAK_GBS, "kernel_ms_avg": r["kernel_s"] * 1e3,
                              "algorithmic_bytes_per_launch": r["bytes_per_launch"],
                              "frac_of_copy": achieved / copy_gbs if copy_gbs else None}}
            ref = bres.get("fp64")
            if ref is not None and p != "fp64":
                common = sorted(set(r["fields"]["area"]) & set(ref["fields"]["area"]))
                e["l2_rel_error_vs_fp64"] = {
                    "speed": l2_rel(r["fields"]["speed"], ref["fields"]["speed"]),
                    "concentration": l2_rel(r["fields"]["conc"], ref["fields"]["conc"]),
                    "cell_area": l2_rel([r["fields"]["area"][d] for d in common],
                                        [ref["fields"]["area"][d] for d in common]),
                    "after_steps": 11 + 10 + args.batched_steps}
            batched[p] = e
        out["batched"] = batched
    mvars = [p for p in args.multispecies.split(",") if p]
    if mvars:
        msz = args.multispecies_size
        ncell = max(1, round(1600 * (msz / 4096) ** 2))
        mnodes, mconns, K = multispecies_case(msz, ncell)
        log(f"multispecies: {msz}x{msz}, {ncell} cells, {len(mnodes)} initial markers")
        mres = {p: measure_multispecies(msz, mnodes, mconns, K, p, args.multispecies_steps, 5, args.run_ahead)
                for p in mvars}
        ms_out = {"workload": f"{msz}x{msz} D2Q9 + 4 coupled D2Q5 species (linear reaction matrix N(0, 1e-4^2), "
                              f"tau_c {MS_TAUS}, source {MS_SOURCE} inside cells for species 0), {ncell} cells "
                              f"r=30+-5 (120 markers, remeshed) in one touching cluster, tension {TENSION}, "
                              f"junctions within {JUNCTION['radius']} lu (k {JUNCTION['k']}, l0 {JUNCTION['l0']}, "
                              f"every {JUNCTION['freq']} steps), box outlet",
                  "reaction_matrix": K.tolist(), "steps": args.multispecies_steps, "sites": msz * msz}
        for p, r in mres.items():
            achieved = r["bytes_per_launch"] / r["kernel_s"] / 1e9
            e = {"value": msz * msz * args.multispecies_steps / r["dt"] / 1e6, "unit": "MLUPS",
                 "ms_per_step": r["dt"] / args.multispecies_steps * 1e3, "dtype": DTYPES[p],
                 "markers_after_remesh": r["markers"], "junction_structs": r["junction_structs"],
                 "lattice_shape": {"variant": r["tuned"][0], "workgroup": r["tuned"][1],
                                   "sites_per_lane": r["tuned"][2]},
                 "autotune_us": {f"{t[0]}/{t[1]}/{t[2]}": round(t[3] * 1e3, 2) for t in r["tried"]},
                 "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": achieved / HBM_PEAK_GBS, "kernel_ms_avg": r["kernel_s"] * 1e3,
                              "algorithmic_bytes_per_launch": r["bytes_per_launch"],
                              "frac_of_copy": achieved / copy_gbs if copy_gbs else None},
                 "host_seconds_per_step": {k: v / max(r["times"]["steps"], 1) for k, v in r["times"].items()
                                           if k not in ("steps", "stops")}}
            ref = mres.get("fp64")
            if ref is not None and p != "fp64":
                common = sorted(set(r["fields"]["area"]) & set(ref["fields"]["area"]))
                e["l2_rel_error_vs_fp64"] = {
                    "speed": l2_rel(r["fields"]["speed"], ref["fields"]["speed"]),
                    "concentration": [l2_rel(a, b) for a, b in zip(r["fields"]["conc"], ref["fields"]["conc"])],
                    "cell_area": l2_rel([r["fields"]["area"][d] for d in common],
                                        [ref["fields"]["area"][d] for d in common]),
                    "after_steps": 5 + args.multispecies_steps}
            ms_out[p] = e
        out["multispecies"] = ms_out
    lw = [k for k in args.long_window.split(",") if k]
    if lw and args.long_steps > 0:
        precs = [p for p in args.long_precisions.split(",") if p]
        out["mixed_precision_long_window"] = {k: long_window_errors(k, precs, args.long_steps) for k in lw}
    if args.adjoint_steps > 0:
        log(f"adjoint: {args.adjoint_steps} steps ...")
        out["adjoint"] = measure_adjoint(args.adjoint_steps, 50)
        log(f"adjoint/forward time ratio (fp32): {out['adjoint']['fp32']['adjoint_to_forward']:.2f}")
    if args.adjoint_ms_steps > 0:
        log(f"four-species adjoint: {args.adjoint_ms_size}^2, {args.adjoint_ms_steps} steps ...")
        out["adjoint_multispecies"] = measure_adjoint_ms(args.adjoint_ms_size, args.adjoint_ms_steps, 50,
                                                         batch=args.adjoint_ms_batch,
                                                         snapshots=args.adjoint_ms_snapshots)
        am = out["adjoint_multispecies"]
        log("four-species adjoint/forward time ratio: " +
            ", ".join(f"{k} {am[k]['adjoint_to_forward']:.2f}" for k in ("fp64_revolve", "fp32", "fp16_revolve")))
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(nx, ny, nodes, conns)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
