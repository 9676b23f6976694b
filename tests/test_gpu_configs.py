"""The five workloads the path is specified on, at their stated sizes:

C1  200^2 tutorial_01-style run (one cell r = 20, 120 nodes, SIGNAL source
    1e-3, tension 0.01, box outlet), 2000 steps: bit-identical to the oracle.
C2  2048^2 tissue (400 cells r = 40 +- 4, 150 nodes, area regulator 1.2 A0):
    the first 3 coupled steps (including the 60k -> 240k node remesh)
    bit-identical to the oracle; the configured run to its stop (the
    reference's edge error in iteration 88: the grown tissue reaches the
    lattice edge), bit-identical to the oracle there; and 5000 steps of the
    same cells laid out 100 sites from the edges with a slower regulator,
    with invariants.
C3  256-domain batched ensemble (sides log-uniform in [64, 512], 1-16 cells
    each, per-domain tau_f, tau_c) stepped in recorded step graphs: 16 sampled
    domains against a standalone oracle run each, 50 steps.
C4  512^2 adjoint, 9 cells, 1000 steps, checkpoints every 50, FP64:
    gradients w.r.t. tau_c, kappa and one cell's rate against central
    differences to 1e-5.
C5  4096^2, four species coupled by a linear reaction matrix N(0, 1e-4^2),
    1600 touching cells with junctions within 2 lu: the first 2 coupled steps
    bit-identical to the oracle."""
import os

import numpy as np
import pytest

from lbibcell_amd import adjoint, capi, runner, workloads

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
SIG = dict(production=1e-3, decay=1e-5, switchoff=5000, init_inside=0.0, prod_domain=-1)
SIG_O = [1e-3, 1e-5, 5000, 0.0, -1]


def _same(sim, o, ns, tag, exact=True):
    ids_d, xy_d, _ = sim.markers()
    ids_o, xy_o, _, _ = o.markers()
    np.testing.assert_array_equal(ids_d, ids_o, err_msg=f"{tag}: marker ids")
    a_d, b_d, d_d = sim.connections()
    a_o, b_o, d_o = o.connections()
    np.testing.assert_array_equal(np.stack([a_d, b_d, d_d]), np.stack([a_o, b_o, d_o]), err_msg=f"{tag}: connections")
    ctx = sim.context()
    dom_o, _ = o.get_domains()
    np.testing.assert_array_equal(ctx.get_domains(), dom_o, err_msg=f"{tag}: domain ids")
    np.testing.assert_array_equal(xy_d, xy_o, err_msg=f"{tag}: marker positions")
    np.testing.assert_array_equal(ctx.get_populations(), o.get_f(), err_msg=f"{tag}: populations")
    ctx.ns = ns
    for s in range(ns):
        np.testing.assert_array_equal(ctx.get_species(s), o.get_g(s), err_msg=f"{tag}: species {s}")


# ------------------------------------------------------------------- C1 ----
def test_c1_tutorial_cell_2000_steps_bit_exact(orc):
    nx = ny = 200
    steps = 2000
    print(f"case C1 {nx}x{ny}, one cell r=20 (120 nodes), {steps} steps", flush=True)
    nodes, conns = workloads.tutorial_cell(nx, ny, 20.0, 120, seed=1)
    sim = runner.Simulation(nx, ny, 1.0)
    sim.add_species("tutorial_01_CDESolverD2Q5_SIGNAL", capi.species_spec(capi.SPECIES_SIGNAL, 0.8, **SIG))
    sim.add_geometry(nodes, conns)
    sim.setup()
    sim.init_solvers()
    sim.add_mass_solver("MassSolverBoxOutlet")
    sim.add_tension(0.01, 10)
    o = orc.Oracle(nx, ny, 1.0)
    o.set_threads(THREADS)
    o.add_species(orc.SP_SIGNAL, 0.8, SIG_O)
    for nd in nodes:
        o.add_node(*nd)
    for c in conns:
        o.add_connection(*c)
    o.cure()
    o.init_solvers()
    o.box_outlet(True)
    o.add_bio(0, [0.01], 10)
    sim.run(steps)
    o.run(steps)
    assert sim.iteration == o.iteration() == steps
    _same(sim, o, 1, "C1")
    print(f"  nodes after {steps} steps: {len(sim.markers()[0])}, perturbations {sim.perturb_counts()}", flush=True)
    sim.close()


# ------------------------------------------------------------------- C2 ----
C2_AREA = (200.0, 1e-5, 0.004, 1.2)


def _c2(orc=None):
    nodes, conns = workloads.hex_tissue(2048, 2048, 20, 20, radius=40.0, rjit=4.0, n=150, pos_jit=2.0, seed=1)
    doms = sorted({c[2] for c in conns})
    sim = runner.Simulation(2048, 2048, 1.0)
    sim.add_species("tutorial_01_CDESolverD2Q5_SIGNAL", capi.species_spec(capi.SPECIES_SIGNAL, 0.8, **SIG))
    sim.add_geometry(nodes, conns)
    for d in doms:
        sim.set_celltype(d, 2)
    sim.setup()
    sim.init_solvers()
    sim.add_mass_solver("MassSolverBoxOutlet")
    sim.add_area_regulator(*C2_AREA, freq=1)
    sim.add_tension(0.01, 10)
    o = None
    if orc is not None:
        o = orc.Oracle(2048, 2048, 1.0)
        o.set_threads(THREADS)
        o.add_species(orc.SP_SIGNAL, 0.8, SIG_O)
        for nd in nodes:
            o.add_node(*nd)
        for c in conns:
            o.add_connection(*c)
        for d in doms:
            o.set_celltype(d, 2)
        o.cure()
        o.init_solvers()
        o.box_outlet(True)
        o.add_bio(1, list(C2_AREA), 1)
        o.add_bio(0, [0.01], 10)
    return sim, o, len(nodes)


def test_c2_tissue_first_steps_bit_exact(orc):
    print("case C2 2048^2 tissue, 400 cells, first 3 steps against the oracle", flush=True)
    sim, o, n0 = _c2(orc)
    sim.run(3)
    o.run(3)
    nm = len(sim.markers()[0])
    print(f"  nodes {n0} -> {nm} after the first remesh", flush=True)
    assert nm > 3.5 * n0
    _same(sim, o, 1, "C2")
    sim.close()


def test_c2_configured_tissue_stops_where_the_oracle_does(orc):
    """The configured tissue cannot run its 10 000 steps under the reference's
    semantics: growth toward 1.2 A0 pushes fluid out through the box outlet
    and carries the outermost cells (4 sites from the edge) into the lattice
    edge, and GeometryHandler::connectGeometryNodesToPhysicalNodes
    (GeometryHandler.cpp:613-636) stops the run.  The device stops in the same
    iteration as the oracle, on the same node, with every node position and
    every population off the outlet sites bit-identical at the stop."""
    print("case C2 configured tissue run to its stop against the oracle", flush=True)
    sim, o, _ = _c2(orc)
    with pytest.raises(RuntimeError, match="too close to the"):
        o.run(200)
    with pytest.raises(capi.LbibError, match="too close to the boundary"):
        sim.run(200)
    st = sim.last_status()
    assert st.bits & capi.ERR_EDGE
    assert st.step_iteration == o.iteration()
    ids_d, xy_d, _ = sim.markers()
    ids_o, xy_o, _, _ = o.markers()
    np.testing.assert_array_equal(ids_d, ids_o)
    np.testing.assert_array_equal(xy_d, xy_o)
    bad = [k for k, (x, y) in enumerate(xy_o) if x - 2 < 0 or y - 2 < 0 or x + 2 > 2048 or y + 2 > 2048]
    assert st.first_bad_marker == bad[0]
    # the device applies the box outlet inside the fused lattice kernel; the
    # reference's mass solver runs after the move that failed, so at the stop
    # only the outlet sites (x = 0, y = 0) differ: there the device holds the
    # populations rescaled to unit density
    fd, fo = sim.context().get_populations(), o.get_f()
    np.testing.assert_array_equal(fd[:, 1:, 1:], fo[:, 1:, 1:])
    np.testing.assert_allclose(fd[:, 0, :].sum(axis=0), 1.0, rtol=1e-14)
    np.testing.assert_allclose(fd[:, :, 0].sum(axis=0), 1.0, rtol=1e-14)
    print(f"  both stop in iteration {o.iteration()} on node {ids_o[bad[0]]}", flush=True)
    sim.close()


def test_c2_tissue_long_run_invariants():
    """A tissue of the same cells that stays clear of the edges for thousands
    of steps (grid 100 sites from the edges, area regulator KP 1e-6, clamp
    4e-4), 5000 steps in FP64: no error, every segment at most 0.5 long, closed
    rings, each cell's site count within one perimeter of its shoelace area,
    every area between 0.8 A0 and 1.6 A0, mean density within 10 % of 1."""
    steps = 5000
    print(f"case C2 tissue (margin 100, slow regulator), {steps} steps in FP64, invariants", flush=True)
    nodes, conns = workloads.hex_tissue(2048, 2048, 20, 20, radius=40.0, rjit=4.0, n=150, pos_jit=2.0, seed=1,
                                        margin=100.0)
    sim = runner.Simulation(2048, 2048, 1.0)
    sim.add_species("tutorial_01_CDESolverD2Q5_SIGNAL", capi.species_spec(capi.SPECIES_SIGNAL, 0.8, **SIG))
    sim.add_geometry(nodes, conns)
    for d in sorted({c[2] for c in conns}):
        sim.set_celltype(d, 2)
    sim.setup()
    sim.init_solvers()
    sim.add_mass_solver("MassSolverBoxOutlet")
    sim.add_area_regulator(200.0, 1e-6, 4e-4, 1.2, freq=1)
    sim.add_tension(0.01, 10)
    sim.run(1)
    a0 = sim.areas()
    sim.run(steps - 1)
    assert sim.iteration == steps
    st = sim.last_status()
    assert st.bits & capi.ERR_MASK == 0
    ctx = sim.context()
    dom = ctx.get_domains()
    areas = sim.areas()
    ids, xy, _ = sim.markers()
    a, b, d = sim.connections()
    pos = {int(i): k for k, i in enumerate(ids)}
    pa = xy[[pos[int(i)] for i in a]]
    pb = xy[[pos[int(i)] for i in b]]
    seglen = np.hypot(*(pb - pa).T)
    assert seglen.max() <= 0.5 + 1e-9
    assert np.array_equal(np.sort(a), np.sort(ids)) and np.array_equal(np.sort(b), np.sort(ids))
    count = np.bincount(dom.reshape(-1), minlength=int(d.max()) + 1)
    perim = np.bincount(d, weights=seglen, minlength=int(d.max()) + 1)
    worst = 0.0
    for c, area in areas.items():
        if c == 0:
            continue
        assert abs(count[c] - area) <= perim[c], (c, count[c], area, perim[c])
        assert 0.8 * a0[c] <= area <= 1.6 * a0[c], (c, a0[c], area)
        worst = max(worst, abs(count[c] - area) / perim[c])
    rho = ctx.get_populations().sum(axis=0)
    assert np.all(np.isfinite(rho)) and abs(rho.mean() - 1.0) < 0.1
    growth = np.mean([areas[c] / a0[c] for c in areas if c])
    print(f"  {len(ids)} nodes, mean area growth {growth:.4f}, worst |sites - area| / perimeter {worst:.3f}, "
          f"mean density {rho.mean():.6f}", flush=True)
    sim.close()


# ------------------------------------------------------------------- C3 ----
def test_c3_batched_ensemble_in_step_graphs_matches_oracles(orc):
    steps = 50
    ens = workloads.batched_ensemble(ndom=256, seed=1)
    print(f"case C3 256-domain atlas {ens['nx']}x{ens['ny']}, {len(ens['nodes'])} nodes, step graphs, "
          f"{steps} steps, 16 domains against the oracle", flush=True)
    sim = runner.Simulation(ens["nx"], ens["ny"], 1.0)
    sim.add_species("tutorial_01_CDESolverD2Q5_SIGNAL", capi.species_spec(capi.SPECIES_SIGNAL, 0.8, **SIG))
    sim.set_batch_layout(ens["rects"])
    sim.add_geometry(ens["nodes"], ens["conns"])
    sim.setup()
    sim.init_solvers()
    sim.add_tension(0.01, 10)
    sim.set_graph_steps(10)
    sim.run(steps)
    g = sim.graph_stats()
    print(f"  graph launches {g}", flush=True)
    assert g[0] == 10 and g[1] >= 4, "the step graphs were not used"
    ctx = sim.context()
    ctx.ns = 1
    dom_a = ctx.get_domains()
    fa, ga = ctx.get_populations(), ctx.get_species(0)
    ids_a, xy_a, _ = sim.markers()
    sample = list(range(0, 256, 16))
    for r in sample:
        x0, y0, L, _, tf, tc = ens["rects"][r]
        o = orc.Oracle(L, L, tf)
        o.set_threads(THREADS)
        o.add_species(orc.SP_SIGNAL, tc[0], SIG_O)
        cells = set(ens["domains"][r])
        conns = [c for c in ens["conns"] if c[2] in cells]
        ids = {c[0] for c in conns}
        for nid, x, y in ens["nodes"]:
            if nid in ids:
                o.add_node(nid, x - x0, y - y0)
        for a, b, d in conns:
            o.add_connection(a, b, d)
        o.cure()
        o.init_solvers()
        o.add_bio(0, [0.01], 10)
        o.run(steps)
        win = (slice(y0, y0 + L), slice(x0, x0 + L))
        dom_o, _ = o.get_domains()
        np.testing.assert_array_equal(dom_a[win], dom_o, err_msg=f"domain {r}: domain ids")
        fo, go = o.get_f(), o.get_g(0)
        ef = np.max(np.abs(fa[(slice(None),) + win] - fo)) / np.max(np.abs(fo))
        eg = np.max(np.abs(ga[(slice(None),) + win] - go)) / max(np.max(np.abs(go)), 1e-300)
        ids_o, xy_o, _, _ = o.markers()
        inside = ((xy_a[:, 0] >= x0) & (xy_a[:, 0] < x0 + L) & (xy_a[:, 1] >= y0) & (xy_a[:, 1] < y0 + L))
        pa = xy_a[inside][np.argsort(ids_a[inside], kind="stable")] - np.array([x0, y0])
        po = xy_o[np.argsort(ids_o, kind="stable")]
        assert pa.shape == po.shape, f"domain {r}: node count {pa.shape} vs {po.shape}"
        ex = np.max(np.abs(pa - po)) / np.max(np.abs(po)) if len(po) else 0.0
        print(f"  domain {r} ({L}^2, {len(ids)} nodes): rel err populations {ef:.1e} species {eg:.1e} "
              f"nodes {ex:.1e}", flush=True)
        if (x0, y0) == (0, 0):
            np.testing.assert_array_equal(fa[(slice(None),) + win], fo)
        assert ef <= 1e-11 and eg <= 1e-11 and ex <= 1e-11
    sim.close()


# ------------------------------------------------------------------- C4 ----
def test_c4_adjoint_512_1000_steps_central_differences():
    steps, K = 1000, 50
    print(f"case C4 adjoint 512^2, 9 cells, {steps} steps, checkpoints every {K}, FP64", flush=True)
    rings, dom, rates = workloads.adjoint_case(512, 512, 3, 3, 40.0, 150, seed=4)
    p = adjoint.AdjointProblem(512, 512, 1.0, rings, dom, decay=1e-5, precision=capi.FP64, checkpoint_every=K)
    _, tgt = p.forward(steps, 0.8, 0.01, rates, want_field=True)
    p.set_target(tgt)
    guess = (0.9, 0.006, rates * 0.7)
    p.forward(steps, *guess)
    g = p.gradient()

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

    errs = [abs(g[0] - fd(0, 1e-5)) / abs(g[0]), abs(g[1] - fd(1, 1e-6)) / abs(g[1]),
            abs(g[2][0] - fd(2, 1e-7)) / abs(g[2][0])]
    fwd_ms, grad_ms = p.timing()
    print(f"  central-difference rel. errors tau_c {errs[0]:.1e} kappa {errs[1]:.1e} rate {errs[2]:.1e}; "
          f"gradient / forward time {grad_ms / fwd_ms:.2f}", flush=True)
    p.close()
    assert max(errs) <= 1e-5


# ------------------------------------------------------------------- C5 ----
MS_TAUS = (0.8, 0.9, 1.0, 1.1)


def test_c5_four_species_junction_tissue_first_steps_bit_exact(orc):
    size, steps = 4096, 2
    nodes, conns = workloads.packed_tissue(size, size, 1600, 30.0, 5.0, 120, seed=1)
    Km = np.random.default_rng(101).normal(0.0, 1e-4, size=(4, 4))
    print(f"case C5 {size}^2, 4 species, {len(set(c[2] for c in conns))} cells, {len(nodes)} nodes, "
          f"{steps} steps against the oracle", flush=True)
    sim = runner.Simulation(size, size, 1.0)
    o = orc.Oracle(size, size, 1.0)
    o.set_threads(THREADS)
    for s in range(4):
        src = 1e-3 if s == 0 else 0.0
        sim.add_species(f"LINEAR_{s}", capi.species_spec(capi.SPECIES_LINEAR, MS_TAUS[s], init_inside=0.1,
                                                          source_inside=src, krow=tuple(Km[s])))
        o.add_species(orc.SP_LINEAR, MS_TAUS[s], [src, 0.1] + list(Km[s]))
    sim.add_geometry(nodes, conns)
    sim.setup()
    sim.init_solvers()
    sim.add_mass_solver("MassSolverBoxOutlet")
    sim.add_tension(0.01, 10)
    sim.add_junctions(radius=2.0, k=0.02, l0=0.5, freq=10)
    for nd in nodes:
        o.add_node(*nd)
    for c in conns:
        o.add_connection(*c)
    o.cure()
    o.init_solvers()
    o.box_outlet(True)
    o.add_bio(0, [0.01], 10)
    o.add_bio(2, [2.0, 0.02, 0.5], 10)
    sim.run(steps)
    o.run(steps)
    _same(sim, o, 4, "C5")
    t, _, _, _ = o.forces()
    print(f"  junction springs {int(np.sum(t == 7))}", flush=True)
    assert np.sum(t == 7) > 1000
    sim.close()
