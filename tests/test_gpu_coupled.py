"""The full coupled step on the GPU (host runtime + C ABI) against the oracle:
tension-driven and growing cells with remeshing, coarsening, boundary-node
regeneration, CDE re-initialisation and the domain scanline.  Bars: membrane
positions and populations within 1e-12 relative (FP64), domain ids and
topology bit-exact."""
import numpy as np
import pytest

from lbibcell_amd import capi, runner, workloads

pytestmark = pytest.mark.gpu


def _build(orc, nx, ny, tau, nodes, conns, species, bio, celltypes=None, box=True, bc="", precision=capi.FP64):
    sim = runner.Simulation(nx, ny, tau, precision=precision)
    o = orc.Oracle(nx, ny, tau)
    for name, kind, t, prm in species:
        sim.add_species(name, capi.species_spec(kind, t, **prm))
    # the runner orders species by name (Parameters::cdeSolvers is a std::map)
    for name, kind, t, prm in sorted(species, key=lambda s: s[0]):
        if kind == capi.SPECIES_SIGNAL:
            o.add_species(orc.SP_SIGNAL, t, [prm["production"], prm["decay"], prm["switchoff"],
                                              prm["init_inside"], prm["prod_domain"]])
        else:
            o.add_species(orc.SP_PLAIN, t)
    names = sorted(s[0] for s in species)
    bc_idx = [names.index(b) for b in bc.split(",") if b]
    for nid, x, y in nodes:
        sim.add_node(nid, x, y)
        o.add_node(nid, x, y)
    for a, b, d in conns:
        sim.add_connection(a, b, d, bc)
        o.add_connection(a, b, d, bc_idx)
    for d, t in (celltypes or {}).items():
        sim.set_celltype(d, t)
        o.set_celltype(d, t)
    sim.setup()
    o.cure()
    sim.init_solvers()
    o.init_solvers()
    if box:
        sim.add_mass_solver("MassSolverBoxOutlet")
        o.box_outlet(True)
    for kind, prm, freq in bio:
        if kind == 0:
            sim.add_tension(prm[0], freq)
        else:
            sim.add_area_regulator(*prm, freq=freq)
        o.add_bio(kind, prm, freq)
    return sim, o


def _compare(sim, o, ns, tag):
    ids_d, xy_d, v_d = sim.markers()
    ids_o, xy_o, v_o, _ = o.markers()
    np.testing.assert_array_equal(ids_d, ids_o, err_msg=f"{tag}: marker ids")
    assert np.max(np.abs(xy_d - xy_o)) <= 1e-12 * np.max(np.abs(xy_o)), f"{tag}: marker positions"
    a_d, b_d, d_d = sim.connections()
    a_o, b_o, d_o = o.connections()
    np.testing.assert_array_equal(np.stack([a_d, b_d, d_d]), np.stack([a_o, b_o, d_o]), err_msg=f"{tag}: connections")
    ctx = sim.context()
    dom_o, _ = o.get_domains()
    np.testing.assert_array_equal(ctx.get_domains(), dom_o, err_msg=f"{tag}: domain ids")
    fd, fo = ctx.get_populations(), o.get_f()
    assert np.max(np.abs(fd - fo)) <= 1e-12 * np.max(np.abs(fo)), f"{tag}: populations"
    ctx.ns = ns
    for s in range(ns):
        gd, go = ctx.get_species(s), o.get_g(s)
        assert np.max(np.abs(gd - go)) <= 1e-12 * max(np.max(np.abs(go)), 1e-300), f"{tag}: species {s}"
    # FP64 path rounds like the oracle: everything is expected to be identical
    np.testing.assert_array_equal(xy_d, xy_o, err_msg=f"{tag}: positions not bit-identical")
    np.testing.assert_array_equal(fd, fo, err_msg=f"{tag}: populations not bit-identical")


SIGNAL = dict(production=1e-3, decay=1e-5, switchoff=5000, init_inside=0.0, prod_domain=1)


@pytest.mark.parametrize("ring2_first", [False, True])
def test_equidistant_cut_candidates_go_to_the_first_connection(orc, ring2_first):
    """Two connections cut the lattice line y = 8 at exactly x = 5.5, so the W
    slot of site (6, 8) gets two candidates at the same distance; the
    connection earlier in the list keeps it (Connection::insertBoundaryNode,
    Connection.cpp:181-222), and the scanline carries the winner's domain
    along row 8 (the oracle pin of this geometry derives the row by hand)."""
    from test_oracle_pins import TIE_SIZE, tie_geometry
    nodes, conns = tie_geometry(ring2_first)
    print(f"case equidistant cut candidates, ring2_first={ring2_first}", flush=True)
    sim, o = _build(orc, *TIE_SIZE, 1.0, nodes, conns, [], [], box=False)
    dom_o, _ = o.get_domains()
    winner = 2 if ring2_first else 1
    np.testing.assert_array_equal(dom_o[8, :14], [0] * 6 + [winner] * 4 + [0] * 4)
    np.testing.assert_array_equal(sim.context().get_domains(), dom_o)


@pytest.mark.parametrize("steps", [1, 12, 150])
def test_single_cell_tension_matches_oracle(orc, steps):
    nx = ny = 64
    nodes, conns = workloads.circle_cell(32.0, 32.0, 8.0, 48, np.random.default_rng(3))
    print(f"case single cell 64x64 steps={steps}", flush=True)
    sim, o = _build(orc, nx, ny, 1.0, nodes, conns,
                    [("tutorial_01_CDESolverD2Q5_SIGNAL", capi.SPECIES_SIGNAL, 0.8, SIGNAL)],
                    [(0, [0.01], 10)])
    sim.run(steps)
    o.run(steps)
    assert sim.iteration == o.iteration() == steps
    _compare(sim, o, 1, f"steps={steps}")
    sim.close()


def test_growing_cells_with_area_regulator_match_oracle(orc):
    nx = ny = 96
    nodes, conns = workloads.grid_cells(nx, ny, 2, 2, 10.0, 40, seed=5)
    print("case 2x2 cells 96x96 with tension + area regulator", flush=True)
    sim, o = _build(orc, nx, ny, 0.9, nodes, conns,
                    [("tutorial_01_CDESolverD2Q5_SIGNAL", capi.SPECIES_SIGNAL, 0.8, SIGNAL)],
                    [(1, [200.0, 1e-5, 0.004, 1.2], 1), (0, [0.005], 10)],
                    celltypes={1: 2, 2: 1, 3: 2, 4: 2})
    sim.run(120)
    o.run(120)
    _compare(sim, o, 1, "area regulator")
    np_, nc_ = o.perturb_counts()
    print("oracle perturbations", np_, nc_, "device", sim.perturb_counts(), flush=True)
    sim.close()


def test_noflux_boundary_two_species_match_oracle(orc):
    nx = ny = 64
    nodes, conns = workloads.circle_cell(30.5, 33.0, 9.0, 56, np.random.default_rng(9))
    print("case no-flux BC with two species", flush=True)
    species = [("CDESolverD2Q5", capi.SPECIES_PLAIN, 0.9, {}),
               ("tutorial_01_CDESolverD2Q5_SIGNAL", capi.SPECIES_SIGNAL, 0.75, SIGNAL)]
    sim, o = _build(orc, nx, ny, 1.1, nodes, conns, species, [(0, [0.01], 10)],
                    bc="tutorial_01_CDESolverD2Q5_SIGNAL,CDESolverD2Q5")
    sim.run(60)
    o.run(60)
    _compare(sim, o, 2, "no-flux BC")
    sim.close()


def test_reference_tutorial_geometry_matches_oracle(orc):
    """The reference's own tutorial_01 geometry fixture (apps/config/vtk/
    Cells_0.vtm, copied as data under tests/golden/vtk: 361 Float32 markers
    on a ring of radius 10 about (200, 200), first and last coinciding, many
    on lattice lines) loaded by Geometry::loadGeometryVTK on the host, run
    with the tutorial_01 pipeline (SIGNAL species, membrane tension 0.01
    renewed every 10 steps, box outlet) against the oracle fed the same
    points: bit-identical, including the perturbation draws of the markers
    that sit on lattice lines or cell diagonals."""
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vtk")
    text = open(os.path.join(gold, "Cells_0", "Cells_0_0.vtp")).read()
    i = text.index('Name="Points"')
    pts = np.array(text[text.index(">", i) + 1: text.index("</DataArray>", i)].split(),
                   dtype=np.float64).reshape(-1, 3)[:, :2].astype(np.float32).astype(np.float64)
    nx = ny = 400
    sig = dict(production=1e-4, decay=1e-5, switchoff=5000, init_inside=0.0, prod_domain=1)
    print("case reference tutorial_01 geometry fixture 400x400", flush=True)
    sim = runner.Simulation(nx, ny, 1.0)
    o = orc.Oracle(nx, ny, 1.0)
    sim.add_species("tutorial_01_CDESolverD2Q5_SIGNAL", capi.species_spec(capi.SPECIES_SIGNAL, 1.0, **sig))
    o.add_species(orc.SP_SIGNAL, 1.0, [sig["production"], sig["decay"], sig["switchoff"], sig["init_inside"],
                                      sig["prod_domain"]])
    sim.load_geometry_vtk(os.path.join(gold, "Cells_0.vtm"))
    m = len(pts)
    for k in range(m):
        o.add_node(k, pts[k, 0], pts[k, 1])
    for k in range(m):
        o.add_connection(k, (k + 1) % m, 1, [])
    sim.setup()
    o.cure()
    sim.init_solvers()
    o.init_solvers()
    sim.add_mass_solver("MassSolverBoxOutlet")
    o.box_outlet(True)
    sim.add_tension(0.01, 10)
    o.add_bio(0, [0.01], 10)
    sim.run(60)
    o.run(60)
    _compare(sim, o, 1, "tutorial_01 fixture")
    print("perturbations oracle", o.perturb_counts(), "device", sim.perturb_counts(), flush=True)
    sim.close()


@pytest.mark.parametrize("window", [1, 5, 16])
def test_run_ahead_resolves_stops_like_the_synchronous_loop(orc, window):
    """Iterations enqueued ahead of the status read: a stop (here a remesh that
    needs more rounds than one geometry pass runs, then perturbations and
    coarsening as they come) skips everything enqueued after it and is
    resolved from the iteration that raised it, so every window gives the
    oracle's result."""
    nx = ny = 64
    # 12 markers on r = 8: 4 lu segments need four remesh rounds in iteration 0
    nodes, conns = workloads.circle_cell(32.0, 32.0, 8.0, 12, np.random.default_rng(4))
    print(f"case run-ahead window={window}", flush=True)
    sim, o = _build(orc, nx, ny, 1.0, nodes, conns,
                    [("tutorial_01_CDESolverD2Q5_SIGNAL", capi.SPECIES_SIGNAL, 0.8, SIGNAL)],
                    [(1, [200.0, 1e-5, 0.004, 1.2], 1), (0, [0.01], 10)], celltypes={1: 2})
    sim.set_run_ahead(window)
    sim.run(45)
    o.run(45)
    assert sim.iteration == o.iteration() == 45
    _compare(sim, o, 1, f"window={window}")
    t = sim.times()
    print("resolved stops", t["stops"], "perturbations", sim.perturb_counts(), flush=True)
    assert t["stops"] >= 1
    sim.close()
