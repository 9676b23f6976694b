"""The reference's own entry points and files on the GPU path:

* apps/tutorial_01.cpp's main() sequence (apps/tutorial_01.cpp:38-152) run by
  lbibcell_amd/lbib_tutorial_01 through the reference's signatures
  (GeometryHandler, ReportHandler, SimulationRunner(GeometryHandler&,
  ReportHandler&), BioSolvers by name) from the reference's own
  apps/config/tutorial_01_parameters.txt (400^2, tau 2.0), force.txt and
  vtk/Cells_0.vtm (copied as data under tests/golden), against the oracle;
  once with the built-in area regulator and once with a BioSolver written
  against applyBioProcess(GeometryHandler&, ForceSolver&) that adds mass
  through getPhysicalNodes();
* the t = 0 VTK image blocks against the reference's Cells_0_2.vti /
  Cells_0_3.vti (coarsening 4);
* the error paths: the velocity guard (LagrangianPoint.cpp:92-103), the edge
  check (GeometryHandler.cpp:613-636), the CRASH dump of runSimulation
  (SimulationRunner.cpp:101-109) and its reload with initSolvers(file);
* force types 1, 3, 4 and 5 (ForceStructs.hpp:116-454) against the oracle;
* uploads the device rejects before any launch, and capacity growth keeping
  the topology."""
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from lbibcell_amd import capi, runner, workloads

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXAMPLE = os.path.join(ROOT, "lbibcell_amd", "lbib_tutorial_01")
SIGNAL_T01 = [1e-4, 1e-5, 5000, 1.0, 1]   # tutorial_01_CDESolverD2Q5_SIGNAL.cpp:39-43


def _parameters(path):
    out, cde = {}, {}
    for line in open(path):
        if not line.strip() or line.startswith("#"):
            continue
        key, *vals = line.split()
        if key == "CDESolvers:":
            cde = dict(zip(vals[0::2], map(float, vals[1::2])))
        else:
            out[key.rstrip(":")] = float(vals[0])
    return out, cde


def _fixture_points():
    text = open(os.path.join(GOLD, "vtk", "Cells_0", "Cells_0_0.vtp")).read()
    i = text.index('Name="Points"')
    return np.array(text[text.index(">", i) + 1: text.index("</DataArray>", i)].split(),
                    dtype=np.float64).reshape(-1, 3)[:, :2].astype(np.float32).astype(np.float64)


def _tutorial01_oracle(orc, steps):
    prm, cde = _parameters(os.path.join(GOLD, "config", "tutorial_01_parameters.txt"))
    nx, ny, tau = int(prm["SizeX"]), int(prm["SizeY"]), prm["tauFluid"]
    o = orc.Oracle(nx, ny, tau)
    for name, tc in sorted(cde.items()):
        assert name == "tutorial_01_CDESolverD2Q5_SIGNAL"
        o.add_species(orc.SP_SIGNAL, tc, SIGNAL_T01)
    pts = _fixture_points()
    m = len(pts)
    for k in range(m):
        o.add_node(k, pts[k, 0], pts[k, 1])
    for k in range(m):
        o.add_connection(k, (k + 1) % m, 1, [])
    o.cure()
    o.init_solvers()
    o.box_outlet(True)
    # tutorial_01_BioSolver{AreaRegulator, MembraneTension, CellJunction,
    # Differentiation, CellDivision} in the order main() adds them
    o.add_bio(1, [200.0, 1e-5, 0.004, 0.0], 1)
    o.add_bio(0, [0.01], 10)
    o.add_bio(2, [1.0, 0.02, 0.5], 10)
    o.add_bio(6, [1.6, 0], 100)
    o.add_bio(7, [380.0], 100)
    o.set_threads(min(16, os.cpu_count() or 1))
    o.run(steps)
    return o, nx, ny


@pytest.mark.parametrize("regulator", ["registry", "handler"])
def test_tutorial01_main_sequence_matches_oracle(orc, tmp_path, regulator):
    steps = 100
    print(f"case tutorial_01 main() from the reference's config files, {steps} steps, {regulator} regulator",
          flush=True)
    assert os.path.exists(EXAMPLE), "build lbib_tutorial_01 first (__graft_entry__.build())"
    out = tmp_path / "out"
    cmd = [EXAMPLE, os.path.join(GOLD, "config", "tutorial_01_parameters.txt"),
           os.path.join(GOLD, "vtk", "Cells_0.vtm"), os.path.join(GOLD, "config", "force.txt"), str(out),
           str(steps), regulator, "report"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    print(r.stdout[-400:], r.stderr[-400:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    o, nx, ny = _tutorial01_oracle(orc, steps)
    n = nx * ny
    ids_o, xy_o, _, _ = o.markers()
    np.testing.assert_array_equal(np.fromfile(out / "ids.bin", dtype=np.uint32), ids_o)
    np.testing.assert_array_equal(np.fromfile(out / "xy.bin").reshape(-1, 2), xy_o)
    a, b, d = o.connections()
    np.testing.assert_array_equal(np.fromfile(out / "conn.bin", dtype=np.uint32).reshape(-1, 3),
                                  np.stack([a, b, d], 1))
    dom_o, _ = o.get_domains()
    np.testing.assert_array_equal(np.fromfile(out / "dom.bin", dtype=np.uint32).reshape(ny, nx), dom_o)
    np.testing.assert_array_equal(np.fromfile(out / "f.bin").reshape(9, ny, nx), o.get_f())
    np.testing.assert_array_equal(np.fromfile(out / "g0.bin").reshape(5, ny, nx), o.get_g(0))
    # the reference's text dump and the reports at every ReportStep (10)
    for f in ("forceOut.txt", "geoOut.txt", "parameterOut.txt", "solverOut.txt"):
        assert (out / f).stat().st_size > 0
    for t in range(10, steps + 1, 10):
        assert (out / f"Cells_{t}.vtm").exists()
    assert not (out / "Cells_5.vtm").exists()
    assert o.iteration() == steps and n == 160000


def _vti(path):
    root = ET.parse(path).getroot()
    img = root.find("ImageData")
    arrays = {}
    for da in img.iter("DataArray"):
        arrays[da.get("Name")] = (int(da.get("NumberOfComponents") or 1),
                                  np.array(da.text.split(), dtype=np.float64))
    return img.get("WholeExtent"), img.get("Origin"), img.get("Spacing"), img.find("Piece").get("Extent"), arrays


def test_vtk_images_at_t0_match_the_reference_fixture(tmp_path):
    """vtkFluidReporter / vtkCDEReporter at t = 0, 400^2, coarsening 4, against
    the reference's apps/config/vtk/Cells_0/Cells_0_3.vti (density 1,
    velocity 0) and Cells_0_2.vti: extent, origin, spacing, array names,
    components and values.  The fixture's CDE block holds two species,
    CDESolverD2Q5BMP (zero everywhere) and CDESolverD2Q5HH (1 inside the cell,
    0 outside), solvers the reference tree no longer ships; they are mirrored
    by two species initialised to 0 and to 1 inside cells, named so that they
    sort in the same order.  The HH values therefore compare the device's
    domain map at the sampled sites with the reference's own output."""
    print("case t=0 VTK images vs reference fixtures", flush=True)
    sim = runner.Simulation(400, 400, 2.0)
    for name, c0 in (("SIGNAL_BMP", 0.0), ("SIGNAL_HH", 1.0)):
        sim.add_species(name, capi.species_spec(capi.SPECIES_SIGNAL, 2.0, production=1e-4, decay=1e-5,
                                                switchoff=5000, init_inside=c0, prod_domain=1))
    sim.load_geometry_vtk(os.path.join(GOLD, "vtk", "Cells_0.vtm"))
    sim.setup()
    sim.init_solvers()
    prefix = str(tmp_path / "Cells")
    sim.write_vtk(prefix, 0, coarsening=4, fluid=False, species=True, cells=False)
    sim.write_vtk(prefix, 0, coarsening=4, fluid=True, species=False, cells=False)
    sim.close()
    mine_cde = _vti(tmp_path / "Cells_0" / "Cells_0_0.vti")
    mine_fluid = _vti(tmp_path / "Cells_0" / "Cells_0_1.vti")
    ref_cde = _vti(os.path.join(GOLD, "vtk", "Cells_0", "Cells_0_2.vti"))
    ref_fluid = _vti(os.path.join(GOLD, "vtk", "Cells_0", "Cells_0_3.vti"))
    for mine, ref in ((mine_cde, ref_cde), (mine_fluid, ref_fluid)):
        assert mine[:4] == ref[:4], (mine[:4], ref[:4])
    assert sorted(mine_fluid[4]) == sorted(ref_fluid[4]) == ["density", "velocity"]
    for name in ("density", "velocity"):
        assert mine_fluid[4][name][0] == ref_fluid[4][name][0]
        np.testing.assert_allclose(mine_fluid[4][name][1], ref_fluid[4][name][1], rtol=0, atol=1e-12)
    assert list(ref_cde[4]) == ["CDESolverD2Q5BMP", "CDESolverD2Q5HH"]
    assert list(mine_cde[4]) == ["SIGNAL_BMP", "SIGNAL_HH"]
    for (mc, mv), (rc, rv) in zip(mine_cde[4].values(), ref_cde[4].values()):
        assert mc == rc and mv.shape == rv.shape == (100 * 100,)
        np.testing.assert_array_equal(mv, rv)
    assert ref_cde[4]["CDESolverD2Q5HH"][1].sum() > 10      # the cell covers sampled sites


# ----------------------------------------------------------- error paths ----
def _cell_case(orc, cx, cy, r, n, forces, nx=64, ny=64, it0=0, seed=3):
    nodes, conns = workloads.circle_cell(cx, cy, r, n, np.random.default_rng(seed))
    sim = runner.Simulation(nx, ny, 1.0)
    o = orc.Oracle(nx, ny, 1.0)
    sig = dict(production=1e-3, decay=1e-5, switchoff=5000, init_inside=0.0, prod_domain=1)
    sim.add_species("tutorial_01_CDESolverD2Q5_SIGNAL", capi.species_spec(capi.SPECIES_SIGNAL, 0.8, **sig))
    o.add_species(orc.SP_SIGNAL, 0.8, [1e-3, 1e-5, 5000, 0.0, 1])
    for nd in nodes:
        sim.add_node(*nd)
        o.add_node(*nd)
    for a, b, d in conns:
        sim.add_connection(a, b, d)
        o.add_connection(a, b, d)
    sim.setup()
    o.cure()
    sim.init_solvers()
    o.init_solvers()
    sim.add_mass_solver("MassSolverBoxOutlet")
    o.box_outlet(True)
    for f in forces:
        sim.add_force(*f)
        o.add_force(*f)
    if it0:
        sim.set_iteration(it0)
        o.set_iteration(it0)
    return sim, o, nodes


def test_velocity_guard_trips_like_the_oracle_and_crash_dump_reloads(orc, tmp_path, monkeypatch):
    """A free force of 20 (type 2, apps/config/force.txt syntax) on one node
    drives the interpolated velocity past the guard |U|^2 > 1/sqrt(3) in the
    first step: the device raises LBIB_ERR_VELOCITY with the oracle's first
    node, runSimulation writes output/CRASH_*_1.txt, the solver dump equals
    the oracle's text of the same state, and initSolvers(file) loads it."""
    print("case velocity guard + CRASH dump", flush=True)
    sim, o, nodes = _cell_case(orc, 32.0, 32.0, 8.0, 48, [(2, 7, 0, 20.0, -3.0)])
    monkeypatch.chdir(tmp_path)
    os.makedirs("output")
    with pytest.raises(RuntimeError, match="magnitude of velocity"):
        o.run(5)
    assert o.iteration() == 0
    sim.run_simulation(5)             # catches, dumps and prints, like the reference
    st = capi.Status()
    sim.L.lbh_last_status(sim.h, st)
    assert st.bits & capi.ERR_VELOCITY
    ids, xy, v, _ = o.markers()
    bad = int(np.argmax((v ** 2).sum(1) > 1 / np.sqrt(3)))
    assert st.first_bad_marker == bad
    for f in ("forceOut", "geoOut", "parameterOut", "solverOut"):
        assert os.path.exists(f"output/CRASH_{f}_1.txt"), f
    o.write_solver_dump(tmp_path / "oracle_solver.txt", ["tutorial_01_CDESolverD2Q5_SIGNAL"])
    assert open("output/CRASH_solverOut_1.txt").read() == open(tmp_path / "oracle_solver.txt").read()
    sim.close()
    # reload: a fresh runner reads the dump; its populations are the dumped text
    sim2, _, _ = _cell_case(orc, 32.0, 32.0, 8.0, 48, [])
    sim2.init_solvers_file("output/CRASH_solverOut_1.txt")
    rows = [ln.split("\t") for ln in open("output/CRASH_solverOut_1.txt").read().split("\n")
            if ln and not ln.startswith("#")]
    fluid = np.array([[float(v) for v in r[2:]] for r in rows[:64 * 64]])
    slot_of = [0, 3, 4, 1, 2, 7, 8, 5, 6]     # slot q stores direction slot_of[q]
    expect = np.empty((9, 64 * 64))
    for q in range(9):
        expect[slot_of[q]] = fluid[:, q]
    np.testing.assert_array_equal(sim2.context().get_populations().reshape(9, -1), expect)
    sim2.close()


def test_edge_check_trips_like_the_oracle(orc):
    """A node pushed west (type 2 force) leaves its cell below x = 2: both the
    oracle and the device stop at the same iteration with the edge error."""
    print("case edge check during a run", flush=True)
    nodes, _ = workloads.circle_cell(8.3, 32.0, 6.0, 40, np.random.default_rng(3))
    west, wx, _ = min(nodes, key=lambda t: t[1])
    cx = 8.3 + (2.25 - wx)          # the westmost node starts at x = 2.25
    sim, o, _ = _cell_case(orc, cx, 32.0, 6.0, 40, [(2, west, 0, -0.5, 0.0)], it0=1)
    fail_o = None
    for t in range(60):
        try:
            o.run(1)
        except RuntimeError as e:
            assert "west boundary" in str(e)
            fail_o = o.iteration()
            break
    assert fail_o is not None, "the oracle never reached the edge"
    with pytest.raises(capi.LbibError):
        sim.run(60)
    st = capi.Status()
    sim.L.lbh_last_status(sim.h, st)
    assert st.bits & capi.ERR_EDGE
    assert st.step_iteration == fail_o
    sim.close()


def test_force_types_1_3_4_5_match_oracle(orc):
    """Anchor (1), horizontal (3) and vertical (4) sliders and a constant
    pair force (5) from the force-file syntax, nine steps (iterations 1-9, no
    deleteAllForces in between): bit-identical to the oracle."""
    print("case force types 1/3/4/5", flush=True)
    forces = [(1, 0, 0, 30.0, 30.0, 0.01, 2.0), (3, 12, 0, 33.0, 0.02, 1.0), (4, 24, 0, 31.0, 0.015, 0.5),
              (5, 36, 18, 0.005)]
    sim, o, _ = _cell_case(orc, 32.0, 32.0, 8.0, 48, forces, it0=1)
    sim.run(9)
    o.run(9)
    ids_d, xy_d, _ = sim.markers()
    ids_o, xy_o, _, _ = o.markers()
    np.testing.assert_array_equal(ids_d, ids_o)
    np.testing.assert_array_equal(xy_d, xy_o)
    np.testing.assert_array_equal(sim.context().get_populations(), o.get_f())
    sim.close()


def test_marker_upload_rejects_edge_positions():
    """lbib_set_markers refuses positions within 2 sites of the lattice edge
    (or of the rectangle's edge in a batched atlas) and non-finite ones before
    anything is launched: status -1, LBIB_ERR_EDGE and the index of the first
    bad marker in the status word."""
    print("case marker upload edge check", flush=True)
    ctx = capi.Context(48, 40, 1.0)
    L = ctx.lib
    good = np.array([[10.0, 10.0], [30.5, 20.25]])
    assert L.lbib_set_markers(ctx.h, 2, capi.dptr(good.reshape(-1)), None, None) == 0
    for k, bad in enumerate(([1.99, 10.0], [10.0, 38.01], [np.nan, 5.0], [46.5, 3.0])):
        xy = np.array([[10.0, 10.0], bad, [20.0, 20.0]])
        assert L.lbib_set_markers(ctx.h, 3, capi.dptr(xy.reshape(-1)), None, None) == -1, bad
        assert "edge" in L.lbib_last_error(ctx.h).decode()
        st = ctx.read_status()
        assert st.bits & capi.ERR_EDGE and st.first_bad_marker == 1
        ctx._chk(L.lbib_clear_status(ctx.h))
    ctx.close()
    # an atlas: a marker in the gap between rectangles, or near its rectangle's edge
    ctx = capi.Context(96, 48, 1.0)
    ctx.set_batch_layout([(0, 0, 40, 48, 1.0, []), (48, 0, 48, 48, 1.0, [])])
    for bad in ([44.0, 20.0], [38.5, 20.0], [49.0, 20.0]):
        xy = np.array([[10.0, 10.0], bad])
        assert ctx.lib.lbib_set_markers(ctx.h, 2, capi.dptr(xy.reshape(-1)), None, None) == -1, bad
    xy = np.array([[10.0, 10.0], [50.0, 20.0], [93.5, 45.0]])
    assert ctx.lib.lbib_set_markers(ctx.h, 3, capi.dptr(xy.reshape(-1)), None, None) == 0
    ctx.close()


def test_reserve_keeps_the_uploaded_topology():
    """Growing the marker and connection capacity keeps positions, ids,
    links and the force table."""
    print("case capacity growth keeps the topology", flush=True)
    nodes, conns = workloads.circle_cell(20.0, 20.0, 6.0, 24, np.random.default_rng(1))
    ctx = capi.Context(48, 48, 1.0, max_markers=32, max_connections=32)
    L = ctx.lib
    xy = np.array([[x, y] for _, x, y in nodes])
    nm = len(xy)
    ctx._chk(L.lbib_set_markers(ctx.h, nm, capi.dptr(xy.reshape(-1)), None, None))
    ids = np.array([i for i, _, _ in nodes], dtype=np.uint32)
    ctx._chk(L.lbib_set_marker_ids(ctx.h, capi.u32ptr(ids)))
    a = np.array([c[0] for c in conns], dtype=np.int32)
    b = np.array([c[1] for c in conns], dtype=np.int32)
    d = np.array([c[2] for c in conns], dtype=np.uint32)
    ctx._chk(L.lbib_set_connections(ctx.h, len(a), capi.i32ptr(a), capi.i32ptr(b), capi.u32ptr(d), None))
    off = np.arange(nm + 1, dtype=np.int32)
    typ = np.full(nm, 6, dtype=np.int32)
    part = ((np.arange(nm) + 1) % nm).astype(np.int32)
    prm = np.zeros((nm, 4))
    prm[:, 0] = 0.01
    ctx._chk(L.lbib_set_forces(ctx.h, nm, capi.i32ptr(off), capi.i32ptr(typ), capi.i32ptr(part), capi.dptr(prm)))
    ctx._chk(L.lbib_reserve(ctx.h, 5000, 6000))
    got = np.empty(2 * nm)
    ctx._chk(L.lbib_get_markers(ctx.h, capi.dptr(got), None, None, None))
    np.testing.assert_array_equal(got, xy.reshape(-1))
    gid = np.empty(nm, dtype=np.uint32)
    ctx._chk(L.lbib_get_marker_ids(ctx.h, capi.u32ptr(gid)))
    np.testing.assert_array_equal(gid, ids)
    ga, gb = np.empty_like(a), np.empty_like(b)
    gd = np.empty_like(d)
    gbc = np.empty_like(a)
    ctx._chk(L.lbib_get_connections(ctx.h, capi.i32ptr(ga), capi.i32ptr(gb), capi.u32ptr(gd), capi.i32ptr(gbc)))
    np.testing.assert_array_equal(np.stack([ga, gb]), np.stack([a, b]))
    np.testing.assert_array_equal(gd, d)
    goff, gt, gp, gprm = ctx.get_forces(nm)
    np.testing.assert_array_equal(goff, off)
    np.testing.assert_array_equal(gt, typ)
    np.testing.assert_array_equal(gp, part)
    np.testing.assert_array_equal(gprm, prm.reshape(gprm.shape))
    ctx.close()
