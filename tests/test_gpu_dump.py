"""SimulationRunner::writeSimulation / initSolvers(filename) (the reference's
text dump, SimulationRunner.cpp:112-180) on the device runner.

The solver file is compared as text with the oracle's restatement of
FluidSolver::writeSolver / CDESolverD2Q5::writeSolver written from its stored
(locally swapped) slots after the same coupled run: identical FP64 states give
identical files, which pins the slot convention, the row order and the
stream's default number formatting."""
import numpy as np
import pytest

from lbibcell_amd import capi, runner, workloads
from test_gpu_coupled import SIGNAL, _build

pytestmark = pytest.mark.gpu

SPECIES = [("CDESolverD2Q5", capi.SPECIES_PLAIN, 0.9, {}),
           ("tutorial_01_CDESolverD2Q5_SIGNAL", capi.SPECIES_SIGNAL, 0.8, SIGNAL)]


def _case(orc):
    nodes, conns = workloads.circle_cell(20.0, 21.0, 6.0, 40, np.random.default_rng(12))
    return _build(orc, 40, 36, 1.0, nodes, conns, SPECIES, [(0, [0.01], 10)])


def test_solver_dump_equals_the_reference_format(orc, tmp_path):
    sim, o = _case(orc)
    sim.run(12)
    o.run(12)
    files = [tmp_path / n for n in ("force.txt", "geometry.txt", "parameters.txt", "solvers.txt")]
    print("case dump 40x36 after 12 steps", flush=True)
    sim.write_simulation(*files)
    o.write_solver_dump(tmp_path / "oracle_solvers.txt", sorted(s[0] for s in SPECIES))
    dev = files[3].read_text()
    ref = (tmp_path / "oracle_solvers.txt").read_text()
    assert dev.splitlines()[:2] == ["#Solvers (xPos\tyPos\tDistributions...)", "#BEGIN FluidSolver"]
    assert dev.count("\n") == 1 + 3 * (2 + 40 * 36)
    assert dev == ref
    geo = files[1].read_text().splitlines()
    assert geo[0].startswith("#Nodes") and any(l.startswith("#Connection") for l in geo)
    assert files[0].read_text().strip() and files[2].read_text().strip()
    sim.close()


def test_solver_dump_round_trip(orc, tmp_path):
    sim, o = _case(orc)
    sim.run(7)
    path = tmp_path / "solvers.txt"
    sim.write_simulation(tmp_path / "f.txt", tmp_path / "g.txt", tmp_path / "p.txt", path)
    # a second runner on the same lattice loads the dump: its populations are
    # the printed values, slot for slot
    b = runner.Simulation(40, 36, 1.0)
    for name, kind, t, prm in SPECIES:
        b.add_species(name, capi.species_spec(kind, t, **prm))
    b.load_geometry_txt(str(tmp_path / "g.txt"))
    b.setup()
    b.init_solvers()
    b.init_solvers_file(path)
    rows = [l.split("\t") for l in path.read_text().splitlines()]
    block, vals = None, {}
    for r in rows:
        if r[0].startswith("#BEGIN"):
            block = r[0].split()[1]
            vals[block] = {}
        elif not r[0].startswith("#"):
            vals[block][(int(r[0]), int(r[1]))] = [float(v) for v in r[2:]]
    slot_of9, slot_of5 = [0, 3, 4, 1, 2, 7, 8, 5, 6], [0, 3, 4, 1, 2]
    ctx = b.context()
    ctx.ns = 2
    f = ctx.get_populations()
    for (x, y), v in list(vals["FluidSolver"].items())[::37]:
        assert [f[slot_of9[q], y, x] for q in range(9)] == v
    for k, name in enumerate(sorted(s[0] for s in SPECIES)):
        g = ctx.get_species(k)
        for (x, y), v in list(vals[name].items())[::29]:
            assert [g[slot_of5[q], y, x] for q in range(5)] == v
    b.close()
    sim.close()


def test_domain_concentration_sums(orc):
    """GeometryHandler::computeAccumulatedDomainConcentrations on the device:
    the per-domain sum of getC() over x >= 1, against the oracle's stored
    species summed in (y, x) order."""
    sim, o = _case(orc)
    sim.run(9)
    o.run(9)
    names = sorted(s[0] for s in SPECIES)
    dom_o, _ = o.get_domains()
    for k, name in enumerate(names):
        got = sim.domain_concentrations(name)
        C = o.get_g(k).sum(axis=0)
        want = np.zeros(got.shape)
        for d in range(len(want)):
            want[d] = C[:, 1:][dom_o[:, 1:] == d].sum()
        print(f"case domain sums {name}: {got}", flush=True)
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)
    with pytest.raises(capi.LbibError):
        sim.domain_concentrations("NoSuchSolver")
    sim.close()


def _vti_arrays(path):
    import re
    text = open(path).read()
    ext = re.search(r'WholeExtent="0 (\d+) 0 (\d+) 0 0"', text)
    out = {}
    for m in re.finditer(r'<DataArray type="Float64" Name="([^"]+)"[^>]*>(.*?)</DataArray>', text, re.S):
        out[m.group(1)] = np.array(m.group(2).split(), dtype=np.float64)
    return int(ext.group(1)) + 1, int(ext.group(2)) + 1, out


def test_vtk_reporters_write_the_device_fields(orc, tmp_path):
    """vtkFluidReporter / vtkCDEReporter / vtkCellReporter without VTK: the
    written density (slot-order sum), velocity, concentrations and membrane
    are the device state; the cell block loads back to the same markers."""
    sim, o = _case(orc)
    sim.run(11)
    prefix = str(tmp_path / "Cells")
    sim.write_vtk(prefix, 11)
    sim.write_vtk(prefix, 11, coarsening=2, cells=False)
    blocks = sorted(p.name for p in (tmp_path / "Cells_11").iterdir())
    assert blocks == ["Cells_11_0.vti", "Cells_11_1.vti", "Cells_11_2.vtp", "Cells_11_3.vti", "Cells_11_4.vti"]
    assert (tmp_path / "Cells_11.vtm").read_text().count("<DataSet") == 5
    ctx = sim.context()
    ctx.ns = 2
    f = ctx.get_populations()
    rho = sum(f[q] for q in [0, 3, 4, 1, 2, 7, 8, 5, 6])     # getRho over the stored slots
    ux, uy = ctx.get_velocity()
    cx, cy, arr = _vti_arrays(tmp_path / "Cells_11" / "Cells_11_0.vti")
    assert (cx, cy) == (40, 36)
    np.testing.assert_array_equal(arr["density"], rho.reshape(-1))
    vel = arr["velocity"].reshape(-1, 3)
    np.testing.assert_array_equal(vel[:, 0], ux.reshape(-1))
    np.testing.assert_array_equal(vel[:, 1], uy.reshape(-1))
    _, _, cde = _vti_arrays(tmp_path / "Cells_11" / "Cells_11_1.vti")
    for k, name in enumerate(sorted(s[0] for s in SPECIES)):
        g = ctx.get_species(k)
        np.testing.assert_array_equal(cde[name], sum(g[q] for q in [0, 3, 4, 1, 2]).reshape(-1))
    cx2, cy2, arr2 = _vti_arrays(tmp_path / "Cells_11" / "Cells_11_3.vti")
    assert (cx2, cy2) == (20, 18)
    np.testing.assert_array_equal(arr2["density"], rho[::2, ::2].reshape(-1))
    # the membrane block loads back
    b = runner.Simulation(40, 36, 1.0)
    b.load_geometry_vtk(str(tmp_path / "Cells_11.vtm"))
    b.write_geometry_vtk(str(tmp_path / "again"), 0)
    import re
    def pts(p):
        t = open(p).read()
        i = t.index('Name="Points"')
        return np.array(t[t.index(">", i) + 1: t.index("</DataArray>", i)].split(), dtype=np.float64).reshape(-1, 3)
    p1 = pts(tmp_path / "Cells_11" / "Cells_11_2.vtp")
    p2 = np.roll(pts(tmp_path / "again_0" / "again_0_0.vtp"), 1, axis=0)
    assert p1.shape == p2.shape
    moved = np.any(p1 != p2, axis=1)   # reloaded nodes on a lattice line are perturbed again
    assert np.max(np.abs(p1 - p2)) <= 2e-6 and moved.sum() <= 2
    ids, xy, _ = sim.markers()
    assert len(xy) == len(p1)
    assert sorted(map(tuple, xy)) == sorted(map(tuple, p1[:, :2]))
    del re
    b.close()
    sim.close()
