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
