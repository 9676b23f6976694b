"""Step graphs: blocks of coupled steps recorded into one HIP graph and
replayed with one launch each.  The replayed run must be bit-identical to the
step-by-step run (FP64 and FP16 storage, one domain and a batched atlas),
including blocks in which an iteration stops for a host-resolved event
(perturbation draws, capacity growth) and is finished step by step."""
import numpy as np
import pytest

from lbibcell_amd import capi, runner, workloads

pytestmark = pytest.mark.gpu

SIGNAL = dict(production=1e-3, decay=1e-5, switchoff=25, init_inside=0.0, prod_domain=-1)


def _tissue(prec, graph, steps, rects=None, nodes=None, conns=None, nx=96, ny=96):
    sim = runner.Simulation(nx, ny, 0.9, precision=prec)
    sim.add_species("tutorial_01_CDESolverD2Q5_SIGNAL", capi.species_spec(capi.SPECIES_SIGNAL, 0.8, **SIGNAL))
    if rects:
        sim.set_batch_layout(rects)
    if nodes is None:
        nodes, conns = workloads.grid_cells(nx, ny, 2, 2, 10.0, 40, seed=5)
    sim.add_geometry(nodes, conns)
    sim.setup()
    sim.init_solvers()
    sim.add_mass_solver("MassSolverBoxOutlet")
    sim.add_area_regulator(200.0, 1e-5, 0.004, 1.2, freq=1)
    sim.add_tension(0.005, 10)
    sim.set_run_ahead(40)
    if graph:
        sim.set_graph_steps(graph)
    sim.run(steps)
    ctx = sim.context()
    ctx.ns = 1
    out = dict(xy=sim.markers()[1], f=ctx.get_populations(), g=ctx.get_species(0), dom=ctx.get_domains(),
               stats=sim.graph_stats(), stops=sim.times()["stops"], it=sim.iteration, note=sim.graph_note())
    sim.close()
    return out


@pytest.mark.parametrize("prec,block", [(capi.FP64, 10), (capi.FP64, 20), (capi.FP16_DEV, 10)])
def test_graph_blocks_equal_step_by_step(prec, block):
    steps = 123   # ragged: a partial block at either end
    print(f"case graph prec={prec} block={block}", flush=True)
    ref = _tissue(prec, 0, steps)
    got = _tissue(prec, block, steps)
    print("  graph stats (steps, launches, recordings)", got["stats"], "stops", got["stops"], ref["stops"],
          repr(got["note"]), flush=True)
    assert got["it"] == ref["it"] == steps
    assert got["stats"][0] == block and got["stats"][1] >= (steps - 2 * block) // block
    np.testing.assert_array_equal(got["xy"], ref["xy"])
    np.testing.assert_array_equal(got["f"], ref["f"])
    np.testing.assert_array_equal(got["g"], ref["g"])
    np.testing.assert_array_equal(got["dom"], ref["dom"])


def test_graph_blocks_on_a_batched_atlas():
    ens = workloads.batched_ensemble(ndom=6, lmin=64, lmax=160, cells=(1, 3), width=512, seed=3)
    print("case graph atlas", len(ens["rects"]), "domains", ens["nx"], ens["ny"], flush=True)
    kw = dict(rects=ens["rects"], nodes=ens["nodes"], conns=ens["conns"], nx=ens["nx"], ny=ens["ny"])
    ref = _tissue(capi.FP16_DEV, 0, 61, **kw)
    got = _tissue(capi.FP16_DEV, 10, 61, **kw)
    print("  graph stats", got["stats"], flush=True)
    assert got["stats"][1] >= 4
    np.testing.assert_array_equal(got["xy"], ref["xy"])
    np.testing.assert_array_equal(got["f"], ref["f"])
    np.testing.assert_array_equal(got["g"], ref["g"])


def test_graph_steps_must_be_a_coarsening_multiple():
    sim = runner.Simulation(32, 32, 1.0)
    nodes, conns = workloads.circle_cell(16.0, 16.0, 5.0, 20, np.random.default_rng(1))
    sim.add_geometry(nodes, conns)
    sim.setup()
    with pytest.raises(capi.LbibError):
        sim.set_graph_steps(15)
    sim.close()
