"""The Schnakenberg pair tutorial_02_CDESolverD2Q5_L / _R (…_L.cpp, …_R.cpp)
on the GPU against the oracle: the coupled reaction where R reads L after
L's collide at the same site (L comes first in the name-ordered registry),
the deltaT-scaled reaction terms, and the re-initialisation of sites that
change domain.  FP64 results are bit-identical to the oracle; the mixed
storage formats stay within the bars of the other mixed-precision tests."""
import numpy as np
import pytest

from lbibcell_amd import capi, runner, workloads
from test_gpu_coupled import _compare

pytestmark = pytest.mark.gpu

NAMES = ("tutorial_02_CDESolverD2Q5_L", "tutorial_02_CDESolverD2Q5_R")


def _specs(tau=0.9):
    # index 0 = L (partner R at 1), index 1 = R (partner L at 0)
    return [capi.species_spec(capi.SPECIES_SCHNAKENBERG_L, tau, production=100.0, source_inside=0.9, decay=1e-4,
                              prod_domain=1, init_inside=1.0),
            capi.species_spec(capi.SPECIES_SCHNAKENBERG_R, tau, production=100.0, source_inside=0.1, decay=1e-4,
                              prod_domain=0, init_inside=1.0)]


def _oracle_pair(orc, o, tau=0.9):
    o.add_species(orc.SP_SCHNAKENBERG_L, tau, [100.0, 0.9, 1e-4, 1, 1.0])
    o.add_species(orc.SP_SCHNAKENBERG_R, tau, [100.0, 0.1, 1e-4, 0, 1.0])


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp16", "fp16r"])
def test_schnakenberg_lattice_steps_match_oracle(orc, prec):
    nx, ny, tau = 52, 38, 0.85
    precision = {"fp64": capi.FP64, "fp32": capi.FP32, "fp16": capi.FP16_DEV, "fp16r": capi.FP16_REM}[prec]
    print(f"case schnakenberg lattice {prec} {nx}x{ny}", flush=True)
    ctx = capi.Context(nx, ny, tau, species=_specs(), precision=precision)
    o = orc.Oracle(nx, ny, tau)
    _oracle_pair(orc, o)
    dom = np.zeros((ny, nx), dtype=np.uint32)
    dom[6:30, 9:40] = 1
    dom[12:20, 20:31] = 2
    ctx.set_domains(dom)
    o.set_domains(dom)
    ctx.init_solvers()
    o.init_solvers()
    rng = np.random.default_rng(21)
    inside = (dom != 0)[None]
    # R, L around the homogeneous steady state (1.0, 0.9) with 5 % noise
    gL = np.where(inside, 0.9 / 5 * (1 + 0.05 * rng.standard_normal((5, ny, nx))), 0.0)
    gR = np.where(inside, 1.0 / 5 * (1 + 0.05 * rng.standard_normal((5, ny, nx))), 0.0)
    for k, g in enumerate((gL, gR)):
        ctx.set_species(k, g)
        o.set_g(k, g)
    Fx = 1e-4 * rng.standard_normal((ny, nx))
    Fy = 1e-4 * rng.standard_normal((ny, nx))
    ctx.set_force_field(Fx, Fy)
    o.set_force_field(Fx, Fy)
    for it in range(25):
        ctx.step_lattice_only(it)
        o.set_iteration(it)
        o.lattice_step(1)
    for k in range(2):
        gd, go = ctx.get_species(k), o.get_g(k)
        err = float(np.max(np.abs(gd - go)))
        print(f"  species {NAMES[k][-1]}: max abs error {err:.2e}", flush=True)
        if prec == "fp64":
            np.testing.assert_array_equal(gd, go)
        else:
            # storage roundoff of values near 0.2 per step, 25 steps
            # fp16r: remainders of the 5 % noise (about 0.01) rounded to 11 bits
            assert err <= {"fp32": 1e-5, "fp16": 5e-3, "fp16r": 2e-4}[prec]
    ctx.close()


def test_schnakenberg_coupled_run_matches_oracle(orc):
    """Two moving cells (tension, area regulator): sites that change domain
    are re-initialised and collided with the partner's value at that site."""
    nx = ny = 80
    nodes, conns = workloads.grid_cells(nx, ny, 2, 1, 11.0, 40, seed=7)
    print("case schnakenberg coupled 80x80, 2 cells", flush=True)
    sim = runner.Simulation(nx, ny, 1.0)
    o = orc.Oracle(nx, ny, 1.0)
    for name in NAMES:
        sim.add_species(name, tau=0.9)          # the registry's constants
    _oracle_pair(orc, o)
    for nid, x, y in nodes:
        sim.add_node(nid, x, y)
        o.add_node(nid, x, y)
    for a, b, d in conns:
        sim.add_connection(a, b, d)
        o.add_connection(a, b, d, [])
    sim.setup()
    o.cure()
    sim.init_solvers()
    o.init_solvers()
    ctx = sim.context()
    ctx.ns = 2
    dom = ctx.get_domains()
    for k in range(2):
        g = ctx.get_species(k)
        # the registry's init: 0 outside, (1 + U(-0.01, 0.01)) / 5 inside
        assert np.all(g[:, dom == 0] == 0.0)
        assert np.all(np.abs(g[:, dom != 0] - 0.2) <= 0.002 + 1e-15)
        assert np.std(g[:, dom != 0]) > 1e-4
        o.set_g(k, g)
    sim.add_mass_solver("MassSolverBoxOutlet")
    o.box_outlet(True)
    for kind, prm, freq in [(1, [260.0, 1e-5, 0.004, 1.2], 1), (0, [0.008], 10)]:
        if kind == 0:
            sim.add_tension(prm[0], freq)
        else:
            sim.add_area_regulator(*prm, freq=freq)
        o.add_bio(kind, prm, freq)
    sim.run(90)
    o.run(90)
    _compare(sim, o, 2, "schnakenberg coupled")
    st = ctx.read_status()
    print(f"  re-initialisation levels last step: {st.n_reinit_levels}", flush=True)
    sim.close()


def test_schnakenberg_partner_must_be_registered():
    sim = runner.Simulation(16, 16, 1.0)
    sim.add_species("tutorial_02_CDESolverD2Q5_R", tau=0.9)
    nodes, conns = workloads.circle_cell(8.0, 8.0, 3.0, 12, np.random.default_rng(1))
    sim.add_geometry(nodes, conns)
    with pytest.raises(capi.LbibError, match="tutorial_02_CDESolverD2Q5_L"):
        sim.setup()
    sim.close()


def test_reference_tutorial02_geometry_matches_oracle(orc):
    """The reference's tutorial_02 geometry (apps/config/tutorial_02_geometry.txt,
    copied as data under tests/golden/txt: 361 markers on a ring about
    (200, 200), every connection carrying BoundarySolverNoFluxD2Q5 for
    tutorial_02_CDESolverD2Q5_R) read by Geometry::loadGeometry's TXT
    format, with the Schnakenberg pair, tension and the box outlet, against
    the oracle fed the same nodes and descriptors: bit-identical."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "txt", "tutorial_02_geometry.txt")
    nodes, conns, mode = [], [], None
    for line in open(path):
        if line.startswith("#Nodes"):
            mode = "n"
            continue
        if line.startswith("#Connection"):
            mode = "c"
            continue
        t = line.split()
        if not t:
            continue
        if mode == "n":
            nodes.append((int(t[0]), float(t[1]), float(t[2])))
        else:
            conns.append((int(t[0]), int(t[1]), int(t[2]), t[3:]))
    assert len(nodes) == 361 and len(conns) == 361
    nx = ny = 400
    print("case reference tutorial_02 geometry fixture 400x400", flush=True)
    sim = runner.Simulation(nx, ny, 1.0)
    o = orc.Oracle(nx, ny, 1.0)
    for name in NAMES:
        sim.add_species(name, tau=0.9)          # the registry's constants
    _oracle_pair(orc, o)
    sim.load_geometry_txt(path)
    for nid, x, y in nodes:
        o.add_node(nid, x, y)
    for a, b, d, desc in conns:
        # species in name order: L = 0, R = 1; the descriptors name R only
        o.add_connection(a, b, d, [1] if "tutorial_02_CDESolverD2Q5_R" in desc else [])
    sim.setup()
    o.cure()
    sim.init_solvers()
    o.init_solvers()
    ctx = sim.context()
    ctx.ns = 2
    for k in range(2):
        o.set_g(k, ctx.get_species(k))
    sim.add_mass_solver("MassSolverBoxOutlet")
    o.box_outlet(True)
    sim.add_tension(0.01, 10)
    o.add_bio(0, [0.01], 10)
    sim.run(40)
    o.run(40)
    _compare(sim, o, 2, "tutorial_02 fixture")
    sim.close()
