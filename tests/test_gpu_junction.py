"""Cell junctions (tutorial_02_BioSolverCellJunction) on the device force
table against the oracle: the rebuilt table of a hand-derived geometry struct
for struct, and coupled runs of touching cells with tension and junctions
(step by step, run ahead and replayed from step graphs) bit-identical in FP64,
force tables included."""
import numpy as np
import pytest

from lbibcell_amd import capi, runner, workloads
from test_oracle_junction import EXPECTED, K_IN, K_TXT, L0_IN, NODES, RINGS

pytestmark = pytest.mark.gpu


def _device_rows(sim):
    ids, _, _ = sim.markers()
    off, t, q, prm = sim.context().get_forces(len(ids))
    rows = {}
    for k in range(len(ids)):
        for j in range(off[k], off[k + 1]):
            rows.setdefault(int(ids[k]), []).append((int(t[j]), int(ids[q[j]]) if q[j] >= 0 else -1,
                                                     *prm[j].tolist()))
    return rows


def _oracle_rows(o):
    t, a, b, prm = o.forces()
    rows = {}
    for j in range(len(t)):
        pair = int(t[j]) in (0, 5, 6, 7)
        rows.setdefault(int(a[j]), []).append((int(t[j]), int(b[j]) if pair else -1, *prm[j].tolist()))
    return rows


def test_device_junction_table_matches_the_oracle_struct_for_struct(orc):
    sim = runner.Simulation(32, 32, 1.0)
    o = orc.Oracle(32, 32, 1.0)
    for nid, x, y in NODES:
        sim.add_node(nid, x, y)
        o.add_node(nid, x, y)
    for ring, dom in RINGS:
        for k in range(len(ring)):
            sim.add_connection(ring[k], ring[(k + 1) % len(ring)], dom)
            o.add_connection(ring[k], ring[(k + 1) % len(ring)], dom)
    sim.setup()
    o.cure()
    sim.init_solvers()
    o.init_solvers()
    pre = [(0, 1, 3, 0.1, 1.0), (7, 2, 8, 9.0, 9.0), (2, 2, 0, 0.5, 0.25), (0, 4, 1, 0.1, 1.0)]
    for t, a, b, p0, p1 in pre:
        sim.add_force(t, a, b, p0, p1)
        o.add_force(t, a, b, p0, p1)
    print("case junction table, hand geometry", flush=True)
    np.testing.assert_array_equal(sim.markers()[1], o.markers()[1])
    # the BioSolver: host-edited table uploaded, then the device op
    sim.add_junctions(radius=1.01, k=K_IN, l0=L0_IN, freq=10)
    sim.apply_bio()
    o.add_bio(2, [1.01, K_IN, L0_IN], 10)
    o.set_iteration(0)
    o.apply_bio()
    # oracle rows: type-2 structs carry no partner
    got, want = _device_rows(sim), _oracle_rows(o)
    for k in got:
        got[k] = [(r[0], r[1] if r[0] in (0, 5, 6, 7) else -1, *r[2:]) for r in got[k]]
    assert got == want
    assert sorted(got) == sorted(EXPECTED)
    # the C ABI op applied again (K already rounded as the text descriptor
    # does): the same table, old junctions deleted first
    sim.context().forces_junctions(1.01, K_TXT, L0_IN)
    got2 = _device_rows(sim)
    for k in got2:
        got2[k] = [(r[0], r[1] if r[0] in (0, 5, 6, 7) else -1, *r[2:]) for r in got2[k]]
    assert got2 == want
    sim.close()


def _touching_cells(seed):
    rng = np.random.default_rng(seed)
    nodes, conns = [], []
    # three rings about 0.6 lu apart: markers with several finders each
    for i, (cx, cy) in enumerate([(20.0, 22.0), (40.7, 22.3), (30.4, 40.3)]):
        nd, cn = workloads.circle_cell(cx, cy, 10.0, 64, rng, jitter=0.2, first_id=1000 * i, domain=i + 1)
        nodes += nd
        conns += cn
    return nodes, conns


def _coupled(orc, steps, run_ahead, graph, seed=4):
    nx = ny = 64
    nodes, conns = _touching_cells(seed)
    sim = runner.Simulation(nx, ny, 1.0)
    o = orc.Oracle(nx, ny, 1.0)
    for nid, x, y in nodes:
        sim.add_node(nid, x, y)
        o.add_node(nid, x, y)
    for a, b, d in conns:
        sim.add_connection(a, b, d)
        o.add_connection(a, b, d)
    sim.setup()
    o.cure()
    sim.init_solvers()
    o.init_solvers()
    # tension first, junctions second: rows hold two type-6 structs, then junctions
    sim.add_tension(0.002, 10)
    o.add_bio(0, [0.002], 10)
    sim.add_junctions(radius=1.0, k=0.02, l0=0.5, freq=10)
    o.add_bio(2, [1.0, 0.02, 0.5], 10)
    sim.set_run_ahead(run_ahead)
    if graph:
        sim.set_graph_steps(graph)
    sim.run(steps)
    o.run(steps)
    return sim, o


@pytest.mark.parametrize("steps,run_ahead,graph", [(1, 1, 0), (25, 1, 0), (61, 16, 0), (61, 40, 10)])
def test_touching_cells_with_junctions_match_oracle(orc, steps, run_ahead, graph):
    print(f"case junction cells steps={steps} run_ahead={run_ahead} graph={graph}", flush=True)
    sim, o = _coupled(orc, steps, run_ahead, graph)
    assert sim.iteration == o.iteration() == steps
    ids_d, xy_d, _ = sim.markers()
    ids_o, xy_o, _, _ = o.markers()
    np.testing.assert_array_equal(ids_d, ids_o)
    np.testing.assert_array_equal(xy_d, xy_o, err_msg="marker positions not bit-identical")
    ctx = sim.context()
    np.testing.assert_array_equal(ctx.get_populations(), o.get_f(), err_msg="populations not bit-identical")
    dom_o, _ = o.get_domains()
    np.testing.assert_array_equal(ctx.get_domains(), dom_o)
    got, want = _device_rows(sim), _oracle_rows(o)
    # every row opens with its two tension structs; the device writes them in
    # a fixed order and the reference in connection order.  The force sum
    # starts 0 + t1 + t2, which IEEE addition gives identically either way, so
    # that pair is compared as a set and everything after it in order.
    def canon(rows):
        out = {}
        for k, v in rows.items():
            assert [r[0] for r in v[:2]] == [6, 6], f"row {k}"
            out[k] = sorted(v[:2]) + v[2:]
        return out
    got, want = canon(got), canon(want)
    n7 = sum(1 for v in want.values() for r in v if r[0] == 7)
    print(f"  junction structs {n7}, graph stats {sim.graph_stats()} note {sim.graph_note()!r}", flush=True)
    assert n7 > 0, "the cells were placed to touch"
    assert got == want
    if graph:
        assert sim.graph_stats()[1] > 0, "no block was replayed from a graph"
    sim.close()
