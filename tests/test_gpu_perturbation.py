"""Perturbations resolved from the flagged markers alone.  Node perturbations
after the move (GeometryNode::setPos -> perturbatePosition,
GeometryNode.cpp:153-214): lbib_perturbation_candidates finds exactly the
markers that meet one of the conditions and lbib_set_marker_positions writes
only the given markers.  Connection perturbations at the start of the cure
(GeometryHandler::perturbConnections, GeometryHandler.cpp:326-356):
lbib_perturbation_rings lists the rings holding a lined-up connection in the
order that walk visits them.  The runner's resolution of either halt through
these keeps runs bit-identical to the oracle, which walks every node and
every ring."""
import ctypes as C

import numpy as np
import pytest

from lbibcell_amd import capi, runner

pytestmark = pytest.mark.gpu


def _ctx_with_markers(xy, ids):
    ctx = capi.Context(64, 48, 1.0, max_markers=512, max_connections=512)
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    ctx._chk(ctx.lib.lbib_set_markers(ctx.h, len(xy), capi.dptr(xy), None, None))
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    ctx._chk(ctx.lib.lbib_set_marker_ids(ctx.h, capi.u32ptr(ids)))
    return ctx


def _candidates(ctx, cap):
    cnt = C.c_int(-1)
    ix = np.full(max(cap, 1), -1, dtype=np.int32)
    ids = np.zeros(max(cap, 1), dtype=np.uint32)
    xy = np.zeros((max(cap, 1), 2))
    ctx._chk(ctx.lib.lbib_perturbation_candidates(ctx.h, cap, C.byref(cnt), ix.ctypes.data_as(capi._I32),
                                                   capi.u32ptr(ids), capi.dptr(xy)))
    w = min(cnt.value, cap)
    return cnt.value, ix[:w], ids[:w], xy[:w]


def _markers(ctx, nm):
    xy = np.zeros((nm, 2))
    ctx._chk(ctx.lib.lbib_get_markers(ctx.h, capi.dptr(xy), None, None, None))
    return xy


def test_candidates_are_exactly_the_nodes_on_lattice_lines_and_diagonals():
    rng = np.random.default_rng(11)
    nm = 300
    xy = np.column_stack([rng.uniform(3.0, 60.0, nm), rng.uniform(3.0, 44.0, nm)])
    # placed cases: x on an integer, y on an integer, the cell diagonal, the
    # anti-diagonal, and x within 5e-10 above an integer (all flagged); x
    # within 5e-10 below an integer is not flagged: the condition truncates
    # with static_cast<int>, so X - int(X) is then almost 1
    xy[7] = (10.0, 20.37)
    xy[31] = (33.61, 17.0)
    xy[64] = (12.25, 30.25)
    xy[100] = (40.3, 8.7)
    xy[150] = (12.0 + 5e-10, 22.41)
    xy[151] = (13.0 - 5e-10, 22.41)
    flagged = [7, 31, 64, 100, 150]
    ids = 1000 + 3 * np.arange(nm)
    print("case perturbation candidates: 300 markers, 5 flagged", flush=True)
    ctx = _ctx_with_markers(xy, ids)
    cnt, ix, got_ids, got_xy = _candidates(ctx, 64)
    assert cnt == len(flagged)
    order = np.argsort(ix)
    np.testing.assert_array_equal(ix[order], flagged)
    np.testing.assert_array_equal(got_ids[order], ids[flagged])
    np.testing.assert_array_equal(got_xy[order], xy[flagged])
    # fewer slots than flagged markers: the count is still the total
    cnt2, ix2, _, _ = _candidates(ctx, 2)
    assert cnt2 == len(flagged) and len(ix2) == 2 and set(ix2.tolist()) <= set(flagged)
    # writing two positions changes those two markers only
    new = np.array([[10.0000004, 20.37], [12.2500003, 30.2499998]])
    at = np.array([7, 64], dtype=np.int32)
    ctx._chk(ctx.lib.lbib_set_marker_positions(ctx.h, 2, at.ctypes.data_as(capi._I32), capi.dptr(new)))
    after = _markers(ctx, nm)
    expect = xy.copy()
    expect[at] = new
    np.testing.assert_array_equal(after, expect)
    cnt3, ix3, _, _ = _candidates(ctx, 64)
    assert cnt3 == 3 and sorted(ix3.tolist()) == [31, 100, 150]
    ctx.close()


def test_set_marker_positions_rejects_bad_indices_and_edge_positions():
    rng = np.random.default_rng(12)
    nm = 40
    xy = np.column_stack([rng.uniform(4.0, 58.0, nm), rng.uniform(4.0, 42.0, nm)])
    ctx = _ctx_with_markers(xy, np.arange(nm))
    print("case set_marker_positions error paths", flush=True)
    one = np.array([[20.5, 20.5]])
    for bad in (-1, nm):
        at = np.array([bad], dtype=np.int32)
        with pytest.raises(capi.LbibError, match="outside the live markers"):
            ctx._chk(ctx.lib.lbib_set_marker_positions(ctx.h, 1, at.ctypes.data_as(capi._I32), capi.dptr(one)))
    at = np.array([3, 9], dtype=np.int32)
    edge = np.array([[20.5, 20.5], [1.5, 20.5]])
    with pytest.raises(capi.LbibError, match="within 2 sites"):
        ctx._chk(ctx.lib.lbib_set_marker_positions(ctx.h, 2, at.ctypes.data_as(capi._I32), capi.dptr(edge)))
    st = ctx.read_status()
    assert st.bits & capi.ERR_EDGE and st.first_bad_marker == 9
    # nothing was written
    np.testing.assert_array_equal(_markers(ctx, nm), xy)
    ctx.close()


def _ring():
    """Ten nodes on r = 6 in a uniform flow of ux = 0.05; nodes 0, 5 and 8
    land, after the first move, on x = 31, on the cell diagonal and on the
    anti-diagonal, so iteration 0 halts for three perturbation draws."""
    nodes = []
    for k in range(10):
        a = 2 * np.pi * k / 10 + 0.13
        nodes.append([k, 24.37 + 6.0 * np.cos(a), 24.61 + 6.0 * np.sin(a)])
    nodes[0][1] = 30.95
    nodes[5][1], nodes[5][2] = 18.75, 23.8
    nodes[8][1], nodes[8][2] = 26.3, 19.65
    return [tuple(n) for n in nodes], [(k, (k + 1) % 10, 1) for k in range(10)]


@pytest.mark.parametrize("window", [1, 8])
def test_perturbation_halt_resolved_from_flagged_nodes_matches_oracle(orc, window):
    nx = ny = 48
    nodes, conns = _ring()
    u0 = np.zeros((ny, nx, 2))
    u0[..., 0] = 0.05
    print(f"case perturbation halt in a uniform flow, run-ahead window {window}", flush=True)
    sim = runner.Simulation(nx, ny, 1.0)
    o = orc.Oracle(nx, ny, 1.0)
    for nid, x, y in nodes:
        sim.add_node(nid, x, y)
        o.add_node(nid, x, y)
    for a, b, d in conns:
        sim.add_connection(a, b, d)
        o.add_connection(a, b, d, [])
    sim.setup()
    o.cure()
    sim.init_solvers(u0)
    o.init_solvers(u0)
    sim.add_tension(0.002, 10)
    o.add_bio(0, [0.002], 10)
    sim.set_run_ahead(window)
    sim.run(24)
    o.run(24)
    assert sim.iteration == o.iteration() == 24
    assert o.perturb_counts()[0] >= 3
    assert sim.perturb_counts() == o.perturb_counts()
    t = sim.times()
    print("perturbations", sim.perturb_counts(), "sparse", t["sparse_perturbations"], "full",
          t["full_perturbations"], flush=True)
    assert t["sparse_perturbations"] >= 1 and t["full_perturbations"] == 0
    ids_d, xy_d, _ = sim.markers()
    ids_o, xy_o, _, _ = o.markers()
    np.testing.assert_array_equal(ids_d, ids_o)
    np.testing.assert_array_equal(xy_d, xy_o)
    ctx = sim.context()
    np.testing.assert_array_equal(ctx.get_populations(), o.get_f())
    dom_o, _ = o.get_domains()
    np.testing.assert_array_equal(ctx.get_domains(), dom_o)
    sim.close()


def test_perturbation_rings_lists_flagged_rings_in_walk_order():
    """Three rings (domains 5, 2, 9 in vector order, connections of each
    shuffled in the vector); rings 5 and 9 each hold one connection whose
    markers share x within 1e-10.  The entry point lists domains 5 and 9, in
    increasing id, each ring's markers from the first marker of its first
    connection in vector order along second -> outgoing connection."""
    rng = np.random.default_rng(21)
    n = 12
    centres = {5: (16.3, 20.2), 2: (40.1, 20.7), 9: (28.4, 34.9)}
    xy, conns = [], []
    for d, (cx, cy) in centres.items():
        base = len(xy)
        for k in range(n):
            a = 2 * np.pi * k / n + 0.1
            xy.append([cx + 5.0 * np.cos(a), cy + 5.0 * np.sin(a)])
        ring = [(base + k, base + (k + 1) % n, d) for k in range(n)]
        conns += [ring[i] for i in rng.permutation(n)]
    xy = np.array(xy)
    # line up connection 3 -> 4 of domain 5 and 7 -> 8 of domain 9
    xy[4, 0] = xy[3, 0] + 1e-10
    xy[24 + 8, 0] = xy[24 + 7, 0] - 1e-10
    ctx = capi.Context(64, 56, 1.0, max_markers=256, max_connections=256)
    print("case perturbation rings: 3 rings of 12, 2 flagged", flush=True)
    ctx._chk(ctx.lib.lbib_set_markers(ctx.h, len(xy), capi.dptr(np.ascontiguousarray(xy)), None, None))
    ids = np.arange(len(xy), dtype=np.uint32)
    ctx._chk(ctx.lib.lbib_set_marker_ids(ctx.h, capi.u32ptr(ids)))
    a = np.array([c[0] for c in conns], dtype=np.int32)
    b = np.array([c[1] for c in conns], dtype=np.int32)
    dm = np.array([c[2] for c in conns], dtype=np.uint32)
    bc = np.full(len(conns), -1, dtype=np.int32)
    ctx._chk(ctx.lib.lbib_set_connections(ctx.h, len(conns), a.ctypes.data_as(capi._I32),
                                          b.ctypes.data_as(capi._I32), capi.u32ptr(dm), bc.ctypes.data_as(capi._I32)))
    nr, nn = C.c_int(-1), C.c_int(-1)
    dom = np.zeros(8, dtype=np.uint32)
    off = np.zeros(9, dtype=np.int32)
    mk = np.full(64, -1, dtype=np.int32)
    out = np.zeros((64, 2))
    ctx._chk(ctx.lib.lbib_perturbation_rings(ctx.h, 8, 64, C.byref(nr), C.byref(nn), capi.u32ptr(dom),
                                             off.ctypes.data_as(capi._I32), mk.ctypes.data_as(capi._I32),
                                             capi.dptr(out)))
    assert nr.value == 2 and nn.value == 2 * n
    np.testing.assert_array_equal(dom[:2], [5, 9])
    np.testing.assert_array_equal(off[:3], [0, n, 2 * n])
    nxt = {int(a[i]): i for i in range(len(conns))}          # connection leaving each marker
    for r, d in enumerate((5, 9)):
        first = next(i for i in range(len(conns)) if dm[i] == d)
        walk, c = [], first
        for _ in range(n):
            walk.append(int(a[c]))
            c = nxt[int(b[c])]
        assert c == first
        np.testing.assert_array_equal(mk[off[r]:off[r + 1]], walk)
        np.testing.assert_array_equal(out[off[r]:off[r + 1]], xy[walk])
    # too few slots: the counts still come back, nothing else is promised
    ctx._chk(ctx.lib.lbib_perturbation_rings(ctx.h, 1, 64, C.byref(nr), C.byref(nn), capi.u32ptr(dom),
                                             off.ctypes.data_as(capi._I32), mk.ctypes.data_as(capi._I32),
                                             capi.dptr(out)))
    assert nr.value == 2
    ctx._chk(ctx.lib.lbib_perturbation_rings(ctx.h, 8, 10, C.byref(nr), C.byref(nn), capi.u32ptr(dom),
                                             off.ctypes.data_as(capi._I32), mk.ctypes.data_as(capi._I32),
                                             capi.dptr(out)))
    assert nr.value == 2 and nn.value == 2 * n
    ctx.close()


def _two_rings():
    """Rings of 96 markers (spacing 0.39: no remeshing or coarsening) in domains
    1 and 2 in a shear flow ux = 0.001 (y - 32); marker 1 of each ring is placed
    so that its connection from marker 0 is vertical within 1e-9 after the
    first move (positions found by fixed-point iteration on the oracle's first
    step), so iteration 0 halts for two connection perturbations."""
    n = 96
    nodes, conns = [], []
    for r, (cx, cy) in enumerate([(18.37, 30.61), (45.13, 33.29)]):
        for k in range(n):
            a = 2 * np.pi * k / n
            nodes.append([r * n + k, cx + 6.0 * np.cos(a), cy + 6.0 * np.sin(a)])
        conns += [(r * n + k, r * n + (k + 1) % n, r + 1) for k in range(n)]
    nodes[1][1] = 24.369619897552514
    nodes[n + 1][1] = 51.129570333185086
    return [tuple(v) for v in nodes], conns


@pytest.mark.parametrize("window", [1, 8])
def test_connection_perturbation_halt_resolved_from_flagged_rings_matches_oracle(orc, window):
    nx = ny = 64
    nodes, conns = _two_rings()
    u0 = np.zeros((ny, nx, 2))
    u0[..., 0] = 0.001 * (np.arange(ny)[:, None] - 32.0)
    print(f"case connection perturbation halt in a shear flow, run-ahead window {window}", flush=True)
    sim = runner.Simulation(nx, ny, 1.0)
    o = orc.Oracle(nx, ny, 1.0)
    for nid, x, y in nodes:
        sim.add_node(nid, x, y)
        o.add_node(nid, x, y)
    for a, b, d in conns:
        sim.add_connection(a, b, d)
        o.add_connection(a, b, d, [])
    sim.setup()
    o.cure()
    sim.init_solvers(u0)
    o.init_solvers(u0)
    sim.add_tension(0.002, 10)
    o.add_bio(0, [0.002], 10)
    sim.set_run_ahead(window)
    sim.run(24)
    o.run(24)
    assert sim.iteration == o.iteration() == 24
    assert o.perturb_counts()[1] >= 2
    assert sim.perturb_counts() == o.perturb_counts()
    t = sim.times()
    print("perturbations", sim.perturb_counts(), "sparse", t["sparse_perturbations"], "full",
          t["full_perturbations"], flush=True)
    assert t["sparse_perturbations"] >= 1 and t["full_perturbations"] == 0
    ids_d, xy_d, _ = sim.markers()
    ids_o, xy_o, _, _ = o.markers()
    np.testing.assert_array_equal(ids_d, ids_o)
    np.testing.assert_array_equal(xy_d, xy_o)
    ctx = sim.context()
    np.testing.assert_array_equal(ctx.get_populations(), o.get_f())
    dom_o, _ = o.get_domains()
    np.testing.assert_array_equal(ctx.get_domains(), dom_o)
    sim.close()
