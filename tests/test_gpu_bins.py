"""Marker bins past their capacity.  The spread, the junction search and the
forces bin markers into 16x16 tiles of fixed capacity (256 entries), and
entries past it go to one overflow list that every reader also scans.  A ring
of 600 markers of radius 5.3 lies in one tile, so more than half of its
markers take the overflow path in the ordered FP64 bins and in the push bins
of the order-free spread.  Bars: FP64 bit-identical to the oracle; the FP32
store's force field within FP32 rounding of the FP64 one."""
import math

import numpy as np
import pytest

from lbibcell_amd import capi, workloads
from test_gpu_coupled import SIGNAL, _build, _compare

pytestmark = pytest.mark.gpu

NX = NY = 64


def _dense_ring():
    nodes, conns = workloads.circle_cell(24.37, 24.61, 5.3, 600, np.random.default_rng(5), jitter=0.01)
    # a constant outward force on every marker (ForceStructs type 2) from the
    # first step; the run starts at iteration 1, so no coarsening thins the
    # ring (it runs at iterations that are multiples of 10)
    forces = []
    for nid, x, y in nodes:
        th = math.atan2(y - 24.61, x - 24.37)
        forces.append((2, nid, 0, 2e-4 * math.cos(th), 2e-4 * math.sin(th)))
    return nodes, conns, forces


def _dense_case(orc, precision=capi.FP64):
    nodes, conns, forces = _dense_ring()
    assert all(16 <= x < 32 and 16 <= y < 32 for _, x, y in nodes), "the ring must sit in one tile"
    sim, o = _build(orc, NX, NY, 1.0, nodes, conns,
                    [("tutorial_01_CDESolverD2Q5_SIGNAL", capi.SPECIES_SIGNAL, 0.8, SIGNAL)], [],
                    precision=precision)
    for t, a, b, fx, fy in forces:
        sim.add_force(t, a, b, fx, fy)
        o.add_force(t, a, b, fx, fy)
    sim.set_iteration(1)
    o.set_iteration(1)
    return sim, o


def test_overflowing_bins_fp64_match_oracle(orc):
    print("case dense ring 600 markers in one tile, FP64, 6 steps", flush=True)
    sim, o = _dense_case(orc)
    sim.run(6)
    o.run(6)
    _compare(sim, o, 1, "dense ring")
    fd = sim.context().get_force_field()
    fo = o.get_force_field()
    np.testing.assert_array_equal(fd[0], fo[0])
    np.testing.assert_array_equal(fd[1], fo[1])
    assert np.max(np.abs(fo[0])) > 0.0, "the spread saw no force"
    sim.close()


def test_overflowing_push_bins_fp32_spread(orc):
    print("case dense ring 600 markers in one tile, FP32 store vs FP64, 1 step", flush=True)
    s64, o = _dense_case(orc)
    s32, _ = _dense_case(orc, precision=capi.FP32)
    s64.run(1)
    s32.run(1)
    f64 = s64.context().get_force_field()
    f32 = s32.context().get_force_field()
    for k in range(2):
        scale = np.max(np.abs(f64[k]))
        assert scale > 0.0
        assert np.max(np.abs(f32[k] - f64[k])) <= 1e-6 * scale, f"force component {k}"
    s64.close()
    s32.close()
