"""Synthetic workload generators: the packed tissue of the four-species
junction workload keeps cells apart (no membrane crossings) while putting
neighbouring membranes within junction reach."""
import numpy as np

from lbibcell_amd import workloads


def test_packed_tissue_cells_touch_without_overlap():
    nx = ny = 512
    nodes, conns = workloads.packed_tissue(nx, ny, 40, radius=20.0, rjit=4.0, n=80, seed=3)
    xy = np.array([[n[1], n[2]] for n in nodes])
    dom = np.array([c[2] for c in conns])     # connection k starts at marker k
    assert len(set(dom)) == 40 and len(nodes) == 40 * 80
    assert xy.min() > 2.0 and xy.max() < nx - 2.0
    centres = np.array([xy[dom == d].mean(axis=0) for d in range(1, 41)])
    radii = np.array([np.hypot(*(xy[dom == d] - centres[d - 1]).T).mean() for d in range(1, 41)])
    # markers of one cell never sit inside another cell's polygon
    for d in range(1, 41):
        other = dom != d
        dist = np.hypot(*(xy[other] - centres[d - 1]).T)
        assert np.all(dist > radii[d - 1] - 0.5), f"cell {d} overlaps a neighbour"
    # every cell has a membrane of another cell within 2 lattice units
    close = 0
    for d in range(1, 41):
        a, b = xy[dom == d], xy[dom != d]
        dmin = np.min(np.hypot(a[:, None, 0] - b[None, :, 0], a[:, None, 1] - b[None, :, 1]))
        close += dmin < 2.0
    assert close == 40
