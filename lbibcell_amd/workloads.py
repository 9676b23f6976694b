"""Synthetic membrane geometries for the configurations the path is measured
and checked on.  Every generator is seeded and returns plain arrays:

    nodes: list of (id, x, y)            GeometryNode ids and positions
    conns: list of (id1, id2, domain)    Connection order (counter-clockwise rings,
                                          so the inside is on the left, which is the
                                          orientation Connection::generateBoundaryNodes
                                          assumes, Connection.cpp:251-305)
"""
from __future__ import annotations

import math

import numpy as np


def circle_cell(cx, cy, radius, n, rng, jitter=0.5, first_id=0, domain=1):
    """Closed polygon of n markers, radius_k = radius + U(-jitter, jitter)."""
    nodes, conns = [], []
    r = radius + rng.uniform(-jitter, jitter, size=n)
    for k in range(n):
        th = 2.0 * math.pi * k / n
        nodes.append((first_id + k, cx + r[k] * math.cos(th), cy + r[k] * math.sin(th)))
    for k in range(n):
        conns.append((first_id + k, first_id + (k + 1) % n, domain))
    return nodes, conns


def tutorial_cell(nx=200, ny=200, radius=20.0, n=120, seed=1):
    """One round cell centred in the domain (tutorial_01-style run)."""
    rng = np.random.default_rng(seed)
    return circle_cell(nx / 2.0, ny / 2.0, radius, n, rng)


def hex_tissue(nx, ny, cols, rows, radius=40.0, rjit=4.0, n=150, pos_jit=2.0, seed=1, margin=0.0):
    """cols x rows cells on a hex grid (odd rows shifted half a pitch), radius
    radius +- rjit, centre jitter +- pos_jit; domains 1..cols*rows.  The grid
    fills [margin, n - margin) in x and y: growing cells push fluid out
    through the box outlet on the lattice edges, and the tissue drifts
    towards them."""
    rng = np.random.default_rng(seed)
    px = (nx - 2 * margin) / (cols + 0.5)
    py = (ny - 2 * margin) / rows
    nodes, conns = [], []
    dom = 1
    nid = 0
    for j in range(rows):
        for i in range(cols):
            cx = margin + (i + 0.5 + 0.5 * (j % 2)) * px + rng.uniform(-pos_jit, pos_jit)
            cy = margin + (j + 0.5) * py + rng.uniform(-pos_jit, pos_jit)
            r = radius + rng.uniform(-rjit, rjit)
            nd, cn = circle_cell(cx, cy, r, n, rng, jitter=0.5, first_id=nid, domain=dom)
            nodes += nd
            conns += cn
            nid += n
            dom += 1
    return nodes, conns


def packed_tissue(nx, ny, ncell, radius=30.0, rjit=5.0, n=120, gap=(0.8, 1.6), margin=12.0, seed=1):
    """ncell cells of radius radius +- rjit grown as one cluster from the
    lattice centre: each new cell sits at a gap U(gap) from the membrane of a
    cell on the cluster front and at least gap[0] from every other, so
    neighbouring membranes are within junction reach.  Marker radii are
    jittered by +-0.2.  Domains 1..ncell."""
    rng = np.random.default_rng(seed)
    r = radius + rng.uniform(-rjit, rjit, size=ncell)
    cx = np.empty(ncell)
    cy = np.empty(ncell)
    cx[0], cy[0] = nx / 2.0, ny / 2.0
    fails = np.zeros(ncell, dtype=np.int64)
    front = [0]
    placed = 1
    while placed < ncell:
        if not front:
            raise RuntimeError("packed_tissue: no room left for the next cell")
        j = front[rng.integers(len(front))]
        th = rng.uniform(0.0, 2.0 * math.pi)
        d = r[j] + r[placed] + rng.uniform(*gap)
        x, y = cx[j] + d * math.cos(th), cy[j] + d * math.sin(th)
        ok = margin + r[placed] <= x <= nx - margin - r[placed] and margin + r[placed] <= y <= ny - margin - r[placed]
        if ok:
            sep = np.hypot(cx[:placed] - x, cy[:placed] - y) - r[:placed] - r[placed]
            ok = sep.min() >= gap[0] - 1e-9
        if ok:
            cx[placed], cy[placed] = x, y
            front.append(placed)
            placed += 1
        else:
            fails[j] += 1
            if fails[j] > 40:
                front.remove(j)
    nodes, conns = [], []
    for c in range(ncell):
        nd, cn = circle_cell(cx[c], cy[c], r[c], n, rng, jitter=0.2, first_id=c * n, domain=c + 1)
        nodes += nd
        conns += cn
    return nodes, conns


def grid_cells(nx, ny, cols, rows, radius, n, seed=1):
    """cols x rows cells on a square grid (used for small multi-cell cases)."""
    rng = np.random.default_rng(seed)
    px, py = nx / cols, ny / rows
    nodes, conns = [], []
    nid, dom = 0, 1
    for j in range(rows):
        for i in range(cols):
            nd, cn = circle_cell((i + 0.5) * px, (j + 0.5) * py, radius, n, rng, jitter=0.5,
                                 first_id=nid, domain=dom)
            nodes += nd
            conns += cn
            nid += n
            dom += 1
    return nodes, conns


def _ceil16(v):
    return (int(v) + 15) // 16 * 16


def shelf_pack(sides, width):
    """Origins (x0, y0) for square domains of the given sides in an atlas of
    the given width: tallest first, left to right on shelves, every origin on
    the 16-site tile grid.  Returns (origins in input order, atlas height)."""
    order = sorted(range(len(sides)), key=lambda k: -sides[k])
    origins = [None] * len(sides)
    x = y = shelf_h = 0
    for k in order:
        w = _ceil16(sides[k])
        if w > width:
            raise ValueError("domain wider than the atlas")
        if x + w > width:
            x, y, shelf_h = 0, y + shelf_h, 0
        origins[k] = (x, y)
        x += w
        shelf_h = max(shelf_h, _ceil16(sides[k]))
    return origins, y + shelf_h


def place_cells(side, count, rmin, rmax, rng, margin=6.0, gap=3.0, tries=400):
    """Up to `count` non-overlapping circles (cx, cy, r) inside a side x side
    square, at least `margin` from its edges, by rejection sampling."""
    cells = []
    for _ in range(tries):
        if len(cells) >= count:
            break
        r = rng.uniform(rmin, rmax)
        lo, hi = margin + r, side - margin - r
        if hi <= lo:
            continue
        cx, cy = rng.uniform(lo, hi), rng.uniform(lo, hi)
        if all(math.hypot(cx - x, cy - y) > r + q + gap for x, y, q in cells):
            cells.append((cx, cy, r))
    if not cells:
        # the smallest cell always fits a side >= 2 (margin + rmin)
        r = rmin
        cells.append((side / 2.0, side / 2.0, min(r, side / 2.0 - margin)))
    return cells


def batched_ensemble(ndom=256, lmin=64, lmax=512, cells=(1, 16), radius=(8.0, 30.0), tau_f=(0.6, 1.5),
                     tau_c=(0.55, 1.2), nspecies=1, width=4096, seg=1.0, seed=1):
    """Parameter-sweep ensemble: ndom independent periodic square domains with
    sides drawn log-uniform in [lmin, lmax] (rounded to even sides), each
    holding 1-16 cells of radius U(radius) placed by rejection, tau_f and
    tau_c uniform in the given ranges; markers about `seg` lattice units
    apart.  Domains are shelf-packed into a `width`-wide atlas.

    Returns dict(nx, ny, rects=[(x0, y0, L, L, tau_f, [tau_c]*nspecies)],
    nodes, conns, domains=[list of cell domain ids per rectangle],
    cells=[(rect, cx, cy, r)] in local coordinates)."""
    rng = np.random.default_rng(seed)
    sides = [2 * int(round(math.exp(rng.uniform(math.log(lmin), math.log(lmax))) / 2.0)) for _ in range(ndom)]
    origins, height = shelf_pack(sides, width)
    nodes, conns, rects, domains, allcells = [], [], [], [], []
    nid, dom = 0, 1
    for k, L in enumerate(sides):
        x0, y0 = origins[k]
        tf = rng.uniform(*tau_f)
        tc = [rng.uniform(*tau_c) for _ in range(nspecies)]
        rects.append((x0, y0, L, L, tf, tc))
        want = int(rng.integers(cells[0], cells[1] + 1))
        ids = []
        for cx, cy, r in place_cells(L, want, radius[0], radius[1], rng):
            n = max(12, int(round(2.0 * math.pi * r / seg)))
            nd, cn = circle_cell(x0 + cx, y0 + cy, r, n, rng, jitter=0.5, first_id=nid, domain=dom)
            nodes += nd
            conns += cn
            allcells.append((k, cx, cy, r))
            ids.append(dom)
            nid += n
            dom += 1
        domains.append(ids)
    return dict(nx=width, ny=_ceil16(height), rects=rects, nodes=nodes, conns=conns, domains=domains,
                cells=allcells)


def inside_polygon(px, py, poly):
    """Even-odd test of points (px, py) (arrays) against a closed polygon (m, 2)."""
    x, y = poly[:, 0], poly[:, 1]
    xj, yj = np.roll(x, 1), np.roll(y, 1)
    inside = np.zeros(np.shape(px), dtype=bool)
    for k in range(len(x)):
        cond = (y[k] > py) != (yj[k] > py)
        with np.errstate(divide="ignore", invalid="ignore"):
            xc = (xj[k] - x[k]) * (py - y[k]) / (yj[k] - y[k]) + x[k]
        inside ^= cond & (px < xc)
    return inside


def adjoint_case(nx=512, ny=512, cols=3, rows=3, radius=40.0, n=150, rate=(2e-4, 1e-3), seed=1):
    """cols x rows round cells for the adjoint: marker rings (counter-
    clockwise), the cell map of the sites inside each ring (cells 1..), and
    one hidden source rate per cell drawn uniformly from `rate`."""
    rng = np.random.default_rng(seed)
    px, py = nx / cols, ny / rows
    rings = []
    yy, xx = np.mgrid[0:ny, 0:nx]
    dom = np.zeros((ny, nx), dtype=np.uint32)
    for j in range(rows):
        for i in range(cols):
            nodes, _ = circle_cell((i + 0.5) * px, (j + 0.5) * py, radius, n, rng, jitter=0.5)
            ring = np.array([[x, y] for _, x, y in nodes])
            rings.append(ring)
            dom[inside_polygon(xx, yy, ring)] = len(rings)
    rates = rng.uniform(rate[0], rate[1], size=len(rings))
    return rings, dom, rates


def cell_map(rings, nx, ny):
    """Cell index of every site (0 outside, ring c -> c + 1), each ring tested
    only over its bounding box."""
    dom = np.zeros((ny, nx), dtype=np.uint32)
    for c, ring in enumerate(rings):
        x0, y0 = max(int(np.floor(ring[:, 0].min())), 0), max(int(np.floor(ring[:, 1].min())), 0)
        x1, y1 = min(int(np.ceil(ring[:, 0].max())) + 1, nx), min(int(np.ceil(ring[:, 1].max())) + 1, ny)
        yy, xx = np.mgrid[y0:y1, x0:x1]
        sub = dom[y0:y1, x0:x1]
        sub[inside_polygon(xx, yy, ring)] = c + 1
    return dom


def junction_pairs(rings, radius):
    """Marker pairs for junction springs: every marker with its closest marker
    of another ring within `radius` (the pairing the cell-junction BioSolver
    makes), each pair once, as global indices of the concatenated rings."""
    from scipy.spatial import cKDTree
    xy = np.concatenate(rings)
    owner = np.concatenate([np.full(len(r), c) for c, r in enumerate(rings)])
    k = min(16, len(xy))
    dist, idx = cKDTree(xy).query(xy, k=k, distance_upper_bound=radius)
    pairs = set()
    for a in range(len(xy)):
        for d, b in zip(dist[a], idx[a]):
            if not np.isfinite(d):
                break
            if owner[b] != owner[a]:
                pairs.add((min(a, int(b)), max(a, int(b))))
                break
    return np.array(sorted(pairs), dtype=np.int32).reshape(-1, 2)


def tissue_adjoint_case(nx, ny, ncell, radius=30.0, rjit=5.0, n=120, rate=(2e-4, 1e-3), seed=1):
    """The packed touching tissue of packed_tissue as an adjoint problem:
    marker rings, the cell map and one hidden source rate per cell."""
    nodes, _ = packed_tissue(nx, ny, ncell, radius, rjit, n, seed=seed)
    xy = np.array([[x, y] for _, x, y in nodes])
    rings = [xy[c * n:(c + 1) * n] for c in range(ncell)]
    rates = np.random.default_rng(seed + 7).uniform(rate[0], rate[1], size=ncell)
    return rings, cell_map(rings, nx, ny), rates
