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


def hex_tissue(nx, ny, cols, rows, radius=40.0, rjit=4.0, n=150, pos_jit=2.0, seed=1):
    """cols x rows cells on a hex grid (odd rows shifted half a pitch), radius
    radius +- rjit, centre jitter +- pos_jit; domains 1..cols*rows."""
    rng = np.random.default_rng(seed)
    px = nx / (cols + 0.5)
    py = ny / rows
    nodes, conns = [], []
    dom = 1
    nid = 0
    for j in range(rows):
        for i in range(cols):
            cx = (i + 0.5 + 0.5 * (j % 2)) * px + rng.uniform(-pos_jit, pos_jit)
            cy = (j + 0.5) * py + rng.uniform(-pos_jit, pos_jit)
            r = radius + rng.uniform(-rjit, rjit)
            nd, cn = circle_cell(cx, cy, r, n, rng, jitter=0.5, first_id=nid, domain=dom)
            nodes += nd
            conns += cn
            nid += n
            dom += 1
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
