"""ctypes binding of the CPU oracle (oracle/liblbib_oracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker, never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbib_oracle.so")

SP_PLAIN, SP_SIGNAL, SP_LINEAR, SP_SCHNAKENBERG_R, SP_SCHNAKENBERG_L = 0, 1, 2, 3, 4
_D = C.POINTER(C.c_double)
_U32 = C.POINTER(C.c_uint32)
_I32 = C.POINTER(C.c_int32)
_P = C.c_void_p

_SIG = {
    "orc_create": (_P, [C.c_int, C.c_int, C.c_double, C.c_int]),
    "orc_destroy": (None, [_P]),
    "orc_last_error": (C.c_char_p, [_P]),
    "orc_set_threads": (None, [_P, C.c_int]),
    "orc_set_iteration": (None, [_P, C.c_uint]),
    "orc_iteration": (C.c_uint, [_P]),
    "orc_add_species": (C.c_int, [_P, C.c_int, C.c_double, _D, C.c_int]),
    "orc_add_node": (None, [_P, C.c_uint, C.c_double, C.c_double]),
    "orc_add_connection": (C.c_int, [_P, C.c_uint, C.c_uint, C.c_uint, _I32, C.c_int]),
    "orc_cure_lattice": (C.c_int, [_P]),
    "orc_init_solvers": (None, [_P, _D]),
    "orc_add_force": (None, [_P, C.c_int, C.c_uint, C.c_uint, C.c_double, C.c_double, C.c_double, C.c_double]),
    "orc_set_celltype": (None, [_P, C.c_uint, C.c_uint]),
    "orc_enable_box_outlet": (None, [_P, C.c_int]),
    "orc_add_bio": (None, [_P, C.c_int, _D, C.c_uint]),
    "orc_run": (C.c_int, [_P, C.c_int]),
    "orc_apply_bio": (C.c_int, [_P]),
    "orc_get_mass": (None, [_P, _D]),
    "orc_copy_cell_types": (None, [_P]),
    "orc_num_forces": (C.c_int, [_P]),
    "orc_get_forces": (None, [_P, _I32, _U32, _U32, _D]),
    "orc_set_force_field": (None, [_P, _D, _D]),
    "orc_set_mass": (None, [_P, _D]),
    "orc_set_domains": (None, [_P, _U32]),
    "orc_lattice_step": (None, [_P, C.c_int]),
    "orc_get_f": (None, [_P, _D]),
    "orc_write_solver_dump": (C.c_int, [_P, C.c_char_p, C.c_char_p]),
    "orc_set_f": (None, [_P, _D]),
    "orc_get_g": (None, [_P, C.c_int, _D]),
    "orc_set_g": (None, [_P, C.c_int, _D]),
    "orc_get_velocity": (None, [_P, _D, _D]),
    "orc_get_force_field": (None, [_P, _D, _D]),
    "orc_get_domains": (None, [_P, _U32, _U32]),
    "orc_get_bn_domains": (None, [_P, _I32]),
    "orc_num_markers": (C.c_int, [_P]),
    "orc_num_connections": (C.c_int, [_P]),
    "orc_get_markers": (None, [_P, _U32, _D, _D, _D]),
    "orc_set_marker_forces": (None, [_P, _D]),
    "orc_get_connections": (None, [_P, _U32, _U32, _U32]),
    "orc_num_areas": (C.c_int, [_P]),
    "orc_get_areas": (None, [_P, _U32, _D]),
    "orc_spread": (None, [_P]),
    "orc_calculate_force": (None, [_P]),
    "orc_interpolate": (C.c_int, [_P]),
    "orc_set_velocity": (None, [_P, _D, _D]),
    "orc_fastcos": (C.c_double, [C.c_double]),
    "orc_delta": (C.c_double, [C.c_double, C.c_double]),
    "orc_perturb_counts": (C.c_long, [_P, C.c_int]),
    "orc_rng_draws": (None, [C.c_int, _D]),
    "orc_mt19937_default_10000": (C.c_uint, []),
    "orc_reinit_count": (C.c_int, [_P]),
    "orc_rect_within": (C.c_int, [C.c_double] * 6),
    "orc_num_bnodes": (C.c_int, [_P]),
    "orc_stage_times": (None, [_P, _D]),
    "orc_get_bnodes": (None, [_P, _I32, _I32, _U32, _I32, _D]),
    "orc_set_marker_velocities": (None, [_P, _D]),
    "orc_stream_bc": (None, [_P]),
    "orc_get_g_slots": (None, [_P, C.c_int, _D]),
    "orc_mark_boundary": (None, [_P, C.c_long, C.c_int, C.c_uint]),
    "orc_reinit_site": (None, [_P, C.c_int, C.c_long]),
    "orc_remesh": (C.c_int, [_P]),
    "orc_coarsen": (None, [_P]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        for k, (r, a) in _SIG.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_D)


class Oracle:
    """One single-threaded reference-semantics simulation."""

    def __init__(self, nx, ny, tau_f, force_model=0):
        self.L = lib()
        self.h = self.L.orc_create(nx, ny, tau_f, force_model)
        self.nx, self.ny, self.n = nx, ny, nx * ny
        self.ns = 0

    def __del__(self):
        try:
            if self.h:
                self.L.orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error(self.h).decode())

    def add_species(self, kind, tau, params=()):
        p = np.ascontiguousarray(params, dtype=np.float64)
        self.ns += 1
        return self.L.orc_add_species(self.h, kind, tau, _d(p) if p.size else None, int(p.size))

    def add_node(self, nid, x, y):
        self.L.orc_add_node(self.h, nid, x, y)

    def add_connection(self, a, b, dom, bc=()):
        arr = np.ascontiguousarray(bc, dtype=np.int32)
        self._chk(self.L.orc_add_connection(self.h, a, b, dom,
                                            arr.ctypes.data_as(_I32) if arr.size else None, int(arr.size)))

    def cure(self):
        self._chk(self.L.orc_cure_lattice(self.h))

    def init_solvers(self, u0=None):
        self.L.orc_init_solvers(self.h, _d(u0) if u0 is not None else None)

    def add_force(self, t, id1, id2=0, a=0.0, b=0.0, c=0.0, d=0.0):
        self.L.orc_add_force(self.h, t, id1, id2, a, b, c, d)

    def set_celltype(self, dom, t):
        self.L.orc_set_celltype(self.h, dom, t)

    def box_outlet(self, on=True):
        self.L.orc_enable_box_outlet(self.h, 1 if on else 0)

    def add_bio(self, kind, prm, freq):
        p = np.zeros(4)
        p[:len(prm)] = prm
        self.L.orc_add_bio(self.h, kind, _d(p), freq)

    def apply_bio(self):
        self._chk(self.L.orc_apply_bio(self.h))

    def copy_cell_types(self):
        self.L.orc_copy_cell_types(self.h)

    def get_mass(self):
        m = np.empty((self.ny, self.nx))
        self.L.orc_get_mass(self.h, _d(m))
        return m

    def forces(self):
        """ForceSolver::forcemap_ in map order: (type, id1, id2, prm (n, 4))."""
        n = self.L.orc_num_forces(self.h)
        t = np.empty(max(n, 1), dtype=np.int32)
        a = np.empty(max(n, 1), dtype=np.uint32)
        b = np.empty(max(n, 1), dtype=np.uint32)
        prm = np.empty((max(n, 1), 4))
        self.L.orc_get_forces(self.h, t.ctypes.data_as(_I32), a.ctypes.data_as(_U32), b.ctypes.data_as(_U32),
                              _d(prm))
        return t[:n], a[:n], b[:n], prm[:n]

    def run(self, steps):
        self._chk(self.L.orc_run(self.h, steps))

    def set_threads(self, t):
        self.L.orc_set_threads(self.h, t)

    def set_iteration(self, it):
        self.L.orc_set_iteration(self.h, it)

    def iteration(self):
        return self.L.orc_iteration(self.h)

    def set_force_field(self, fx, fy):
        fx = np.ascontiguousarray(fx, dtype=np.float64)
        fy = np.ascontiguousarray(fy, dtype=np.float64)
        self.L.orc_set_force_field(self.h, _d(fx), _d(fy))

    def set_mass(self, m):
        m = np.ascontiguousarray(m, dtype=np.float64)
        self.L.orc_set_mass(self.h, _d(m))

    def set_domains(self, d):
        d = np.ascontiguousarray(d, dtype=np.uint32)
        self.L.orc_set_domains(self.h, d.ctypes.data_as(_U32))

    def lattice_step(self, steps=1):
        self.L.orc_lattice_step(self.h, steps)

    def write_solver_dump(self, path, species_names):
        """The reference's solver text dump of the current state."""
        rc = self.L.orc_write_solver_dump(self.h, str(path).encode(), "\n".join(species_names).encode())
        if rc != 0:
            raise RuntimeError(f"cannot write {path}")

    def get_f(self):
        f = np.empty((9, self.ny, self.nx))
        self.L.orc_get_f(self.h, _d(f))
        return f

    def set_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        self.L.orc_set_f(self.h, _d(f))

    def get_g(self, k):
        g = np.empty((5, self.ny, self.nx))
        self.L.orc_get_g(self.h, k, _d(g))
        return g

    def set_g(self, k, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        self.L.orc_set_g(self.h, k, _d(g))

    def get_velocity(self):
        ux = np.empty((self.ny, self.nx))
        uy = np.empty((self.ny, self.nx))
        self.L.orc_get_velocity(self.h, _d(ux), _d(uy))
        return ux, uy

    def set_velocity(self, ux, uy):
        ux = np.ascontiguousarray(ux, dtype=np.float64)
        uy = np.ascontiguousarray(uy, dtype=np.float64)
        self.L.orc_set_velocity(self.h, _d(ux), _d(uy))

    def get_force_field(self):
        fx = np.empty((self.ny, self.nx))
        fy = np.empty((self.ny, self.nx))
        self.L.orc_get_force_field(self.h, _d(fx), _d(fy))
        return fx, fy

    def get_domains(self):
        d = np.empty((self.ny, self.nx), dtype=np.uint32)
        ct = np.empty((self.ny, self.nx), dtype=np.uint32)
        self.L.orc_get_domains(self.h, d.ctypes.data_as(_U32), ct.ctypes.data_as(_U32))
        return d, ct

    def get_bn_domains(self):
        b = np.empty((self.ny, self.nx, 4), dtype=np.int32)
        self.L.orc_get_bn_domains(self.h, b.ctypes.data_as(_I32))
        return b

    def markers(self):
        m = self.L.orc_num_markers(self.h)
        ids = np.empty(m, dtype=np.uint32)
        xy = np.empty((m, 2))
        v = np.empty((m, 2))
        f = np.empty((m, 2))
        self.L.orc_get_markers(self.h, ids.ctypes.data_as(_U32), _d(xy), _d(v), _d(f))
        return ids, xy, v, f

    def set_marker_forces(self, fxy):
        fxy = np.ascontiguousarray(fxy, dtype=np.float64)
        self.L.orc_set_marker_forces(self.h, _d(fxy))

    def connections(self):
        c = self.L.orc_num_connections(self.h)
        a = np.empty(c, dtype=np.uint32)
        b = np.empty(c, dtype=np.uint32)
        d = np.empty(c, dtype=np.uint32)
        self.L.orc_get_connections(self.h, a.ctypes.data_as(_U32), b.ctypes.data_as(_U32), d.ctypes.data_as(_U32))
        return a, b, d

    def areas(self):
        k = self.L.orc_num_areas(self.h)
        d = np.empty(k, dtype=np.uint32)
        a = np.empty(k)
        self.L.orc_get_areas(self.h, d.ctypes.data_as(_U32), _d(a))
        return dict(zip(d.tolist(), a.tolist()))

    def spread(self):
        self.L.orc_spread(self.h)

    def calculate_force(self):
        self.L.orc_calculate_force(self.h)

    def interpolate(self):
        self._chk(self.L.orc_interpolate(self.h))

    def perturb_counts(self):
        return self.L.orc_perturb_counts(self.h, 0), self.L.orc_perturb_counts(self.h, 1)

    def reinit_count(self):
        return self.L.orc_reinit_count(self.h)

    STAGES = ("forces_and_spread", "boundary_and_streaming", "collide", "interpolate_move_cut_links_domains",
              "remesh_coarsen", "mass_and_biosolvers")

    def stage_times(self):
        """Seconds spent in each stage of the coupled steps run so far."""
        out = np.zeros(6)
        self.L.orc_stage_times(self.h, _d(out))
        return dict(zip(self.STAGES, out.tolist()))

    # single stages for the hand-derived tests
    def bnodes(self):
        """Alive boundary nodes: site, dir (to the physical neighbour), domain,
        generating connection index, (x, y, vx, vy)."""
        k = self.L.orc_num_bnodes(self.h)
        site = np.empty(max(k, 1), dtype=np.int32)
        d = np.empty(max(k, 1), dtype=np.int32)
        dom = np.empty(max(k, 1), dtype=np.uint32)
        conn = np.empty(max(k, 1), dtype=np.int32)
        xyv = np.empty((max(k, 1), 4))
        self.L.orc_get_bnodes(self.h, site.ctypes.data_as(_I32), d.ctypes.data_as(_I32), dom.ctypes.data_as(_U32),
                              conn.ctypes.data_as(_I32), _d(xyv))
        return site[:k], d[:k], dom[:k], conn[:k], xyv[:k]

    def set_marker_velocities(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        self.L.orc_set_marker_velocities(self.h, _d(v))

    def stream_bc(self):
        self.L.orc_stream_bc(self.h)

    def get_g_slots(self, k):
        g = np.empty((5, self.ny, self.nx))
        self.L.orc_get_g_slots(self.h, k, _d(g))
        return g

    def mark_boundary(self, x, y, slot, dom):
        self.L.orc_mark_boundary(self.h, y * self.nx + x, slot, dom)

    def reinit_site(self, k, x, y):
        self.L.orc_reinit_site(self.h, k, y * self.nx + x)

    def remesh(self):
        return self.L.orc_remesh(self.h)

    def coarsen(self):
        self.L.orc_coarsen(self.h)


def fastcos(x):
    return lib().orc_fastcos(x)


def delta(dx, dy):
    return lib().orc_delta(dx, dy)


def rng_draws(n):
    out = np.empty(n)
    lib().orc_rng_draws(n, _d(out))
    return out


def mt19937_default_10000():
    return lib().orc_mt19937_default_10000()


def rect_within(n, s, w, e, x, y):
    return bool(lib().orc_rect_within(n, s, w, e, x, y))
