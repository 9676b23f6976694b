"""ctypes binding of the C++ host runtime (liblbib_host.so): the reference's
SimulationRunner sequence (apps/tutorial_01.cpp:59-145) driven from Python."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import capi

_HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(_HERE, "liblbib_host.so")

_P = C.c_void_p
_D = C.POINTER(C.c_double)
_U32 = C.POINTER(C.c_uint32)
_SIG = {
    "lbh_create": (_P, [C.c_int, C.c_int, C.c_double, C.c_int, C.c_int]),
    "lbh_destroy": (None, [_P]),
    "lbh_last_error": (C.c_char_p, [_P]),
    "lbh_add_species": (C.c_int, [_P, C.c_char_p, C.POINTER(capi.SpeciesSpec)]),
    "lbh_set_batch_layout": (C.c_int, [_P, C.c_int, C.POINTER(capi.Rect)]),
    "lbh_set_graph_steps": (C.c_int, [_P, C.c_uint]),
    "lbh_graph_stats": (None, [_P, C.POINTER(C.c_long)]),
    "lbh_graph_note": (C.c_int, [_P, C.c_char_p, C.c_int]),
    "lbh_domain_concentrations": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_double), C.c_int]),
    "lbh_load_geometry_vtk": (C.c_int, [_P, C.c_char_p]),
    "lbh_write_geometry_txt": (C.c_int, [_P, C.c_char_p]),
    "lbh_write_geometry_vtk": (C.c_int, [_P, C.c_char_p, C.c_uint]),
    "lbh_write_vtk": (C.c_int, [_P, C.c_char_p, C.c_uint, C.c_int, C.c_int]),
    "lbh_load_parameters": (C.c_int, [_P, C.c_char_p]),
    "lbh_load_geometry_txt": (C.c_int, [_P, C.c_char_p]),
    "lbh_add_node": (C.c_int, [_P, C.c_uint, C.c_double, C.c_double]),
    "lbh_add_connection": (C.c_int, [_P, C.c_uint, C.c_uint, C.c_uint, C.c_char_p]),
    "lbh_set_celltype": (C.c_int, [_P, C.c_uint, C.c_uint]),
    "lbh_setup": (C.c_int, [_P]),
    "lbh_init_solvers": (C.c_int, [_P, _D]),
    "lbh_add_force": (C.c_int, [_P, C.c_int, C.c_uint, C.c_uint, C.c_double, C.c_double, C.c_double, C.c_double]),
    "lbh_load_force_file": (C.c_int, [_P, C.c_char_p]),
    "lbh_add_mass_solver": (C.c_int, [_P, C.c_char_p]),
    "lbh_add_bio_solver": (C.c_int, [_P, C.c_char_p]),
    "lbh_add_bio_solver_params": (C.c_int, [_P, C.c_int, _D, C.c_uint]),
    "lbh_apply_bio": (C.c_int, [_P]),
    "lbh_run": (C.c_int, [_P, C.c_uint]),
    "lbh_run_simulation": (C.c_int, [_P, C.c_uint]),
    "lbh_write_simulation": (C.c_int, [_P, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p]),
    "lbh_init_solvers_file": (C.c_int, [_P, C.c_char_p]),
    "lbh_sync": (C.c_int, [_P]),
    "lbh_context": (_P, [_P]),
    "lbh_iteration": (C.c_uint, [_P]),
    "lbh_set_iteration": (C.c_int, [_P, C.c_uint]),
    "lbh_num_markers": (C.c_int, [_P]),
    "lbh_num_connections": (C.c_int, [_P]),
    "lbh_get_markers": (C.c_int, [_P, _U32, _D, _D]),
    "lbh_get_connections": (C.c_int, [_P, _U32, _U32, _U32]),
    "lbh_perturb_counts": (C.c_long, [_P, C.c_int]),
    "lbh_times": (None, [_P, _D]),
    "lbh_set_run_ahead": (C.c_int, [_P, C.c_uint]),
    "lbh_reset_times": (None, [_P]),
    "lbh_last_status": (C.c_int, [_P, C.POINTER(capi.Status)]),
    "lbh_num_areas": (C.c_int, [_P]),
    "lbh_get_areas": (C.c_int, [_P, _U32, _D]),
}
_lib = None


def load():
    global _lib
    if _lib is None:
        capi.load()
        if not os.path.exists(HOST_LIB):
            raise RuntimeError(f"{HOST_LIB} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(HOST_LIB)
        for k, (r, a) in _SIG.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


class Simulation:
    """One coupled simulation: geometry + lattice + plugins, stepped on the GPU."""

    def __init__(self, nx, ny, tau_fluid, force_model=capi.FORCE_LUO, precision=capi.FP64):
        self.L = load()
        self.h = self.L.lbh_create(nx, ny, tau_fluid, force_model, precision)
        self.nx, self.ny = nx, ny
        self._ctx = None

    def close(self):
        if self.h:
            self.L.lbh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise capi.LbibError(self.L.lbh_last_error(self.h).decode())

    def add_species(self, name, spec=None, tau=1.0):
        """Register a CDE solver by name (CDESolverFactory names).  Without a
        spec the registry's constants for that name are used (tau from `tau`)."""
        if spec is None:
            spec = capi.species_spec(tau=tau)
            spec.kind = -1
        self._chk(self.L.lbh_add_species(self.h, name.encode(), C.byref(spec)))

    def domain_concentrations(self, name, cap=65536):
        """GeometryHandler::computeAccumulatedDomainConcentrations: array indexed
        by domain id (0 for ids absent from the lattice)."""
        out = np.zeros(cap)
        w = self.L.lbh_domain_concentrations(self.h, name.encode(), capi.dptr(out), cap)
        if w < 0:
            raise capi.LbibError(self.L.lbh_last_error(self.h).decode())
        return out[:w]

    def load_geometry_vtk(self, path):
        """Geometry::loadGeometryVTK: the polygon block of an ASCII VTK XML
        .vtm (or a .vtp) with domain_identifier, cell_type and
        boundarydescriptors point arrays."""
        self._chk(self.L.lbh_load_geometry_vtk(self.h, str(path).encode()))

    def write_geometry_txt(self, path):
        """Geometry::writeGeometry of the host geometry (no device needed)."""
        self._chk(self.L.lbh_write_geometry_txt(self.h, str(path).encode()))

    def write_geometry_vtk(self, prefix, time=0):
        """The host geometry as a vtkCellReporter block (no device needed)."""
        self._chk(self.L.lbh_write_geometry_vtk(self.h, str(prefix).encode(), int(time)))

    def write_vtk(self, prefix, time, coarsening=1, fluid=True, species=True, cells=True):
        """vtkFluidReporter / vtkCDEReporter / vtkCellReporter: append this
        step's blocks to <prefix>_<time>.vtm (ASCII VTK XML)."""
        which = (1 if fluid else 0) | (2 if species else 0) | (4 if cells else 0)
        self._chk(self.L.lbh_write_vtk(self.h, str(prefix).encode(), int(time), int(coarsening), which))

    def set_graph_steps(self, n):
        """Record blocks of n steps (a multiple of 10) into a HIP graph and
        replay them with one launch each; 0 enqueues step by step."""
        self._chk(self.L.lbh_set_graph_steps(self.h, int(n)))

    def graph_stats(self):
        """(steps per graph, 0 once graph mode turned itself off; launches; recordings)"""
        out = (C.c_long * 3)()
        self.L.lbh_graph_stats(self.h, out)
        return tuple(out)

    def graph_note(self):
        """Why graph mode turned itself off ('' while on)."""
        buf = C.create_string_buffer(512)
        self.L.lbh_graph_note(self.h, buf, 512)
        return buf.value.decode()

    def set_batch_layout(self, rects):
        """Batched ensemble: rects = [(x0, y0, w, h, tau_fluid, [tau_species...]), ...],
        one periodic domain each (x0, y0 multiples of 16); before any node is
        added (each domain draws its own position perturbations)."""
        arr = capi.make_rects(rects)
        self._chk(self.L.lbh_set_batch_layout(self.h, len(arr), arr if len(arr) else None))

    def add_node(self, nid, x, y):
        self._chk(self.L.lbh_add_node(self.h, nid, x, y))

    def add_connection(self, a, b, dom, bc=""):
        self._chk(self.L.lbh_add_connection(self.h, a, b, dom, bc.encode()))

    def add_geometry(self, nodes, conns, bc=""):
        for nid, x, y in nodes:
            self.add_node(nid, x, y)
        for a, b, d in conns:
            self.add_connection(a, b, d, bc)

    def load_geometry_txt(self, path):
        self._chk(self.L.lbh_load_geometry_txt(self.h, path.encode()))

    def set_celltype(self, dom, t):
        self._chk(self.L.lbh_set_celltype(self.h, dom, t))

    def setup(self):
        self._chk(self.L.lbh_setup(self.h))

    def init_solvers(self, u0=None):
        self._chk(self.L.lbh_init_solvers(self.h, capi.dptr(u0) if u0 is not None else None))

    def add_force(self, t, id1, id2=0, a=0.0, b=0.0, c=0.0, d=0.0):
        self._chk(self.L.lbh_add_force(self.h, t, id1, id2, a, b, c, d))

    def add_mass_solver(self, name):
        self._chk(self.L.lbh_add_mass_solver(self.h, name.encode()))

    def add_bio_solver(self, name):
        self._chk(self.L.lbh_add_bio_solver(self.h, name.encode()))

    def add_tension(self, force=0.01, freq=10):
        p = np.array([force, 0, 0, 0], dtype=np.float64)
        self._chk(self.L.lbh_add_bio_solver_params(self.h, 0, capi.dptr(p), freq))

    def add_junctions(self, radius=1.0, k=0.02, l0=0.5, freq=10):
        """tutorial_02_BioSolverCellJunction (tutorial_02_BioSolverCellJunction.cpp:37-99)
        with its constants as defaults: every freq iterations the type-7 springs
        are rebuilt between each node and its closest node of another cell
        within radius."""
        p = np.array([radius, k, l0, 0], dtype=np.float64)
        self._chk(self.L.lbh_add_bio_solver_params(self.h, 2, capi.dptr(p), freq))

    def add_bio_solver_with(self, kind, prm, freq):
        """A BioSolver with its constants given (kinds: 3 Hookean tension (k, l0),
        4 growth (source), 5 threshold growth (maxsource, threshold),
        6 differentiation (threshold, species index in name order),
        7 / 8 cell division along the longest axis / a random direction
        (max cell size), 9 PI area regulator (target, target of cell 1, kp,
        maxsource))."""
        p = np.zeros(4)
        p[:len(prm)] = prm
        self._chk(self.L.lbh_add_bio_solver_params(self.h, int(kind), capi.dptr(p), freq))

    def add_area_regulator(self, target=200.0, kp=1e-5, maxsource=0.004, factor=0.0, freq=1):
        p = np.array([target, kp, maxsource, factor], dtype=np.float64)
        self._chk(self.L.lbh_add_bio_solver_params(self.h, 1, capi.dptr(p), freq))

    def run(self, steps):
        self._chk(self.L.lbh_run(self.h, steps))

    def apply_bio(self):
        """The BioSolvers of the current iteration alone, outside a step."""
        self._chk(self.L.lbh_apply_bio(self.h))

    def run_simulation(self, iterations):
        """SimulationRunner::runSimulation: `iterations` steps; a failure is
        caught and dumped to output/CRASH_*_<t>.txt under the working directory."""
        self._chk(self.L.lbh_run_simulation(self.h, int(iterations)))

    def write_simulation(self, force, geometry, parameters, solvers):
        """SimulationRunner::writeSimulation: the reference's four text files."""
        self._chk(self.L.lbh_write_simulation(self.h, str(force).encode(), str(geometry).encode(),
                                              str(parameters).encode(), str(solvers).encode()))

    def init_solvers_file(self, solvers):
        """SimulationRunner::initSolvers(filename): load a solver dump."""
        self._chk(self.L.lbh_init_solvers_file(self.h, str(solvers).encode()))

    def set_run_ahead(self, n):
        """Iterations enqueued between device status reads (1 = every iteration)."""
        self._chk(self.L.lbh_set_run_ahead(self.h, n))

    def sync(self):
        self._chk(self.L.lbh_sync(self.h))

    @property
    def iteration(self):
        return self.L.lbh_iteration(self.h)

    def set_iteration(self, it):
        self.L.lbh_set_iteration(self.h, it)

    def context(self):
        """A capi.Context view of the device state (not owning)."""
        if self._ctx is None:
            ptr = self.L.lbh_context(self.h)
            ctx = capi.Context.__new__(capi.Context)
            ctx.lib = capi.load()
            ctx.h = C.c_void_p(ptr)
            ctx.nx, ctx.ny, ctx.n = self.nx, self.ny, self.nx * self.ny
            ctx.ns = 0
            ctx.close = lambda: None
            self._ctx = ctx
        return self._ctx

    def markers(self):
        m = self.L.lbh_num_markers(self.h)
        ids = np.empty(m, dtype=np.uint32)
        xy = np.empty((m, 2))
        v = np.empty((m, 2))
        self._chk(self.L.lbh_get_markers(self.h, capi.u32ptr(ids), capi.dptr(xy), capi.dptr(v)))
        return ids, xy, v

    def connections(self):
        c = self.L.lbh_num_connections(self.h)
        a = np.empty(c, dtype=np.uint32)
        b = np.empty(c, dtype=np.uint32)
        d = np.empty(c, dtype=np.uint32)
        self.L.lbh_get_connections(self.h, capi.u32ptr(a), capi.u32ptr(b), capi.u32ptr(d))
        return a, b, d

    def perturb_counts(self):
        return self.L.lbh_perturb_counts(self.h, 0), self.L.lbh_perturb_counts(self.h, 1)

    def times(self):
        out = np.zeros(9)
        self.L.lbh_times(self.h, capi.dptr(out))
        return dict(lattice=out[0], geometry=out[1], host=out[2], bio=out[3], wait=out[5], steps=int(out[4]),
                    stops=int(out[6]), sparse_perturbations=int(out[7]), full_perturbations=int(out[8]))

    def reset_times(self):
        self.L.lbh_reset_times(self.h)

    def last_status(self):
        s = capi.Status()
        self.L.lbh_last_status(self.h, C.byref(s))
        return s

    def areas(self):
        k = self.L.lbh_num_areas(self.h)
        d = np.empty(k, dtype=np.uint32)
        a = np.empty(k)
        self._chk(self.L.lbh_get_areas(self.h, capi.u32ptr(d), capi.dptr(a)))
        return dict(zip(d.tolist(), a.tolist()))
