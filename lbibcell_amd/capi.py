"""ctypes binding of the liblbib_hip C ABI (include/lbib.h).

The shared library is built in-tree (``lbibcell_amd/liblbib_hip.so``).  Loading
fails loudly when it is missing: there is no CPU fallback for the device path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbib_hip.so")

MAX_SPECIES = 4
FP64, FP32, FP16_DEV, BF16_DEV, FP16_REM = 0, 1, 2, 3, 4
FORCE_LUO, FORCE_GUO_AS_WRITTEN, FORCE_GUO = 0, 1, 2
SPECIES_PLAIN, SPECIES_SIGNAL, SPECIES_LINEAR = 0, 1, 2
SPECIES_SCHNAKENBERG_R, SPECIES_SCHNAKENBERG_L = 3, 4

ERR_VELOCITY = 0x1
ERR_EDGE = 0x2
ERR_SCANLINE = 0x4
ERR_SLOPE = 0x8
ERR_CAPACITY = 0x10
ERR_TOPOLOGY = 0x20
ERR_MASK = 0xFF
HALT_PERTURB_NODE = 0x100
HALT_PERTURB_CONN = 0x200
HALT_PERTURB_REMESH = 0x400
HALT_GROW = 0x800
HALT_MASK = 0xF00
FLAG_REMESH = 0x1000
FLAG_COARSEN = 0x2000
SKIPPED = 0x4000


class SpeciesSpec(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("tau", C.c_double),
        ("production", C.c_double),
        ("decay", C.c_double),
        ("init_inside", C.c_double),
        ("switchoff", C.c_uint32),
        ("prod_domain", C.c_int32),
        ("source_inside", C.c_double),
        ("krow", C.c_double * MAX_SPECIES),
    ]


class Config(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("tau_fluid", C.c_double),
        ("force_model", C.c_int32),
        ("precision", C.c_int32),
        ("nspecies", C.c_int32),
        ("species", SpeciesSpec * MAX_SPECIES),
        ("box_outlet", C.c_int32),
        ("max_markers", C.c_int32),
        ("max_connections", C.c_int32),
    ]


class Rect(C.Structure):
    """lbib_rect: one periodic domain [x0, x0+w) x [y0, y0+h) of a batched atlas."""
    _fields_ = [
        ("x0", C.c_int32), ("y0", C.c_int32), ("w", C.c_int32), ("h", C.c_int32),
        ("tau_fluid", C.c_double),
        ("tau_species", C.c_double * MAX_SPECIES),
    ]


class LatticeShape(C.Structure):
    _fields_ = [
        ("variant", C.c_int32), ("workgroup", C.c_int32), ("sites_per_lane", C.c_int32),
        ("ms", C.c_double),
    ]

    def as_tuple(self):
        return (self.variant, self.workgroup, self.sites_per_lane)


class Status(C.Structure):
    _fields_ = [
        ("bits", C.c_uint32),
        ("first_bad_marker", C.c_int32),
        ("n_touched_sites", C.c_int32),
        ("n_boundary_nodes", C.c_int32),
        ("n_reinit", C.c_int32),
        ("n_candidates", C.c_int32),
        ("n_markers", C.c_int32),
        ("n_connections", C.c_int32),
        ("n_remeshed", C.c_int32),
        ("n_coarsened", C.c_int32),
        ("step_iteration", C.c_uint32),
        ("n_reinit_levels", C.c_int32),
    ]


_lib = None

# name -> (restype, argtypes)
_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_U32 = C.POINTER(C.c_uint32)
SIGNATURES = {
    "lbib_create": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    "lbib_destroy": (None, [_P]),
    "lbib_last_error": (C.c_char_p, [_P]),
    "lbib_synchronize": (C.c_int, [_P]),
    "lbib_device_bytes": (C.c_size_t, [_P]),
    "lbib_init_solvers": (C.c_int, [_P, _D]),
    "lbib_set_populations": (C.c_int, [_P, _D]),
    "lbib_get_populations": (C.c_int, [_P, _D]),
    "lbib_set_species": (C.c_int, [_P, C.c_int, _D]),
    "lbib_get_species": (C.c_int, [_P, C.c_int, _D]),
    "lbib_get_velocity": (C.c_int, [_P, _D, _D]),
    "lbib_get_domains": (C.c_int, [_P, _U32]),
    "lbib_set_domains": (C.c_int, [_P, _U32]),
    "lbib_set_force_field": (C.c_int, [_P, _D, _D]),
    "lbib_get_force_field": (C.c_int, [_P, _D, _D]),
    "lbib_set_mass_field": (C.c_int, [_P, _D]),
    "lbib_set_markers": (C.c_int, [_P, C.c_int, _D, _D, _D]),
    "lbib_get_markers": (C.c_int, [_P, _D, _D, _D, _D]),
    "lbib_perturbation_candidates": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int), _I32, _U32, _D]),
    "lbib_set_marker_positions": (C.c_int, [_P, C.c_int, _I32, _D]),
    "lbib_perturbation_rings": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), _U32, _I32,
                                          _I32, _D]),
    "lbib_set_connections": (C.c_int, [_P, C.c_int, _I32, _I32, _U32, _I32]),
    "lbib_set_bc_descriptors": (C.c_int, [_P, C.c_int, _I32, _I32]),
    "lbib_reserve": (C.c_int, [_P, C.c_int, C.c_int]),
    "lbib_get_marker_ids": (C.c_int, [_P, _U32]),
    "lbib_set_marker_ids": (C.c_int, [_P, _U32]),
    "lbib_get_connections": (C.c_int, [_P, _I32, _I32, _U32, _I32]),
    "lbib_get_forces": (C.c_int, [_P, _I32, _I32, _I32, _D]),
    "lbib_set_celltypes": (C.c_int, [_P, C.c_int, _U32]),
    "lbib_set_forces": (C.c_int, [_P, C.c_int, _I32, _I32, _I32, _D]),
    "lbib_step_lattice": (C.c_int, [_P, C.c_uint32]),
    "lbib_step_geometry": (C.c_int, [_P, C.c_uint32, C.c_int]),
    "lbib_step_topology": (C.c_int, [_P, C.c_int]),
    "lbib_forces_tension": (C.c_int, [_P, C.c_double]),
    "lbib_forces_junctions": (C.c_int, [_P, C.c_double, C.c_double, C.c_double]),
    "lbib_forces_springs": (C.c_int, [_P, C.c_double, C.c_double]),
    "lbib_mass_growth": (C.c_int, [_P, C.c_int, C.c_int, C.c_double, C.c_double]),
    "lbib_forces_clear": (C.c_int, [_P]),
    "lbib_cure_lattice": (C.c_int, [_P]),
    "lbib_set_box_outlet": (C.c_int, [_P, C.c_int]),
    "lbib_step_lattice_only": (C.c_int, [_P, C.c_uint32]),
    "lbib_read_status": (C.c_int, [_P, C.POINTER(Status)]),
    "lbib_clear_status": (C.c_int, [_P]),
    "lbib_step_begin": (C.c_int, [_P, C.c_uint32]),
    "lbib_rewind": (C.c_int, [_P, C.c_uint32]),
    "lbib_compute_areas": (C.c_int, [_P, _D]),
    "lbib_area_regulator": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_double]),
    "lbib_add_mass_by_domain": (C.c_int, [_P, C.c_int, _D]),
    "lbib_add_mass_sites": (C.c_int, [_P, C.c_int64, C.POINTER(C.c_int64), _D]),
    "lbib_profile_enable": (C.c_int, [_P, C.c_int]),
    "lbib_profile_read": (C.c_int, [_P, _D, C.POINTER(C.c_int64), _D]),
    "lbib_profile_copy": (C.c_int, [_P, C.c_int64, C.c_int, _D]),
    "lbib_set_batch_layout": (C.c_int, [_P, C.c_int, C.POINTER(Rect)]),
    "lbib_autotune": (C.c_int, [_P, C.c_int, C.POINTER(LatticeShape), C.c_int, C.POINTER(C.c_int),
                                C.POINTER(LatticeShape)]),
    "lbib_set_lattice_shape": (C.c_int, [_P, C.POINTER(LatticeShape)]),
    "lbib_domain_concentrations": (C.c_int, [_P, C.c_int, C.c_int, _D]),
    "lbib_domain_concentrations_noalloc": (C.c_int, [_P, C.c_int, C.c_int, C.c_void_p]),
    "lbib_capture_begin": (C.c_int, [_P]),
    "lbib_capture_end": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "lbib_graph_launch": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_int)]),
    "lbib_graph_reset": (C.c_int, [_P]),
}


def load(path: str = LIB_PATH):
    """Load liblbib_hip.so and bind every declared entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with __graft_entry__.build() (no CPU fallback exists)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def i32ptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I32)


def u32ptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_U32)


class LbibError(RuntimeError):
    pass


def make_rects(rects):
    """(x0, y0, w, h, tau_fluid, taus) tuples -> ctypes array of lbib_rect."""
    rects = list(rects)
    arr = (Rect * max(len(rects), 1))()
    for k, r in enumerate(rects):
        x0, y0, w, h, tf = r[:5]
        taus = list(r[5]) if len(r) > 5 else []
        arr[k].x0, arr[k].y0, arr[k].w, arr[k].h, arr[k].tau_fluid = int(x0), int(y0), int(w), int(h), float(tf)
        for j in range(MAX_SPECIES):
            arr[k].tau_species[j] = float(taus[j]) if j < len(taus) else 1.0
    return arr if rects else (Rect * 0)()


def species_spec(kind=SPECIES_PLAIN, tau=1.0, production=1e-4, decay=1e-5, init_inside=1.0,
                 switchoff=5000, prod_domain=1, source_inside=0.0, krow=(0.0, 0.0, 0.0, 0.0)):
    s = SpeciesSpec()
    s.kind = kind
    s.tau = tau
    s.production = production
    s.decay = decay
    s.init_inside = init_inside
    s.switchoff = switchoff
    s.prod_domain = prod_domain
    s.source_inside = source_inside
    for j in range(MAX_SPECIES):
        s.krow[j] = krow[j] if j < len(krow) else 0.0
    return s


class Context:
    """Thin owner of one lbib_ctx (one HIP stream, all device memory)."""

    def __init__(self, nx, ny, tau_fluid, species=(), force_model=FORCE_LUO, precision=FP64,
                 box_outlet=False, max_markers=0, max_connections=0):
        self.lib = load()
        cfg = Config()
        cfg.nx, cfg.ny, cfg.tau_fluid = nx, ny, tau_fluid
        cfg.force_model, cfg.precision = force_model, precision
        cfg.nspecies = len(species)
        for k, s in enumerate(species):
            cfg.species[k] = s
        cfg.box_outlet = 1 if box_outlet else 0
        cfg.max_markers, cfg.max_connections = max_markers, max_connections
        h = C.c_void_p()
        rc = self.lib.lbib_create(C.byref(cfg), C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.lib.lbib_last_error(h).decode() if h.value else "lbib_create failed"
            if h.value:
                self.lib.lbib_destroy(h)
            self.h = None
            raise LbibError(msg)
        self.nx, self.ny, self.n, self.ns = nx, ny, nx * ny, len(species)

    def close(self):
        if self.h is not None:
            self.lib.lbib_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise LbibError(self.lib.lbib_last_error(self.h).decode())

    # lattice state
    def init_solvers(self, u0=None):
        self._chk(self.lib.lbib_init_solvers(self.h, dptr(u0) if u0 is not None else None))

    def set_populations(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        self._chk(self.lib.lbib_set_populations(self.h, dptr(f)))

    def get_populations(self):
        f = np.empty((9, self.ny, self.nx))
        self._chk(self.lib.lbib_get_populations(self.h, dptr(f)))
        return f

    def set_species(self, s, g):
        g = np.ascontiguousarray(g, dtype=np.float64)
        self._chk(self.lib.lbib_set_species(self.h, s, dptr(g)))

    def get_species(self, s):
        g = np.empty((5, self.ny, self.nx))
        self._chk(self.lib.lbib_get_species(self.h, s, dptr(g)))
        return g

    def get_velocity(self):
        ux = np.empty((self.ny, self.nx))
        uy = np.empty((self.ny, self.nx))
        self._chk(self.lib.lbib_get_velocity(self.h, dptr(ux), dptr(uy)))
        return ux, uy

    def get_domains(self):
        d = np.empty((self.ny, self.nx), dtype=np.uint32)
        self._chk(self.lib.lbib_get_domains(self.h, u32ptr(d)))
        return d

    def set_domains(self, d):
        d = np.ascontiguousarray(d, dtype=np.uint32)
        self._chk(self.lib.lbib_set_domains(self.h, u32ptr(d)))

    def get_forces(self, nm):
        """The device force table of nm markers as CSR by dense marker index:
        (off[nm + 1], type[nf], partner[nf], prm[nf, 4])."""
        off = np.empty(nm + 1, dtype=np.int32)
        self._chk(self.lib.lbib_get_forces(self.h, i32ptr(off), None, None, None))
        nf = int(off[nm])
        t = np.empty(max(nf, 1), dtype=np.int32)
        q = np.empty(max(nf, 1), dtype=np.int32)
        prm = np.empty((max(nf, 1), 4), dtype=np.float64)
        self._chk(self.lib.lbib_get_forces(self.h, i32ptr(off), i32ptr(t), i32ptr(q), dptr(prm)))
        return off, t[:nf], q[:nf], prm[:nf]

    def forces_junctions(self, radius=1.0, k=0.02, l0=0.5):
        """deleteForceType(7) and the cell-junction springs (lbib_forces_junctions)."""
        self._chk(self.lib.lbib_forces_junctions(self.h, float(radius), float(k), float(l0)))

    def set_force_field(self, fx, fy):
        fx = np.ascontiguousarray(fx, dtype=np.float64)
        fy = np.ascontiguousarray(fy, dtype=np.float64)
        self._chk(self.lib.lbib_set_force_field(self.h, dptr(fx), dptr(fy)))

    def get_force_field(self):
        fx = np.empty((self.ny, self.nx))
        fy = np.empty((self.ny, self.nx))
        self._chk(self.lib.lbib_get_force_field(self.h, dptr(fx), dptr(fy)))
        return fx, fy

    def set_mass_field(self, m):
        m = np.ascontiguousarray(m, dtype=np.float64)
        self._chk(self.lib.lbib_set_mass_field(self.h, dptr(m)))

    def step_lattice_only(self, it):
        self._chk(self.lib.lbib_step_lattice_only(self.h, it))

    def synchronize(self):
        self._chk(self.lib.lbib_synchronize(self.h))

    def read_status(self):
        s = Status()
        self._chk(self.lib.lbib_read_status(self.h, C.byref(s)))
        return s

    def step_begin(self, it):
        self._chk(self.lib.lbib_step_begin(self.h, it))

    def rewind(self, it):
        self._chk(self.lib.lbib_rewind(self.h, it))

    def profile_enable(self, on=True, every=1):
        """Time every `every`-th fused lattice launch with HIP events."""
        self._chk(self.lib.lbib_profile_enable(self.h, max(int(every), 1) if on else 0))

    def profile_read(self):
        """(summed lattice-kernel ms, launches, algorithmic bytes per launch)"""
        ms = C.c_double()
        n = C.c_int64()
        b = C.c_double()
        self._chk(self.lib.lbib_profile_read(self.h, C.byref(ms), C.byref(n), C.byref(b)))
        return ms.value, n.value, b.value

    def profile_copy(self, nbytes=1 << 30, reps=10):
        """Average ms of one plain copy of nbytes (read + write) on the context stream."""
        ms = C.c_double()
        self._chk(self.lib.lbib_profile_copy(self.h, int(nbytes), int(reps), C.byref(ms)))
        return ms.value

    def device_bytes(self):
        return int(self.lib.lbib_device_bytes(self.h))

    def domain_concentrations(self, s, ndom):
        """Sum of species s's concentration per domain 0..ndom-1 (x from 1, as
        GeometryHandler::computeAccumulatedDomainConcentrations)."""
        out = np.zeros(ndom)
        self._chk(self.lib.lbib_domain_concentrations(self.h, int(s), int(ndom), out.ctypes.data_as(_D)))
        return out

    def set_batch_layout(self, rects):
        """rects: iterable of (x0, y0, w, h, tau_fluid, [tau_species...])."""
        arr = make_rects(rects)
        self._chk(self.lib.lbib_set_batch_layout(self.h, len(arr), arr if len(arr) else None))

    def autotune(self, reps=5):
        """Times every applicable lattice launch shape on the current state and
        keeps the fastest; returns (chosen, [(variant, wg, spt, ms), ...])."""
        tried = (LatticeShape * 32)()
        n = C.c_int()
        best = LatticeShape()
        self._chk(self.lib.lbib_autotune(self.h, int(reps), tried, 32, C.byref(n), C.byref(best)))
        rows = [(t.variant, t.workgroup, t.sites_per_lane, t.ms) for t in tried[: n.value]]
        return (best.variant, best.workgroup, best.sites_per_lane, best.ms), rows

    def set_lattice_shape(self, variant, workgroup, sites_per_lane=1):
        s = LatticeShape(variant, workgroup, sites_per_lane, 0.0)
        self._chk(self.lib.lbib_set_lattice_shape(self.h, C.byref(s)))
