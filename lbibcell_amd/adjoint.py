"""Reverse-mode gradients of the coupled D2Q9 + D2Q5 + IB step (C ABI
lbib_adj_*): the loss L = sum_x (C_T(x) - C*(x))^2 of the final species
concentration against a target, differentiated with respect to the species
relaxation time tau_c, the membrane stiffness kappa (type-6 tension,
ForceStructs.hpp:455-514) and one source rate per cell; with several species
(set_species), also the reaction matrix K, and with junction springs between
cells (set_junctions, ForceType7 of ForceStructs.hpp:521-588), their
stiffness.

`AdjointProblem` owns one device context; `CoupledLoss` is a thin
torch.autograd.Function over it, there to check the hand-written adjoint
against finite differences and FP64 references."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi

_P = C.c_void_p
_D = C.POINTER(C.c_double)


class AdjConfig(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("precision", C.c_int32),
        ("tau_fluid", C.c_double), ("decay", C.c_double),
        ("n_markers", C.c_int32), ("n_cells", C.c_int32), ("checkpoint_every", C.c_int32),
    ]


SIGNATURES = {
    "lbib_adj_create": (C.c_int, [C.POINTER(AdjConfig), C.POINTER(_P)]),
    "lbib_adj_destroy": (None, [_P]),
    "lbib_adj_last_error": (C.c_char_p, [_P]),
    "lbib_adj_set_problem": (C.c_int, [_P, _D, _D, _D, C.POINTER(C.c_int32), C.POINTER(C.c_uint32), _D]),
    "lbib_adj_forward": (C.c_int, [_P, C.c_int, C.c_double, C.c_double, _D, _D, _D]),
    "lbib_adj_gradient": (C.c_int, [_P, _D, _D, _D]),
    "lbib_adj_timing": (C.c_int, [_P, _D, _D]),
    "lbib_adj_set_species": (C.c_int, [_P, C.c_int, _D, _D, _D]),
    "lbib_adj_set_junctions": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int32), C.c_double]),
    "lbib_adj_forward_ms": (C.c_int, [_P, C.c_int, _D, C.c_double, C.c_double, _D, _D, _D, _D]),
    "lbib_adj_gradient_ms": (C.c_int, [_P, _D, _D, _D, _D, _D]),
    "lbib_adj_set_batch_layout": (C.c_int, [_P, C.c_int, C.POINTER(C.c_int32)]),
    "lbib_adj_set_storage": (C.c_int, [_P, C.c_int]),
    "lbib_adj_set_checkpointing": (C.c_int, [_P, C.c_int, C.c_int]),
    "lbib_adj_stats": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
}

# checkpointing schemes (lbib_adj_set_checkpointing)
CHECKPOINT_EVERY = 0
REVOLVE = 1

W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def _lib():
    lib = capi.load()
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    return lib


def _dp(a):
    return a.ctypes.data_as(_D)


class AdjointProblem:
    """A fixed problem (lattice, initial state, marker rings, cell map,
    target) whose loss and gradient are evaluated for any parameters.

    rings: list of marker-position arrays (m_r, 2), one closed ring per cell;
    dom: (ny, nx) cell index per site (0 outside, cells 1..ncell)."""

    def __init__(self, nx, ny, tau_fluid, rings, dom, target=None, f0=None, g0=None, decay=1e-5,
                 precision=capi.FP32, checkpoint_every=50, storage=None, revolve_snapshots=None):
        self.L = _lib()
        self.serial = 0   # counts forwards and gradients (the autograd wrappers check it)
        self.nx, self.ny = nx, ny
        self.ncell = int(np.max(dom)) if np.size(dom) else 0
        xy = np.concatenate([np.asarray(r, dtype=np.float64) for r in rings]) if rings else np.zeros((0, 2))
        nxt = []
        base = 0
        for r in rings:
            m = len(r)
            nxt += [base + (k + 1) % m for k in range(m)]
            base += m
        self.nm = base
        cfg = AdjConfig(nx, ny, precision, tau_fluid, decay, self.nm, self.ncell, checkpoint_every)
        self.h = _P()
        rc = self.L.lbib_adj_create(C.byref(cfg), C.byref(self.h))
        if rc != 0:
            msg = self.L.lbib_adj_last_error(self.h).decode() if self.h else "lbib_adj_create failed"
            self.close()
            raise capi.LbibError(msg)
        if storage is not None:
            self._chk(self.L.lbib_adj_set_storage(self.h, int(storage)))
        if revolve_snapshots is not None:
            self.set_checkpointing(REVOLVE, revolve_snapshots)
        n = nx * ny
        f0 = np.tile(W[:, None], (1, n)) if f0 is None else np.ascontiguousarray(f0, dtype=np.float64).reshape(9, n)
        g0 = np.zeros((5, n)) if g0 is None else np.ascontiguousarray(g0, dtype=np.float64).reshape(5, n)
        self._xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1)
        self._next = np.ascontiguousarray(nxt, dtype=np.int32)
        self._dom = np.ascontiguousarray(dom, dtype=np.uint32).reshape(-1)
        self._f0, self._g0 = np.ascontiguousarray(f0), np.ascontiguousarray(g0)
        self.set_target(np.zeros(n) if target is None else target)

    def _chk(self, rc):
        if rc != 0:
            raise capi.LbibError(self.L.lbib_adj_last_error(self.h).decode())

    def set_target(self, target):
        self._target = np.ascontiguousarray(target, dtype=np.float64).reshape(-1)
        self._chk(self.L.lbib_adj_set_problem(
            self.h, _dp(self._f0), _dp(self._g0), _dp(self._xy) if self.nm else None,
            self._next.ctypes.data_as(C.POINTER(C.c_int32)) if self.nm else None,
            self._dom.ctypes.data_as(C.POINTER(C.c_uint32)), _dp(self._target)))

    def forward(self, steps, tau_c, kappa, rates, want_field=False):
        """Loss (and the final concentration (ny, nx) when want_field)."""
        r = np.zeros(self.ncell + 1)
        r[1:] = np.asarray(rates, dtype=np.float64).reshape(-1)[: self.ncell]
        loss = C.c_double()
        field = np.empty(self.nx * self.ny) if want_field else None
        self.serial += 1
        self._chk(self.L.lbib_adj_forward(self.h, int(steps), float(tau_c), float(kappa), _dp(r), C.byref(loss),
                                          _dp(field) if want_field else None))
        return (loss.value, field.reshape(self.ny, self.nx)) if want_field else loss.value

    def gradient(self):
        """(dL/dtau_c, dL/dkappa, dL/drates[ncell]) of the last forward."""
        dt, dk = C.c_double(), C.c_double()
        dr = np.zeros(self.ncell + 1)
        self.serial += 1   # the adjoint consumes the forward's state
        self._chk(self.L.lbib_adj_gradient(self.h, C.byref(dt), C.byref(dk), _dp(dr)))
        return dt.value, dk.value, dr[1:].copy()

    def set_species(self, g0, target, decay):
        """Several species: g0 (ns, 5, ny, nx), target (ns, ny, nx), decay (ns)
        outside every cell; replaces the species given at construction."""
        g0 = np.ascontiguousarray(g0, dtype=np.float64)
        ns = g0.shape[0]
        self._gms = g0.reshape(-1)
        self._tms = np.ascontiguousarray(target, dtype=np.float64).reshape(-1)
        self._dec = np.ascontiguousarray(decay, dtype=np.float64).reshape(-1)
        if self._tms.size != ns * self.nx * self.ny or self._dec.size != ns:
            raise capi.LbibError("set_species: target and decay must cover every species")
        self._chk(self.L.lbib_adj_set_species(self.h, ns, _dp(self._gms), _dp(self._tms), _dp(self._dec)))
        self.ns = ns

    def set_junctions(self, pairs, l0=0.5):
        """Junction springs between marker pairs (P, 2) (global marker indices,
        rings concatenated in order), rest length l0."""
        self._pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1)
        self._chk(self.L.lbib_adj_set_junctions(self.h, self._pairs.size // 2,
                                                self._pairs.ctypes.data_as(C.POINTER(C.c_int32)), float(l0)))

    def set_batch_layout(self, rects):
        """A batch of independent periodic problems in this lattice: rects =
        [(x0, y0, w, h)] on the 16-site tile grid; shared parameters, summed
        loss."""
        self._rects = np.ascontiguousarray(rects, dtype=np.int32).reshape(-1)
        self._chk(self.L.lbib_adj_set_batch_layout(self.h, self._rects.size // 4,
                                                   self._rects.ctypes.data_as(C.POINTER(C.c_int32))))

    def forward_ms(self, steps, tau_c, kappa, k_junction=0.0, K=None, rates=None, want_field=False):
        """Loss with every parameter: tau_c (ns), kappa, k_junction, K (ns, ns),
        rates (ncell); the final concentrations (ns, ny, nx) when want_field."""
        ns = getattr(self, "ns", 1)
        tc = np.ascontiguousarray(np.broadcast_to(np.asarray(tau_c, dtype=np.float64), (ns,)))
        Km = None if K is None else np.ascontiguousarray(K, dtype=np.float64).reshape(-1)
        r = np.zeros(self.ncell + 1)
        if rates is not None:
            r[1:] = np.asarray(rates, dtype=np.float64).reshape(-1)[: self.ncell]
        loss = C.c_double()
        field = np.empty(ns * self.nx * self.ny) if want_field else None
        self.serial += 1
        self._chk(self.L.lbib_adj_forward_ms(self.h, int(steps), _dp(tc), float(kappa), float(k_junction),
                                             _dp(Km) if Km is not None else None, _dp(r), C.byref(loss),
                                             _dp(field) if want_field else None))
        return (loss.value, field.reshape(ns, self.ny, self.nx)) if want_field else loss.value

    def gradient_ms(self):
        """dict(tau_c (ns), kappa, k_junction, K (ns, ns), rates (ncell)) of the
        last forward."""
        ns = getattr(self, "ns", 1)
        dt = np.zeros(ns)
        dk, dj = C.c_double(), C.c_double()
        dK = np.zeros(ns * ns)
        dr = np.zeros(self.ncell + 1)
        self.serial += 1
        self._chk(self.L.lbib_adj_gradient_ms(self.h, _dp(dt), C.byref(dk), C.byref(dj), _dp(dK), _dp(dr)))
        return dict(tau_c=dt, kappa=dk.value, k_junction=dj.value, K=dK.reshape(ns, ns), rates=dr[1:].copy())

    def set_checkpointing(self, scheme, param):
        """CHECKPOINT_EVERY with K steps, or REVOLVE with that many snapshot
        slots (binomial checkpointing); the gradient does not depend on it."""
        self._chk(self.L.lbib_adj_set_checkpointing(self.h, int(scheme), int(param)))

    def stats(self):
        """dict(forward_steps, gradient_steps (recomputed steps included),
        store_bytes of the checkpoints) of the last forward / gradient"""
        f, g, b = C.c_int64(), C.c_int64(), C.c_int64()
        self._chk(self.L.lbib_adj_stats(self.h, C.byref(f), C.byref(g), C.byref(b)))
        return dict(forward_steps=f.value, gradient_steps=g.value, store_bytes=b.value)

    def timing(self):
        """(device ms of the last forward, of the last gradient)"""
        f, g = C.c_double(), C.c_double()
        self._chk(self.L.lbib_adj_timing(self.h, C.byref(f), C.byref(g)))
        return f.value, g.value

    def close(self):
        if getattr(self, "h", None):
            self.L.lbib_adj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _needs_gradient(ctx):
    return any(ctx.needs_input_grad)


def coupled_loss_function():
    """torch.autograd.Function over AdjointProblem (imported lazily so that
    the module does not need torch).  The forward runs the device forward
    only; backward runs the adjoint when a gradient is asked for (again
    running the forward first if the problem has seen another forward in the
    meantime) and returns the gradients on the inputs' device and dtype."""
    import torch

    class CoupledLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, tau_c, kappa, rates, problem, steps):
            params = (float(tau_c), float(kappa), rates.detach().cpu().numpy().astype(np.float64))
            loss = problem.forward(steps, *params)
            ctx.problem, ctx.steps, ctx.params = problem, steps, params
            ctx.serial = problem.serial
            ctx.like = (tau_c, kappa, rates)
            return torch.tensor(loss, dtype=torch.float64, device=rates.device)

        @staticmethod
        def backward(ctx, grad_out):
            if not _needs_gradient(ctx):
                return None, None, None, None, None
            p = ctx.problem
            if p.serial != ctx.serial:   # another forward or gradient ran since: redo this forward
                p.forward(ctx.steps, *ctx.params)
            dt, dk, dr = p.gradient()
            tc, kp, rt = ctx.like
            g = grad_out.to(torch.float64)
            out = (g * dt, g * dk, g * torch.from_numpy(dr).to(g.device))
            return (out[0].to(tc.dtype) if ctx.needs_input_grad[0] else None,
                    out[1].to(kp.dtype) if ctx.needs_input_grad[1] else None,
                    out[2].to(rt.dtype) if ctx.needs_input_grad[2] else None, None, None)

    return CoupledLoss


def multispecies_loss_function():
    """torch.autograd.Function over AdjointProblem.forward_ms / gradient_ms:
    apply(tau_c (ns), kappa, k_junction, K (ns, ns), rates (ncell), problem,
    steps) -> loss."""
    import torch

    class MultiSpeciesLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, tau_c, kappa, k_junction, K, rates, problem, steps):
            params = dict(tau_c=tau_c.detach().cpu().numpy().astype(np.float64), kappa=float(kappa),
                          k_junction=float(k_junction), K=K.detach().cpu().numpy().astype(np.float64),
                          rates=rates.detach().cpu().numpy().astype(np.float64))
            loss = problem.forward_ms(steps, **params)
            ctx.problem, ctx.steps, ctx.params = problem, steps, params
            ctx.serial = problem.serial
            ctx.like = (tau_c, kappa, k_junction, K, rates)
            return torch.tensor(loss, dtype=torch.float64, device=rates.device)

        @staticmethod
        def backward(ctx, grad_out):
            if not _needs_gradient(ctx):
                return (None,) * 7
            p = ctx.problem
            if p.serial != ctx.serial:
                p.forward_ms(ctx.steps, **ctx.params)
            gr = p.gradient_ms()
            g = grad_out.to(torch.float64)
            vals = [gr["tau_c"], gr["kappa"], gr["k_junction"], gr["K"], gr["rates"]]
            out = []
            for need, like, v in zip(ctx.needs_input_grad[:5], ctx.like, vals):
                if not need:
                    out.append(None)
                    continue
                t = torch.as_tensor(np.asarray(v, dtype=np.float64), device=g.device) * g
                out.append(t.to(like.dtype).reshape(like.shape))
            return (*out, None, None)

    return MultiSpeciesLoss
