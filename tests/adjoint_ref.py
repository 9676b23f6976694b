"""FP64 restatement of the differentiable coupled step in torch (CPU), used as
the gradient reference of the hand-written adjoint: torch.autograd
differentiates this forward, written independently of the HIP kernels with
the same model (type-6 tension on closed rings, fastCos-delta spreading and
interpolation, pull streaming, BGK with Luo forcing and no mass source, D2Q5
on the post-collision velocity with rate s[c] inside cell c and -decay*C
outside, X += U).  step_ms / loss_ms add several species coupled by a linear
reaction matrix and Hookean junction springs between marker pairs."""
import math

import torch

HALFPI = math.pi / 2.0
EX = [0, 1, 0, -1, 0, 1, -1, -1, 1]
EY = [0, 0, 1, 0, -1, 1, 1, -1, -1]
WQ = [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4


def fast_cos(x):
    """GeometryNode.cpp:55-81"""
    t = x + HALFPI
    t = torch.where(t > math.pi, t - 2.0 * math.pi, t)
    c1 = torch.where(t < 0, 1.27323954 * t + 0.405284735 * t * t, 1.27323954 * t - 0.405284735 * t * t)
    return torch.where(c1 < 0, 0.225 * (c1 * -c1 - c1) + c1, 0.225 * (c1 * c1 - c1) + c1)


def _stencil(X):
    s0 = torch.floor(X.detach()).long() - 1
    q = torch.arange(4)
    pos = (s0[:, None] + q[None, :]).to(torch.float64)
    return s0, 1.0 + fast_cos(HALFPI * (pos - X[:, None]))


def _weights(X, nx, ny):
    sx, hx = _stencil(X[:, 0])
    sy, hy = _stencil(X[:, 1])
    w = hx[:, :, None] * hy[:, None, :] / 16.0                      # [marker, qx, qy]
    q = torch.arange(4)
    ix = (sx[:, None] + q[None, :]) % nx
    iy = (sy[:, None] + q[None, :]) % ny
    site = iy[:, None, :] * nx + ix[:, :, None]                      # [marker, qx, qy]
    return w, site


def step(f, g, X, nxt, prv, dom, tau_f, tau_c, kappa, rates_full, decay):
    """One step; f (9, ny, nx), g (5, ny, nx), X (m, 2)."""
    ny, nx = f.shape[1:]
    n = nx * ny
    Fs = torch.zeros(2, n, dtype=torch.float64)
    if X.shape[0]:
        dn = X[nxt] - X
        dp = X[prv] - X
        Fm = kappa * dn / torch.sqrt((dn * dn).sum(1, keepdim=True)) + \
            kappa * dp / torch.sqrt((dp * dp).sum(1, keepdim=True))
        w, site = _weights(X, nx, ny)
        contrib = w[..., None] * Fm[:, None, None, :]
        Fs = Fs.index_add(1, site.reshape(-1), contrib.reshape(-1, 2).T)
    Fx, Fy = Fs[0].reshape(ny, nx), Fs[1].reshape(ny, nx)
    fh = torch.stack([torch.roll(f[q], shifts=(EY[q], EX[q]), dims=(0, 1)) for q in range(9)])
    gh = torch.stack([torch.roll(g[q], shifts=(EY[q], EX[q]), dims=(0, 1)) for q in range(5)])
    rho = fh.sum(0)
    ux = sum(EX[q] * fh[q] for q in range(9)) / rho
    uy = sum(EY[q] * fh[q] for q in range(9)) / rho
    u2 = ux * ux + uy * uy
    fn = []
    for q in range(9):
        eu = EX[q] * ux + EY[q] * uy
        feq = WQ[q] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * u2)
        fc = 3.0 * WQ[q] * rho * (EX[q] * Fx + EY[q] * Fy)
        fn.append(fh[q] - fh[q] / tau_f + feq / tau_f + fc)
    fn = torch.stack(fn)
    uox = sum(EX[q] * fn[q] for q in range(9)) / rho
    uoy = sum(EY[q] * fn[q] for q in range(9)) / rho
    C = gh.sum(0)
    R = torch.where(dom > 0, rates_full[dom], -decay * C)
    gn = []
    for q in range(5):
        geq = C / 3.0 if q == 0 else C / 6.0 * (1.0 + 3.0 * (EX[q] * uox + EY[q] * uoy))
        gn.append(gh[q] - gh[q] / tau_c + geq / tau_c + R * (1.0 / 3.0 if q == 0 else 1.0 / 6.0))
    gn = torch.stack(gn)
    if X.shape[0]:
        w, site = _weights(X, nx, ny)
        ur = torch.stack([uox.reshape(-1), uoy.reshape(-1)], 1)     # (n, 2)
        U = (w[..., None] * ur[site]).sum((1, 2))
        X = X + U
    return fn, gn, X


def loss(problem, steps, tau_c, kappa, rates):
    """problem: dict(nx, ny, tau_f, decay, f0 (9,ny,nx), g0 (5,ny,nx), xy (m,2),
    nxt, prv, dom (ny,nx) long, target (ny,nx)); parameters are tensors."""
    f = torch.as_tensor(problem["f0"], dtype=torch.float64)
    g = torch.as_tensor(problem["g0"], dtype=torch.float64)
    X = torch.as_tensor(problem["xy"], dtype=torch.float64)
    nxt = torch.as_tensor(problem["nxt"], dtype=torch.long)
    prv = torch.as_tensor(problem["prv"], dtype=torch.long)
    dom = torch.as_tensor(problem["dom"], dtype=torch.long)
    rates_full = torch.cat([torch.zeros(1, dtype=torch.float64), rates])
    for _ in range(steps):
        f, g, X = step(f, g, X, nxt, prv, dom, problem["tau_f"], tau_c, kappa, rates_full, problem["decay"])
    C = g.sum(0)
    return ((C - torch.as_tensor(problem["target"], dtype=torch.float64)) ** 2).sum()


def step_ms(f, g, X, nxt, prv, pairs, dom, tau_f, tau_c, kappa, kj, l0, K, rates_full, decay):
    """One step with ns species g (ns, 5, ny, nx): R_s = sum_j K_sj C_j, plus
    rates[c] inside cell c for species 0, plus -decay_s C_s outside cells;
    junction springs kj (|d| - l0) d/|d| on pairs (P, 2), d = X_b - X_a."""
    ny, nx = f.shape[1:]
    n = nx * ny
    ns = g.shape[0]
    Fs = torch.zeros(2, n, dtype=torch.float64)
    if X.shape[0]:
        dn = X[nxt] - X
        dp = X[prv] - X
        Fm = kappa * dn / torch.sqrt((dn * dn).sum(1, keepdim=True)) + \
            kappa * dp / torch.sqrt((dp * dp).sum(1, keepdim=True))
        if pairs.shape[0]:
            d = X[pairs[:, 1]] - X[pairs[:, 0]]
            dist = torch.sqrt((d * d).sum(1, keepdim=True))
            fj = kj * (dist - l0) * d / dist
            Fm = Fm.index_add(0, pairs[:, 0], fj).index_add(0, pairs[:, 1], -fj)
        w, site = _weights(X, nx, ny)
        contrib = w[..., None] * Fm[:, None, None, :]
        Fs = Fs.index_add(1, site.reshape(-1), contrib.reshape(-1, 2).T)
    Fx, Fy = Fs[0].reshape(ny, nx), Fs[1].reshape(ny, nx)
    fh = torch.stack([torch.roll(f[q], shifts=(EY[q], EX[q]), dims=(0, 1)) for q in range(9)])
    gh = torch.stack([torch.stack([torch.roll(g[s, q], shifts=(EY[q], EX[q]), dims=(0, 1)) for q in range(5)])
                      for s in range(ns)])
    rho = fh.sum(0)
    ux = sum(EX[q] * fh[q] for q in range(9)) / rho
    uy = sum(EY[q] * fh[q] for q in range(9)) / rho
    u2 = ux * ux + uy * uy
    fn = []
    for q in range(9):
        eu = EX[q] * ux + EY[q] * uy
        feq = WQ[q] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * u2)
        fc = 3.0 * WQ[q] * rho * (EX[q] * Fx + EY[q] * Fy)
        fn.append(fh[q] - fh[q] / tau_f + feq / tau_f + fc)
    fn = torch.stack(fn)
    uox = sum(EX[q] * fn[q] for q in range(9)) / rho
    uoy = sum(EY[q] * fn[q] for q in range(9)) / rho
    Cs = gh.sum(1)                                                   # (ns, ny, nx)
    inside = dom > 0
    gn = []
    for s in range(ns):
        R = sum(K[s, j] * Cs[j] for j in range(ns))
        own = rates_full[dom] if s == 0 else torch.zeros_like(Cs[s])
        R = R + torch.where(inside, own, -decay[s] * Cs[s])
        gs = []
        for q in range(5):
            geq = Cs[s] / 3.0 if q == 0 else Cs[s] / 6.0 * (1.0 + 3.0 * (EX[q] * uox + EY[q] * uoy))
            gs.append(gh[s, q] - gh[s, q] / tau_c[s] + geq / tau_c[s] + R * (1.0 / 3.0 if q == 0 else 1.0 / 6.0))
        gn.append(torch.stack(gs))
    gn = torch.stack(gn)
    if X.shape[0]:
        w, site = _weights(X, nx, ny)
        ur = torch.stack([uox.reshape(-1), uoy.reshape(-1)], 1)
        U = (w[..., None] * ur[site]).sum((1, 2))
        X = X + U
    return fn, gn, X


def loss_ms(problem, steps, tau_c, kappa, kj, K, rates):
    """problem: dict(nx, ny, tau_f, decay (ns), l0, f0, g0 (ns,5,ny,nx), xy, nxt,
    prv, pairs (P,2), dom, target (ns,ny,nx)); parameters are tensors."""
    f = torch.as_tensor(problem["f0"], dtype=torch.float64)
    g = torch.as_tensor(problem["g0"], dtype=torch.float64)
    X = torch.as_tensor(problem["xy"], dtype=torch.float64)
    nxt = torch.as_tensor(problem["nxt"], dtype=torch.long)
    prv = torch.as_tensor(problem["prv"], dtype=torch.long)
    pairs = torch.as_tensor(problem["pairs"], dtype=torch.long).reshape(-1, 2)
    dom = torch.as_tensor(problem["dom"], dtype=torch.long)
    rates_full = torch.cat([torch.zeros(1, dtype=torch.float64), rates])
    for _ in range(steps):
        f, g, X = step_ms(f, g, X, nxt, prv, pairs, dom, problem["tau_f"], tau_c, kappa, kj, problem["l0"], K,
                          rates_full, problem["decay"])
    C = g.sum(1)
    return ((C - torch.as_tensor(problem["target"], dtype=torch.float64)) ** 2).sum()
