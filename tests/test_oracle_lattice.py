"""Pins the oracle's lattice arithmetic (D2Q9 + D2Q5) against what the
mathematics fixes: equilibrium fixed point, conservation laws, the textbook
BGK step with pull streaming, viscous decay of a shear wave, and the Gaussian
spreading of the reference's own DiracD2Q5 closed form
(LbmLib/src/solver/CDESolver/DiracD2Q5.cpp:40-50)."""
import numpy as np
import pytest

EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def pull(f):
    """Pull streaming, periodic: f_q(x) <- f_q(x - e_q)."""
    out = np.empty_like(f)
    for q in range(f.shape[0]):
        out[q] = np.roll(np.roll(f[q], EX[q], axis=1), EY[q], axis=0)
    return out


def textbook_step(f, tau, Fx, Fy, m):
    """BGK with Luo forcing 3 w_q rho0 (e_q . F) and mass injection rho=rho0+m."""
    g = pull(f)
    rho0 = g.sum(0)
    rho = rho0 + m
    ux = (g * EX[:, None, None]).sum(0) / rho
    uy = (g * EY[:, None, None]).sum(0) / rho
    out = np.empty_like(g)
    usq = ux * ux + uy * uy
    for q in range(9):
        eu = EX[q] * ux + EY[q] * uy
        feq = W[q] * rho * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * usq)
        out[q] = g[q] - g[q] / tau + feq / tau + 3 * W[q] * rho0 * (EX[q] * Fx + EY[q] * Fy)
    return out


def test_rest_state_is_a_fixed_point(orc):
    o = orc.Oracle(12, 10, 1.0)
    o.init_solvers()
    f0 = o.get_f()
    np.testing.assert_allclose(f0, W[:, None, None] * np.ones((9, 10, 12)), rtol=0, atol=1e-16)
    o.lattice_step(5)
    np.testing.assert_allclose(o.get_f(), f0, rtol=0, atol=4e-16)  # rounding of rho = sum w_q


def test_init_is_one_relaxation_from_rest_spike(orc):
    # FluidSolver.cpp:50-60: f = (1,0,...,0) relaxed once toward w_q
    tau = 1.7
    o = orc.Oracle(6, 5, tau)
    o.init_solvers()
    f = o.get_f()
    expect = W / tau
    expect[0] += 1 - 1 / tau
    np.testing.assert_allclose(f[:, 2, 3], expect, rtol=1e-15)


@pytest.mark.parametrize("tau", [0.6, 1.0, 1.37])
def test_matches_textbook_bgk_with_pull_streaming(orc, tau):
    rng = np.random.default_rng(7)
    nx, ny = 13, 9
    o = orc.Oracle(nx, ny, tau)
    o.init_solvers()
    f = W[:, None, None] * (1 + 0.05 * rng.standard_normal((9, ny, nx)))
    Fx = 1e-3 * rng.standard_normal((ny, nx))
    Fy = 1e-3 * rng.standard_normal((ny, nx))
    m = 1e-3 * rng.standard_normal((ny, nx))
    o.set_f(f)
    o.set_force_field(Fx, Fy)
    o.set_mass(m)
    o.lattice_step(1)
    ref = textbook_step(f, tau, Fx, Fy, m)
    np.testing.assert_allclose(o.get_f(), ref, rtol=1e-13, atol=1e-16)


def test_mass_and_momentum_balance(orc):
    rng = np.random.default_rng(3)
    nx, ny, tau = 16, 11, 0.8
    o = orc.Oracle(nx, ny, tau)
    o.init_solvers()
    f = W[:, None, None] * (1 + 0.02 * rng.standard_normal((9, ny, nx)))
    o.set_f(f)
    Fx = np.full((ny, nx), 2e-4)
    Fy = np.full((ny, nx), -1e-4)
    m = np.zeros((ny, nx))
    m[3, 4] = 0.01
    o.set_force_field(Fx, Fy)
    o.set_mass(m)
    mass0 = f.sum()
    px0 = (f * EX[:, None, None]).sum()
    py0 = (f * EY[:, None, None]).sum()
    o.lattice_step(1)
    g = o.get_f()
    # collide adds m/tau of mass (FluidSolver.cpp:106-110: rho = rho0 + m only in feq)
    assert abs(g.sum() - (mass0 + 0.01 / tau)) < 1e-12
    rho0 = pull(f).sum(0)
    # momentum: sum feq(rho, u) moments = rho*u = momentum, plus Luo rho0*F
    assert abs((g * EX[:, None, None]).sum() - (px0 + (rho0 * Fx).sum())) < 1e-12
    assert abs((g * EY[:, None, None]).sum() - (py0 + (rho0 * Fy).sum())) < 1e-12


def test_shear_wave_decay_rate(orc):
    nx, ny, tau = 4, 64, 0.9
    o = orc.Oracle(nx, ny, tau)
    y = np.arange(ny)
    k = 2 * np.pi / ny
    U = 1e-4
    ux = U * np.sin(k * y)[:, None] * np.ones((ny, nx))
    u0 = np.zeros((ny, nx, 2))
    u0[..., 0] = ux
    o.init_solvers(np.ascontiguousarray(u0.reshape(-1)))
    steps = 400
    o.lattice_step(steps)
    vx, _ = o.get_velocity()
    amp = 2 * (vx[:, 0] * np.sin(k * y)).sum() / ny
    nu = (tau - 0.5) / 3
    expect = U * np.exp(-nu * k * k * steps)
    assert abs(amp / expect - 1) < 0.01


def test_d2q5_gaussian_spreading_matches_dirac_closed_form(orc):
    # DiracD2Q5.cpp:40-50: D = 0.04, t0 = 50, initial C = exp(-r^2/4Dt)/(4 pi D t)
    D, t0, pos = 0.04, 50.0, 32.0
    tau_c = 3 * D + 0.5
    nx = ny = 64
    o = orc.Oracle(nx, ny, 1.0)
    o.add_species(orc.SP_PLAIN, tau_c)
    o.init_solvers()
    yy, xx = np.mgrid[0:ny, 0:nx]
    G = lambda t: np.exp(-((xx - pos) ** 2 + (yy - pos) ** 2) / (4 * D * t)) / (4 * np.pi * D * t)
    g0 = np.repeat((G(t0) * 0.2)[None], 5, axis=0)
    o.set_g(0, g0)
    mass0 = g0.sum()
    steps = 200
    o.lattice_step(steps)
    C = o.get_g(0).sum(0)
    assert abs(C.sum() - mass0) < 1e-12 * mass0
    ref = G(t0 + steps)
    err = np.linalg.norm(C - ref) / np.linalg.norm(ref)
    assert err < 0.03, err


def test_signal_species_source_and_decay(orc):
    # ..._SIGNAL.cpp:92-100: production inside domain 1 while it < switchoff, decay elsewhere
    nx, ny = 8, 6
    d = np.zeros((ny, nx), dtype=np.uint32)
    d[2:4, 3:6] = 1
    o = orc.Oracle(nx, ny, 1.0)
    o.add_species(orc.SP_SIGNAL, 0.8, [1e-3, 0.0, 3, 0.0, 1])
    o.set_domains(d)
    o.init_solvers()
    assert o.get_g(0).sum() == 0.0
    o.lattice_step(5)      # it = 0, 1, 2 produce; 3, 4 have zero decay
    np.testing.assert_allclose(o.get_g(0).sum(), 3 * 6 * 1e-3, rtol=1e-13)
    # pure decay: initial C = 1 inside cells, switched off from the start
    o = orc.Oracle(nx, ny, 1.0)
    o.add_species(orc.SP_SIGNAL, 0.8, [1e-3, 1e-5, 0, 1.0, 1])
    o.set_domains(d)
    o.init_solvers()
    c0 = o.get_g(0).sum()
    np.testing.assert_allclose(c0, 6.0, rtol=1e-15)
    o.lattice_step(4)
    np.testing.assert_allclose(o.get_g(0).sum(), c0 * (1 - 1e-5) ** 4, rtol=1e-12)


def test_box_outlet_sets_unit_density_on_edges(orc):
    rng = np.random.default_rng(5)
    nx, ny = 10, 7
    o = orc.Oracle(nx, ny, 1.0)
    o.add_species(orc.SP_PLAIN, 0.7)
    o.init_solvers()
    o.set_f(W[:, None, None] * (1 + 0.1 * rng.random((9, ny, nx))))
    o.box_outlet(True)
    o.lattice_step(1)
    rho = o.get_f().sum(0)
    np.testing.assert_allclose(rho[:, 0], 1.0, rtol=1e-14)
    np.testing.assert_allclose(rho[0, :], 1.0, rtol=1e-14)
    assert np.abs(rho[1:, 1:] - 1).max() > 1e-3


def _schnakenberg_pair(orc, nx, ny, dom, tau=0.9, init=1.0):
    # registry order is by name: L (index 0) before R (index 1)
    o = orc.Oracle(nx, ny, 1.0)
    o.add_species(orc.SP_SCHNAKENBERG_L, tau, [100.0, 0.9, 1e-4, 1, init])
    o.add_species(orc.SP_SCHNAKENBERG_R, tau, [100.0, 0.1, 1e-4, 0, init])
    o.set_domains(dom)
    o.init_solvers()
    return o


def test_schnakenberg_uniform_cell_follows_the_sequential_recurrence(orc):
    """tutorial_02_CDESolverD2Q5_{L,R}::reaction (…_L.cpp:94-106, …_R.cpp:94-105)
    on a lattice that is one cell: the fluid stays at rest, streaming moves
    nothing, and each step is the explicit map
        L' = L + dt g (b - R^2 L),  R' = R + dt g (a - R + R^2 L')
    where R reads L after L's collide (L is collided first at every site).
    Reading L before its collide (L instead of L') fails this check."""
    nx, ny = 6, 5
    o = _schnakenberg_pair(orc, nx, ny, np.ones((ny, nx), dtype=np.uint32))
    gam, a, b, dt = 100.0, 0.1, 0.9, 1e-4
    L = R = 1.0
    for _ in range(40):
        L = L + dt * gam * (b - R * R * L)
        R = R + dt * gam * (a - R + R * R * L)
    o.lattice_step(40)
    cL, cR = o.get_g(0).sum(axis=0), o.get_g(1).sum(axis=0)
    np.testing.assert_allclose(cL, L, rtol=1e-13)
    np.testing.assert_allclose(cR, R, rtol=1e-13)
    # the Jacobi variant (both read the other's pre-collision value) differs
    Lj = Rj = 1.0
    for _ in range(40):
        Lj, Rj = Lj + dt * gam * (b - Rj * Rj * Lj), Rj + dt * gam * (a - Rj + Rj * Rj * Lj)
    assert abs(Rj - R) > 1e-6


def test_schnakenberg_steady_state_and_no_reaction_outside_cells(orc):
    # homogeneous steady state R* = a + b, L* = b / (a + b)^2 is a fixed point
    nx, ny = 8, 6
    o = _schnakenberg_pair(orc, nx, ny, np.ones((ny, nx), dtype=np.uint32))
    Rs, Ls = 1.0, 0.9
    o.set_g(0, np.full((5, ny, nx), Ls / 5.0))
    o.set_g(1, np.full((5, ny, nx), Rs / 5.0))
    o.lattice_step(10)
    np.testing.assert_allclose(o.get_g(0).sum(axis=0), Ls, rtol=1e-14)
    np.testing.assert_allclose(o.get_g(1).sum(axis=0), Rs, rtol=1e-14)
    # outside cells alone: pure diffusion conserves each species exactly
    o = _schnakenberg_pair(orc, nx, ny, np.zeros((ny, nx), dtype=np.uint32))
    rng = np.random.default_rng(3)
    gL, gR = rng.uniform(0.1, 0.3, (5, ny, nx)), rng.uniform(0.1, 0.3, (5, ny, nx))
    o.set_g(0, gL)
    o.set_g(1, gR)
    o.lattice_step(6)
    np.testing.assert_allclose(o.get_g(0).sum(), gL.sum(), rtol=1e-13)
    np.testing.assert_allclose(o.get_g(1).sum(), gR.sum(), rtol=1e-13)
