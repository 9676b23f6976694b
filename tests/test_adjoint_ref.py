"""The FP64 torch restatement used as the adjoint's gradient reference,
pinned on the CPU: without markers its step is the oracle's lattice step;
the closed-ring tension is internal (zero net force) and the interpolation
of a uniform velocity returns that velocity times the stencil weight sum."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import adjoint_ref  # noqa: E402

W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def test_step_without_markers_is_the_oracle_lattice_step(orc):
    nx, ny, tau_f, tau_c, prod, decay = 12, 10, 0.9, 0.8, 1e-3, 1e-3
    rng = np.random.default_rng(4)
    dom = np.zeros((ny, nx), dtype=np.uint32)
    dom[2:6, 3:8] = 1
    f0 = W[:, None, None] * (1 + 0.05 * rng.standard_normal((9, ny, nx)))
    g0 = 0.1 * (1 + 0.1 * rng.standard_normal((5, ny, nx)))
    o = orc.Oracle(nx, ny, tau_f)
    o.add_species(orc.SP_SIGNAL, tau_c, [prod, decay, 10 ** 6, 0.0, -1])
    o.set_domains(dom)
    o.init_solvers()
    o.set_f(f0)
    o.set_g(0, g0)
    f, g = torch.tensor(f0), torch.tensor(g0)
    X = torch.zeros((0, 2), dtype=torch.float64)
    e = torch.zeros(0, dtype=torch.long)
    rates = torch.tensor([0.0, prod], dtype=torch.float64)
    for it in range(5):
        o.set_iteration(it)
        o.lattice_step(1)
        f, g, X = adjoint_ref.step(f, g, X, e, e, torch.tensor(dom.astype(np.int64)), tau_f, tau_c, 0.0, rates,
                                   decay)
    np.testing.assert_allclose(g.sum(0).numpy(), o.get_g(0).sum(axis=0), rtol=0, atol=1e-14)
    np.testing.assert_allclose(f.sum(0).numpy(), o.get_f().sum(axis=0), rtol=0, atol=1e-14)


def test_ring_tension_is_internal_and_interpolation_weights_sum_near_one():
    m = 24
    th = 2 * np.pi * np.arange(m) / m
    X = torch.tensor(np.stack([15.3 + 5 * np.cos(th), 14.7 + 5 * np.sin(th)], 1))
    nxt = torch.tensor([(k + 1) % m for k in range(m)])
    prv = torch.tensor([(k - 1) % m for k in range(m)])
    kappa = 0.03
    dn, dp = X[nxt] - X, X[prv] - X
    Fm = kappa * dn / dn.norm(dim=1, keepdim=True) + kappa * dp / dp.norm(dim=1, keepdim=True)
    assert float(Fm.sum(0).abs().max()) < 1e-15
    # inward: a regular polygon's net tension per node is 2 kappa sin(pi / m) toward the centre
    np.testing.assert_allclose(Fm.norm(dim=1).numpy(), 2 * kappa * np.sin(np.pi / m), rtol=1e-12)
    w, site = adjoint_ref._weights(X, 32, 32)
    # the fastCos delta integrates to 1 within its approximation error
    assert float((w.sum((1, 2)) - 1).abs().max()) < 5e-3


def test_multispecies_step_reduces_to_the_single_species_step():
    """step_ms with one species, K = 0 and no junction pairs is step; with
    several species and K = 0 the species evolve independently."""
    nx, ny, tau_f = 14, 12, 0.9
    rng = np.random.default_rng(8)
    dom = np.zeros((ny, nx), dtype=np.int64)
    dom[3:8, 4:10] = 1
    f0 = torch.tensor(W[:, None, None] * (1 + 0.05 * rng.standard_normal((9, ny, nx))))
    g0 = torch.tensor(0.1 * (1 + 0.1 * rng.standard_normal((2, 5, ny, nx))))
    m = 16
    th = 2 * np.pi * np.arange(m) / m
    X = torch.tensor(np.stack([7.2 + 2.5 * np.cos(th), 5.9 + 2.5 * np.sin(th)], 1))
    nxt = torch.tensor([(k + 1) % m for k in range(m)])
    prv = torch.tensor([(k - 1) % m for k in range(m)])
    none = torch.zeros((0, 2), dtype=torch.long)
    rates = torch.tensor([0.0, 1e-3], dtype=torch.float64)
    dt = torch.tensor(dom)
    f1, g1, X1 = adjoint_ref.step(f0, g0[0], X, nxt, prv, dt, tau_f, 0.8, 0.01, rates, 1e-3)
    f2, g2, X2 = adjoint_ref.step_ms(f0, g0[:1], X, nxt, prv, none, dt, tau_f, [0.8], 0.01, 0.0, 0.5,
                                     torch.zeros((1, 1), dtype=torch.float64), rates, [1e-3])
    assert torch.equal(f1, f2) and torch.equal(g1, g2[0]) and torch.equal(X1, X2)
    # species 1 with zero rate inside: its own decay outside only
    _, g3, _ = adjoint_ref.step_ms(f0, g0, X, nxt, prv, none, dt, tau_f, [0.8, 0.8], 0.01, 0.0, 0.5,
                                   torch.zeros((2, 2), dtype=torch.float64), rates, [1e-3, 1e-3])
    assert torch.equal(g3[0], g2[0])
    r0 = torch.zeros(2, dtype=torch.float64)
    _, g4, _ = adjoint_ref.step(f0, g0[1], X, nxt, prv, dt, tau_f, 0.8, 0.01, r0, 1e-3)
    assert torch.equal(g3[1], g4)

