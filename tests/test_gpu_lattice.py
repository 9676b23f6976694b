"""Device D2Q9 + D2Q5 step (through the C ABI) against the oracle on the same
seeded inputs: prescribed force and mass fields, domain map, every species
kind, box outlet.  FP64 bar: 1e-12 relative on populations (the kernels are
written to round identically, so the observed difference is expected to be 0)."""
import numpy as np
import pytest

from lbibcell_amd import capi

pytestmark = pytest.mark.gpu

W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _species_sets():
    return {
        "none": [],
        "plain": [(capi.SPECIES_PLAIN, 0.8, {})],
        "signal": [(capi.SPECIES_SIGNAL, 0.8, dict(production=1e-3, decay=1e-5, switchoff=7,
                                                    init_inside=0.5, prod_domain=1))],
        "linear3": [(capi.SPECIES_LINEAR, 0.7, dict(source_inside=1e-4, init_inside=0.2,
                                                     krow=(-1e-3, 2e-4, 0.0))),
                    (capi.SPECIES_LINEAR, 0.9, dict(source_inside=0.0, init_inside=0.1,
                                                     krow=(1e-4, -2e-3, 3e-4))),
                    (capi.SPECIES_PLAIN, 1.1, {})],
    }


def _oracle_params(kind, prm, ns):
    if kind == capi.SPECIES_SIGNAL:
        return [prm["production"], prm["decay"], prm["switchoff"], prm["init_inside"], prm["prod_domain"]]
    if kind == capi.SPECIES_LINEAR:
        k = list(prm["krow"]) + [0.0] * (ns - len(prm["krow"]))
        return [prm["source_inside"], prm["init_inside"]] + k[:ns]
    return []


@pytest.mark.parametrize("force_model", [capi.FORCE_LUO, capi.FORCE_GUO_AS_WRITTEN, capi.FORCE_GUO])
@pytest.mark.parametrize("spset", ["none", "plain", "signal", "linear3"])
@pytest.mark.parametrize("shape", [(37, 23), (300, 17)])
def test_lattice_step_matches_oracle(orc, force_model, spset, shape):
    nx, ny = shape
    tau = 0.83
    rng = np.random.default_rng(nx * 7 + ny + force_model)
    specs = _species_sets()[spset]
    ns = len(specs)
    print(f"case nx={nx} ny={ny} fm={force_model} species={spset}", flush=True)
    ctx = capi.Context(nx, ny, tau, species=[capi.species_spec(k, t, **p) for k, t, p in specs],
                       force_model=force_model, box_outlet=True)
    o = orc.Oracle(nx, ny, tau, force_model)
    for k, t, p in specs:
        o.add_species(k, t, _oracle_params(k, p, ns))
    o.box_outlet(True)
    dom = np.zeros((ny, nx), dtype=np.uint32)
    dom[ny // 4: 3 * ny // 4, nx // 3: 2 * nx // 3] = 1
    dom[1:3, 1:4] = 2
    ctx.set_domains(dom)
    o.set_domains(dom)
    ctx.init_solvers()
    o.init_solvers()
    np.testing.assert_array_equal(ctx.get_populations(), o.get_f())
    for s in range(ns):
        np.testing.assert_array_equal(ctx.get_species(s), o.get_g(s))
    f = W[:, None, None] * (1 + 0.05 * rng.standard_normal((9, ny, nx)))
    ctx.set_populations(f)
    o.set_f(f)
    for s in range(ns):
        g = 0.1 * (1 + 0.05 * rng.standard_normal((5, ny, nx)))
        ctx.set_species(s, g)
        o.set_g(s, g)
    Fx = 1e-4 * rng.standard_normal((ny, nx))
    Fy = 1e-4 * rng.standard_normal((ny, nx))
    ctx.set_force_field(Fx, Fy)
    o.set_force_field(Fx, Fy)
    for it in range(12):
        if it % 5 == 0:
            m = 1e-3 * rng.standard_normal((ny, nx))
            ctx.set_mass_field(m)
            o.set_mass(m)
        ctx.step_lattice_only(it)
        o.set_iteration(it)
        o.lattice_step(1)
    fd, fo = ctx.get_populations(), o.get_f()
    assert _rel(fd, fo) <= 1e-12
    ud, vd = ctx.get_velocity()
    uo, vo = o.get_velocity()
    assert _rel(ud, uo) <= 1e-12 and _rel(vd, vo) <= 1e-12
    for s in range(ns):
        assert _rel(ctx.get_species(s), o.get_g(s)) <= 1e-12
    # the FP64 kernels round like the oracle: expect exact equality
    np.testing.assert_array_equal(fd, fo)
    np.testing.assert_array_equal(ud, uo)
    for s in range(ns):
        np.testing.assert_array_equal(ctx.get_species(s), o.get_g(s))
    ctx.close()


def test_init_with_velocity_matches_oracle(orc):
    nx, ny = 40, 30
    rng = np.random.default_rng(11)
    u0 = np.ascontiguousarray(1e-3 * rng.standard_normal((ny * nx * 2)))
    ctx = capi.Context(nx, ny, 1.3, species=[capi.species_spec(capi.SPECIES_PLAIN, 0.9)])
    o = orc.Oracle(nx, ny, 1.3)
    o.add_species(orc.SP_PLAIN, 0.9)
    ctx.init_solvers(u0)
    o.init_solvers(u0)
    np.testing.assert_array_equal(ctx.get_populations(), o.get_f())
    np.testing.assert_array_equal(ctx.get_species(0), o.get_g(0))
    ctx.close()


def test_profile_hooks_count_launches_and_time_a_copy():
    ctx = capi.Context(64, 48, 1.0, species=[capi.species_spec(capi.SPECIES_PLAIN, 0.8)])
    try:
        ctx.init_solvers()
        ctx.profile_enable(True)
        ctx.profile_read()
        for it in range(3):
            ctx.step_lattice_only(it)
        ms, n, b = ctx.profile_read()
        assert n == 3 and ms > 0.0
        assert b == (2 * 9 + 2 * 5) * 8 * 64 * 48
        ctx.profile_enable(False)
        assert ctx.profile_copy(1 << 20, 2) > 0.0
        with pytest.raises(capi.LbibError, match="nbytes"):
            ctx.profile_copy(16, 1)
    finally:
        ctx.close()
