"""Vector lattice kernel (launch shape variant 2: V consecutive sites per
lane, 8- or 16-byte plane accesses, shifted pulls through wavefront
shuffles) against the one-site-per-lane kernel (variant 0, spt 1); for the
moments-and-remainders storage (fp16r), variant 2 is the kernel that shares
each loaded moment with the neighbouring lanes by shuffles.

Both kernels evaluate the same FP32 expressions in the same order on the
same decoded values, so every result must be bit-identical: populations,
species, velocity and the mass field, for the Luo and Guo force models, one
to four species (SIGNAL and a linear reaction matrix), the box outlet, a
per-site mass field, rows that are not a multiple of a workgroup's sites,
rows whose width admits only the 8-byte vectors, no-flux cut links in a
coupled run, and batched atlases (the segment table; rectangles of odd
width end inside a lane).  The one-site kernel is itself checked against the
oracle in test_gpu_mixed.py."""
import numpy as np
import pytest

from lbibcell_amd import capi, runner, workloads

pytestmark = pytest.mark.gpu

W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
PRECS = {"fp32": capi.FP32, "fp16": capi.FP16_DEV, "bf16": capi.BF16_DEV, "fp16r": capi.FP16_REM}
SIG = dict(production=1e-3, decay=1e-5, switchoff=7, init_inside=0.5, prod_domain=1)


def vec_shapes(prec, nx):
    if prec == "fp16r":
        return [(2, wg, 1) for wg in (64, 128, 256)]
    vs = (2, 4) if prec == "fp32" else (4, 8)
    return [(2, wg, v) for v in vs if nx % v == 0 for wg in (64, 128, 256)]


def _species(ns):
    if ns == 1:
        return [capi.species_spec(capi.SPECIES_SIGNAL, 0.8, **SIG)]
    rng = np.random.default_rng(12)
    out = []
    for s in range(ns):
        k = tuple(1e-3 * rng.standard_normal(4))
        out.append(capi.species_spec(capi.SPECIES_LINEAR, 0.7 + 0.1 * s, krow=k,
                                     source_inside=1e-3 if s == 0 else 0.0))
    return out


def _lattice_run(prec, fm, nx, ny, ns, shape, steps=6):
    rng = np.random.default_rng(100 + nx + ns)
    ctx = capi.Context(nx, ny, 0.83, species=_species(ns), force_model=fm, precision=PRECS[prec], box_outlet=True)
    dom = np.zeros((ny, nx), dtype=np.uint32)
    dom[ny // 4: 3 * ny // 4, nx // 3: 2 * nx // 3] = 1
    ctx.set_domains(dom)
    ctx.init_solvers()
    ctx.set_populations(W[:, None, None] * (1 + 0.05 * rng.standard_normal((9, ny, nx))))
    for s in range(ns):
        ctx.set_species(s, 0.1 * (1 + 0.05 * rng.standard_normal((5, ny, nx))))
    ctx.set_force_field(1e-4 * rng.standard_normal((ny, nx)), 1e-4 * rng.standard_normal((ny, nx)))
    ctx.set_lattice_shape(*shape)
    for it in range(steps):
        if it == 2:
            ctx.set_mass_field(1e-4 * rng.standard_normal((ny, nx)))
        ctx.step_lattice_only(it)
    out = [ctx.get_populations()] + [ctx.get_species(s) for s in range(ns)] + list(ctx.get_velocity())
    ctx.close()
    return out


# (48, 40): one partial workgroup per row; (600, 9): several workgroups per
# row; (44, 13): 44 sites per row admit only the 8-byte vectors
@pytest.mark.parametrize("size", [(48, 40), (600, 9), (44, 13)])
@pytest.mark.parametrize("prec", list(PRECS))
def test_vector_kernel_equals_one_site_kernel(prec, size):
    nx, ny = size
    for fm in (capi.FORCE_LUO, capi.FORCE_GUO):
        for ns in (1, 4):
            ref = _lattice_run(prec, fm, nx, ny, ns, (0, 64, 1))
            for shape in vec_shapes(prec, nx):
                print(f"case vec lattice {prec} {nx}x{ny} fm={fm} ns={ns} shape={shape}", flush=True)
                got = _lattice_run(prec, fm, nx, ny, ns, shape)
                for k, (a, b) in enumerate(zip(got, ref)):
                    np.testing.assert_array_equal(a, b, err_msg=f"{prec} {size} fm={fm} ns={ns} {shape} field {k}")


def _noflux_run(prec, shape):
    nx = ny = 64
    nodes, conns = workloads.circle_cell(30.5, 33.0, 9.0, 56, np.random.default_rng(9))
    sim = runner.Simulation(nx, ny, 1.1, precision=PRECS[prec])
    sim.add_species("CDESolverD2Q5", capi.species_spec(capi.SPECIES_PLAIN, 0.9))
    sim.add_species("tutorial_01_CDESolverD2Q5_SIGNAL",
                    capi.species_spec(capi.SPECIES_SIGNAL, 0.75, production=1e-3, decay=1e-5, switchoff=5000,
                                      init_inside=0.0, prod_domain=1))
    for nid, x, y in nodes:
        sim.add_node(nid, x, y)
    for a, b, d in conns:
        sim.add_connection(a, b, d, "tutorial_01_CDESolverD2Q5_SIGNAL,CDESolverD2Q5")
    sim.setup()
    sim.init_solvers()
    sim.add_mass_solver("MassSolverBoxOutlet")
    sim.add_tension(0.01, 10)
    ctx = sim.context()
    ctx.set_lattice_shape(*shape)
    sim.run(40)
    ctx.ns = 2
    out = [ctx.get_populations(), ctx.get_species(0), ctx.get_species(1)] + list(ctx.get_velocity())
    out.append(sim.markers()[1])
    sim.close()
    return out


@pytest.mark.parametrize("prec", ["fp32", "fp16", "fp16r"])
def test_vector_kernel_noflux_coupled_run(prec):
    ref = _noflux_run(prec, (0, 64, 1))
    for shape in vec_shapes(prec, 64)[::3]:
        print(f"case vec no-flux coupled {prec} shape={shape}", flush=True)
        got = _noflux_run(prec, shape)
        for k, (a, b) in enumerate(zip(got, ref)):
            np.testing.assert_array_equal(a, b, err_msg=f"{prec} {shape} field {k}")
