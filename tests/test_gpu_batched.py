"""Batched ensemble: many independent periodic domains in one atlas, every
per-step launch serving all of them (lbib_set_batch_layout).

Each domain of the atlas must evolve exactly as the same domain run alone:
- lattice-only stepping is compared bit for bit with one standalone context
  per domain, for every launch shape the tuner can pick (FP64 and the
  mixed-precision storage formats), with per-domain relaxation times, the box
  outlet on each rectangle's own edges and sites between rectangles left
  untouched;
- the coupled step (IB, cut links, scanline, re-initialisation, remeshing,
  tension and area regulator) is compared with the oracle run per domain: bit
  for bit for the domain at the atlas origin (identical coordinates), within
  1e-11 relative elsewhere (marker coordinates carry the rectangle's offset,
  so their last bits round differently), domain ids exact."""
import numpy as np
import pytest

from lbibcell_amd import capi, runner, workloads

pytestmark = pytest.mark.gpu

PRECS = {"fp64": capi.FP64, "fp32": capi.FP32, "fp16": capi.FP16_DEV, "bf16": capi.BF16_DEV,
         "fp16r": capi.FP16_REM}
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
SIG = dict(production=1e-3, decay=1e-5, switchoff=7, init_inside=0.5, prod_domain=-1)

# (x0, y0, w, h, tau_f, [tau_c]): an odd width, a non-square one, a tile-
# aligned one; a gap column and row are left between them
RECTS = [(0, 0, 37, 30, 0.83, [0.7]), (48, 0, 64, 44, 1.27, [1.05]), (0, 48, 26, 36, 0.61, [0.58]),
         (32, 64, 51, 20, 1.0, [0.9])]


def _shapes(prec, even):
    if prec == "fp64":
        return [(0, wg, 1) for wg in (64, 128, 256)]
    if prec == "fp16r":   # one load per pulled moment, or moments shared by shuffles
        return [(v, wg, 1) for v in (0, 2) for wg in (64, 128, 256)]
    s = [(0, wg, spt) for wg in (64, 128, 256) for spt in (1, 2, 4)]
    if prec in ("fp16", "bf16") and even:
        s += [(1, wg, 1) for wg in (64, 128, 256)]
    # vector kernel (every atlas here is 112 or 512 sites wide)
    s += [(2, wg, v) for v in ((2, 4) if prec == "fp32" else (4, 8)) for wg in (64, 128, 256)]
    return s


def _lattice_case(prec, rects, shape, steps=8):
    nx, ny = 112, 96
    rng = np.random.default_rng(31)
    sp = capi.species_spec(capi.SPECIES_SIGNAL, 0.8, **SIG)
    atlas = capi.Context(nx, ny, 1.0, species=[sp], precision=PRECS[prec], box_outlet=True)
    atlas.set_batch_layout(rects)
    dom = np.zeros((ny, nx), dtype=np.uint32)
    f = np.tile(W[:, None, None], (1, ny, nx))
    g = np.zeros((5, ny, nx))
    F = np.zeros((2, ny, nx))
    solo = []
    for k, (x0, y0, w, h, tf, tc) in enumerate(rects):
        win = (slice(y0, y0 + h), slice(x0, x0 + w))
        d = np.zeros((h, w), dtype=np.uint32)
        d[h // 4: 3 * h // 4, w // 3: 2 * w // 3] = k + 1
        dom[win] = d
        f[(slice(None),) + win] = W[:, None, None] * (1 + 0.05 * rng.standard_normal((9, h, w)))
        g[(slice(None),) + win] = 0.1 * (1 + 0.05 * rng.standard_normal((5, h, w)))
        F[(slice(None),) + win] = 1e-4 * rng.standard_normal((2, h, w))
        c = capi.Context(w, h, tf, species=[capi.species_spec(capi.SPECIES_SIGNAL, tc[0], **SIG)],
                         precision=PRECS[prec], box_outlet=True)
        c.set_domains(d)
        c.init_solvers()
        c.set_populations(f[(slice(None),) + win])
        c.set_species(0, g[(slice(None),) + win])
        c.set_force_field(F[0][win], F[1][win])
        solo.append(c)
    atlas.set_domains(dom)
    atlas.init_solvers()
    atlas.set_populations(f)
    atlas.set_species(0, g)
    atlas.set_force_field(F[0], F[1])
    if shape is not None:
        atlas.set_lattice_shape(*shape)
    f_start = atlas.get_populations()
    for it in range(steps):
        atlas.step_lattice_only(it)
        for c in solo:
            c.step_lattice_only(it)
    fa, ga = atlas.get_populations(), atlas.get_species(0)
    live = np.zeros((ny, nx), dtype=bool)
    for k, (x0, y0, w, h, _, _) in enumerate(rects):
        win = (slice(y0, y0 + h), slice(x0, x0 + w))
        live[win] = True
        np.testing.assert_array_equal(fa[(slice(None),) + win], solo[k].get_populations(),
                                      err_msg=f"{prec} {shape}: fluid of domain {k}")
        np.testing.assert_array_equal(ga[(slice(None),) + win], solo[k].get_species(0),
                                      err_msg=f"{prec} {shape}: species of domain {k}")
        solo[k].close()
    # sites outside every rectangle keep what they held (an even number of
    # steps brings the same population buffer back)
    assert steps % 2 == 0
    np.testing.assert_array_equal(fa[:, ~live], f_start[:, ~live])
    atlas.close()


@pytest.mark.parametrize("prec", list(PRECS))
def test_atlas_lattice_steps_equal_standalone_domains(prec):
    print(f"case atlas lattice {prec} default shape", flush=True)
    _lattice_case(prec, RECTS, None)


@pytest.mark.parametrize("prec", list(PRECS))
def test_every_tuned_shape_gives_identical_results(prec):
    # even widths so that the paired-lane kernel applies to 16-bit storage
    rects = [(x0, y0, w + (w % 2), h, tf, tc) for x0, y0, w, h, tf, tc in RECTS]
    for shape in _shapes(prec, True):
        print(f"case atlas lattice {prec} shape={shape}", flush=True)
        _lattice_case(prec, rects, shape, steps=4)


@pytest.mark.parametrize("prec", ["fp64", "fp16"])
def test_autotune_keeps_state_and_picks_the_fastest(prec):
    nx, ny = 512, 320
    rects = [(0, 0, 249, 200, 0.9, [0.8]), (256, 0, 256, 320, 1.2, [0.7]), (0, 208, 128, 112, 0.7, [0.6])]
    sp = capi.species_spec(capi.SPECIES_PLAIN, 0.8)
    a = capi.Context(nx, ny, 1.0, species=[sp], precision=PRECS[prec])
    b = capi.Context(nx, ny, 1.0, species=[sp], precision=PRECS[prec])
    rng = np.random.default_rng(4)
    f = W[:, None, None] * (1 + 0.05 * rng.standard_normal((9, ny, nx)))
    for c in (a, b):
        c.set_batch_layout(rects)
        c.init_solvers()
        c.set_populations(f)
    chosen, rows = a.autotune(reps=3)
    print(f"case autotune {prec}: chosen {chosen[:3]} of {len(rows)}", flush=True)
    for r in rows:
        print("   shape", r[:3], f"{r[3] * 1e3:.1f} us", flush=True)
    assert len(rows) == len(_shapes(prec, False))   # one odd width: no paired-lane shapes
    assert chosen[3] == min(r[3] for r in rows)
    for it in range(3):
        a.step_lattice_only(it)
        b.step_lattice_only(it)
    np.testing.assert_array_equal(a.get_populations(), b.get_populations())
    np.testing.assert_array_equal(a.get_species(0), b.get_species(0))
    with pytest.raises(capi.LbibError):
        a.set_lattice_shape(1, 256, 1)
    a.close()
    b.close()


def test_layout_validation():
    c = capi.Context(64, 64, 1.0)
    for bad in ([(8, 0, 32, 32, 1.0)], [(0, 0, 40, 32, 1.0), (32, 0, 16, 16, 1.0)], [(0, 0, 80, 16, 1.0)],
                [(0, 0, 32, 32, 0.5)]):
        with pytest.raises(capi.LbibError):
            c.set_batch_layout(bad)
    c.set_batch_layout([(0, 0, 30, 30, 1.0), (32, 32, 32, 32, 0.9)])
    c.set_batch_layout([])
    c.close()


def _coupled_atlas(orc, steps):
    # three domains; cells in local coordinates, moved to the atlas by the offset
    dom_specs = [((0, 0), 64, 0.9, 0.8, [(30.0, 33.0, 9.0, 52)]),
                 ((80, 0), 72, 1.2, 0.65, [(22.0, 25.0, 7.0, 40), (48.0, 46.0, 8.0, 48)]),
                 ((0, 80), 56, 0.7, 1.1, [(27.5, 28.0, 10.0, 60)])]
    nx, ny = 160, 144
    rng = np.random.default_rng(17)
    sim = runner.Simulation(nx, ny, 1.0)
    sim.add_species("tutorial_01_CDESolverD2Q5_SIGNAL",
                    capi.species_spec(capi.SPECIES_SIGNAL, 0.8, production=1e-3, decay=1e-5, switchoff=5000,
                                      init_inside=0.0, prod_domain=-1))
    sim.set_batch_layout([(x0, y0, L, L, tf, [tc]) for (x0, y0), L, tf, tc, _ in dom_specs])
    oracles, local = [], []
    nid, dom = 0, 1
    for (x0, y0), L, tf, tc, cells in dom_specs:
        o = orc.Oracle(L, L, tf)
        o.add_species(orc.SP_SIGNAL, tc, [1e-3, 1e-5, 5000, 0.0, -1])
        ids = []
        for cx, cy, r, n in cells:
            nodes, conns = workloads.circle_cell(cx, cy, r, n, rng, first_id=nid, domain=dom)
            for k, x, y in nodes:
                o.add_node(k, x, y)
                sim.add_node(k, x + x0, y + y0)
            for a, b, d in conns:
                o.add_connection(a, b, d, [])
                sim.add_connection(a, b, d)
            ids.append(dom)
            nid += n
            dom += 1
        o.cure()
        o.init_solvers()
        o.box_outlet(True)
        o.add_bio(1, [150.0, 1e-5, 0.004, 1.2], 1)
        o.add_bio(0, [0.008], 10)
        oracles.append(o)
        local.append(((x0, y0), L, ids))
    sim.setup()
    sim.init_solvers()
    sim.add_mass_solver("MassSolverBoxOutlet")
    sim.add_area_regulator(150.0, 1e-5, 0.004, 1.2, freq=1)
    sim.add_tension(0.008, 10)
    sim.run(steps)
    for o in oracles:
        o.run(steps)
    return sim, oracles, local


def test_atlas_coupled_run_matches_oracle_per_domain(orc):
    steps = 40
    print(f"case atlas coupled 3 domains, {steps} steps", flush=True)
    sim, oracles, local = _coupled_atlas(orc, steps)
    print("  perturbation draws: atlas", sim.perturb_counts(), "per-domain oracles",
          [o.perturb_counts() for o in oracles], flush=True)
    ctx = sim.context()
    ctx.ns = 1
    dom_a = ctx.get_domains()
    fa, ga = ctx.get_populations(), ctx.get_species(0)
    ids_a, xy_a, _ = sim.markers()
    for k, (o, ((x0, y0), L, ids)) in enumerate(zip(oracles, local)):
        win = (slice(y0, y0 + L), slice(x0, x0 + L))
        dom_o, _ = o.get_domains()
        np.testing.assert_array_equal(dom_a[win], dom_o, err_msg=f"domain {k}: domain ids")
        fo, go = o.get_f(), o.get_g(0)
        ef = np.max(np.abs(fa[(slice(None),) + win] - fo)) / np.max(np.abs(fo))
        eg = np.max(np.abs(ga[(slice(None),) + win] - go)) / max(np.max(np.abs(go)), 1e-300)
        # markers of this domain (inside its rectangle), in id order on both
        # sides: ids of remeshed markers differ, their creation order does not
        ids_o, xy_o, _, _ = o.markers()
        xy_o = xy_o[np.argsort(ids_o, kind="stable")]
        mine = ((xy_a[:, 0] >= x0) & (xy_a[:, 0] < x0 + L) & (xy_a[:, 1] >= y0) & (xy_a[:, 1] < y0 + L))
        order = np.argsort(ids_a[mine], kind="stable")
        pa = xy_a[mine][order] - np.array([x0, y0])
        assert pa.shape == xy_o.shape, f"domain {k}: marker count {pa.shape} vs {xy_o.shape}"
        ex = np.max(np.abs(pa - xy_o)) / np.max(np.abs(xy_o))
        print(f"  domain {k}: rel err populations {ef:.1e} species {eg:.1e} markers {ex:.1e}", flush=True)
        if (x0, y0) == (0, 0):
            np.testing.assert_array_equal(fa[(slice(None),) + win], fo)
            np.testing.assert_array_equal(xy_a[mine][order], xy_o)
        assert ef <= 1e-11 and eg <= 1e-11 and ex <= 1e-11
    sim.close()
