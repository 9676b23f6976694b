"""Binomial checkpointing (revolve) and 16-bit population storage of the
adjoint (lbib_adj_set_checkpointing, lbib_adj_set_storage):
- revolve with any number of snapshot slots gives the gradient of the
  every-K checkpointing (the states it recomputes are the same), for one
  species and for three species with junction springs;
- the recomputation stays within the binomial bound: with s slots and the
  least r such that C(s + r, s) >= T, no step runs more than r + 1 times in
  the gradient (r recomputations plus the one that stores its forces);
- with FP32 arithmetic, FP16 / BF16 deviation storage gives a gradient close
  to that of FP32 storage, and revolve gives the every-K gradient there too;
- the autograd wrappers (one and several species) return the adjoint's
  gradients, only for the inputs that ask for one;
- storage and checkpointing arguments are validated."""
import math

import numpy as np
import pytest

from lbibcell_amd import adjoint, capi, workloads

pytestmark = pytest.mark.gpu

W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
PARAMS = dict(tau_c=0.85, kappa=0.02)
MS = dict(tau_c=[0.8, 0.95, 1.1], decay=[1e-3, 2e-3, 5e-4], kappa=0.02, k_junction=0.015, l0=0.5)


def _small(seed=3, nx=40, ny=36, n=32, radius=8.0):
    rings, dom, rates = workloads.adjoint_case(nx, ny, 1, 1, radius, n, seed=seed)
    rng = np.random.default_rng(seed)
    f0 = W[:, None, None] * (1 + 0.02 * rng.standard_normal((9, ny, nx)))
    g0 = 0.1 * (1 + 0.1 * rng.standard_normal((5, ny, nx)))
    target = 0.5 * (1 + 0.1 * rng.standard_normal((ny, nx)))
    return dict(nx=nx, ny=ny, rings=rings, dom=dom, rates=rates, f0=f0, g0=g0, target=target)


def _problem(case, precision=capi.FP64, K=5, storage=None, revolve=None):
    p = adjoint.AdjointProblem(case["nx"], case["ny"], 0.9, case["rings"], case["dom"], target=case["target"],
                               f0=case["f0"], g0=case["g0"], decay=1e-3, precision=precision, checkpoint_every=K,
                               storage=storage)
    if revolve is not None:
        p.set_checkpointing(adjoint.REVOLVE, revolve)
    return p


def _reps(s, steps):
    r = 0
    while math.comb(s + r, s) < steps:
        r += 1
    return r


@pytest.mark.parametrize("steps", [23, 40])
def test_revolve_gives_the_every_k_gradient(steps):
    case = _small(seed=7)
    p = _problem(case, K=5)
    p.forward(steps, PARAMS["tau_c"], PARAMS["kappa"], case["rates"])
    ref = p.gradient()
    ref_bytes = p.stats()["store_bytes"]
    p.close()
    for s in (1, 2, 3, 8, 60):
        print(f"case revolve steps={steps} snapshots={s}", flush=True)
        p = _problem(case, revolve=s)
        p.forward(steps, PARAMS["tau_c"], PARAMS["kappa"], case["rates"])
        dt, dk, dr = p.gradient()
        st = p.stats()
        p.close()
        assert dt == pytest.approx(ref[0], rel=1e-12)
        assert dk == pytest.approx(ref[1], rel=1e-10)
        np.testing.assert_allclose(dr, ref[2], rtol=1e-12)
        assert st["forward_steps"] == steps
        r = _reps(s, steps)
        print("   gradient steps", st["gradient_steps"], "bound", (r + 1) * steps, "store", st["store_bytes"],
              "every-K store", ref_bytes, flush=True)
        assert steps <= st["gradient_steps"] <= (r + 1) * steps
        if s < steps // 5:
            assert st["store_bytes"] < ref_bytes


def _ms_case(seed=11, nx=60, ny=36, ns=3):
    rings, dom, rates = workloads.adjoint_case(nx, ny, 2, 1, 10.5, 44, seed=seed)
    pairs = workloads.junction_pairs(rings, 9.5)
    rng = np.random.default_rng(seed)
    f0 = W[:, None, None] * (1 + 0.02 * rng.standard_normal((9, ny, nx)))
    g0 = 0.1 * (1 + 0.1 * rng.standard_normal((ns, 5, ny, nx)))
    target = 0.5 * (1 + 0.1 * rng.standard_normal((ns, ny, nx)))
    K = 1e-2 * rng.standard_normal((ns, ns))
    return dict(nx=nx, ny=ny, rings=rings, dom=dom, rates=rates, pairs=pairs, f0=f0, g0=g0, target=target, K=K,
                ns=ns)


def _ms_problem(case, precision=capi.FP64, K=7, storage=None, revolve=None):
    p = adjoint.AdjointProblem(case["nx"], case["ny"], 0.9, case["rings"], case["dom"], f0=case["f0"],
                               g0=case["g0"][0], decay=MS["decay"][0], precision=precision, checkpoint_every=K,
                               storage=storage)
    p.set_species(case["g0"], case["target"], MS["decay"][: case["ns"]])
    p.set_junctions(case["pairs"], MS["l0"])
    if revolve is not None:
        p.set_checkpointing(adjoint.REVOLVE, revolve)
    return p


def _ms_run(p, case, steps):
    loss = p.forward_ms(steps, MS["tau_c"][: case["ns"]], MS["kappa"], MS["k_junction"], case["K"], case["rates"])
    return loss, p.gradient_ms()


def test_revolve_multispecies_junctions():
    case = _ms_case()
    steps = 31
    p = _ms_problem(case, K=7)
    l0, g0 = _ms_run(p, case, steps)
    p.close()
    for s in (2, 5):
        print(f"case revolve three species + junctions, snapshots={s}", flush=True)
        p = _ms_problem(case, revolve=s)
        l1, g1 = _ms_run(p, case, steps)
        p.close()
        assert l1 == pytest.approx(l0, rel=1e-14)
        for k in g0:
            np.testing.assert_allclose(np.asarray(g1[k]), np.asarray(g0[k]), rtol=1e-10, atol=1e-300, err_msg=k)


@pytest.mark.parametrize("storage", [capi.FP16_DEV, capi.BF16_DEV])
def test_sixteen_bit_storage_gradient(storage):
    case = _ms_case(seed=13)
    steps = 40
    p = _ms_problem(case, precision=capi.FP32, K=10)
    l32, g32 = _ms_run(p, case, steps)
    p.close()
    p = _ms_problem(case, precision=capi.FP32, K=10, storage=storage)
    l16, g16 = _ms_run(p, case, steps)
    p.close()
    p = _ms_problem(case, precision=capi.FP32, storage=storage, revolve=3)
    l16r, g16r = _ms_run(p, case, steps)
    p.close()
    name = "fp16" if storage == capi.FP16_DEV else "bf16"
    errs = {k: float(np.max(np.abs(np.asarray(g16[k]) - np.asarray(g32[k]))) /
                     max(float(np.max(np.abs(np.asarray(g32[k])))), 1e-300)) for k in g32}
    print(f"case {name} storage: loss {l16:.9g} vs {l32:.9g}; gradient rel. errors vs FP32 storage {errs}",
          flush=True)
    # the same states are recomputed: revolve equals every-K up to FP32 summation order
    assert l16r == pytest.approx(l16, rel=1e-6)
    for k in g16:
        np.testing.assert_allclose(np.asarray(g16r[k]), np.asarray(g16[k]), rtol=2e-4, atol=1e-300, err_msg=k)
    # bars: FP16 deviations keep 11 significant bits, BF16 8
    tol = 2e-2 if storage == capi.FP16_DEV else 1e-1
    assert l16 == pytest.approx(l32, rel=tol)
    for k, e in errs.items():
        assert e < tol, (k, e)


def test_multispecies_autograd_wrapper_returns_the_adjoint_gradient():
    import torch
    case = _ms_case(seed=17)
    steps = 12
    p = _ms_problem(case, K=4)
    Fn = adjoint.multispecies_loss_function()
    tc = torch.tensor(MS["tau_c"], dtype=torch.float64, requires_grad=True)
    kp = torch.tensor(MS["kappa"], dtype=torch.float64, requires_grad=True)
    kj = torch.tensor(MS["k_junction"], dtype=torch.float64)   # no gradient asked for
    Km = torch.tensor(case["K"], dtype=torch.float64, requires_grad=True)
    rt = torch.tensor(case["rates"], dtype=torch.float64, requires_grad=True)
    print("case multispecies autograd wrapper", flush=True)
    L = Fn.apply(tc, kp, kj, Km, rt, p, steps)
    # another forward in between: backward runs the forward again
    p.forward_ms(steps, np.array(MS["tau_c"]) * 1.01, MS["kappa"], MS["k_junction"], case["K"], case["rates"])
    (2.0 * L).backward()
    l_ref, g = _ms_run(p, case, steps)
    # the forward spread adds with FP64 atomics: runs agree to rounding
    assert float(L.detach()) == pytest.approx(l_ref, rel=1e-12)
    np.testing.assert_allclose(tc.grad.numpy(), 2.0 * g["tau_c"], rtol=1e-12)
    assert kp.grad.item() == pytest.approx(2.0 * g["kappa"], rel=1e-10)
    assert kj.grad is None
    np.testing.assert_allclose(Km.grad.numpy(), 2.0 * g["K"], rtol=1e-10)
    np.testing.assert_allclose(rt.grad.numpy(), 2.0 * g["rates"], rtol=1e-12)
    p.close()


def test_single_species_wrapper_computes_only_requested_gradients():
    import torch
    case = _small(seed=9)
    p = _problem(case)
    Fn = adjoint.coupled_loss_function()
    tc = torch.tensor(PARAMS["tau_c"], dtype=torch.float64, requires_grad=True)
    kp = torch.tensor(PARAMS["kappa"], dtype=torch.float64)
    rt = torch.tensor(case["rates"], dtype=torch.float32, requires_grad=True)
    print("case single-species autograd wrapper, partial gradients", flush=True)
    L = Fn.apply(tc, kp, rt, p, 11)
    L.backward()
    assert kp.grad is None and rt.grad.dtype == torch.float32
    # the wrapper ran with the float32 rates
    p.forward(11, PARAMS["tau_c"], PARAMS["kappa"], rt.detach().numpy().astype(np.float64))
    dt, dk, dr = p.gradient()
    assert tc.grad.item() == pytest.approx(dt, rel=1e-12)
    np.testing.assert_allclose(rt.grad.numpy(), dr.astype(np.float32), rtol=1e-6)
    # no gradient asked for: nothing runs the adjoint
    with torch.no_grad():
        L2 = Fn.apply(tc.detach(), kp, rt.detach(), p, 11)
    assert float(L2) == pytest.approx(float(L.detach()), rel=1e-12)
    p.close()


def test_storage_and_checkpointing_arguments_are_validated():
    case = _small()
    print("case adjoint storage / checkpointing validation", flush=True)
    # 16-bit storage needs FP32 arithmetic
    with pytest.raises(capi.LbibError, match="FP32 arithmetic"):
        _problem(case, precision=capi.FP64, storage=capi.FP16_DEV)
    p = _problem(case)
    # after the problem is set the storage cannot change
    with pytest.raises(capi.LbibError, match="before lbib_adj_set_problem"):
        p._chk(p.L.lbib_adj_set_storage(p.h, capi.FP64))
    for scheme, param in ((adjoint.REVOLVE, 0), (adjoint.CHECKPOINT_EVERY, 0), (7, 3)):
        with pytest.raises(capi.LbibError, match="snapshot slots"):
            p.set_checkpointing(scheme, param)
    p.close()
