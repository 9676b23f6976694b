"""lbib_domain_concentrations_noalloc, the allocation-free variant of
lbib_domain_concentrations (static device scratch, output in a device array
of the caller), called through the C ABI and checked against the allocating
entry point on two calls in a row, the second on a smaller input; sizes
beyond the scratch are refused."""
import ctypes as C

import numpy as np
import pytest

from lbibcell_amd import capi

pytestmark = pytest.mark.gpu


def _hip():
    L = C.CDLL("libamdhip64.so")
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return L


def _case(nx, ny, ndom, seed):
    rng = np.random.default_rng(seed)
    sp = capi.species_spec(capi.SPECIES_SIGNAL, 0.8, production=1e-3, decay=1e-5, switchoff=5000,
                           init_inside=0.5, prod_domain=-1)
    ctx = capi.Context(nx, ny, 1.0, species=[sp])
    dom = np.zeros((ny, nx), dtype=np.uint32)
    for d in range(1, ndom):
        x0, y0 = rng.integers(1, nx - 8), rng.integers(0, ny - 8)
        dom[y0:y0 + int(rng.integers(2, 8)), x0:x0 + int(rng.integers(2, 8))] = d
    ctx.set_domains(dom)
    ctx.init_solvers()
    ctx.set_species(0, 0.1 + rng.random((5, ny, nx)))
    return ctx


def _check(hip, ctx, ndom):
    ref = ctx.domain_concentrations(0, ndom)           # the allocating entry point
    extra = 8
    buf = C.c_void_p()
    assert hip.hipMalloc(C.byref(buf), (ndom + extra) * 8) == 0
    sentinel = np.full(ndom + extra, np.nan)
    assert hip.hipMemcpy(buf, sentinel.ctypes.data, sentinel.nbytes, 1) == 0     # host to device
    rc = ctx.lib.lbib_domain_concentrations_noalloc(ctx.h, 0, ndom, buf)
    assert rc == 0, ctx.lib.lbib_last_error(ctx.h).decode()
    ctx.synchronize()
    got = np.empty(ndom + extra)
    assert hip.hipMemcpy(got.ctypes.data, buf, got.nbytes, 2) == 0               # device to host
    assert hip.hipFree(buf) == 0
    np.testing.assert_allclose(got[:ndom], ref, rtol=1e-12, atol=0)
    assert np.all(np.isnan(got[ndom:])), "wrote past ndom"
    return got[:ndom]


def test_noalloc_domain_sums_equal_the_allocating_entry_point():
    hip = _hip()
    print("case no-alloc domain sums: 256x192 / 40 domains, then 96x64 / 7 domains", flush=True)
    big = _case(256, 192, 40, seed=1)
    first = _check(hip, big, 40)
    assert np.count_nonzero(first) == 40
    small = _case(96, 64, 7, seed=2)
    second = _check(hip, small, 7)        # the scratch keeps the first call's sums: they must not leak
    assert abs(second.sum() - small.domain_concentrations(0, 7).sum()) <= 1e-12 * second.sum()
    # beyond the static scratch: refused, nothing launched
    buf = C.c_void_p()
    assert hip.hipMalloc(C.byref(buf), 8) == 0
    assert small.lib.lbib_domain_concentrations_noalloc(small.h, 0, 65537, buf) == -1
    assert "65536" in small.lib.lbib_last_error(small.h).decode()
    assert small.lib.lbib_domain_concentrations_noalloc(small.h, 3, 7, buf) == -1
    assert hip.hipFree(buf) == 0
    big.close()
    small.close()
