"""The C ABI libraries load and export every entry point include/lbib.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbib_[a-z0-9_]+)\s*\(", text)))


def test_hip_library_exports_every_declared_symbol():
    from lbibcell_amd import capi
    lib = ctypes.CDLL(capi.LIB_PATH)
    names = declared("lbib.h")
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python bindings cover every declared entry point
    from lbibcell_amd import adjoint
    bound = set(capi.SIGNATURES) | set(adjoint.SIGNATURES)
    assert set(names) <= bound, set(names) - bound


def test_host_runtime_loads():
    from lbibcell_amd import runner
    L = runner.load()
    for name in runner._SIG:
        assert hasattr(L, name), name


def test_status_bits_match_header():
    from lbibcell_amd import capi
    text = open(os.path.join(ROOT, "include", "lbib.h")).read()
    for name in ["ERR_VELOCITY", "ERR_EDGE", "ERR_SCANLINE", "ERR_SLOPE", "ERR_CAPACITY", "ERR_TOPOLOGY",
                 "HALT_PERTURB_NODE", "HALT_PERTURB_CONN", "HALT_PERTURB_REMESH", "HALT_GROW",
                 "FLAG_REMESH", "FLAG_COARSEN", "SKIPPED"]:
        m = re.search(r"#define LBIB_%s\s+(0x[0-9a-fA-F]+)u" % name, text)
        assert m and int(m.group(1), 16) == getattr(capi, name), name


def test_struct_layouts_match_header_sizes(tmp_path):
    """ctypes mirrors of the header structs have the C compiler's layout."""
    import subprocess
    from lbibcell_amd import capi
    src = tmp_path / "sz.c"
    src.write_text('#include "lbib.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                   'int main(void){printf("%zu %zu %zu %zu %zu\\n", sizeof(lbib_species_spec),'
                   ' sizeof(lbib_config), sizeof(lbib_status), offsetof(lbib_config, species),'
                   ' offsetof(lbib_config, box_outlet)); return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == [ctypes.sizeof(capi.SpeciesSpec), ctypes.sizeof(capi.Config), ctypes.sizeof(capi.Status),
                   capi.Config.species.offset, capi.Config.box_outlet.offset]


def test_context_creation_rejects_bad_config_without_gpu_calls():
    # argument validation happens before any device call
    from lbibcell_amd import capi
    lib = capi.load()
    cfg = capi.Config()
    cfg.nx, cfg.ny, cfg.tau_fluid, cfg.nspecies = 2, 2, 1.0, 0
    h = ctypes.c_void_p()
    assert lib.lbib_create(ctypes.byref(cfg), ctypes.byref(h)) != 0
    assert b"invalid configuration" in lib.lbib_last_error(h)
    lib.lbib_destroy(h)
