"""The C-ABI library loads and exports every symbol include/psf/psf.h declares;
host-only entry points work without a GPU."""
import ctypes as ct

import pytest

from parallelspectralfluids3d_amd import _lib as L
from parallelspectralfluids3d_amd import hip
from parallelspectralfluids3d_amd.grid import Grid


def test_exports_every_declared_symbol():
    lib = L.load()
    names = L.declared_symbols()
    assert len(names) >= 24
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.psf_abi_version() == 4


def test_footprint_without_gpu():
    d = hip.make_desc(Grid(64, 64, 64), 1 / 64, 256, L.C_FP32, L.PHI_FP32, 1 << 20)
    b = hip.footprint(d)
    F = 798720
    assert b >= 256 ** 3 * 4 + F * 256 * 4
    assert b < 256 ** 3 * 4 + F * 256 * 4 + 200 * 2 ** 20
    bad = hip.make_desc(Grid(64, 64, 64), 1 / 64, 0)
    with pytest.raises(L.PsfError) as e:
        hip.footprint(bad)
    assert e.value.status == L.PSF_ERR_ARG


def test_context_errors_are_codes():
    lib = L.load()
    h = ct.c_void_p()
    assert lib.psf_ctx_create(-1, ct.byref(h)) in (L.PSF_ERR_ARG, L.PSF_ERR_HIP)
    rc = lib.psf_ctx_create(0, ct.byref(h))
    assert rc in (L.PSF_OK, L.PSF_ERR_HIP)
    if rc == L.PSF_OK:
        lib.psf_ctx_destroy(h)
    assert lib.psf_ctx_destroy(None) == L.PSF_ERR_ARG


def test_cpp_header_compiles(tmp_path):
    """include/psf/solver.hpp and psf.h compile as plain C++ and C (no HIP headers)."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    inc = os.path.join(root, "include")
    src = tmp_path / "t.cpp"
    src.write_text('#include "psf/solver.hpp"\nint main(){ return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", f"-I{inc}", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    csrc = tmp_path / "t.c"
    csrc.write_text('#include "psf/psf.h"\nint main(void){ return 0; }\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", f"-I{inc}", str(csrc)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
