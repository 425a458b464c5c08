"""The C-ABI library loads and exports every symbol include/gelios.h declares
(no device calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT
from gelios_2d_amd import _lib
from gelios_2d_amd.params import GelParams, reference_params


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "gelios.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gel_[a-z0-9_]+)\s*\(", txt)))


def test_library_built_and_exports_header():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
        assert s in _lib.SIGNATURES, s


def test_params_layout_matches_reference_fill():
    """gel_params_reference (C) and reference_params (Python) agree field for field,
    which also checks the ctypes mirror of the struct layout."""
    L = _lib.lib()
    c = GelParams()
    L.gel_params_reference(ctypes.byref(c))
    p = reference_params()
    assert c.as_dict() == p.as_dict()


def test_errors_without_device():
    L = _lib.lib()
    assert L.gel_strerror(0) == b"ok"
    ctx = ctypes.c_void_p()
    g = _lib.GelGrid(2, 2, None, None, None, None)
    p = reference_params()
    assert L.gel_ctx_create(ctypes.byref(ctx), ctypes.byref(g), ctypes.byref(p)) == 1
    assert L.gel_ctx_create(None, None, None) == 1
    assert L.gel_step(None, 1, None) == 1
    # batches: a null entry is rejected before any case is read
    out = ctypes.c_void_p()
    cases = (ctypes.c_void_p * 2)(None, None)
    assert L.gel_batch_create(ctypes.byref(out), 2, cases) == 1
    assert out.value is None
    with pytest.raises(_lib.GelError):
        _lib.check(1)


@pytest.mark.skipif(not os.path.exists("/root/reference/Setka.h"), reason="needs the reference's headers")
def test_integration_binding_compiles_with_reference_headers(tmp_path):
    """The C++ binding of INTEGRATION.md (a GpuSweep over any Setka through
    gel_ctx_create_mesh) compiles against the reference's own headers: include/gelios.h
    uses no name that the reference's macros (Help.h) rewrite."""
    import subprocess
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    i = txt.index("// GpuSweep.cpp -- replaces the OpenMP cell loop")
    src = txt[i:txt.index("```", i)]
    src += ("\nint probe(Setka* S) { GpuSweep g(S); gel_stats st; g.put(S, 0); g.steps(3, &st); "
            "g.step_moving(S, &st); g.get(S, 1); return 0; }\n")
    f = tmp_path / "GpuSweep.cpp"
    f.write_text(src)
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-w", "-I/root/reference",
                        "-I" + os.path.join(ROOT, "include"), str(f)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
