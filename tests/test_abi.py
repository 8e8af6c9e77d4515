"""The C-ABI library loads and exports every symbol include/gtk_hip.h
declares (no GPU call)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "gtk_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gtk_[a-z_]+)\s*\(", src)))


def test_header_symbols_exported(GT):
    L = GT._lib.load()
    names = declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert sorted(GT._lib.EXPORTS) == names


def test_errors_without_device(GT):
    """gtk_create reports an error instead of crashing when no GPU is usable."""
    import pytest
    L = GT._lib.load()
    h = ctypes.c_void_p()
    rc = L.gtk_create(10_000, ctypes.byref(h))
    assert rc != 0
    assert L.gtk_last_error(None)
    assert not h.value


def test_julia_binding_matches_header():
    """julia/GTKHip.jl (the binding a GalerkinToolkit.jl caller loads; no
    Julia in this image, so it is checked as text) calls only entry points
    the header declares, and binds the ones its assembly path needs."""
    src = open(os.path.join(ROOT, "julia", "GTKHip.jl")).read()
    called = set(re.findall(r"ccall\(\(:(gtk_[a-z_]+),\s*lib\)", src))
    names = set(declared())
    assert called and called <= names, sorted(called - names)
    for need in ("gtk_create", "gtk_destroy", "gtk_last_error", "gtk_set_mesh", "gtk_set_space", "gtk_set_reference",
                 "gtk_build_pattern", "gtk_get_pattern", "gtk_set_cell_coefficients", "gtk_assemble_matrix",
                 "gtk_assemble_vector", "gtk_lift_dirichlet", "gtk_solve", "gtk_set_values"):
        assert need in called, need
