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
                 "gtk_assemble_vector", "gtk_lift_dirichlet", "gtk_solve", "gtk_set_values",
                 "gtk_set_reference_faces", "gtk_set_reference_face_gradients", "gtk_add_block", "gtk_gather_block",
                 "gtk_update_coordinates", "gtk_block_multiply", "gtk_get_values"):
        assert need in called, need


def _split_top(s):
    """Split s at the commas outside (), {} and []."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def _balanced(s, i):
    """The text of the parenthesised group opening at s[i]."""
    depth = 0
    for j in range(i, len(s)):
        depth += s[j] == "("
        depth -= s[j] == ")"
        if depth == 0:
            return s[i + 1:j]
    raise ValueError("unbalanced")


def header_arity():
    src = open(os.path.join(ROOT, "include", "gtk_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(gtk_[a-z_]+)\s*\(", src):
        args = _balanced(src, m.end() - 1).strip()
        if m.group(1) in out or args.startswith("*"):
            continue
        out[m.group(1)] = 0 if args in ("", "void") else len(_split_top(args))
    return out


def test_julia_ccall_arity_matches_header():
    """Every ccall of julia/GTKHip.jl passes as many argument types, and as
    many arguments, as the header's prototype has parameters."""
    src = open(os.path.join(ROOT, "julia", "GTKHip.jl")).read()
    arity = header_arity()
    calls = 0
    for m in re.finditer(r"ccall\(\(:(gtk_[a-z_]+),\s*lib\)", src):
        body = _split_top(_balanced(src, m.start() + len("ccall")))
        name = m.group(1)
        types = body[2]
        assert types.startswith("(") and types.endswith(")"), (name, types)
        ntypes = len(_split_top(types[1:-1]))
        assert ntypes == arity[name], (name, ntypes, arity[name])
        assert len(body) - 3 == ntypes, (name, body)
        calls += 1
    assert calls >= 30
