"""The Julia package extension (conservativeregridding.jl_amd/julia/) against the
C ABI, read as text (no Julia on this host): every `ccall` names an entry point
of include/crhip.h with the same number of arguments, each of the same kind
(pointer, 32-bit or 64-bit integer, double) as the library's ctypes signature,
and every name the extension calls is defined in it or imported."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXT = os.path.join(ROOT, "conservativeregridding.jl_amd", "julia", "ConservativeRegriddingMI355XExt.jl")


def _split_top(s):
    """Split a comma list at depth 0 of (), {}."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({":
            depth += 1
        elif ch in ")}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def _balanced(text, i):
    """The text inside the parenthesis that opens at text[i]."""
    assert text[i] == "("
    depth = 0
    for j in range(i, len(text)):
        depth += text[j] == "("
        depth -= text[j] == ")"
        if depth == 0:
            return text[i + 1:j]
    raise ValueError("unbalanced")


def ccalls():
    text = open(EXT).read()
    out = []
    for m in re.finditer(r"ccall\(", text):
        body = _split_top(_balanced(text, m.end() - 1))
        name = re.match(r"\(:(\w+),\s*libcrhip\)", body[0]).group(1)
        argt = body[2]
        assert argt.startswith("(") and argt.endswith(")"), (name, argt)
        types = _split_top(argt[1:-1])
        out.append((name, body[1], types, body[3:]))
    return out


def _kind_julia(t):
    if t.startswith(("Ptr{", "Ref{")) or t == "Cstring":
        return "ptr"
    return {"Int64": "i64", "Int32": "i32", "Cint": "i32", "Float64": "f64", "Cvoid": "void"}[t]


def _kind_ctypes(t):
    if t is None:
        return "void"
    if t in (ctypes.c_int64, ctypes.c_longlong):
        return "i64"
    if t in (ctypes.c_int32, ctypes.c_int):
        return "i32"
    if t is ctypes.c_double:
        return "f64"
    return "ptr"


def test_every_ccall_matches_the_c_abi(cr):
    from importlib import import_module
    sig = import_module("crhip._lib").SIGNATURES
    calls = ccalls()
    assert len(calls) >= 15
    for name, ret, types, args in calls:
        assert name in sig, f"{name} is not in include/crhip.h"
        cret, cargs = sig[name]
        assert _kind_julia(ret) == _kind_ctypes(cret), name
        assert [_kind_julia(t) for t in types] == [_kind_ctypes(t) for t in cargs], name
        # one argument per declared type (a tuple of types is ccall's contract)
        assert len(args) == len(types), f"{name}: {len(args)} arguments for {len(types)} types"


def test_the_device_path_is_one_regrid_per_launch():
    names = [c[0] for c in ccalls()]
    for must in ("cr_regridder_build", "cr_regrid", "cr_regrid_backward", "cr_regridder_normalize",
                 "cr_grid_from_points", "cr_grid_from_polygons", "cr_regridder_from_esmf"):
        assert must in names, must
    text = open(EXT).read()
    # perform_regridding! on the slicers holds exactly one cr_regrid call
    m = re.search(r"function CR\.perform_regridding!\(dst::DeviceSlicer.*?\nend\n", text, re.S)
    assert m and m.group(0).count("ccall((:cr_regrid,") == 1
    for f in ("extract_source_arraylike", "extract_dest_arraylike", "slice_views", "regrid"):
        assert re.search(rf"function CR\.{f}\(", text), f


JULIA_BUILTINS = {
    "ccall", "check", "throw", "ArgumentError", "DimensionMismatch", "ErrorException", "error", "print", "get",
    "unsafe_string", "Ref", "Ptr", "pointer", "finalizer", "new", "size", "zeros", "stride", "ndims", "length",
    "sum", "Tuple", "axes", "view", "ntuple", "CartesianIndices", "isempty", "collect", "reshape", "Array",
    "similar", "Vector", "Int", "Int32", "Int64", "Float64", "Matrix", "Dict", "push!", "append!", "get!",
    "transpose", "unthunk", "ROCArray", "Regridder", "NoTangent", "UnitSphericalPoint", "Colon", "sizeof",
    "Cstring", "Cint", "Cvoid", "string", "GC", "typeof", "nothing",
}


def test_every_called_name_is_defined_or_imported():
    text = open(EXT).read()
    code = re.sub(r'"""[\s\S]*?"""', "", text)
    code = re.sub(r'"(?:[^"\\]|\\.)*"', '""', code)
    code = re.sub(r"#.*", "", code)
    defined = set(re.findall(r"^\s*(?:function\s+|struct\s+|mutable struct\s+|const\s+)([A-Za-z_]\w*!?)", code, re.M))
    defined |= set(re.findall(r"^\s*([a-z_]\w*!?)\(.*?\)\s*(?:where.*?)?=", code, re.M))
    defined |= set(re.findall(r"\b([a-z_]\w*)\(x\)\s*=", code))  # local one-line helpers
    called = set(re.findall(r"(?<![\w.:])([A-Za-z_]\w*!?)\(", code))
    qualified_ok = called - defined - JULIA_BUILTINS
    assert not qualified_ok, f"names called but not defined: {sorted(qualified_ok)}"
