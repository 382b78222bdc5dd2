"""The Julia binding (julia/*.jl) cannot run in this image (no Julia); its
`ccall`s are checked here against the C declarations in include/csg.h: every
called symbol is declared, and its return type and argument count and types
map onto the C prototype.  The context cache bound is checked by reading the
binding's code paths that create contexts.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL_DIR = os.path.join(ROOT, "embedded-spinodal-core-shell-julia_amd", "julia")
JL_FILES = [os.path.join(JL_DIR, f) for f in ("CoreShellHIP.jl", "CoreShellHIPChainRules.jl")]

# C type (normalised) -> the Julia ccall types that pass it correctly
C_TO_JL = {
    "int": {"Cint"},
    "int32_t": {"Int32", "Cint"},
    "int64_t": {"Int64"},
    "double": {"Float64", "Cdouble"},
    "void": {"Cvoid"},
    "const char*": {"Cstring", "Ptr{UInt8}"},
    "double*": {"Ptr{Float64}", "Ref{Float64}"},
    "const double*": {"Ptr{Float64}", "Ref{Float64}"},
    "double* const*": {"Ptr{Ptr{Float64}}"},
    "int32_t*": {"Ptr{Int32}", "Ref{Int32}"},
    "const int32_t*": {"Ptr{Int32}", "Ref{Int32}"},
    "csg_ctx*": {"Ptr{Cvoid}"},
    "const csg_ctx*": {"Ptr{Cvoid}"},
    "csg_batch*": {"Ptr{Cvoid}"},
    "const csg_batch*": {"Ptr{Cvoid}"},
    "csg_ctx**": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"},
    "csg_batch**": {"Ref{Ptr{Cvoid}}", "Ptr{Ptr{Cvoid}}"},
    "csg_params*": {"Ref{Parameters}", "Ptr{Parameters}"},
    "const csg_params*": {"Ref{Parameters}", "Ptr{Parameters}"},
    "void*": {"Ptr{Cvoid}"},
}


def _norm_c(t: str) -> str:
    t = re.sub(r"\s+", " ", t.strip())
    t = re.sub(r"\s*\*\s*", "*", t)
    t = t.replace("*const*", "* const*")
    return t


def _c_decls() -> dict:
    text = open(os.path.join(ROOT, "include", "csg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"#.*", "", text)
    out = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b(csg_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", text):
        ret, name, args = _norm_c(m.group(1)), m.group(2), m.group(3).strip()
        ret = ret.replace("typedef ", "")
        types = []
        if args and args != "void":
            for a in args.split(","):
                a = a.strip()
                # drop the parameter name (the last identifier)
                a = re.sub(r"\b[A-Za-z_]\w*$", "", a).strip()
                types.append(_norm_c(a))
        out[name] = (ret, types)
    return out


def _split_types(tup: str) -> list:
    parts, depth, cur = [], 0, ""
    for ch in tup:
        if ch == "{":
            depth += 1
        elif ch == "}":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur.strip())
    return parts


CCALL = re.compile(r"ccall\(\s*\(:(csg_\w+),\s*libcsg\)\s*,\s*([\w{}]+)\s*,\s*\(([^()]*)\)", re.S)


def _jl_ccalls() -> list:
    calls = []
    for f in JL_FILES:
        text = open(f).read()
        text = re.sub(r"#.*", "", text)
        for m in CCALL.finditer(text):
            calls.append((os.path.basename(f), m.group(1), m.group(2), _split_types(m.group(3))))
        # pointwise(sym, ...) calls through a function pointer with one signature
        pm = re.search(r"function pointwise\(.*?ccall\(fptr,\s*([\w{}]+)\s*,\s*\(([^()]*)\)", text, re.S)
        if pm:
            for sym in re.findall(r"pointwise\(:(csg_\w+)", text):
                calls.append((os.path.basename(f), sym, pm.group(1), _split_types(pm.group(2))))
    return calls


def test_julia_ccalls_match_c_header():
    decls = _c_decls()
    calls = _jl_ccalls()
    assert len(calls) >= 20
    seen = set()
    for fname, sym, ret, args in calls:
        assert sym in decls, f"{fname}: {sym} is not declared in csg.h"
        cret, cargs = decls[sym]
        assert ret in C_TO_JL[cret], f"{fname}: {sym} returns {cret}, ccall says {ret}"
        assert len(args) == len(cargs), f"{fname}: {sym} takes {len(cargs)} arguments, ccall passes {len(args)}"
        for i, (ja, ca) in enumerate(zip(args, cargs)):
            assert ca in C_TO_JL, f"unmapped C type {ca!r} of {sym}"
            assert ja in C_TO_JL[ca], f"{fname}: {sym} argument {i}: C {ca}, ccall {ja}"
        seen.add(sym)
    # the drop-in path is bound
    for sym in ("csg_create", "csg_destroy", "csg_set_fields", "csg_run", "csg_step", "csg_get_field",
                "csg_set_params", "csg_energy", "csg_transport_rate", "csg_chemical_potential",
                "csg_adjoint_forward", "csg_adjoint_backward", "csg_batch_run"):
        assert sym in seen, sym


def test_checker_catches_a_wrong_signature():
    """A planted mismatch (csg_step with an Int32 step count) fails the map."""
    decls = _c_decls()
    cret, cargs = decls["csg_step"]
    assert cargs == ["csg_ctx*", "double", "int64_t"]
    assert "Int32" not in C_TO_JL[cargs[2]]


def test_julia_contexts_are_cached_not_created_per_call():
    """Context(...) is constructed only inside the bounded cache helper; the
    function API and the adjoint rule go through it."""
    main = open(JL_FILES[0]).read()
    rules = open(JL_FILES[1]).read()
    code = re.sub(r"#.*", "", main)
    ctor_calls = [m.start() for m in re.finditer(r"(?<![\w.])Context\(n, spacing", code)]
    assert len(ctor_calls) == 1
    fn = code.rfind("function ", 0, ctor_calls[0])
    assert code[fn:].startswith("function cached_context")
    assert "const MAX_CONTEXTS = " in main and "destroy!(pop!(cache, popfirst!(order)))" in main
    rcode = re.sub(r"#.*", "", rules)
    assert not re.search(r"(?<![\w.])Context\(", rcode.replace("Context,", "")), "the rule builds contexts itself"
    assert "geometry(phi, hi)" not in rcode and "geometry(phi, lo)" not in rcode
