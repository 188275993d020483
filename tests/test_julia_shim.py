"""The Julia binding (QCLHip.jl) calls only declared entry points, each with
the C prototype's arity and argument kinds (no Julia runtime is needed: the
ccall type tuples are read from the source)."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "qclnegf.jl_amd", "julia", "QCLHip.jl")
HEADER = os.path.join(ROOT, "include", "qclhip.h")


def _split_top(text, opens="({", closes=")}"):
    parts, depth, cur = [], 0, ""
    for ch in text:
        if ch in opens:
            depth += 1
        elif ch in closes:
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur.strip())
    return parts


def _balanced(text, start):
    """The text inside the parentheses opening at text[start]."""
    depth = 0
    for i in range(start, len(text)):
        if text[i] == "(":
            depth += 1
        elif text[i] == ")":
            depth -= 1
            if depth == 0:
                return text[start + 1:i]
    raise AssertionError("unbalanced parentheses")


def prototypes():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"\b(qclhip_[a-z0-9_]+)\s*\(", text):
        params = _split_top(_balanced(text, m.end() - 1))
        if params == ["void"]:
            params = []
        out[m.group(1)] = params
    return out


def _c_kind(param):
    if "*" in param:
        return "ptr"
    base = param.replace("const", "").split()[0]
    return {"double": "f64", "int": "i32", "int64_t": "i64"}[base]


def _jl_kind(t):
    if t.startswith(("Ptr", "Ref", "Cstring")):
        return "ptr"
    return {"Cdouble": "f64", "Float64": "f64", "Cint": "i32", "Int32": "i32", "Int64": "i64"}[t]


def ccalls():
    """(symbols, argument-type tuple) of every ccall in the shim."""
    src = open(SHIM).read()
    calls = []
    for m in re.finditer(r"\bccall\(", src):
        args = _split_top(_balanced(src, m.end() - 1))
        target, types = args[0], args[2]
        assert types.startswith("(") and types.endswith(")"), types
        tys = _split_top(types[1:-1])
        sym = re.match(r"\(:(qclhip_[a-z0-9_]+),\s*libqclhip\)", target)
        if sym:
            names = [sym.group(1)]
        else:
            assert target.startswith("_entry("), target
            # the entries chosen at run time are the symbols named in the
            # expression that picks them
            block = src[max(0, m.start() - 800):m.start()]
            names = sorted(set(re.findall(r":(qclhip_[a-z0-9_]+)", block)))
            assert names, "no candidate entries before a run-time ccall"
        calls.append((names, tys))
    return calls


def test_every_shim_call_matches_its_prototype():
    protos = prototypes()
    calls = ccalls()
    assert len(calls) > 60
    for names, tys in calls:
        for name in names:
            assert name in protos, f"{name} is not declared in qclhip.h"
            params = protos[name]
            assert len(params) == len(tys), (name, len(params), len(tys))
            for p, t in zip(params, tys):
                assert _c_kind(p) == _jl_kind(t), (name, p, t)


def test_shim_covers_the_operator_entries():
    called = {n for names, _ in ccalls() for n in names}
    protos = set(prototypes())
    # entries a Julia caller reaches through other means or that are
    # Python/bench instrumentation
    uncovered = sorted(protos - called)
    for required in ("qclhip_current_green", "qclhip_current_green_noalloc", "qclhip_current_green_scalar_lesser",
                     "qclhip_current_green_adjoint", "qclhip_batch_current_green", "qclhip_graph_launch",
                     "qclhip_residual_state", "qclhip_residual_keldysh", "qclhip_mix_fields_residual_accumulate",
                     "qclhip_grid_autotune", "qclhip_seed_weights", "qclhip_embedding_adjoint"):
        assert required in called, required
    assert len(uncovered) < len(protos) // 3, uncovered
