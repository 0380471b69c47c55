"""The committed Julia binding (forceconstants.jl_amd/julia/FC3AMD.jl) on the CPU.

Julia is not installed here, so the file is checked as text:
  - the reference's public signatures appear in it verbatim (src/ifc.jl:3-5,20-22,
    third_order_analytical.jl:185, mcc.jl:29,60,89, io.jl:9,39, asr.jl:43,62), with the
    library's DeviceArray in place of CuArray in the mcc3 methods;
  - every `ccall((:fc3_..., libfc3), Cint, (types...), args...)` names an entry point
    that include/fc3.h declares, with as many argument types as the C declaration has
    parameters, and passes that many arguments.
"""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL_DIR = os.path.join(ROOT, "forceconstants.jl_amd", "julia")
HEADER = os.path.join(ROOT, "include", "fc3.h")

SIGNATURES = [
    "function second_order(sys::SuperCellSystem{D}, pot::Potential, "
    "calc::ForceConstantCalculator; n_threads::Int = Threads.nthreads(), "
    "check_asr = false, check_symmetry = false) where D",
    "function third_order(sys::SuperCellSystem{D}, pot::Potential, "
    "calc::ForceConstantCalculator; n_threads::Int = Threads.nthreads(), "
    "check_asr = false) where D",
    "function mass_weight_third_order!(Ψ::ThirdOrderForceConstants, masses::AbstractVector)",
    "function mcc3(Ψ::DeviceArray{Float32, 3}, phi::DeviceArray{Float32, 2})",
    "function mcc3!(Ψ::DeviceArray{Float32, 3}, phi::DeviceArray{Float32, 2})",
    "function mcc3(Ψ::DeviceArray{Float32, 3}, phi::DeviceArray{Float32, 2}, block_size::Int)",
    "function save_second_order(Φ::SecondOrderForceConstants, N_atoms, D, outpath; "
    "filename = \"second_order\", fmt = :JLD2)",
    "function save_third_order(Ψ::ThirdOrderForceConstants, N_atoms, D, outpath; "
    "filename = \"third_order\", fmt = :JLD2)",
    "asr_satisfied(ifc2::AbstractArray{T,2}, N_atoms, D, tol; verbose = false) where {T}",
    "asr_satisfied(ifc3::Array{T,3}, N_atoms, D, tol; verbose = false) where {T}",
]


def _norm(s):
    return re.sub(r"\s+", " ", s)


def _jl_sources():
    return {f: open(os.path.join(JL_DIR, f), encoding="utf-8").read()
            for f in sorted(os.listdir(JL_DIR)) if f.endswith(".jl")}


def _split_top(s):
    """Split at commas outside (), [], {}."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return [x.strip() for x in out if x.strip()]


def _balanced(s, i):
    """s[i] == '(': index just past its matching ')'."""
    depth = 0
    for j in range(i, len(s)):
        if s[j] == "(":
            depth += 1
        elif s[j] == ")":
            depth -= 1
            if depth == 0:
                return j + 1
    raise AssertionError("unbalanced parentheses")


def _header_params():
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"\b(fc3_[a-z0-9_]+)\s*\(", txt):
        end = _balanced(txt, m.end() - 1)
        params = txt[m.end():end - 1].strip()
        out[m.group(1)] = 0 if params in ("", "void") else len(_split_top(params))
    return out


def test_reference_signatures_verbatim():
    text = _norm(_jl_sources()["FC3AMD.jl"])
    for sig in SIGNATURES:
        assert _norm(sig) in text, sig


def test_ccalls_match_the_header():
    decl = _header_params()
    n_calls = 0
    for name, src in _jl_sources().items():
        for m in re.finditer(r"ccall\(\(:(fc3_[a-z0-9_]+), libfc3\)", src):
            sym = m.group(1)
            assert sym in decl, f"{name}: {sym} is not declared in fc3.h"
            start = m.start() + len("ccall")
            body = src[start + 1:_balanced(src, start) - 1]
            parts = _split_top(body)
            # parts: (:sym, libfc3), return type, (arg types...), args...
            types = parts[2]
            assert types.startswith("(") and types.endswith(")"), (name, sym, types)
            n_types = len(_split_top(types[1:-1]))
            n_args = len(parts) - 3
            assert n_types == decl[sym], f"{name}: {sym} has {n_types} types, fc3.h {decl[sym]}"
            assert n_args == n_types, f"{name}: {sym} passes {n_args} arguments for {n_types}"
            n_calls += 1
    assert n_calls >= 20


@pytest.mark.parametrize("name", ["FC3AMD.jl", "FC3AMDChainRules.jl"])
def test_balanced_blocks(name):
    """Every `function`/`module`/`struct`/`if`/`for`/`begin` has its `end` (a cheap
    syntax check without a Julia parser: keywords counted outside strings and comments)."""
    src = _jl_sources()[name]
    src = re.sub(r'"""[\s\S]*?"""', '""', src)
    src = re.sub(r'"(?:\\.|[^"\\])*"', '""', src)
    src = re.sub(r"#.*", "", src)
    opens = len(re.findall(r"(?m)^\s*(?:mutable struct|struct|function|module|if|for|while|"
                           r"let|begin|try)\b", src))
    opens += len(re.findall(r"\bdo\s*(?:\w+\s*)?$", src, flags=re.M))
    ends = len(re.findall(r"\bend\b", src))
    assert opens == ends, (opens, ends)
