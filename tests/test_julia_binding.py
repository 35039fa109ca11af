"""The committed Julia binding (paulioperators.jl_amd/julia/PauliPropAMD.jl)
against the C header: every `ccall` names an entry point that
include/pauliprop.h declares, with the header's number of parameters and a
matching return type, and every declared entry point is bound.  Julia is not
installed here, so this checks the binding's text (balanced brackets, the
ccall signatures); the reference-side semantics it forwards are covered by the
Python mirror's tests."""
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HEADER = os.path.join(ROOT, "include", "pauliprop.h")
BINDING = os.path.join(ROOT, "paulioperators.jl_amd", "julia", "PauliPropAMD.jl")


def strip_c_comments(s):
    return re.sub(r"/\*.*?\*/", " ", s, flags=re.S)


def header_prototypes():
    """name -> (return type, parameter count) for every function declaration."""
    s = strip_c_comments(open(HEADER).read())
    out = {}
    for m in re.finditer(r"\b(int|double|const char\s*\*)\s*(pp_\w+)\s*\(([^;{]*?)\)\s*;", s, flags=re.S):
        ret, name, params = m.group(1), m.group(2), m.group(3).strip()
        n = 0 if params in ("", "void") else params.count(",") + 1
        out[name] = (re.sub(r"\s+", "", ret), n)
    return out


def split_top(s):
    """Split s at commas outside (), {}, []."""
    parts, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({[":
            depth += 1
        elif ch in ")}]":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur)
    return [p.strip() for p in parts]


def matching(s, i):
    """Index of the bracket closing the one at s[i]."""
    pairs = {"(": ")", "{": "}", "[": "]"}
    stack = [pairs[s[i]]]
    j = i + 1
    while stack:
        ch = s[j]
        if ch in pairs:
            stack.append(pairs[ch])
        elif ch in ")}]":
            assert ch == stack.pop(), f"unbalanced at {j}"
        j += 1
    return j - 1


def julia_code():
    """The binding with comments and string literals blanked."""
    s = open(BINDING).read()
    s = re.sub(r'"(?:[^"\\]|\\.)*"', '""', s)
    s = re.sub(r"#[^\n]*", "", s)
    return s


def ccalls():
    """(symbol, return type, [argument types], [arguments]) per ccall."""
    s = julia_code()
    out = []
    for m in re.finditer(r"ccall\(", s):
        i = m.end() - 1
        body = s[i + 1:matching(s, i)]
        parts = split_top(body)
        sym = re.fullmatch(r"\(:(\w+),\s*LIB\)", parts[0])
        assert sym, f"ccall target {parts[0]!r} is not (:symbol, LIB)"
        tup = parts[2]
        assert tup.startswith("(") and tup.endswith(")"), tup
        types = split_top(tup[1:-1])
        out.append((sym.group(1), parts[1], types, parts[3:]))
    return out


JULIA_RET = {"int": "Cint", "double": "Float64", "constchar*": "Cstring"}


def test_binding_brackets_balance():
    s = julia_code()
    stack = []
    pairs = {"(": ")", "{": "}", "[": "]"}
    for ch in s:
        if ch in pairs:
            stack.append(pairs[ch])
        elif ch in ")}]":
            assert stack and stack.pop() == ch
    assert not stack
    assert re.search(r"^module PauliPropAMD\b", s, flags=re.M) and s.rstrip().endswith("end")


def test_every_ccall_matches_the_header():
    protos = header_prototypes()
    calls = ccalls()
    assert len(calls) >= len(protos)
    for sym, ret, types, args in calls:
        assert sym in protos, f"{sym} is not declared in include/pauliprop.h"
        hret, nparams = protos[sym]
        assert len(types) == nparams, f"{sym}: {len(types)} argument types, header has {nparams}"
        assert len(args) == nparams, f"{sym}: {len(args)} arguments, header has {nparams}"
        assert ret == JULIA_RET[hret], f"{sym}: returns {ret}, header {hret}"


def test_every_entry_point_is_bound():
    protos = header_prototypes()
    bound = {c[0] for c in ccalls()}
    missing = sorted(set(protos) - bound)
    assert not missing, f"entry points without a Julia binding: {missing}"


def test_reference_interface_surface():
    """The methods the reference's users call on a SparsePauliVector exist
    with the reference's keywords (spv_evolve.jl:312-331, channels.jl:203-265)."""
    s = julia_code()
    sig = re.search(r"function evolve!\(d::DeviceSPV\{N,W\}, generators.*?\) where \{N,W\}", s, flags=re.S)
    assert sig
    for kw in ("window::Int=1", "truncation::TruncationStrategy=NoTruncation()",
               "local_truncation::TruncationStrategy=NoTruncation()",
               "correction::CorrectionAccumulator=NoCorrection()",
               "counters::Union{Nothing,WindowCounters}=nothing"):
        assert kw in sig.group(0), kw
    body = s[sig.end():s.index("\nend", sig.end())]
    assert "pp_evolve_ex" in body and "_strategy_ops!" in body  # non-compilable truncation route
    for ch in ("depolarizing_channel!", "dephasing_channel!", "bit_flip_channel!", "bit_phase_flip_channel!",
               "pauli_channel!"):
        assert re.search(rf"function {re.escape(ch)}\(d::DeviceSPV", s), ch
    assert s.count("_spv_mask(W, N, qubits)") >= 2
    for name in ("truncate!", "stochastic_clip!", "expectation_value", "Base.getindex", "Base.get",
                 "Base.haskey", "Base.setindex!", "Base.delete!", "Base.sum!", "LinearAlgebra.mul!",
                 "LinearAlgebra.norm", "inner_product", "Base.isapprox", "check_spv", "commutator",
                 "anticommutator"):
        assert re.search(rf"(function {re.escape(name)}\(|^{re.escape(name)}\()", s, flags=re.M), name
