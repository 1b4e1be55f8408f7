"""The Julia binding (draugr.jl_amd/julia/DraugrHIP) cannot run here (no Julia
toolchain), so its ccall signatures are checked as text against the C
declarations of include/draugr_hip.h: every ccall names an exported entry
point, with the header's argument count, argument types and return type."""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "draugr.jl_amd", "julia", "DraugrHIP", "src", "DraugrHIP.jl")
HEADER = os.path.join(ROOT, "include", "draugr_hip.h")

# Julia ccall type -> C type as written in the header
JL2C = {
    "Int32": "int32_t", "Int64": "int64_t", "Float64": "double", "Cint": "int", "Cvoid": "void",
    "Ptr{Float64}": "double*", "Ptr{Int32}": "int32_t*", "Ptr{Int64}": "int64_t*", "Ptr{Cvoid}": "void*",
    "Cstring": "char*", "Ptr{UInt8}": "char*", "Ref{Float64}": "double*", "Ref{Int32}": "int32_t*", "Ptr{Ptr{Int32}}": "int32_t**", "Ptr{Ptr{Float64}}": "double**",
}


def header_decls():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    decls = {}
    for m in re.finditer(r"([\w\s\*]+?)\b(\w+)\s*\(([^)]*)\)\s*;", text):
        ret, name, args = m.group(1), m.group(2), m.group(3)
        norm = lambda s: re.sub(r"\bconst\b", "", s).replace(" ", "")  # noqa: E731
        types = []
        for a in args.split(","):
            a = a.strip()
            if not a or a == "void":
                continue
            t = re.sub(r"\b\w+$", "", a)  # drop the parameter name
            types.append(norm(t))
        decls[name] = (norm(ret.strip().split("\n")[-1]), types)
    return decls


def julia_ccalls():
    text = open(JL).read()
    calls = []
    for m in re.finditer(r"ccall\(\(:(\w+),\s*libdraugr_hip\),\s*([\w{}]+),\s*\(([^()]*)\)", text):
        name, ret, args = m.group(1), m.group(2), m.group(3)
        types = [a.strip() for a in args.split(",") if a.strip()]
        calls.append((name, ret, types))
    return calls


def test_every_julia_ccall_matches_the_header():
    decls = header_decls()
    calls = julia_ccalls()
    assert len(calls) >= 15
    for name, ret, types in calls:
        assert name in decls, name
        cret, ctypes_ = decls[name]
        assert len(types) == len(ctypes_), (name, types, ctypes_)
        for jt, ct in zip(types, ctypes_):
            assert JL2C.get(jt) == ct, (name, jt, ct)
        assert JL2C.get(ret) == cret, (name, ret, cret)
