"""The Julia binding (fomo.jl_amd/julia/FomoHIP.jl) against the C-ABI it calls: the FomoDesc and
FomoShot structs mirror include/fomo_hip.h field by field (names, order, types, sizes) and agree
with the ctypes mirror in fomo/_lib.py; every ccall names an exported symbol of the library with
the header's parameter count and compatible parameter types; and the public functions keep the
reference's names, positional arguments and keyword defaults (README.md:274-346).  Julia itself is
not in this image, so the file is checked as text."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "fomo.jl_amd", "julia", "FomoHIP.jl")
HDR = os.path.join(ROOT, "include", "fomo_hip.h")

JL_SIZE = {"Int32": 4, "Float32": 4, "Float64": 8}
JL_TO_C = {"Int32": "int32_t", "Float32": "float", "Float64": "double", "Ptr{Int32}": "const int32_t *",
           "Ptr{Float32}": "const float *", "Ptr{Float64}": "double *"}


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def jl_struct(name):
    src = open(JL).read()
    m = re.search(r"^struct %s\n(.*?)^end" % name, src, re.S | re.M)
    assert m, name
    fields = []
    for line in m.group(1).splitlines():
        line = line.split("#")[0].strip()
        if line:
            n, t = line.split("::")
            fields.append((n.strip(), t.strip()))
    return fields


def c_struct(name):
    src = _strip_comments(open(HDR).read())
    m = re.search(r"typedef struct %s \{(.*?)\} %s;" % (name, name), src, re.S)
    assert m, name
    fields = []
    for decl in m.group(1).split(";"):
        decl = " ".join(decl.split())
        if not decl:
            continue
        mm = re.match(r"(const )?(int32_t|float|double)\s*(.+)", decl)
        assert mm, decl
        base = (mm.group(1) or "") + mm.group(2)
        for var in mm.group(3).split(","):
            var = var.strip()
            star = var.startswith("*")
            var = var.lstrip("*").strip()
            arr = re.match(r"(\w+)\[(\d+)\]", var)
            t = base + (" *" if star else "")
            if arr:
                fields.append((arr.group(1), f"{t}[{arr.group(2)}]"))
            else:
                fields.append((var, t))
    return fields


def jl_type_as_c(t):
    m = re.match(r"NTuple\{(\d+),(\w+)\}", t)
    if m:
        return f"{JL_TO_C[m.group(2)]}[{m.group(1)}]"
    return JL_TO_C[t]


def test_fomodesc_mirrors_header_and_ctypes():
    import fomo._lib as L
    jl, hd = jl_struct("FomoDesc"), c_struct("fomo_desc")
    assert [n for n, _ in jl] == [n for n, _ in hd]
    assert [jl_type_as_c(t) for _, t in jl] == [t for _, t in hd]
    assert [n for n, _ in L.Desc._fields_] == [n for n, _ in jl]
    size = sum(4 * int(re.match(r"NTuple\{(\d+)", t).group(1)) if t.startswith("NTuple") else JL_SIZE[t] for _, t in jl)
    assert size == C.sizeof(L.Desc) == 104


def test_fomoshot_mirrors_header_and_ctypes():
    import fomo._lib as L
    jl, hd = jl_struct("FomoShot"), c_struct("fomo_shot")
    assert [n for n, _ in jl] == [n for n, _ in hd] == [n for n, _ in L.Shot._fields_]
    for (n, tj), (_, tc) in zip(jl, hd):
        assert jl_type_as_c(tj).replace("const ", "") == tc.replace("const ", ""), n
    # natural C alignment (Julia lays out isbits structs the same way): 3 x 4 B, pad, 3 pointers,
    # 4 B + pad, 2 pointers, 4 B + pad
    assert C.sizeof(L.Shot) == 72


def _balanced(text, start):
    depth = 0
    for k in range(start, len(text)):
        if text[k] == "(":
            depth += 1
        elif text[k] == ")":
            depth -= 1
            if depth == 0:
                return text[start + 1:k]
    raise ValueError("unbalanced")


def _split_top(s):
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


def header_functions():
    src = _strip_comments(open(HDR).read())
    out = {}
    for m in re.finditer(r"\b(?:int|const char \*)\s*(fomo_\w+)\s*\(", src):
        params = _balanced(src, m.end() - 1)
        out[m.group(1)] = [] if params.strip() == "void" else [" ".join(p.split()) for p in _split_top(params)]
    return out


JL_PARAM_OK = {
    "Ref{FomoDesc}": lambda c: c.startswith("const fomo_desc *"),
    "Ptr{Float32}": lambda c: re.match(r"(const )?float \*", c) is not None,
    "Ptr{Int32}": lambda c: re.match(r"(const )?int32_t \*", c) is not None,
    "Ptr{Float64}": lambda c: c.startswith("double *"),
    "Ref{Float64}": lambda c: c.startswith("double *"),
    "Int32": lambda c: c.startswith("int32_t "),
    "Ptr{Cvoid}": lambda c: c.startswith("fomo_plan *"),
}


def test_every_ccall_matches_the_header_and_the_library():
    import fomo._lib as L
    src = open(JL).read()
    hdr = header_functions()
    calls = list(re.finditer(r"ccall\(\(:(\w+), LIB\), (Cint|Cstring),", src))
    names = {m.group(1) for m in calls}
    assert {"fomo_acoustic2d", "fomo_elastic2d", "fomo_scalar2d", "fomo_coupled2d", "fomo_acoustic2d_batch",
            "fomo_elastic2d_batch", "fomo_acoustic3d", "fomo_elastic3d", "fomo_last_error",
            "fomo_trace_misfit"} <= names
    so = C.CDLL(L.LIB_PATH)
    for m in calls:
        name = m.group(1)
        assert name in hdr, name
        getattr(so, name)  # exported
        k = src.index("(", m.end())
        types = _split_top(_balanced(src, k))
        assert len(types) == len(hdr[name]), (name, types, hdr[name])
        for tj, tc in zip(types, hdr[name]):
            assert JL_PARAM_OK[tj](tc), (name, tj, tc)


@pytest.mark.parametrize("fn,positional,kw", [
    ("acoustic2d", "vp::Matrix{Float32}, rho::Matrix{Float32}, dh::Float32, dt::Float32, nt::Int, f0::Float32",
     {"nbc": "50", "fd_order": "8", "snap_interval": "0", "boundary": ":habc", "sponge_factor": "0.015f0",
      "wavelet": "nothing", "scale_source_by_dt": "false", "use_cuda_graph": "true", "verbose": "true"}),
    ("elastic2d", "vp::Matrix{Float32}, vs::Matrix{Float32}, rho::Matrix{Float32}, dh::Float32, dt::Float32,",
     {"nbc": "50", "fd_order": "8", "snap_interval": "0", "boundary": ":habc", "sponge_factor": "0.015f0",
      "wavelet": "nothing", "scale_source_by_dt": "false", "use_cuda_graph": "true", "verbose": "true"}),
    ("scalar2d", "vp::AbstractMatrix{Float32}, dh::Float32, dt::Float32, nt::Int, f0::Float32",
     {"nbc": "50", "fd_order": "8", "wavelet": "nothing", "source_scale": ":v2dt2", "verbose": "true"}),
    ("coupled2d", "vp::Matrix{Float32}, vs::Matrix{Float32}, dh::Float32, dt::Float32, nt::Int, f0::Float32",
     {"nbc": "50", "fd_order": "8", "snap_interval": "0", "smooth_sigma": "3.0", "wavelet": "nothing",
      "verbose": "true"}),
    ("acoustic2d_batch", "vp::AbstractMatrix{Float32}, rho::AbstractMatrix{Float32}, dh::Float32, dt::Float32,",
     {"nbc": "50", "fd_order": "8", "wavelet": "nothing", "scale_source_by_dt": "false", "verbose": "true"}),
    ("elastic2d_batch", "vp::AbstractMatrix{Float32}, vs::AbstractMatrix{Float32}, rho::AbstractMatrix{Float32},",
     {"nbc": "50", "fd_order": "8", "wavelet": "nothing", "scale_source_by_dt": "false", "verbose": "true"}),
    ("acoustic3d", "vp::Array{Float32,3}, rho::Array{Float32,3}, dh::Float32, dt::Float32, nt::Int, f0::Float32",
     {"nbc": "50", "fd_order": "8", "snap_interval": "0", "snap_plane": ":xz", "snap_index": "0"}),
    ("elastic3d", "vp::Array{Float32,3}, vs::Array{Float32,3}, rho::Array{Float32,3}, dh::Float32, dt::Float32,",
     {"nbc": "50", "fd_order": "8", "snap_interval": "0", "snap_plane": ":xz", "snap_index": "0"}),
])
def test_public_signatures_keep_the_reference_contract(fn, positional, kw):
    src = open(JL).read()
    m = re.search(r"^function %s\(" % fn, src, re.M)
    assert m, fn
    sig = " ".join(_balanced(src, m.end() - 1).split())
    assert sig.startswith(" ".join(positional.split())), sig[:120]
    for k, v in kw.items():
        assert re.search(r"\b%s::[^=]+=%s(,|$)" % (k, re.escape(v)), sig), (fn, k)
    # the returns the reference documents
    body = src[m.end():src.index("\nend\n", m.end())]
    ret = body[body.rindex("return"):]
    want = {"acoustic2d": "(seis_p=seis_p, seis_vx=seis_vx, seis_vz=seis_vz, snaps=snaps, stats=(kernel_time_s=t[],))",
            "elastic2d": "(seis_vx=seis_vx, seis_vz=seis_vz, snaps=snaps, stats=(kernel_time_s=t[],))",
            "scalar2d": "(seis_u=seis_u, stats=(kernel_time_s=t[],))",
            "coupled2d": "(seis_P=seis_P, seis_S=seis_S, snaps_P=snaps_P, snaps_S=snaps_S, stats=(kernel_time_s=t[],))",
            "acoustic2d_batch": "(seis_p=seis_p, seis_vx=seis_vx, seis_vz=seis_vz, stats=(kernel_time_s=t[], n_shots=ns))",
            "elastic2d_batch": "(seis_vx=seis_vx, seis_vz=seis_vz, stats=(kernel_time_s=t[], n_shots=ns))",
            "acoustic3d": "seis_vx, seis_vy, seis_vz, snaps",
            "elastic3d": "seis_vx, seis_vy, seis_vz, snaps"}[fn]
    assert want in ret, ret
