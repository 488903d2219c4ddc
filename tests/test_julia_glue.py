"""The Julia glue (julia/AvSVD.jl) against the C header, without Julia.

Julia is not installed here, so the glue is checked as text: every ccall names
a function include/avsvd.h declares with the same number of arguments, and the
AvsOpts mirror lists avs_opts's fields in order with matching widths.
"""
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
JL = os.path.join(ROOT, "avocado-sandwich_amd", "julia", "AvSVD.jl")
HDR = os.path.join(ROOT, "include", "avsvd.h")


def _header_prototypes():
    src = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    protos = {}
    for m in re.finditer(r"\b(?:avs_status|void|const char \*)\s*(avs_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", src):
        args = [a for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        protos[m.group(1)] = len(args)
    return protos


def _ccalls(src):
    """(symbol, number of argument types) of every ccall((:sym, lib), Ret, (types...), ...)"""
    out = []
    for m in re.finditer(r"ccall\(\((:?)(\w+|sym), libavsvd\),\s*[\w{}]+,\s*\(", src):
        i = m.end()
        depth, j = 1, i
        while depth:
            depth += {"(": 1, ")": -1}.get(src[j], 0)
            j += 1
        types = src[i:j - 1]
        n = len([t for t in re.split(r",(?![^{]*})", types) if t.strip()])
        out.append((m.group(2), n))
    return out


def test_every_ccall_matches_a_header_prototype():
    protos = _header_prototypes()
    src = open(JL).read()
    calls = _ccalls(src)
    assert len(calls) >= 15
    for sym, n in calls:
        if sym == "sym":
            continue  # symbols passed in a variable are checked below
        assert sym in protos, sym
        assert protos[sym] == n, (sym, protos[sym], n)
    # symbols the glue passes through a variable
    for sym in re.findall(r":(avs_[a-z0-9_]+)\b", src):
        assert sym in protos, sym
    # the streamed OOC entry points and the device-array path are bound
    for sym in ("avs_ooc_blockbidiag", "avs_ooc_tiled_qr", "avs_set_matrix_device", "avs_get_matrix_device"):
        assert ":" + sym in src, sym


def test_avsopts_mirror_matches_the_header_struct():
    hdr = re.sub(r"/\*.*?\*/", "", open(HDR).read(), flags=re.S)
    body = re.search(r"typedef struct avs_opts\s*{(.*?)}\s*avs_opts;", hdr, re.S).group(1)
    cfields = re.findall(r"(void \*|int32_t)\s*(\w+)(\[\d+\])?;", body)
    jl = open(JL).read()
    jbody = re.search(r"struct AvsOpts\n(.*?)\nend", jl, re.S).group(1)
    jfields = re.findall(r"(\w+)::([\w{},]+)", jbody)
    assert [f[1] for f in cfields] == [f[0] for f in jfields]
    for (ct, name, arr), (jname, jt) in zip(cfields, jfields):
        if ct == "void *":
            assert jt == "Ptr{Cvoid}", name
        elif arr:
            assert jt == "NTuple{%s,Int32}" % arr.strip("[]"), name
        else:
            assert jt == "Int32", name
