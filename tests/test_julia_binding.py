"""The Julia front end (julia/PushPullHIP.jl) against the C ABI it binds.

Julia is not installed here, so the module is checked as text: every library symbol it
calls is declared in include/pushpull_hip.h (and exported by the built library), and its
`Desc` / `Volume` structs, passed by reference to the library, have the field order, types
and byte layout of `pp_desc` / `pp_volume` (checked against the ctypes mirrors that the
GPU tests use, whose sizes the header fixes).
"""
from __future__ import annotations

import ctypes as ct
import os
import re

import pytest

from pushpull import _lib
from pushpull.batch import BATCH_CHAN_BND, PPVolume

JL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "julia", "PushPullHIP.jl")


def _source() -> str:
    with open(JL) as fh:
        return fh.read()


def _called_symbols(src: str) -> set[str]:
    # sym(:name) calls, and the entry-point symbols handed to the shared pull/push wrappers
    syms = set(re.findall(r"sym\(:(\w+)\)", src))
    syms |= set(re.findall(r"_family\(:(\w+)", src))
    return syms


def test_julia_calls_only_declared_symbols():
    syms = _called_symbols(_source())
    assert len(syms) >= 20, sorted(syms)
    declared = set(_lib.declared_functions())
    missing = sorted(s for s in syms if s not in declared)
    assert not missing, f"PushPullHIP.jl calls symbols the header does not declare: {missing}"


def test_julia_symbols_exported_by_library():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = ct.CDLL(_lib.LIB_PATH)
    for s in sorted(_called_symbols(_source())):
        assert hasattr(L, s), s


# Julia field types -> (size, alignment) in bytes, as laid out for ccall (C layout)
_SCALAR = {"Int64": 8, "Int32": 4, "UInt64": 8, "Float32": 4, "Float64": 8, "Cint": 4}


def _jl_field_layout(src: str, name: str, consts: dict[str, int]) -> list[tuple[str, int, int]]:
    m = re.search(rf"^struct {name}\n(.*?)^end", src, flags=re.S | re.M)
    assert m, f"struct {name} not found"
    out = []
    for line in m.group(1).splitlines():
        line = line.split("#")[0].strip()
        if not line:
            continue
        fname, ftype = (t.strip() for t in line.split("::"))
        if ftype.startswith("Ptr{"):
            size = align = 8
        elif ftype.startswith("NTuple{"):
            n, elem = re.match(r"NTuple\{(.+),(\w+)\}", ftype).groups()
            n = eval(n, {}, dict(consts))  # e.g. 3 * PP_MAX_CHAN_BND
            align = _SCALAR[elem]
            size = n * align
        else:
            size = align = _SCALAR[ftype]
        out.append((fname, size, align))
    return out


def _offsets(layout):
    off, res = 0, []
    for fname, size, align in layout:
        off = (off + align - 1) // align * align
        res.append((fname, off, size))
        off += size
    maxal = max(a for _, _, a in layout)
    return res, (off + maxal - 1) // maxal * maxal


@pytest.mark.parametrize("jl_name,ctype", [("Desc", _lib.PPDesc), ("Volume", PPVolume)])
def test_julia_structs_match_c_layout(jl_name, ctype):
    src = _source()
    consts = {k: int(re.search(rf"const {k} = (\d+)", src).group(1)) for k in ("PP_MAX_CHAN_BND", "PP_BATCH_CHAN_BND")}
    assert consts["PP_MAX_CHAN_BND"] == _lib.PP_MAX_CHAN_BND
    assert consts["PP_BATCH_CHAN_BND"] == BATCH_CHAN_BND
    offs, size = _offsets(_jl_field_layout(src, jl_name, consts))
    cfields = [(f[0], getattr(ctype, f[0]).offset, getattr(ctype, f[0]).size) for f in ctype._fields_]
    assert [o[0] for o in offs] == [c[0] for c in cfields], "field order differs from the C struct"
    assert offs == cfields
    assert size == ct.sizeof(ctype)
