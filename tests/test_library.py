"""The C-ABI library: it loads, it exports every entry point the header
declares, and its argument checks answer without touching a device."""
import ctypes
import os
import re

import pytest

import funicular_hip
from funicular_hip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "funicular_hip.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(fhip_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_what_the_binding_binds():
    assert set(declared()) == set(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared():
        assert hasattr(lib, name), name
    assert b"gfx950" in lib.fhip_version()


def test_argument_errors_need_no_device():
    lib = _lib.load()
    n = ctypes.c_size_t()
    assert lib.fhip_gram_workspace_bytes(ctypes.byref(n), 10, 4, 4, 7, 7, 1) == _lib.FHIP_E_UNSUPPORTED
    assert b"no kernel" in lib.fhip_last_error()
    assert lib.fhip_gram_workspace_bytes(ctypes.byref(n), -1, 4, 4, 3, 3, 1) == _lib.FHIP_E_ARGUMENT
    assert lib.fhip_gram_workspace_bytes(ctypes.byref(n), 1000, 64, 64, 3, 3, 1) == 0
    assert n.value > 0
    assert lib.fhip_trsm_runn(None, 1, 1, 4, None, 4, 3, 1, None, 0, None) == _lib.FHIP_E_UNSUPPORTED
    assert lib.fhip_trsm_runn(None, 1, 1, 4, None, 4, 3, 2, None, 0, None) == _lib.FHIP_E_ARGUMENT
    assert lib.fhip_ghost_randn(None, 3, 10, 0, 10, 0, 1, 0, None) == _lib.FHIP_E_ARGUMENT
    # The allocation-free solve refuses what its fixed scratch cannot hold,
    # and bad arguments, before it looks for a device.
    assert lib.fhip_trsm_runn_noalloc(None, 8, 8, 1025, None, 1025, 3, 3, None) == _lib.FHIP_E_ARGUMENT
    assert b"scratch" in lib.fhip_last_error()
    assert lib.fhip_trsm_runn_noalloc(None, 8, 8, 16, None, 16, 3, 3, None) == _lib.FHIP_E_ARGUMENT
    assert b"null matrix" in lib.fhip_last_error()
    assert lib.fhip_trsm_runn_noalloc(None, 8, 8, 16, None, 16, 3, 1, None) == _lib.FHIP_E_UNSUPPORTED
    assert lib.fhip_trsm_runn_noalloc(None, 8, 8, 0, None, 16, 3, 3, None) == 0
    with pytest.raises(_lib.LibraryError) as e:
        _lib.call("fhip_convert_copy", None, 0, 1, None, 3, 1, 1, 1, None)
    assert "null pointer" in str(e.value)
    # The elementwise updates: empty blocks are no-ops, bad shapes are errors.
    assert lib.fhip_affine(None, 1, 0, 5, 2.0, 0.0, 1, 1.0, 0.0, 3, None) == 0
    assert lib.fhip_affine(None, 4, 4, 1, 2.0, 0.0, 1, 1.0, 0.0, 3, None) == _lib.FHIP_E_ARGUMENT
    assert lib.fhip_affine(None, 4, -1, 1, 2.0, 0.0, 1, 1.0, 0.0, 3, None) == _lib.FHIP_E_ARGUMENT
    assert lib.fhip_affine(None, 4, 4, 1, 2.0, 0.0, 1, 1.0, 0.0, 9, None) == _lib.FHIP_E_UNSUPPORTED
    # A leading dimension short of the rows is refused before any device work
    # (the address is never used).
    assert lib.fhip_scale(ctypes.c_void_p(256), 2, 4, 1, 2.0, 0.0, 1, 3, None) == _lib.FHIP_E_ARGUMENT
    assert b"leading dimension" in lib.fhip_last_error()
    assert lib.fhip_affine(ctypes.c_void_p(256), 2, 4, 1, 2.0, 0.0, 1, 1.0, 0.0, 3, None) == _lib.FHIP_E_ARGUMENT
    assert b"leading dimension" in lib.fhip_last_error()


def test_workspace_depends_on_shape_only():
    # The Gram's chunking is fixed by (m, ka, kb): the sums never depend on
    # the workspace a caller happens to pass.
    lib = _lib.load()
    a, b = ctypes.c_size_t(), ctypes.c_size_t()
    lib.fhip_gram_workspace_bytes(ctypes.byref(a), 629856, 512, 512, 3, 3, 1)
    lib.fhip_gram_workspace_bytes(ctypes.byref(b), 629856, 512, 512, 3, 3, 1)
    assert a.value == b.value
    lib.fhip_gram_workspace_bytes(ctypes.byref(b), 629855, 512, 512, 3, 3, 1)
    assert b.value <= a.value


def test_ghost_tables_match_the_oracle():
    from oracle import ghost as OG
    out = ctypes.c_uint64()
    _lib.call("fhip_ghost_tables_checksum", ctypes.byref(out))
    assert out.value == OG.tables_checksum()


def _split_top(text):
    parts, depth, cur = [], 0, ""
    for ch in text:
        if ch in "([{":
            depth += 1
        elif ch in ")]}":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    parts.append(cur)
    return [p for p in (q.strip() for q in parts) if p]


def _closing(text, i):
    """Index of the parenthesis that closes the one at text[i]."""
    depth = 0
    for j in range(i, len(text)):
        if text[j] == "(":
            depth += 1
        elif text[j] == ")":
            depth -= 1
            if depth == 0:
                return j
    raise AssertionError("unbalanced parentheses")


def test_julia_binding_matches_the_header():
    # The Julia extension cannot run here (no Julia), so its ccalls are held to
    # the C prototypes: each names a declared entry point, with as many argument
    # types as the prototype has parameters, and as many arguments as types.
    import re
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    header = open(os.path.join(root, "include", "funicular_hip.h")).read()
    protos = {}
    for m in re.finditer(r"\b(?:int|const char \*)\s*(fhip_\w+)\s*\(([^;]*?)\)\s*;", header, re.S):
        params = [p for p in _split_top(m.group(2).replace("\n", " ")) if p != "void"]
        protos[m.group(1)] = len(params)
    jl = open(os.path.join(root, "funicular.jl_amd", "julia", "FunicularHIPExt.jl")).read()
    calls = list(re.finditer(r"ccall\(\(:(fhip_\w+),\s*LIB\[\]\),\s*\w+,\s*\(", jl))
    assert len(calls) > 20
    for m in calls:
        name = m.group(1)
        t0 = m.end() - 1
        t1 = _closing(jl, t0)
        types = _split_top(jl[t0 + 1:t1])
        a1 = _closing(jl, m.start() + len("ccall"))
        args = _split_top(jl[t1 + 1:a1])
        assert name in protos, name
        assert len(types) == protos[name], (name, len(types), protos[name])
        assert len(args) == len(types), (name, len(args), len(types))


def _julia(name):
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    return open(os.path.join(root, "funicular.jl_amd", "julia", name)).read()


def test_julia_binding_honours_the_contract():
    # Textual checks of what the contract asks of a backend and the binding
    # could not show without Julia: ka_backend returns a handle whose
    # synchronize drains the device (src/panels.jl:365-371 calls it), h2d!/d2h!
    # convert eltypes through a device staging buffer (src/device_api.jl:31-44)
    # while the byte copy refuses mixed eltypes, and the entry points the
    # tests run in Python (the mixed pass, the adjoint's solve, the rowdot)
    # have Julia methods.
    import re
    jl = _julia("FunicularHIPExt.jl")
    assert re.search(r"^ka_backend\(::HIPBackend\)\s*=\s*HIPKernels\(\)", jl, re.M)
    assert "nothing" not in re.search(r"^ka_backend.*$", jl, re.M).group(0)
    assert re.search(r"KA\.synchronize\(::HIPKernels\)\s*=.*fhip_device_sync", jl)
    raw = jl[jl.index("function raw_copy!"):jl.index("function copy2d!")]
    assert "eltype(dst) === eltype(src) ||" in raw and "ArgumentError" in raw
    for fn in ("function h2d!", "function d2h!"):
        body = jl[jl.index(fn):jl.index("\nend", jl.index(fn))]
        assert "converting_copy!" in body, fn
    conv = jl[jl.index("function converting_copy!"):jl.index("\nend", jl.index("function converting_copy!"))]
    assert "alloc_device" in conv and "convert_copy!" in conv and "fhip_stream_sync" in conv
    for name in ("cholqr2_mixed!", "rdiv_upper_adjoint!", "rowdot!", "gram_fp32", "rdiv_rows_compute!",
                 "cholesky_upper_fp32"):
        assert re.search(r"^function " + re.escape(name) + r"\(", jl, re.M), name


def test_julia_rrules_cover_cholqr2_and_project():
    # The ChainRulesCore rules sit in their own extension; their ccalls go
    # through the HIP extension, whose functions they name must exist there.
    import re
    rules = _julia("FunicularHIPChainRulesExt.jl")
    ext = _julia("FunicularHIPExt.jl")
    for target in ("cholqr2!", "cholqr2_qr", "project"):
        assert re.search(r"function ChainRulesCore\.rrule\(::typeof\(" + re.escape(target) + r"\)", rules), target
    for used in re.findall(r"hip\(\)\.(\w+!?)", rules):
        assert re.search(r"^function " + re.escape(used) + r"\(", ext, re.M), used
    assert "ccall" not in rules
