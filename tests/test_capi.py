"""The C ABI library: it loads, exports every declared entry point, and the scalar-block layout
of the Python layer matches the header. No compute call (runs without a GPU)."""

import ctypes
import os
import re

from micmec_amd import _layout, _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "micmec_amd.h")


def declared_functions():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(mm_[a-z0-9_]+)\s*\(", src,
                                 flags=re.M)))


def test_library_loads_and_exports_all_declarations():
    lib = _native.lib()
    names = declared_functions()
    assert len(names) >= 12
    for name in names:
        assert hasattr(lib, name), name
    assert lib.mm_version().startswith(b"micmec_amd")


def test_scalar_layout_matches_header():
    src = open(HEADER).read()
    defs = dict(re.findall(r"#define (MM_S_[A-Z0-9_]+) (\d+)", src))
    assert defs
    for k, v in defs.items():
        assert getattr(_layout, k[3:]) == int(v), k
    assert int(defs["MM_S_COUNT"]) > max(int(v) for k, v in defs.items() if k != "MM_S_COUNT")


def test_argument_validation_without_gpu():
    lib = _native.lib()
    # invalid arguments are rejected before any device work
    handle = ctypes.c_void_p()
    rc = lib.mm_field_create(ctypes.byref(handle), 0, 7, 0, 0, None, None, None, None, 1, 1,
                             None, None, None, None, None, 1.0)
    assert rc == _native.MM_ERR_ARG
    assert b"model" in lib.mm_last_error()
    assert lib.mm_field_compute(None, None, None, None, None, None, 0, None) == _native.MM_ERR_ARG
    assert lib.mm_cg_trial(None, None, None, None, 0.0, 0, 0.0, None, None, None, None,
                           None, None) == _native.MM_ERR_ARG
    assert lib.mm_cg_stats(None, None, None, None, None, None, None,
                           None) == _native.MM_ERR_ARG
    assert lib.mm_field_set_params_device(None, None, None, None, None, None, 0.0,
                                          None) == _native.MM_ERR_ARG
    assert lib.mm_adjoint_vjp(None, None, None, None, None, None, None, None, None, None,
                              None) == _native.MM_ERR_ARG
