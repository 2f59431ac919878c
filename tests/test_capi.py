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


def test_record_range_rejected_without_gpu():
    """A batch whose (row, system) groups, each padded to a whole cell block, need more work
    slots than 32-bit records address is refused before any device work: 2^20 one-cell
    systems (each cell's eight corners on one node) need 2^28 slots of 256-cell blocks."""
    import numpy as np
    lib = _native.lib()
    P, i64, i32, f64 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
    lib.mm_batch_field_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.c_int, i32,
                                          i64, i64, P, P, P, P, P, i32, i32, P, P, P, P, P, f64]
    n = 1 << 20
    sn = np.repeat(np.arange(n, dtype=np.int64), 8)
    sc = np.repeat(np.arange(n, dtype=np.int64), 8)
    wb = np.zeros(n, np.uint8)
    row = np.zeros(n, np.int32)
    sys_of = np.arange(n, dtype=np.int32)
    ns = np.ones(1, np.int32)
    h0 = np.identity(3).ravel().copy()
    c0 = np.zeros(81)
    fe = np.zeros(1)
    te = np.ones(1)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    handle = ctypes.c_void_p()
    rc = lib.mm_batch_field_create(ctypes.byref(handle), 0, 0, n, n, n, vp(sn), vp(sc), vp(wb),
                                   vp(row), vp(sys_of), 1, 1, vp(ns), vp(h0), vp(c0), vp(fe),
                                   vp(te), 1.0)
    assert rc == _native.MM_ERR_ARG
    assert b"32-bit record range" in lib.mm_last_error()
    assert not handle.value
