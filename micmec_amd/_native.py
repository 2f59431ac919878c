"""Loader of the HIP library ``libmicmec_amd.so`` and its ctypes signatures.

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C micmec_amd/csrc``).
There is no CPU fallback: importing a compute entry point without the library raises
``ImportError``. torch is imported first so that the library binds to the HIP runtime that
torch has already loaded (both carry the soname libamdhip64.so.7).
"""

import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime first)

__all__ = ["lib", "library_path", "check", "MMError", "MM_OK", "MM_ERR_ARG", "MM_ERR_NAN",
           "MM_ERR_HIP", "MM_ERR_NOMEM", "MODELS"]

MM_OK = 0
MM_ERR_ARG = 1
MM_ERR_NAN = 2
MM_ERR_HIP = 3
MM_ERR_NOMEM = 4
MODELS = {"corner8": 0, "original": 1}

library_path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmicmec_amd.so")

_lib = None


class MMError(RuntimeError):
    """A HIP-side failure reported by the library."""


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(library_path):
        raise ImportError(
            "libmicmec_amd.so is not built (expected at %s); run "
            "`python -c 'import __graft_entry__ as g; g.build()'` or `make -C micmec_amd/csrc`"
            % library_path)
    L = ctypes.CDLL(library_path, mode=ctypes.RTLD_GLOBAL)
    P = ctypes.c_void_p
    i32, i64, f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    L.mm_version.restype = ctypes.c_char_p
    L.mm_last_error.restype = ctypes.c_char_p
    L.mm_field_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.c_int, i64, i64, P, P,
                                  P, P, i32, i32, P, P, P, P, P, f64]
    L.mm_field_compute.argtypes = [P, P, P, P, P, P, ctypes.c_int, P]
    L.mm_field_check.argtypes = [P, P, ctypes.POINTER(i32)]
    L.mm_field_launch_cells.argtypes = [P, P, P, P]
    L.mm_cg_trial.argtypes = [P, P, P, P, f64, ctypes.c_int, f64, P, P, P, P, P, P]
    L.mm_cg_stats.argtypes = [P, P, P, P, P, P, P, P]
    L.mm_field_set_params_device.argtypes = [P, P, P, P, P, P, f64, P]
    L.mm_field_info.argtypes = [P, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(i64)]
    L.mm_field_destroy.argtypes = [P]
    L.mm_topology_build.argtypes = [ctypes.POINTER(P), ctypes.c_int, P, i32, i32, i32, i32, P,
                                    i32, P, P]
    L.mm_topology_sizes.argtypes = [P, ctypes.POINTER(i64), ctypes.POINTER(i64),
                                    ctypes.POINTER(i64)]
    L.mm_topology_copy.argtypes = [P, P, P, P, P, P, P, P, P, P]
    L.mm_topology_destroy.argtypes = [P]
    L.mm_field_destroy.restype = None
    L.mm_md_destroy.argtypes = [P]
    L.mm_md_destroy.restype = None
    for name, argtypes in _EXTRA_SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is not None:
            fn.argtypes = argtypes
            fn.restype = ctypes.c_int
    _lib = L
    return L


# Signatures of the integrator / batch / adjoint entry points, filled in by their modules.
_EXTRA_SIGNATURES = {}


def register_signature(name, argtypes):
    _EXTRA_SIGNATURES[name] = argtypes
    if _lib is not None:
        fn = getattr(_lib, name)
        fn.argtypes = argtypes
        fn.restype = ctypes.c_int


def lib():
    return _load()


def last_error():
    return _load().mm_last_error().decode(errors="replace")


def check(rc, what=""):
    """Turn a status code into the exception the reference would raise."""
    if rc == MM_OK:
        return
    msg = "%s: %s" % (what, last_error()) if what else last_error()
    if rc == MM_ERR_ARG:
        raise ValueError(msg)
    if rc == MM_ERR_NAN:
        raise ValueError(msg)
    if rc == MM_ERR_NOMEM:
        raise MemoryError(msg)
    raise MMError(msg)


def ptr(t):
    """Raw pointer of a torch tensor (or None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
