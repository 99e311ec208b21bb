"""ctypes binding of the HIP engine library (include/meep_fdtd.h).

The library is built in-tree (``meep_gpu._build.build``) as ``meep_gpu/_lib/libmeep_fdtd.so``.
There is no other compute path: if the library is missing or its call fails, the caller gets
an exception naming the reason, never a CPU substitute.
"""

from __future__ import annotations

import contextlib
import ctypes
import os
from typing import Dict, Optional

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libmeep_fdtd.so")
#: The library per subnormal policy: kept (the default) or flushed (float32 and float64).
LIB_PATHS = {"keep": LIB_PATH, "flush": os.path.join(_HERE, "_lib", "libmeep_fdtd_flush.so")}
SUBNORMAL_MODES = ("keep", "flush")

_FP = ctypes.POINTER(ctypes.c_float)
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int32)
_U8P = ctypes.POINTER(ctypes.c_uint8)

FDTD_MAX_POLES = 8

#: Every entry point include/meep_fdtd.h declares, with its ctypes signature.
SIGNATURES = {
    "fdtd_last_error": (ctypes.c_char_p, []),
    "fdtd_subnormal_mode": (ctypes.c_int, []),
    "fdtd_host_flush": (ctypes.c_int, [ctypes.c_int32]),
    "fdtd_device_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "fdtd_device_memory": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint64),
                                          ctypes.POINTER(ctypes.c_uint64)]),
    "fdtd_sim_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "fdtd_sim_destroy": (None, [ctypes.c_void_p]),
    "fdtd_sim_bytes": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]),
    "fdtd_set_layer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _FP, _FP,
                                      _FP, _U8P]),
    "fdtd_set_inverse_epsilon": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _FP]),
    "fdtd_pole_rows": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_int64)]),
    "fdtd_inverse_epsilon_bytes": (ctypes.c_int, [ctypes.c_void_p, _IP]),
    "fdtd_work_split": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _IP]),
    "fdtd_set_offdiagonal": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                            _FP]),
    "fdtd_add_pole": (ctypes.c_int, [ctypes.c_void_p, _FP, ctypes.POINTER(_FP), _FP, _IP, _IP]),
    "fdtd_add_source": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, _IP, _IP,
                                       _DP, _DP, _DP, ctypes.c_int32, ctypes.c_double]),
    "fdtd_upload": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, _FP, ctypes.c_size_t]),
    "fdtd_download": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, _FP, ctypes.c_size_t]),
    "fdtd_get_step": (ctypes.c_int, [ctypes.c_void_p, _IP]),
    "fdtd_set_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "fdtd_prime_e": (ctypes.c_int, [ctypes.c_void_p]),
    "fdtd_step_h": (ctypes.c_int, [ctypes.c_void_p]),
    "fdtd_step_d": (ctypes.c_int, [ctypes.c_void_p]),
    "fdtd_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "fdtd_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "fdtd_flush": (ctypes.c_int, [ctypes.c_void_p]),
    "fdtd_add_flux_plane": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _IP, _IP,
                                           ctypes.c_int32, _DP, ctypes.c_double, ctypes.c_int32,
                                           _IP]),
    "fdtd_flux_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _DP, ctypes.c_size_t]),
    "fdtd_flux_spectrum": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _DP, _DP,
                                          ctypes.c_size_t]),
    "fdtd_flux_spectrum_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                                  ctypes.c_void_p, ctypes.c_void_p,
                                                  ctypes.c_size_t]),
    "fdtd_flux_write": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _DP, ctypes.c_size_t]),
    "fdtd_set_boundary": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "fdtd_set_source_currents": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                                ctypes.POINTER(ctypes.c_double), ctypes.c_int32]),
    "fdtd_batch_step_h": (ctypes.c_int, [ctypes.c_void_p]),
    "fdtd_batch_step_d": (ctypes.c_int, [ctypes.c_void_p]),
    "fdtd_link_parts": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _IP,
                                       ctypes.POINTER(ctypes.c_float)]),
    "fdtd_sim_set_config": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "fdtd_sim_get_config": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "fdtd_batch_set_config": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "fdtd_batch_get_config": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    "fdtd_graph_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "fdtd_batch_graph_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "fdtd_set_fold": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.c_int32]),
    "fdtd_set_conductivity": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                             _FP]),
    "fdtd_set_conductive_cells": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                                 ctypes.c_int32, _U8P]),
    "fdtd_add_dft_region": (ctypes.c_int, [ctypes.c_void_p, _IP, _IP, ctypes.c_int32, _IP,
                                           ctypes.c_int32, ctypes.c_int32, _DP, ctypes.c_double,
                                           ctypes.c_int32, _IP]),
    "fdtd_flux_scale": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double,
                                       ctypes.c_double]),
    "fdtd_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_void_p)]),
    "fdtd_batch_destroy": (None, [ctypes.c_void_p]),
    "fdtd_batch_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "fdtd_batch_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "fdtd_set_launch": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32]),
    "fdtd_get_launch": (ctypes.c_int, [_IP]),
    "fdtd_set_launch_pair": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_int32, ctypes.c_int32]),
    "fdtd_get_launch_pair": (ctypes.c_int, [ctypes.c_int32, _IP]),
    "fdtd_set_launch_shell": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32]),
    "fdtd_get_launch_shell": (ctypes.c_int, [ctypes.c_int32, _IP]),
    "fdtd_set_launch_tile": (ctypes.c_int, [ctypes.c_int32]),
    "fdtd_get_launch_tile": (ctypes.c_int, [_IP]),
    "fdtd_set_graphs": (ctypes.c_int, [ctypes.c_int32]),
    "fdtd_set_flux_transform": (ctypes.c_int, [ctypes.c_int32]),
    "fdtd_batch_flush": (ctypes.c_int, [ctypes.c_void_p]),
    "fdtd_time_pairs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _FP]),
    "fdtd_run_timed": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, _FP]),
    "fdtd_copy_probe": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, _FP]),
    "fdtd_stream_probe": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                         ctypes.c_uint64, ctypes.c_int32, _FP]),
}


class FdtdConfig(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32 * 3), ("invariant", ctypes.c_int32 * 3),
                ("dtdx", ctypes.c_float), ("pml", ctypes.c_int32), ("dt", ctypes.c_double)]


class EngineError(RuntimeError):
    """A failed engine call, with the engine's own reason."""


_libs: Dict[str, ctypes.CDLL] = {}


def load(path: Optional[str] = None, subnormals: str = "keep") -> ctypes.CDLL:
    """Load the engine library of one subnormal policy ("keep", the default, or "flush") and
    bind every declared symbol (no GPU call is made)."""
    if subnormals not in SUBNORMAL_MODES:
        raise ValueError(f"subnormals must be one of {SUBNORMAL_MODES}, got {subnormals!r}")
    path = path or LIB_PATHS[subnormals]
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise EngineError(
            f"the HIP engine library is not built ({path} is missing); run "
            f"`python -c 'import __graft_entry__ as g; g.build()'` from the repository root")
    lib = ctypes.CDLL(path)
    for name, (restype, argtypes) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    want = SUBNORMAL_MODES.index(subnormals)
    if path == LIB_PATHS[subnormals] and lib.fdtd_subnormal_mode() != want:
        raise EngineError(f"{path} was not built with subnormals {subnormals!r}")
    _libs[path] = lib
    return lib


@contextlib.contextmanager
def host_flush(on: bool = True):
    """Host arithmetic of this thread (NumPy included) under x86 FTZ|DAZ, cleared after."""
    lib = load()
    lib.fdtd_host_flush(1 if on else 0)
    try:
        yield
    finally:
        lib.fdtd_host_flush(0)


def check(rc: int, what: str, lib: Optional[ctypes.CDLL] = None) -> None:
    """Raise ``EngineError`` with the engine's reason when ``rc`` is not 0; ``lib`` is the
    library the call went to (the default build when None)."""
    if rc != 0:
        text = (lib or load()).fdtd_last_error().decode("utf-8", "replace")
        raise EngineError(f"{what}: {text} (code {rc})")


def fptr(array) -> ctypes.POINTER(ctypes.c_float):
    return array.ctypes.data_as(_FP)


def dptr(array) -> ctypes.POINTER(ctypes.c_double):
    return array.ctypes.data_as(_DP)


def iptr(array) -> ctypes.POINTER(ctypes.c_int32):
    return array.ctypes.data_as(_IP)


def u8ptr(array) -> ctypes.POINTER(ctypes.c_uint8):
    return array.ctypes.data_as(_U8P)
