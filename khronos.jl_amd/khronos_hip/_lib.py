"""ctypes binding of libkhronos_hip (include/khronos_hip.h).

The library is built in-tree (khronos.jl_amd/lib/libkhronos_hip.so).  There is
no fallback: if the shared library is missing, importing the time-step engine
raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libkhronos_hip.so")

KH_F32, KH_F64 = 0, 1
KH_MODE_FAST, KH_MODE_EXACT = 0, 1


class KhRasterObject(C.Structure):
    _fields_ = [("shape", C.c_int32), ("axis", C.c_int32), ("center", C.c_double * 3), ("size", C.c_double * 3),
                ("value", C.c_double * 6), ("n_vertices", C.c_int32), ("reserved", C.c_int32),
                ("vertices", C.c_void_p)]


class KhSimDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("precision", C.c_int32), ("mode", C.c_int32), ("dft_accum_f64", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("sigma_b", C.c_void_p * 3), ("sigma_d", C.c_void_p * 3),
        ("eps_inv", C.c_void_p * 3), ("mu_inv", C.c_void_p * 3),
        ("eps_inv_scalar", C.c_double), ("mu_inv_scalar", C.c_double),
        ("source_cutoff_time", C.c_double),
        ("complex_fields", C.c_int32), ("periodic", C.c_int32 * 3),
        ("bloch_fwd", (C.c_double * 2) * 3), ("bloch_rev", (C.c_double * 2) * 3),
        ("sigma_mat", C.c_void_p * 6),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "kh_last_error": (C.c_char_p, []),
    "kh_version": (C.c_int, []),
    "kh_stream_rate": (C.c_int, [C.c_int32, C.c_int32, C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "kh_mem_footprint": (C.c_int, [C.POINTER(KhSimDesc), C.POINTER(C.c_size_t)]),
    "kh_create": (C.c_int, [C.POINTER(KhSimDesc), C.POINTER(C.c_void_p)]),
    "kh_destroy": (C.c_int, [C.c_void_p]),
    "kh_add_source": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                C.c_void_p, C.POINTER(C.c_int32)]),
    "kh_set_waveform": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p]),
    "kh_add_dft": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                             C.c_int32, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    "kh_set_eps_inv": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "kh_mode_overlap": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_int32, C.c_void_p, C.c_double, C.c_void_p]),
    "kh_near2far": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                              C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_void_p,
                              C.c_int32, C.c_double, C.c_double, C.c_void_p, C.c_int32, C.c_void_p]),
    "kh_near2far_spectrum": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                       C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                       C.c_int32, C.c_void_p]),
    "kh_raster_component": (C.c_int, [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]),
    "kh_raster_component_noalloc": (C.c_int, [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                              C.c_void_p, C.c_void_p]),
    "kh_server_create": (C.c_int, [C.c_double, C.c_int32, C.c_int32, C.c_void_p]),
    "kh_server_destroy": (C.c_int, [C.c_void_p]),
    "kh_server_admit": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p]),
    "kh_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_double, C.c_int64]),
    "kh_server_cancel": (C.c_int, [C.c_void_p, C.c_double]),
    "kh_server_poll": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "kh_set_chi3": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kh_add_pole": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(C.c_int32)]),
    "kh_read_pole": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_size_t]),
    "kh_write_pole": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_size_t]),
    "kh_step": (C.c_int, [C.c_void_p, C.c_int64]),
    "kh_set_graph": (C.c_int, [C.c_void_p, C.c_int32]),
    "kh_prepare_graphs": (C.c_int, [C.c_void_p, C.c_int64]),
    "kh_get_launch_counts": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kh_query": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "kh_sync": (C.c_int, [C.c_void_p]),
    "kh_get_timestep": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "kh_sources_active": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "kh_read_field": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t]),
    "kh_write_field": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t]),
    "kh_read_dft": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t]),
    "kh_flux": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                          C.c_double, C.c_void_p]),
    "kh_dft_norm": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "kh_dft_norms": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "kh_reset": (C.c_int, [C.c_void_p]),
    "kh_set_checkpoints": (C.c_int, [C.c_void_p, C.c_int32]),
    "kh_adjoint_gradient": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    "kh_set_timing": (C.c_int, [C.c_void_p, C.c_int32]),
    "kh_get_timing": (C.c_int, [C.c_void_p, C.c_void_p]),
    "kh_get_stream": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
}

_lib = None


class KhronosError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libkhronos_hip error {code}: {msg}")
        self.code = code


ABI_VERSION = 8  # KH_ABI_VERSION of include/khronos_hip.h that these signatures describe


def load(path: str = LIB_PATH):
    """Load (once) and return the shared library with typed entry points."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(
            f"libkhronos_hip not built: {path} is missing "
            "(run `python khronos.jl_amd/build_lib.py` or __graft_entry__.build())")
    # torch ships its own libamdhip64 under the same soname; loading torch first
    # makes the library bind to that one, so a process never holds two HIP
    # runtimes (torch's and /opt/rocm's) at once.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.kh_version() != ABI_VERSION:
        raise ImportError(f"{path} has ABI version {lib.kh_version()}, these bindings need {ABI_VERSION} "
                          "(rebuild with __graft_entry__.build())")
    _lib = lib
    return lib


def check(code: int):
    if code != 0:
        msg = _lib.kh_last_error().decode() if _lib is not None else "?"
        raise KhronosError(code, msg)
