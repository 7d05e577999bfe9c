"""The C-ABI library loads and exports every symbol include/khronos_hip.h declares.

CPU-only: calls only entry points that do no device work.
"""
import ctypes as C
import os
import re

import numpy as np

from khronos_hip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "khronos_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(kh_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), "ctypes signature table out of sync with the header"


def test_version_and_footprint_without_device():
    lib = _lib.load()
    assert lib.kh_version() == 8
    assert C.sizeof(_lib.KhSimDesc) == 336  # layout of kh_sim_desc in include/khronos_hip.h
    d = _lib.KhSimDesc()
    d.nx, d.ny, d.nz = 384, 192, 160
    d.precision = _lib.KH_F32
    d.dx = d.dy = d.dz = 0.05
    d.dt = 0.025
    n = C.c_size_t()
    assert lib.kh_mem_footprint(C.byref(d), C.byref(n)) == 0
    cells = 386 * 194 * 162
    assert 18 * cells * 4 <= n.value <= 24 * cells * 4  # FAST: E, H, U, W x 3
    d.mode = _lib.KH_MODE_EXACT
    assert lib.kh_mem_footprint(C.byref(d), C.byref(n)) == 0
    assert 24 * cells * 4 <= n.value <= 32 * cells * 4  # EXACT adds D and B
    d.complex_fields = 1
    assert lib.kh_mem_footprint(C.byref(d), C.byref(n)) == 0
    assert 48 * cells * 4 <= n.value <= 64 * cells * 4  # complex: two parts
    d.nz = 0
    assert lib.kh_mem_footprint(C.byref(d), C.byref(n)) == -1
    assert b"extents" in lib.kh_last_error()


def test_adjoint_entry_points_reject_bad_arguments_without_device():
    lib = _lib.load()
    lo = (C.c_int32 * 9)(*([1] * 9))
    n = (C.c_int32 * 9)(*([0] * 9))
    g = (C.c_double * 1)()
    assert lib.kh_set_checkpoints(None, 100) != 0
    assert lib.kh_adjoint_gradient(None, 0, None, None, lo, n, g) != 0
    assert lib.kh_last_error().decode() != ""


def test_stream_rate_rejects_bad_arguments_without_device():
    lib = _lib.load()
    out = C.c_double()
    for args in ((0, 1, 1024, 1), (7, 1, 1024, 1), (1, 0, 1024, 1), (1, 4, 1024, 1), (1, 1, 1023, 1),
                 (1, 1, 256, 1), (1, 1, 1024, 0)):
        assert lib.kh_stream_rate(*args, C.byref(out)) != 0, args
    assert lib.kh_stream_rate(1, 1, 1024, 1, None) != 0
    assert lib.kh_last_error().decode() != ""
