"""Device rasterisation entry point, host side: the ctypes struct matches the
C layout and bad arguments come back as errors before any device work."""
import ctypes as C

import numpy as np
import pytest

from khronos_hip import _lib


def test_raster_object_layout():
    # int32 shape, int32 axis, double center[3], size[3], value[6], int32
    # n_vertices, int32 reserved, const double *vertices
    assert C.sizeof(_lib.KhRasterObject) == 4 + 4 + 24 + 24 + 48 + 4 + 4 + 8
    assert _lib.KhRasterObject.center.offset == 8 and _lib.KhRasterObject.value.offset == 56
    assert _lib.KhRasterObject.n_vertices.offset == 104 and _lib.KhRasterObject.vertices.offset == 112


def test_raster_rejects_bad_arguments_without_a_device():
    lib = _lib.load()
    cnt = (C.c_int32 * 3)(2, 2, 2)
    x = np.zeros(2)
    out = np.zeros(8)
    obj = (_lib.KhRasterObject * 1)()
    p = lambda a: a.ctypes.data_as(C.c_void_p)
    half = (C.c_double * 3)(0.05, 0.05, 0.05)
    # component out of range
    assert lib.kh_raster_component(1, 7, cnt, p(x), p(x), p(x), C.cast(obj, C.c_void_p), 1, 0, half, p(out)) != 0
    # unknown shape
    obj[0].shape = 5
    assert lib.kh_raster_component(1, 0, cnt, p(x), p(x), p(x), C.cast(obj, C.c_void_p), 1, 0, half, p(out)) != 0
    assert "shape" in lib.kh_last_error().decode()
    # a prism without vertices
    obj[0].shape = 3
    obj[0].n_vertices = 2
    assert lib.kh_raster_component(1, 0, cnt, p(x), p(x), p(x), C.cast(obj, C.c_void_p), 1, 0, half, p(out)) != 0
    assert "prism" in lib.kh_last_error().decode()
    # smoothing mode out of range
    obj[0].shape = 0
    assert lib.kh_raster_component(1, 0, cnt, p(x), p(x), p(x), C.cast(obj, C.c_void_p), 1, 3, half, p(out)) != 0
    # empty array
    z = (C.c_int32 * 3)(0, 2, 2)
    assert lib.kh_raster_component(1, 0, z, p(x), p(x), p(x), C.cast(obj, C.c_void_p), 1, 0, half, p(out)) != 0
