"""Geometry rasterisation and subpixel smoothing on the device
(kh_raster_component; src/Geometry.jl:450-663 painting, :792-955 smoothing).

Same arrays as ``geometry.rasterize`` followed by
``smoothing.apply_subpixel_smoothing``, computed by the HIP library: the
host passes each component's grid coordinates (the reference's
GridVolume snapping, src/Geometry.jl:353-366) and the object list, the device
paints every voxel (first object in list order wins) and smooths the
electric components at interfaces.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from .geometry import Box, Cylinder, Prism, Sphere
from .grid import Ex, Ey, Ez, Hx, Hy, Hz, component_voxel_count, full_grid_volume, gv_coords
from .smoothing import AnisotropicSmoothing, NoSmoothing, VolumeAveraging


def _objects(geometry):
    arr = (_lib.KhRasterObject * max(1, len(geometry)))()
    keep = []  # prism vertex arrays, alive as long as the object array
    for o, obj in enumerate(geometry):
        sh = obj.shape
        r = arr[o]
        r.axis = 0
        if isinstance(sh, Box):
            r.shape = 0
            r.center[:] = [float(v) for v in sh.center]
            r.size[:] = [float(v) for v in sh.size]
        elif isinstance(sh, Sphere):
            r.shape = 1
            r.center[:] = [float(v) for v in sh.center]
            r.size[:] = [float(sh.radius), 0.0, 0.0]
        elif isinstance(sh, Cylinder):
            r.shape = 2
            r.axis = int(sh.axis)
            r.center[:] = [float(v) for v in sh.center]
            r.size[:] = [float(sh.radius), float(sh.height), 0.0]
        elif isinstance(sh, Prism):
            r.shape = 3
            r.center[:] = [float(v) for v in sh.center]
            r.size[:] = [float(sh.height), 0.0, 0.0]
            v = np.ascontiguousarray(np.asarray(sh.vertices, dtype=np.float64).reshape(-1, 2))
            keep.append(v)
            r.n_vertices = len(v)
            r.vertices = v.ctypes.data_as(C.c_void_p)
        else:
            raise NotImplementedError(f"device rasterisation of {type(sh).__name__}")
        vals = []
        for electric in (True, False):
            for a in range(3):
                p = obj.material.perm(electric, a)
                vals.append(math.nan if p is None else float(p))
        r.value[:] = vals
    arr._vertex_arrays = keep
    return arr


def raster_device(sim, geometry):
    """(eps_inv[3], mu_inv[3]) like rasterize + apply_subpixel_smoothing; an
    entry is None when no object sets that component."""
    lib = _lib.load()
    R = sim.real
    prec = _lib.KH_F32 if R is np.float32 else _lib.KH_F64
    objs = _objects(geometry)
    mode = getattr(sim, "subpixel_smoothing", None)
    smoothing = 0
    if geometry and sim.ndims == 3 and mode is not None and not isinstance(mode, NoSmoothing):
        smoothing = 2 if isinstance(mode, AnisotropicSmoothing) else 1
    half = (C.c_double * 3)(*[float(sim.delta_n[a]) / 2 for a in range(3)])
    out = [], []
    for electric, comps in ((True, (Ex, Ey, Ez)), (False, (Hx, Hy, Hz))):
        for a, Cc in enumerate(comps):
            if not any(o.material.perm(electric, a) is not None for o in geometry):
                out[0 if electric else 1].append(None)
                continue
            c = Cc()
            cnt = component_voxel_count(sim.Nx, sim.Ny, sim.Nz, c)
            xs, ys, zs = gv_coords(sim, full_grid_volume(sim, c))
            xs = np.ascontiguousarray(xs[:cnt[0]], dtype=np.float64)
            ys = np.ascontiguousarray(ys[:cnt[1]], dtype=np.float64)
            zs = np.ascontiguousarray(zs[:cnt[2]], dtype=np.float64)
            arr = np.empty(cnt, dtype=R, order="F")
            cn = (C.c_int32 * 3)(*cnt)
            comp = a if electric else 3 + a
            _lib.check(lib.kh_raster_component(prec, comp, cn, xs.ctypes.data_as(C.c_void_p),
                                               ys.ctypes.data_as(C.c_void_p), zs.ctypes.data_as(C.c_void_p),
                                               C.cast(objs, C.c_void_p), len(geometry),
                                               smoothing if electric else 0, half, arr.ctypes.data_as(C.c_void_p)))
            out[0 if electric else 1].append(arr)
    return out[0], out[1]
