"""The allocation-free rasteriser (kh_raster_component_noalloc) through the
C ABI: its staging lives in the library's fixed device array instead of
device memory allocated per call, and it writes into the caller's device
array on the caller's stream.  Its values must equal kh_raster_component's
(same kernels) value for value, on two calls in a row of which the second is
smaller, painted and smoothed, FP32 and FP64; a call whose staging would not
fit, or whose output is not device memory of the right size, is refused
before any device work."""
import ctypes as C

import numpy as np
import pytest
import torch

import khronos_hip as kh
from khronos_hip import _lib
from khronos_hip.grid import Ex, Ey, Ez, Hx, Hy, Hz, component_voxel_count, full_grid_volume, gv_coords
from khronos_hip.raster import _objects

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20  # KH_NOALLOC_SCRATCH_BYTES


def _scene(precision, smoothing, resolution):
    geo = [
        kh.Object(kh.Box([0.1, -0.05, 0.02], [0.6, 0.3, 0.45]), kh.Material(ε=4.0)),
        kh.Object(kh.Sphere([-0.3, 0.2, -0.1], 0.33), kh.Material(εx=2.0, εy=3.0, εz=5.0)),
        kh.Object(kh.Cylinder([0.25, 0.3, -0.2], 0.2, 0.7, axis=0), kh.Material(ε=9.0, μ=1.5)),
        kh.Object(kh.Prism([-0.1, 0.1, 0.3], [[-0.2, -0.1], [0.25, -0.15], [0.1, 0.2], [-0.15, 0.15]], 0.3),
                  kh.Material(ε=6.0)),
        kh.Object(kh.Box([0.0, 0.0, -0.45], [2.0, 2.0, 0.3]), kh.Material(ε=2.1)),
    ]
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.4), component=kh.Ez(),
                            center=[0, 0, 0], size=[0, 0, 0])]
    return kh.Simulation(cell_size=[1.6, 1.4, 1.3], resolution=resolution, boundaries=[[0.2, 0.2]] * 3,
                         sources=src, geometry=geo, precision=precision, subpixel_smoothing=smoothing)


def _args(sim, c):
    cnt = component_voxel_count(sim.Nx, sim.Ny, sim.Nz, c)
    xs, ys, zs = gv_coords(sim, full_grid_volume(sim, c))
    xs, ys, zs = (np.ascontiguousarray(v[:n], dtype=np.float64) for v, n in zip((xs, ys, zs), cnt))
    return cnt, xs, ys, zs


def _both(lib, sim, objs, comp, smoothing, stream):
    """(kh_raster_component, kh_raster_component_noalloc) values of one component."""
    R = sim.real
    prec = _lib.KH_F32 if R is np.float32 else _lib.KH_F64
    c = (Ex, Ey, Ez, Hx, Hy, Hz)[comp]()
    cnt, xs, ys, zs = _args(sim, c)
    cn = (C.c_int32 * 3)(*cnt)
    half = (C.c_double * 3)(*[float(sim.delta_n[a]) / 2 for a in range(3)])
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    host = np.empty(cnt, dtype=R, order="F")
    _lib.check(lib.kh_raster_component(prec, comp, cn, ptr(xs), ptr(ys), ptr(zs), C.cast(objs, C.c_void_p),
                                       len(sim.geometry), smoothing, half, ptr(host)))
    n = cnt[0] * cnt[1] * cnt[2]
    out = torch.full((n,), float("nan"), dtype=torch.float32 if R is np.float32 else torch.float64, device="cuda")
    torch.cuda.synchronize()
    _lib.check(lib.kh_raster_component_noalloc(prec, comp, cn, ptr(xs), ptr(ys), ptr(zs),
                                               C.cast(objs, C.c_void_p), len(sim.geometry), smoothing, half,
                                               C.c_void_p(out.data_ptr()), C.c_void_p(stream)))
    dev = out.cpu().numpy().reshape(cnt, order="F")
    return host, dev


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("smoothing", [0, 1, 2])
def test_noalloc_raster_equals_allocating_raster(precision, smoothing):
    lib = _lib.load()
    mode = {0: None, 1: kh.VolumeAveraging(), 2: kh.AnisotropicSmoothing()}[smoothing]
    side = torch.cuda.Stream()  # a caller's stream (default priority)
    for call, (res, stream) in enumerate(((31, side.cuda_stream), (17, 0))):
        sim = _scene(precision, mode, res)
        objs = _objects(sim.geometry)
        for comp in range(6):
            sm = smoothing if comp < 3 else 0
            print(f"[noalloc] {precision} smoothing {smoothing} call {call} resolution {res} component {comp}",
                  flush=True)
            host, dev = _both(lib, sim, objs, comp, sm, stream)
            assert not np.any(np.isnan(dev))
            np.testing.assert_array_equal(dev, host)


def test_noalloc_refuses_what_does_not_fit():
    lib = _lib.load()
    sim = _scene("f64", kh.AnisotropicSmoothing(), 17)
    objs = _objects(sim.geometry)
    xs = np.linspace(-1.0, 1.0, 1100)
    ys = np.linspace(-1.0, 1.0, 100)
    zs = np.linspace(-1.0, 1.0, 100)
    cn = (C.c_int32 * 3)(1100, 100, 100)
    half = (C.c_double * 3)(0.01, 0.01, 0.01)
    ptr = lambda a: a.ctypes.data_as(C.c_void_p)
    n = 1100 * 100 * 100
    assert n * 8 > SCRATCH_BYTES  # the smoothed FP64 call stages an unsmoothed raster of n doubles
    out = torch.zeros(n, dtype=torch.float64, device="cuda")
    rc = lib.kh_raster_component_noalloc(_lib.KH_F64, 0, cn, ptr(xs), ptr(ys), ptr(zs), C.cast(objs, C.c_void_p),
                                         len(sim.geometry), 2, half, C.c_void_p(out.data_ptr()), None)
    assert rc == -1 and b"scratch" in lib.kh_last_error()
    # without smoothing nothing but the inputs is staged: the same size runs
    rc = lib.kh_raster_component_noalloc(_lib.KH_F64, 0, cn, ptr(xs), ptr(ys), ptr(zs), C.cast(objs, C.c_void_p),
                                         len(sim.geometry), 0, half, C.c_void_p(out.data_ptr()), None)
    assert rc == 0, lib.kh_last_error()
    assert float(out.min()) > 0
    # an output that is host memory, or too small, is refused before any device work
    host = np.zeros(n)
    rc = lib.kh_raster_component_noalloc(_lib.KH_F64, 0, cn, ptr(xs), ptr(ys), ptr(zs), C.cast(objs, C.c_void_p),
                                         len(sim.geometry), 0, half, ptr(host), None)
    assert rc == -1 and b"device memory" in lib.kh_last_error()
    small = torch.zeros(1000, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    free_before = torch.cuda.mem_get_info()[0]
    rc = lib.kh_raster_component_noalloc(_lib.KH_F64, 0, cn, ptr(xs), ptr(ys), ptr(zs), C.cast(objs, C.c_void_p),
                                         len(sim.geometry), 0, half, C.c_void_p(small.data_ptr()), None)
    assert rc == -1 and b"fewer" in lib.kh_last_error()
    assert torch.cuda.mem_get_info()[0] == free_before
