"""Geometry rasterisation and subpixel smoothing on the device
(kh_raster_component) against the host setup (geometry.rasterize +
smoothing.apply_subpixel_smoothing, pinned in tests/test_smoothing.py and
tests/test_host_setup.py): painted arrays identical, smoothed ones within
1e-12 (FP64) or one FP32 ulp, on mixed scenes and at the strip waveguide's
full size."""
import time

import numpy as np
import pytest

import khronos_hip as kh
from khronos_hip.geometry import rasterize
from khronos_hip.raster import raster_device
from khronos_hip.smoothing import apply_subpixel_smoothing

pytestmark = pytest.mark.gpu


def _scene(precision, smoothing):
    geo = [
        kh.Object(kh.Box([0.1, -0.05, 0.02], [0.6, 0.3, 0.45]), kh.Material(ε=4.0)),
        kh.Object(kh.Sphere([-0.3, 0.2, -0.1], 0.33), kh.Material(εx=2.0, εy=3.0, εz=5.0)),
        kh.Object(kh.Cylinder([0.25, 0.3, -0.2], 0.2, 0.7, axis=0), kh.Material(ε=9.0, μ=1.5)),
        kh.Object(kh.Cylinder([-0.2, -0.3, 0.25], 0.17, 0.5, axis=1), kh.Material(ε=2.25)),
        kh.Object(kh.Cylinder([0.0, 0.0, 0.0], 0.12, 1.2, axis=2), kh.Material(ε=12.0)),
        kh.Object(kh.Box([0.0, 0.0, -0.45], [2.0, 2.0, 0.3]), kh.Material(ε=2.1)),
    ]
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.4), component=kh.Ez(),
                            center=[0, 0, 0], size=[0, 0, 0])]
    return kh.Simulation(cell_size=[1.6, 1.4, 1.3], resolution=23, boundaries=[[0.2, 0.2]] * 3, sources=src,
                         geometry=geo, precision=precision, subpixel_smoothing=smoothing)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("smoothing", [None, kh.VolumeAveraging(), kh.AnisotropicSmoothing()])
def test_device_raster_matches_host(precision, smoothing):
    sim = _scene(precision, smoothing)
    he, hm = rasterize(sim, sim.geometry)
    he = apply_subpixel_smoothing(sim, sim.geometry, he)
    de, dm = raster_device(sim, sim.geometry)
    name = type(smoothing).__name__
    for h, d in zip(hm, dm):
        assert (h is None) == (d is None)
        if h is not None:
            np.testing.assert_array_equal(np.asarray(d), h)
    for a, (h, d) in enumerate(zip(he, de)):
        assert h is not None and d is not None
        h = np.asarray(h)
        d = np.asarray(d)
        assert h.shape == d.shape
        diff = np.abs(d.astype(np.float64) - h.astype(np.float64)) / np.abs(h.astype(np.float64))
        changed = int(np.count_nonzero(h != np.asarray(rasterize(sim, sim.geometry)[0][a])))
        print(f"[raster] {precision} {name} E{'xyz'[a]}: {d.size} voxels, {changed} smoothed, "
              f"max rel diff {diff.max():.2e}, {int(np.count_nonzero(diff))} differ", flush=True)
        tol = 1e-12 if precision == "f64" else 1.2e-7
        assert diff.max() <= tol
        if smoothing is None or isinstance(smoothing, type(None)):
            np.testing.assert_array_equal(d, h)


def test_device_raster_strip_waveguide_full_size():
    """The strip waveguide workload (384 x 192 x 160, anisotropic smoothing)
    set up on the device gives the host's arrays; both setup times printed."""
    from khronos_hip.workloads import strip_waveguide
    host = strip_waveguide(precision="f32", mode="fast")
    dev = strip_waveguide(precision="f32", mode="fast")
    dev.raster = "device"
    t0 = time.perf_counter()
    de, _ = raster_device(dev, dev.geometry)
    t_dev = time.perf_counter() - t0
    t0 = time.perf_counter()
    he, _ = rasterize(host, host.geometry)
    he = apply_subpixel_smoothing(host, host.geometry, he)
    t_host = time.perf_counter() - t0
    print(f"[raster] strip waveguide {host.Nx}x{host.Ny}x{host.Nz}: host {t_host:.3f} s, device {t_dev:.3f} s",
          flush=True)
    for h, d in zip(he, de):
        if h is None:
            assert d is None
            continue
        h = np.asarray(h, dtype=np.float64)
        d = np.asarray(d, dtype=np.float64)
        diff = np.abs(d - h) / np.abs(h)
        assert diff.max() <= 1.2e-7
        assert np.count_nonzero(diff) <= 1e-4 * diff.size


def _prism_scene(precision, smoothing):
    from khronos_hip import gds as GD
    geo = GD.polygon_to_objects([(-0.5, -0.4), (0.3, -0.5), (-0.1, 0.2)], -0.3, 0.25, kh.Material(ε=6.0))
    geo += GD.polygon_to_objects([(0.1, 0.0), (0.5, 0.05), (0.6, 0.35), (0.35, 0.55), (0.05, 0.35)], -0.1, 0.4,
                                 kh.Material(εx=3.0, εy=4.0, εz=5.0))
    geo += GD.polygon_to_objects([(-0.6, 0.1), (-0.2, 0.1), (-0.2, 0.3), (-0.4, 0.3), (-0.4, 0.5), (-0.6, 0.5)],
                                 0.0, 0.3, kh.Material(ε=2.5, μ=1.3))
    geo += GD.polygon_to_objects(GD.expand_path([(-0.6, -0.55), (0.0, -0.2), (0.55, -0.45)], 0.06, 2), -0.4, 0.0,
                                 kh.Material(ε=9.0))
    geo += [kh.Object(kh.Box([0.0, 0.0, -0.45], [2.0, 2.0, 0.3]), kh.Material(ε=2.1))]
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.4), component=kh.Ez(),
                            center=[0, 0, 0], size=[0, 0, 0])]
    return kh.Simulation(cell_size=[1.6, 1.4, 1.3], resolution=23, boundaries=[[0.2, 0.2]] * 3, sources=src,
                         geometry=geo, precision=precision, subpixel_smoothing=smoothing)


@pytest.mark.parametrize("precision", ["f64", "f32"])
@pytest.mark.parametrize("smoothing", [None, kh.VolumeAveraging(), kh.AnisotropicSmoothing()])
def test_device_raster_prisms_match_host(precision, smoothing):
    """GDS-made prisms (a triangle, a convex pentagon, an ear-clipped L, a
    mitred path) under a box: painted identical, smoothed within 1e-12
    (FP64) or one FP32 ulp."""
    sim = _prism_scene(precision, smoothing)
    he, hm = rasterize(sim, sim.geometry)
    raw = [np.asarray(a) for a in he]
    he = apply_subpixel_smoothing(sim, sim.geometry, he)
    de, dm = raster_device(sim, sim.geometry)
    for h, d in zip(hm, dm):
        assert (h is None) == (d is None)
        if h is not None:
            np.testing.assert_array_equal(np.asarray(d), h)
    for a, (h, d) in enumerate(zip(he, de)):
        h = np.asarray(h)
        d = np.asarray(d)
        diff = np.abs(d.astype(np.float64) - h.astype(np.float64)) / np.abs(h.astype(np.float64))
        print(f"[raster] prisms {precision} {type(smoothing).__name__} E{'xyz'[a]}: "
              f"{int(np.count_nonzero(h != raw[a]))} smoothed, max rel diff {diff.max():.2e}", flush=True)
        assert diff.max() <= (1e-12 if precision == "f64" else 1.2e-7)
        if smoothing is None:
            np.testing.assert_array_equal(d, h)


def test_gds_layout_simulation_device_raster_equals_host():
    """A layout written as GDSII (a mitred waveguide path and a rotated,
    reflected taper cell on a substrate) read back, extruded and simulated:
    device rasterisation and smoothing give the host's permittivity, so the
    fields after 60 steps are identical."""
    import os, sys
    sys.path.insert(0, os.path.dirname(__file__))
    from gds_writer import Writer
    from khronos_hip import gds as GD
    w = Writer()
    w.begin("TAPER")
    w.boundary([(0.0, -0.1), (0.4, -0.25), (0.4, 0.25), (0.0, 0.1)], layer=1)
    w.end()
    w.begin("TOP")
    w.path([(-0.7, -0.2), (-0.1, -0.2), (0.2, 0.15)], 0.16, layer=1, pathtype=0)
    w.sref("TAPER", (0.15, -0.3), angle=30.0, reflect=True)
    w.boundary([(-0.9, -0.8), (0.9, -0.8), (0.9, 0.8), (-0.9, 0.8)], layer=0)
    w.end()
    lib = GD.read_gds(w.bytes())
    geo = GD.gds_to_objects(lib, "TOP", {(1, 0): (-0.11, 0.11, kh.Material(ε=12.0))})
    geo += GD.gds_to_objects(lib, "TOP", {(0, 0): (-0.6, -0.11, kh.Material(ε=2.1))})
    out = []
    for raster in ("host", "device"):
        src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.4), component=kh.Ez(),
                                center=[-0.5, -0.2, 0.0], size=[0.0, 0.2, 0.2])]
        sim = kh.Simulation(cell_size=[1.8, 1.6, 1.0], resolution=25, boundaries=[[0.2, 0.2]] * 3, sources=src,
                            geometry=geo, precision="f32", subpixel_smoothing=kh.AnisotropicSmoothing(),
                            raster=raster)
        sim.step(60)
        out.append([sim.engine.read_field(c) for c in range(6)])
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)
    assert np.abs(out[0][2]).max() > 0
