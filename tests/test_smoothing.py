"""Subpixel smoothing (src/Geometry.jl:792-955): pins of the restated
GeometryPrimitives pieces and of the smoothing formulas.

* volfrac: exact fraction of a box on one side of a plane, against the
  axis-aligned closed form, the half-by-symmetry plane through the centre,
  complements, and Monte Carlo for random planes;
* surface points: sphere radial projection, box nearest face inside and
  clamp outside;
* a planar interface that cuts voxels at a known fraction gives the
  closed-form values 1/<eps> (component normal to the interface) and <1/eps>
  (tangential) under AnisotropicSmoothing, and 1/<eps> for every component
  under VolumeAveraging;
* the reference's own assertions on a smoothed sphere (test/test_subpixel.jl:
  default off, binary without smoothing, intermediate values only at the
  interface, centre and corner unchanged, bounded, components differ).
"""
import numpy as np
import pytest

import khronos_hip as kh
from khronos_hip.smoothing import surfpt_box, surfpt_sphere, volfrac
from oracle_engine import OracleEngine


def test_volfrac_closed_forms():
    lo, hi = np.zeros((1, 3)), np.array([[1.0, 2.0, 0.5]])
    for ax in range(3):
        for frac in (0.0, 0.25, 0.6, 1.0):
            n = np.zeros((1, 3))
            n[0, ax] = 1.0
            r0 = np.zeros((1, 3))
            r0[0, ax] = frac * hi[0, ax]
            assert abs(volfrac(lo, hi, n, r0)[0] - frac) < 1e-14
            assert abs(volfrac(lo, hi, -n, r0)[0] - (1 - frac)) < 1e-14
    rng = np.random.default_rng(4)
    n = rng.standard_normal((50, 3))
    n /= np.linalg.norm(n, axis=1)[:, None]
    ctr = np.tile((lo + hi) / 2, (50, 1))
    f = volfrac(np.tile(lo, (50, 1)), np.tile(hi, (50, 1)), n, ctr)
    assert np.max(np.abs(f - 0.5)) < 1e-12          # any plane through the centre halves the box
    r0 = rng.random((50, 3)) * hi
    a = volfrac(np.tile(lo, (50, 1)), np.tile(hi, (50, 1)), n, r0)
    b = volfrac(np.tile(lo, (50, 1)), np.tile(hi, (50, 1)), -n, r0)
    assert np.max(np.abs(a + b - 1)) < 1e-12


def test_volfrac_matches_monte_carlo():
    rng = np.random.default_rng(9)
    hi = np.array([1.0, 2.0, 0.5])
    X = rng.random((400_000, 3)) * hi
    for _ in range(12):
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        r0 = rng.random(3) * hi
        mc = np.mean((X - r0) @ n <= 0)
        f = volfrac(np.zeros((1, 3)), hi[None], n[None], r0[None])[0]
        assert abs(f - mc) < 3e-3, (f, mc)


def test_surface_points():
    P = np.array([[0.3, 0.1, -0.2], [2.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    sp, n = surfpt_sphere(P, [0.0, 0.0, 0.0], 1.0)
    assert np.allclose(np.linalg.norm(sp, axis=1), 1.0)
    assert np.allclose(n[:2], P[:2] / np.linalg.norm(P[:2], axis=1)[:, None])
    Q = np.array([[0.1, 0.45, 0.0], [0.8, 0.2, 0.0], [0.8, 0.8, 0.8]])
    sp, n = surfpt_box(Q, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0])
    assert np.allclose(sp[0], [0.1, 0.5, 0.0]) and np.allclose(n[0], [0, 1, 0])
    assert np.allclose(sp[1], [0.5, 0.2, 0.0]) and np.allclose(n[1], [1, 0, 0])
    assert np.allclose(sp[2], [0.5, 0.5, 0.5]) and np.allclose(n[2], np.ones(3) / np.sqrt(3))


def _slab_sim(smoothing, face):
    eps_s, eps_b = 9.0, 2.0
    geo = [kh.Object(kh.Box([face + 0.5, 0.0, 0.0], [1.0, 100.0, 100.0]), kh.Material(ε=eps_s)),
           kh.Object(kh.Box([0.0, 0.0, 0.0], [100.0, 100.0, 100.0]), kh.Material(ε=eps_b))]
    src = [kh.UniformSource(time_profile=kh.ContinuousWaveSource(fcen=1.0), component=kh.Ez(),
                            center=[0.0, 0.0, 0.0], size=[0.0, 0.0, 0.0])]
    sim = kh.Simulation(cell_size=[2.0, 2.0, 2.0], resolution=10, geometry=geo, sources=src,
                        boundaries=[[0.5, 0.5]] * 3, subpixel_smoothing=smoothing, engine=OracleEngine)
    return sim, eps_s, eps_b


@pytest.mark.parametrize("smoothing", [kh.AnisotropicSmoothing(), kh.VolumeAveraging()])
def test_planar_interface_closed_form(smoothing):
    # Ex samples sit at x = -0.95 + 0.1 i (0-based i); Ey/Ez at x = -1.0 + 0.1 i.
    # A face at x = -0.95 + 0.1 k + 0.02 leaves 0.3 of Ex voxel k inside the
    # slab and 0.8 of Ey/Ez voxel k + 1.
    k = 7
    face = -0.95 + 0.1 * k + 0.02
    sim, es, eb = _slab_sim(smoothing, face)
    spec = sim.build_spec()
    ex, ey, ez = spec.eps_inv
    normal_val = 1.0 / (0.3 * es + 0.7 * eb)
    assert np.allclose(ex[k], normal_val, rtol=1e-12)
    assert np.allclose(ex[k + 1], 1.0 / es, rtol=1e-12) and np.allclose(ex[k - 1], 1.0 / eb, rtol=1e-12)
    for arr in (ey, ez):
        if isinstance(smoothing, kh.AnisotropicSmoothing):
            want = 0.8 / es + 0.2 / eb          # tangential: <1/eps>
        else:
            want = 1.0 / (0.8 * es + 0.2 * eb)
        assert np.allclose(arr[k + 1], want, rtol=1e-12)
        assert np.allclose(arr[k], 1.0 / eb, rtol=1e-12) and np.allclose(arr[k + 2], 1.0 / es, rtol=1e-12)


def _sphere_sim(smoothing):
    geo = [kh.Object(kh.Sphere([0.0, 0.0, 0.0], 1.0), kh.Material(ε=12.0)),
           kh.Object(kh.Box([0.0, 0.0, 0.0], [100.0, 100.0, 100.0]), kh.Material(ε=1.0))]
    src = [kh.UniformSource(time_profile=kh.ContinuousWaveSource(fcen=1.0), component=kh.Ez(),
                            center=[0.0, 0.0, 0.0], size=[0.0, 0.0, 0.0])]
    kw = {} if smoothing is None else {"subpixel_smoothing": smoothing}
    return kh.Simulation(cell_size=[4.0, 4.0, 4.0], resolution=10, geometry=geo, sources=src,
                         boundaries=[[1.0, 1.0]] * 3, engine=OracleEngine, **kw)


def test_reference_sphere_assertions():
    inv_s, inv_b = 1 / 12.0, 1.0
    close = lambda v, w: np.isclose(v, w, rtol=1e-4)
    none = _sphere_sim(None)
    assert isinstance(none.subpixel_smoothing, kh.NoSmoothing)
    ex = none.build_spec().eps_inv[0]
    assert np.all(close(ex, inv_s) | close(ex, inv_b))
    for sm in (kh.VolumeAveraging(), kh.AnisotropicSmoothing()):
        e = _sphere_sim(sm).build_spec().eps_inv
        ex = e[0]
        mid = ~close(ex, inv_s) & ~close(ex, inv_b)
        assert np.count_nonzero(mid) > 0
        c = tuple(s // 2 for s in ex.shape)
        assert close(ex[c], inv_s) and close(ex[0, 0, 0], inv_b)
        assert np.all((ex >= inv_s - 1e-5) & (ex <= inv_b + 1e-5))
        if isinstance(sm, kh.AnisotropicSmoothing):
            n = [min(a.shape[i] for a in e) for i in range(3)]
            sl = tuple(slice(0, v) for v in n)
            x, y, z = (a[sl] for a in e)
            inter = ~close(x, inv_s) & ~close(x, inv_b)
            assert np.count_nonzero(inter & (~close(x, y) | ~close(x, z))) > 0
