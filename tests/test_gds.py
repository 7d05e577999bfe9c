"""GDSII import (src/GDS.jl) and the prism shape it produces.

Pins: the reference's own unit tests (test/test_gds.jl:14-131: excess-64
reals, path caps, convexity and ear clipping, SREF transforms); a GDS file
written here and read back (cells, layers, paths, SREF with reflection and
rotation, AREF instances at their pitches, units); a square prism equals the
box of the same extent (membership and nearest surface point, the box's
rule being pinned in tests/test_smoothing.py); painted prism volume against
the polygon's area; prisms through the host smoothing.  The reference's GDS
fixture files are not in the checkout, so its file-based tests are
restated on files written by tests/gds_writer.py.
"""
import math

import numpy as np
import pytest

import khronos_hip as kh
from khronos_hip import gds as GD
from khronos_hip.geometry import Box, Prism, rasterize
from khronos_hip.smoothing import surfpt_box, surfpt_prism
from gds_writer import Writer


def test_decode_gdsii_real_reference_values():
    assert GD.decode_gdsii_real(bytes(8)) == 0.0
    assert GD.decode_gdsii_real(bytes([0x41, 0x10, 0, 0, 0, 0, 0, 0])) == pytest.approx(1.0)
    assert GD.decode_gdsii_real(bytes([0xC1, 0x10, 0, 0, 0, 0, 0, 0])) == pytest.approx(-1.0)
    assert GD.decode_gdsii_real(bytes([0x3E, 0x41, 0x89, 0x37, 0x4B, 0xC6, 0xA7, 0xF0])) == pytest.approx(0.001,
                                                                                                         rel=1e-10)
    v = GD.decode_gdsii_real(bytes([0x3A, 0x44, 0xB8, 0x2F, 0xA0, 0x9B, 0x5A, 0x54]))
    assert 0 < v < 1e-6
    # the usual 1 nm database unit is the same fraction one base-16 exponent lower
    nm = GD.decode_gdsii_real(bytes([0x39, 0x44, 0xB8, 0x2F, 0xA0, 0x9B, 0x5A, 0x54]))
    assert nm == pytest.approx(1e-9, rel=1e-15) and v == 16.0 * nm
    from gds_writer import real8
    for x in (1e-3, 1e-9, 0.5, 90.0, -2.25):
        assert GD.decode_gdsii_real(real8(x)) == pytest.approx(x, rel=1e-15)


def test_expand_path_reference_cases():
    c = [(0.0, 0.0), (10.0, 0.0)]
    p = GD.expand_path(c, 1.0, 0)
    xs, ys = [q[0] for q in p], [q[1] for q in p]
    assert len(p) == 4 and min(xs) == pytest.approx(0.0) and max(xs) == pytest.approx(10.0)
    assert min(ys) == pytest.approx(-1.0) and max(ys) == pytest.approx(1.0)
    xs2 = [q[0] for q in GD.expand_path(c, 1.0, 2)]
    assert min(xs2) == pytest.approx(-1.0) and max(xs2) == pytest.approx(11.0)
    xs4 = [q[0] for q in GD.expand_path(c, 1.0, 4, 0.5, 2.0)]
    assert min(xs4) == pytest.approx(-0.5) and max(xs4) == pytest.approx(12.0)
    L = GD.expand_path([(0.0, 0.0), (5.0, 0.0), (5.0, 5.0)], 0.5, 0)
    assert len(L) >= 4
    # the miter corner of an L of half-width 0.5: outer (5.5, -0.5), inner (4.5, 0.5);
    # area = centre length 10 x width 1
    assert (5.5, -0.5) in [(round(x, 12), round(y, 12)) for x, y in L]
    assert (4.5, 0.5) in [(round(x, 12), round(y, 12)) for x, y in L]
    assert abs(GD.signed_area(L)) == pytest.approx(10.0)


def test_convexity_and_ear_clipping():
    square = [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)]
    lshape = [(0.0, 0.0), (2.0, 0.0), (2.0, 1.0), (1.0, 1.0), (1.0, 2.0), (0.0, 2.0)]
    assert GD.is_convex(square) and not GD.is_convex(lshape)
    tris = GD.ear_clip_triangulate(lshape)
    assert len(tris) == 4 and all(len(t) == 3 and abs(GD.signed_area(t)) > 0 for t in tris)
    assert sum(abs(GD.signed_area(t)) for t in tris) == pytest.approx(abs(GD.signed_area(lshape)), rel=1e-10)
    # clockwise input gives the same cover
    cw = GD.ear_clip_triangulate(list(reversed(lshape)))
    assert sum(abs(GD.signed_area(t)) for t in cw) == pytest.approx(3.0)


def test_transform_gds_point_reference_values():
    T = GD.transform_gds_point
    assert T(1.0, 2.0, 1.0, False, 0.0, (0.0, 0.0)) == pytest.approx((1.0, 2.0))
    assert T(1.0, 2.0, 1.0, False, 0.0, (10.0, 20.0)) == pytest.approx((11.0, 22.0))
    x, y = T(1.0, 0.0, 1.0, False, math.pi / 2, (0.0, 0.0))
    assert abs(x) < 1e-12 and y == pytest.approx(1.0)
    assert T(1.0, 2.0, 1.0, True, 0.0, (0.0, 0.0)) == pytest.approx((1.0, -2.0))
    assert T(1.0, 2.0, 3.0, False, 0.0, (0.0, 0.0)) == pytest.approx((3.0, 6.0))
    x, y = T(1.0, 1.0, 2.0, True, math.pi / 2, (5.0, 10.0))
    assert x == pytest.approx(7.0, abs=1e-12) and y == pytest.approx(12.0, abs=1e-12)


def _library():
    w = Writer()
    arm = [(0.0, 0.5), (3.0, 0.5), (3.0, 1.0), (0.0, 1.0)]
    w.begin("ARM")
    w.boundary(arm, layer=1)
    w.end()
    w.begin("TOP")
    w.boundary([(-5.0, -3.0), (5.0, -3.0), (5.0, 3.0), (-5.0, 3.0)], layer=0)
    w.sref("ARM", (1.0, 0.0))
    w.sref("ARM", (1.0, 0.0), reflect=True)
    w.sref("ARM", (0.0, 0.0), angle=90.0, mag=2.0)
    w.aref("ARM", (-4.0, -2.0), 2, 3, (0.5, 0.0), (0.0, 0.25))
    w.path([(0.0, 2.0), (4.0, 2.0), (4.0, 2.5)], 0.2, layer=2, pathtype=2)
    w.end()
    return GD.read_gds(w.bytes()), arm


def test_written_library_reads_back(tmp_path):
    lib, arm = _library()
    path = tmp_path / "t.gds"
    w = Writer()
    w.begin("ONE")
    w.boundary([(0.0, 0.0), (1.0, 0.0), (0.0, 1.0)], layer=7, datatype=3)
    w.end()
    path.write_bytes(w.bytes())
    one = GD.read_gds(str(path))
    assert one.cells["ONE"].polygons[0].layer == 7 and one.cells["ONE"].polygons[0].datatype == 3
    assert one.precision == pytest.approx(1e-9) and one.unit == pytest.approx(1e-9)  # the reference's field
    assert lib.name == "library" and sorted(lib.cells) == ["ARM", "TOP"]
    a = lib.cells["ARM"]
    assert len(a.polygons) == 1 and len(a.polygons[0].vertices) == 4   # closing point dropped
    assert a.polygons[0].vertices == pytest.approx(arm)
    top = lib.cells["TOP"]
    assert len(top.polygons) == 2 and len(top.references) == 4
    assert GD.get_layers(top) == {(0, 0), (2, 0)}
    flat = GD.flatten_cell(lib, "TOP")
    assert len(flat) == 2 + 3 + 6
    arms = GD.flatten_cell(lib, "TOP", 1, 0)
    assert len(arms) == 9 and len(GD.flatten_cell(lib, "TOP", layers=[(0, 0), (2, 0)])) == 2
    ys0 = [v[1] for v in arms[0].vertices]
    ys1 = [v[1] for v in arms[1].vertices]
    assert min(ys0) * min(ys1) < 0                       # the reflected copy
    rot = np.array(arms[2].vertices)                     # rotated 90 degrees, magnified 2
    assert rot == pytest.approx(np.array([(-2 * y, 2 * x) for x, y in arm]))
    inst = sorted((round(p.vertices[0][0], 9), round(p.vertices[0][1], 9)) for p in arms[3:])
    want = sorted((round(-4.0 + 0.5 * c + arm[0][0], 9), round(-2.0 + 0.25 * r + arm[0][1], 9))
                  for c in range(2) for r in range(3))
    assert inst == want
    path_poly = GD.flatten_cell(lib, "TOP", 2, 0)[0]
    xs = [v[0] for v in path_poly.vertices]
    assert min(xs) == pytest.approx(-0.1) and abs(GD.signed_area(path_poly.vertices)) == pytest.approx(
        0.2 * (4.0 + 0.5) + 0.2 * 0.2, rel=1e-9)


def test_gds_to_objects_prisms():
    lib, _ = _library()
    objs = GD.gds_to_objects(lib, "TOP", {(0, 0): (-1.0, 0.0, kh.Material(ε=2.1)),
                                          (2, 0): (0.0, 0.22, kh.Material(ε=12.0))})
    assert len(objs) >= 2
    for o in objs:
        lo, hi = o.shape.bounds()
        assert np.all(np.isfinite(lo)) and np.all(np.isfinite(hi)) and np.all(lo < hi)
        assert GD.signed_area(list(o.shape.vertices)) > 0            # counter-clockwise
    # axis places only the centre (the reference's identity-axes prisms)
    o3 = GD.polygon_to_objects([(0, 0), (1, 0), (0, 1)], 2.0, 4.0, kh.Material(ε=2.0), axis=3)[0]
    o2 = GD.polygon_to_objects([(0, 0), (1, 0), (0, 1)], 2.0, 4.0, kh.Material(ε=2.0), axis=2)[0]
    assert o3.shape.center == pytest.approx((1 / 3, 1 / 3, 3.0)) and o2.shape.center == pytest.approx((1 / 3, 3.0, 1 / 3))


def test_square_prism_is_the_box():
    rng = np.random.default_rng(0)
    c, s, h = np.array([0.3, -0.2, 0.1]), np.array([0.8, 0.6, 0.5]), 0.5
    pr = Prism(c, [(-0.4, -0.3), (0.4, -0.3), (0.4, 0.3), (-0.4, 0.3)], h)
    bx = Box(c, s)
    P = c + rng.uniform(-1.0, 1.0, size=(4000, 3))
    assert np.array_equal(pr.contains(P[:, 0], P[:, 1], P[:, 2]), bx.contains(P[:, 0], P[:, 1], P[:, 2]))
    sp1, n1 = surfpt_prism(P, c, pr.vertices, h)
    sp2, n2 = surfpt_box(P, c, s)
    # outside, the prism's normal is (P - nearest boundary point) / distance:
    # rounding of the nearest point along the edge grows as 1 / distance
    assert np.allclose(sp1, sp2, atol=1e-14) and np.allclose(n1, n2, atol=1e-11)


def test_prism_normal_on_a_vertex_is_an_edge_normal():
    # A point exactly on a vertex is on the boundary; rounding can put it a
    # hair outside (largest edge distance > 0) while its distance to the
    # polygon is 0.  The normal must then be an adjacent edge's unit normal,
    # not 0/0 (before this rule, 801 of these 2000 triangles gave NaN).
    rng = np.random.default_rng(0)
    n_out = 0
    for _ in range(2000):
        v = rng.uniform(-1.0, 1.0, (3, 2))
        if (v[1, 0] - v[0, 0]) * (v[2, 1] - v[0, 1]) - (v[1, 1] - v[0, 1]) * (v[2, 0] - v[0, 0]) < 0:
            v = v[::-1].copy()
        P = np.array([[v[k, 0], v[k, 1], 0.0] for k in range(3)])
        sp, nrm = surfpt_prism(P, [0.0, 0.0, 0.0], v, 1.0)
        assert np.all(np.isfinite(nrm)) and np.all(np.isfinite(sp))
        for k in range(3):
            edges = []
            for i in ((k - 1) % 3, k):
                e = v[(i + 1) % 3] - v[i]
                edges.append(np.array([e[1], -e[0], 0.0]) / np.hypot(*e))
            assert min(np.abs(nrm[k] - e).max() for e in edges) < 1e-15
            assert np.allclose(sp[k], P[k], atol=1e-15)
        n_out += 1
    assert n_out == 2000


def test_prism_raster_volume_and_smoothing():
    tri = [(0.0, 0.0), (1.2, 0.1), (0.3, 0.9)]
    objs = GD.polygon_to_objects(tri, -0.2, 0.2, kh.Material(ε=4.0))
    res = 40
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.4), component=kh.Ez(),
                            center=[0, 0, 0], size=[0, 0, 0])]
    sim = kh.Simulation(cell_size=[2.0, 2.0, 1.0], cell_center=[0.5, 0.4, 0.0], resolution=res, sources=src,
                        geometry=objs, boundaries=[[0.2, 0.2]] * 3)
    e, _ = rasterize(sim, sim.geometry)
    inside = np.count_nonzero(np.isclose(e[2], 0.25)) / res ** 3
    assert inside == pytest.approx(abs(GD.signed_area(tri)) * 0.4, rel=0.05)
    from khronos_hip.smoothing import apply_subpixel_smoothing
    sim2 = kh.Simulation(cell_size=[2.0, 2.0, 1.0], cell_center=[0.5, 0.4, 0.0], resolution=res, sources=src,
                         geometry=objs, boundaries=[[0.2, 0.2]] * 3, subpixel_smoothing=kh.VolumeAveraging())
    es = apply_subpixel_smoothing(sim2, sim2.geometry, rasterize(sim2, sim2.geometry)[0])
    v = np.asarray(es[2])
    assert v.min() >= 0.25 - 1e-12 and v.max() <= 1.0 + 1e-12
    assert np.count_nonzero((v > 0.25 + 1e-9) & (v < 1.0 - 1e-9)) > 0   # interface cells averaged
