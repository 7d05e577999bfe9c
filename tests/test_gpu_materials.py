"""GPU parity of the material pass: chi3 Kerr and Lorentz/Drude poles.

Same bar as the linear path: EXACT mode equal to the oracle value for value
(fields, polarization arrays) with DFTs within 1e-13; FAST mode within rel-L2
1e-12 (float64) / 1e-5 (float32) of the oracle, DFTs within 1e-10 / 1e-4.
The scene has a Lorentz + chi3 block with per-voxel permittivity (the chi1
correction lands where the reference's linear indexing puts it), a Drude
sphere on the scalar permittivity, a chi3 slab that reaches into the PML, and
an E source inside the nonlinear block.
"""
import math
import sys

import numpy as np
import pytest

import khronos_hip as kh
from oracle_engine import OracleEngine

pytestmark = pytest.mark.gpu

COMPS = [kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz]


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / (nb if nb > 0 else 1.0))


def _mat_sim(precision, mode, engine=None, n=(26, 22, 20), res=10, amp=3.0):
    cell = [n[0] / res, n[1] / res, n[2] / res]
    lor = kh.LorentzianSusceptibility(1.1, 0.08, 0.9)
    dru = kh.DrudeSusceptibility(0.3, 1.5)
    geo = [kh.Object(kh.Box([0.0, 0.0, 0.0], [0.7, 0.6, 0.5]),
                     kh.Material(ε=2.0, chi3=30.0, susceptibilities=[lor])),
           kh.Object(kh.Sphere([-0.55, 0.35, 0.2], 0.3), kh.Material(susceptibilities=[dru, lor])),
           kh.Object(kh.Box([0.0, 0.0, -0.85], [3.0, 3.0, 0.2]), kh.Material(chi3=20.0))]
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.4), component=kh.Ez(),
                            center=[0.05, -0.05, 0.0], size=[0.0, 0.0, 0.0], amplitude=amp),
           kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=0.9, fwidth=0.3), component=kh.Ex(),
                            center=[-0.3, 0.1, -0.6], size=[0.0, 0.4, 0.0], amplitude=amp)]
    mons = [kh.FluxMonitor(center=[0.5, 0.0, 0.0], size=[0.0, 0.8, 0.6], frequencies=[0.8, 1.0, 1.2]),
            kh.DFTMonitor(component=kh.Ey(), center=[0.0, 0.0, 0.1], size=[0.6, 0.6, 0.0],
                          frequencies=[0.9, 1.1], decimation=2)]
    return kh.Simulation(cell_size=cell, resolution=res, sources=src, geometry=geo, monitors=mons,
                         boundaries=[[0.3, 0.3]] * 3, precision=precision, mode=mode, engine=engine)


def _pair(precision, mode, steps):
    g = _mat_sim(precision, mode)
    o = _mat_sim(precision, "exact", engine=OracleEngine)
    _note(f"[materials] {precision} {mode} grid {g.Nx}x{g.Ny}x{g.Nz} steps {steps}")
    g.prepare()
    o.prepare()
    assert g.spec.chi3 is not None and len(g.spec.poles) == 2
    g.step(steps)
    o.step(steps)
    g.engine.sync()
    return g, o


def _poles(sim):
    return [[sim.engine.read_pole(k, w, a) for w in (0, 1) for a in range(3)] for k in range(len(sim.spec.poles))]


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_materials_exact_mode_value_for_value(precision):
    g, o = _pair(precision, "exact", 80)
    for C in COMPS:
        a, b = g.read_field(C()), o.read_field(C())
        assert np.array_equal(a, b), (C.__name__, float(np.max(np.abs(a - b))))
    for pg, po in zip(_poles(g), _poles(o)):
        for a, b in zip(pg, po):
            assert np.array_equal(a, b)
    assert any(np.any(a != 0) for a in _poles(g)[0])
    for i in range(len(g.spec.dfts)):
        assert rel_l2(g.engine.read_dft(i), o.engine.read_dft(i)) < 1e-13


@pytest.mark.parametrize("precision,tol,dtol", [("f64", 1e-12, 1e-10), ("f32", 1e-5, 1e-4)])
def test_materials_fast_mode_matches_oracle(precision, tol, dtol):
    g, o = _pair(precision, "fast", 150)
    fa = [g.read_field(C()).astype(np.float64) for C in COMPS]
    fb = [o.read_field(C()).astype(np.float64) for C in COMPS]
    nb = math.sqrt(sum(float(np.sum(b * b)) for b in fb))
    tot = math.sqrt(sum(float(np.sum((a - b) ** 2)) for a, b in zip(fa, fb))) / nb
    assert tot < tol, tot
    for pg, po in zip(_poles(g), _poles(o)):
        num = math.sqrt(sum(float(np.sum((a.astype(np.float64) - b) ** 2)) for a, b in zip(pg, po)))
        den = math.sqrt(sum(float(np.sum(b.astype(np.float64) ** 2)) for b in po))
        assert num / den < tol, num / den
    for i in range(len(g.spec.dfts)):
        assert rel_l2(g.engine.read_dft(i), o.engine.read_dft(i)) < dtol


def test_materials_change_the_fields():
    """The material pass is live: the same scene without chi3 and poles differs."""
    g = _mat_sim("f64", "fast")
    g.prepare()
    g.step(120)
    lin = _mat_sim("f64", "fast")
    lin.prepare()
    spec = lin.spec
    spec.chi3 = None
    spec.poles = []
    from khronos_hip.engine import HipEngine
    lin.engine = HipEngine(spec)
    lin.step(120)
    d = rel_l2(g.read_field(kh.Ez()), lin.read_field(kh.Ez()))
    assert d > 1e-3, d


def test_materials_reset_reproduces_bit_for_bit():
    g = _mat_sim("f32", "fast")
    g.step(50)
    first = [g.read_field(C()) for C in COMPS] + _poles(g)[0]
    kh.reset_fields(g)
    g.step(50)
    second = [g.read_field(C()) for C in COMPS] + _poles(g)[0]
    for a, b in zip(first, second):
        assert np.array_equal(a, b)


def test_pole_write_read_roundtrip():
    g = _mat_sim("f64", "exact")
    g.prepare()
    rng = np.random.default_rng(1)
    shp = g.engine.pole_shape()
    X = rng.standard_normal(shp)
    g.engine.write_pole(0, 1, 2, X)
    Y = g.engine.read_pole(0, 1, 2)
    nz = np.nonzero(g.spec.poles[0].sigma)
    box = tuple(slice(int(i.min()) + 1, int(i.max()) + 2) for i in nz)  # grid index = array index + 1
    mask = np.zeros(shp, dtype=bool)
    mask[box] = True
    assert np.array_equal(Y[mask], X[mask])
    assert np.all(Y[~mask] == 0)


@pytest.mark.parametrize("precision,mode", [("f32", "fast"), ("f64", "fast"), ("f64", "exact")])
def test_inactive_material_cells_skipped(precision, mode):
    """The material pass skips the cells of its box where no pole's sigma and
    no chi3 is non-zero: P stays zero there, so fPD, the FAST corrections and
    the chi3 delta are zero and the pass would store what it read.  A host
    write of a polarization array turns the skip off for good (it may put a
    non-zero P where sigma is zero).  Both runs give the same fields,
    polarization and DFTs (equal as numbers: only the sign of a zero may
    differ, where the full pass adds a zero correction to -0)."""
    skip = _mat_sim(precision, mode)
    full = _mat_sim(precision, mode)
    skip.prepare()
    full.prepare()
    full.engine.write_pole(0, 0, 0, full.engine.read_pole(0, 0, 0))  # zeros back: skip off
    _note(f"[materials] skip vs full pass, {precision} {mode}")
    for chunk in (40, 80):
        skip.step(chunk)
        full.step(chunk)
        for C in COMPS:
            assert np.array_equal(skip.read_field(C()), full.read_field(C())), C.__name__
        for pa, pb in zip(_poles(skip), _poles(full)):
            for x, y in zip(pa, pb):
                assert np.array_equal(x, y)
        for i in range(len(skip.spec.dfts)):
            assert np.array_equal(skip.engine.read_dft(i), full.engine.read_dft(i))
    assert any(np.any(skip.read_field(C()) != 0) for C in COMPS)
