"""GPU parity with periodic and Bloch boundaries (complex fields).

Bar as for the other paths: EXACT mode equal to the oracle value for value
(real and imaginary parts, polarization), DFTs within 1e-13; FAST mode within
rel-L2 1e-12 (float64) / 1e-5 (float32), DFTs within 1e-10 / 1e-4.  The
Bloch scene wraps x and y with different k, keeps PML on z and carries a
Lorentz + chi3 block (complex polarization; the reference's real P arrays
cannot hold a complex field there) and a source on a wrapped boundary plane.
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


def _bloch_sim(precision, mode, engine=None, bloch=True, materials=True):
    if bloch:
        bcs = [[kh.Bloch(k=1.7), kh.Bloch(k=1.7)], [kh.Bloch(k=-0.9), kh.Bloch(k=-0.9)], [kh.PML(), kh.PML()]]
    else:
        bcs = [[kh.Periodic(), kh.Periodic()], [kh.Periodic(), kh.Periodic()], [kh.PML(), kh.PML()]]
    geo = [kh.Object(kh.Cylinder([0.1, -0.1, 0.0], 0.35, 0.4, axis=2), kh.Material(ε=3.0))]
    if materials:
        geo.insert(0, kh.Object(kh.Box([-0.2, 0.2, 0.1], [0.5, 0.4, 0.3]),
                                kh.Material(ε=2.0, chi3=15.0,
                                            susceptibilities=[kh.LorentzianSusceptibility(1.2, 0.1, 0.7)])))
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.4), component=kh.Ex(),
                            center=[0.0, 0.0, -0.5], size=[1.2, 1.0, 0.0], amplitude=2.0),
           kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=0.9, fwidth=0.3), component=kh.Ez(),
                            center=[-0.55, 0.25, 0.2], size=[0.0, 0.0, 0.0], amplitude=3.0)]
    mons = [kh.FluxMonitor(center=[0.0, 0.0, 0.55], size=[1.2, 1.0, 0.0], frequencies=[0.8, 1.0, 1.2]),
            kh.DFTMonitor(component=kh.Ey(), center=[0.0, 0.0, 0.2], size=[1.2, 1.0, 0.0],
                          frequencies=[0.9, 1.1], decimation=2)]
    return kh.Simulation(cell_size=[1.2, 1.0, 1.8], resolution=10, sources=src, geometry=geo, monitors=mons,
                         boundaries=[[0.0, 0.0], [0.0, 0.0], [0.3, 0.3]], boundary_conditions=bcs,
                         precision=precision, mode=mode, engine=engine)


def _pair(precision, mode, steps, **kw):
    g = _bloch_sim(precision, mode, **kw)
    o = _bloch_sim(precision, "exact", engine=OracleEngine, **kw)
    _note(f"[periodic] {precision} {mode} {kw} grid {g.Nx}x{g.Ny}x{g.Nz} steps {steps}")
    g.step(steps)
    o.step(steps)
    g.engine.sync()
    return g, o


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_bloch_exact_mode_value_for_value(precision):
    g, o = _pair(precision, "exact", 70)
    assert g.spec.complex_fields and g.spec.periodic == (True, True, False)
    for C in COMPS:
        a, b = g.read_field(C()), o.read_field(C())
        assert np.array_equal(a, b), (C.__name__, float(np.max(np.abs(a - b))))
    assert np.any(g.read_field(kh.Hz()).imag != 0)
    for k in range(len(g.spec.poles)):
        for w in (0, 1):
            for a in range(3):
                assert np.array_equal(g.engine.read_pole(k, w, a), o.engine.read_pole(k, w, a))
    for i in range(len(g.spec.dfts)):
        assert rel_l2(g.engine.read_dft(i), o.engine.read_dft(i)) < 1e-13


@pytest.mark.parametrize("precision,tol,dtol", [("f64", 1e-12, 1e-10), ("f32", 1e-5, 1e-4)])
def test_bloch_fast_mode_matches_oracle(precision, tol, dtol):
    g, o = _pair(precision, "fast", 140)
    fa = [g.read_field(C()).astype(np.complex128) for C in COMPS]
    fb = [o.read_field(C()).astype(np.complex128) for C in COMPS]
    nb = math.sqrt(sum(float(np.sum(np.abs(b) ** 2)) for b in fb))
    tot = math.sqrt(sum(float(np.sum(np.abs(a - b) ** 2)) for a, b in zip(fa, fb))) / nb
    assert tot < tol, tot
    for i in range(len(g.spec.dfts)):
        assert rel_l2(g.engine.read_dft(i), o.engine.read_dft(i)) < dtol


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_periodic_real_fields_match_oracle(mode):
    g, o = _pair("f64", mode, 100, bloch=False)
    assert not g.spec.complex_fields
    tol = 0.0 if mode == "exact" else 1e-12
    fa = [g.read_field(C()) for C in COMPS]
    fb = [o.read_field(C()) for C in COMPS]
    nb = math.sqrt(sum(float(np.sum(b * b)) for b in fb))
    tot = math.sqrt(sum(float(np.sum((a - b) ** 2)) for a, b in zip(fa, fb))) / nb
    assert tot <= tol, tot


def test_bloch_flux_on_device_matches_host():
    g, _ = _pair("f64", "fast", 80, materials=False)
    from khronos_hip.monitors import flux_from_dft
    fm = g.monitors[0]
    md = fm.monitor_data
    mons = md.tangential_E_monitors + md.tangential_H_monitors
    arrs = [g.read_dft(m.monitor_data) for m in mons]
    host = flux_from_dft(md, *arrs, [m.monitor_data.gv for m in mons])
    dev = kh.get_flux(fm)
    np.testing.assert_allclose(dev, host, rtol=1e-12, atol=1e-14 * np.max(np.abs(host)))


def test_complex_field_write_read_roundtrip():
    g = _bloch_sim("f32", "fast", materials=False)
    g.prepare()
    rng = np.random.default_rng(2)
    for C in COMPS:
        code = kh.grid.component_code(C())
        shp = g.engine.field_shape(code)
        F = (rng.standard_normal(shp) + 1j * rng.standard_normal(shp)).astype(np.complex64)
        g.engine.write_field(code, F)
        back = g.engine.read_field(code)
        n = (g.Nx + 2, g.Ny + 2, g.Nz + 2)
        box = tuple(slice(0, min(shp[a], n[a])) for a in range(3))
        assert np.array_equal(back[box], F[box])


@pytest.mark.parametrize("cell_x,cell_y", [(1.2, 1.0), (1.3, 1.0), (1.2, 1.1)])
@pytest.mark.parametrize("bloch", [True, False])
def test_folded_x_wrap_equals_ghost_fill(cell_x, cell_y, bloch):
    """FAST with a periodic x axis: the updates read the x-neighbour across
    the boundary from the opposite interior plane (times the Bloch phase)
    and the x ghost planes are filled only before a field is read; the DFT
    samples of monitors that span the ghost plane are formed from the
    opposite plane.  A context whose fields were written keeps filling the
    ghosts every half step.  Both give the same fields (ghosts included),
    polarization and DFTs value for value, for even and odd Nx and Ny, with a
    source on the wrapped boundary planes."""
    from khronos_hip.grid import component_code

    def mk():
        s = _bloch_sim("f32", "fast", bloch=bloch)
        s2 = kh.Simulation(cell_size=[cell_x, cell_y, 1.8], resolution=10, sources=s.sources, geometry=s.geometry,
                           monitors=[kh.FluxMonitor(center=[0.0, 0.0, 0.55], size=[cell_x, cell_y, 0.0],
                                                    frequencies=[0.8, 1.0, 1.2]),
                                     kh.DFTMonitor(component=kh.Ey(), center=[0.0, 0.0, 0.2],
                                                   size=[cell_x, cell_y, 0.0], frequencies=[0.9, 1.1],
                                                   decimation=2)],
                           boundaries=[[0.0, 0.0], [0.0, 0.0], [0.3, 0.3]], boundary_conditions=s.boundary_conditions,
                           precision="f32", mode="fast")
        s2.prepare()
        return s2

    fold, fill = mk(), mk()
    z = fill.read_field(kh.Ez())
    fill.engine.write_field(component_code(kh.Ez()), z)  # a field write keeps the per-step ghost fill
    _note(f"[periodic] folded x wrap vs ghost fill, bloch={bloch}, grid {fold.Nx}x{fold.Ny}x{fold.Nz}")
    assert fold.Nx % 2 == (1 if cell_x == 1.3 else 0) and fold.Ny % 2 == (1 if cell_y == 1.1 else 0)
    done = 0
    for chunk in (23, 40, 17, 60):
        fold.step(chunk)
        fill.step(chunk)
        done += chunk
        for C in COMPS:
            a, b = fold.read_field(C()), fill.read_field(C())
            assert np.array_equal(a, b), (done, C.__name__)
        for w in (0, 1):
            for a in range(3):
                assert np.array_equal(fold.engine.read_pole(0, w, a), fill.engine.read_pole(0, w, a))
        for i in range(len(fold.spec.dfts)):
            assert np.array_equal(fold.engine.read_dft(i), fill.engine.read_dft(i)), (done, i)
    assert np.any(fold.engine.read_dft(0) != 0)
