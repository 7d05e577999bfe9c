"""GPU parity of the conductive (C) stage: material sigma_D / sigma_B and
adiabatic absorbers.

Bar as for the other paths: EXACT mode equal to the oracle value for value,
DFTs within 1e-13; FAST mode within rel-L2 1e-12 (float64) / 1e-5 (float32),
DFTs within 1e-10 / 1e-4.  Scenes: a PML box with a conductive sphere
(sigma_D), a magnetic-conductor slab (sigma_B only), an absorber on one x side
and a lossless dielectric; the same conductors with zero PML thickness (no PML
at all: FAST mode then runs the C stage in the whole-grid kernel); and a Bloch
scene whose conductive block sits on the wrapped axes (complex fields).
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


def _rel(fa, fb):
    nb = math.sqrt(sum(float(np.sum(np.abs(b) ** 2)) for b in fb))
    return math.sqrt(sum(float(np.sum(np.abs(a - b) ** 2)) for a, b in zip(fa, fb))) / nb


def _cond_sim(precision, mode, engine=None, scene="pml"):
    geo = [kh.Object(kh.Sphere([0.15, -0.1, 0.0], 0.3), kh.Material(ε=2.5, σD=1.8)),
           kh.Object(kh.Box([-0.3, 0.2, 0.3], [0.4, 0.3, 0.2]), kh.Material(σB=0.9, σDx=0.4)),
           kh.Object(kh.Box([0.0, 0.0, -0.4], [0.6, 0.6, 0.2]), kh.Material(ε=4.0))]
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.5), component=kh.Ez(),
                            center=[-0.35, -0.05, -0.1], size=[0.0, 0.0, 0.0], amplitude=2.0),
           kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.2, fwidth=0.6), component=kh.Hy(),
                            center=[0.1, 0.1, 0.35], size=[0.4, 0.0, 0.0])]
    mons = [kh.FluxMonitor(center=[0.0, 0.0, 0.5], size=[1.0, 0.8, 0.0], frequencies=[0.8, 1.0, 1.2]),
            kh.DFTMonitor(component=kh.Ex(), center=[0.0, 0.0, 0.0], size=[0.8, 0.8, 0.0],
                          frequencies=[0.9, 1.1], decimation=2)]
    bcs, absorbers = None, None
    if scene == "pml":
        bnd = [[0.3, 0.3]] * 3
        absorbers = [[kh.Absorber(num_layers=4, sigma_order=2), None], None, None]
    elif scene == "nopml":
        bnd = [[0.0, 0.0]] * 3
        absorbers = [None, [None, kh.Absorber(num_layers=3)], None]
    else:  # bloch
        bnd = [[0.0, 0.0], [0.0, 0.0], [0.3, 0.3]]
        bcs = [[kh.Bloch(k=1.3), kh.Bloch(k=1.3)], [kh.Periodic(), kh.Periodic()], [kh.PML(), kh.PML()]]
    return kh.Simulation(cell_size=[1.4, 1.2, 1.6], resolution=10, sources=src, geometry=geo, monitors=mons,
                         boundaries=bnd, boundary_conditions=bcs, absorbers=absorbers, precision=precision,
                         mode=mode, engine=engine)


def _pair(precision, mode, scene, steps):
    g = _cond_sim(precision, mode, scene=scene)
    o = _cond_sim(precision, "exact", engine=OracleEngine, scene=scene)
    _note(f"[conductivity] {precision} {mode} {scene} grid {g.Nx}x{g.Ny}x{g.Nz} steps {steps}")
    g.prepare()
    o.prepare()
    assert all(c is not None for c in g.spec.cond[:3]) and g.spec.cond[3] is not None
    g.step(steps)
    o.step(steps)
    g.engine.sync()
    return g, o


@pytest.mark.parametrize("scene", ["pml", "nopml", "bloch"])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_conductivity_exact_mode_value_for_value(precision, scene):
    g, o = _pair(precision, "exact", scene, 90)
    for C in COMPS:
        a, b = g.read_field(C()), o.read_field(C())
        assert np.array_equal(a, b), (C.__name__, float(np.max(np.abs(a - b))))
    for i in range(len(g.spec.dfts)):
        a, b = g.engine.read_dft(i), o.engine.read_dft(i)
        assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(b)


@pytest.mark.parametrize("scene", ["pml", "nopml", "bloch"])
@pytest.mark.parametrize("precision,tol,dtol", [("f64", 1e-12, 1e-10), ("f32", 1e-5, 1e-4)])
def test_conductivity_fast_mode_matches_oracle(precision, tol, dtol, scene):
    g, o = _pair(precision, "fast", scene, 160)
    fa = [g.read_field(C()).astype(np.complex128) for C in COMPS]
    fb = [o.read_field(C()).astype(np.complex128) for C in COMPS]
    err = _rel(fa, fb)
    _note(f"[conductivity] fast {precision} {scene} rel {err:.3e}")
    assert err < tol, err
    for i in range(len(g.spec.dfts)):
        a, b = g.engine.read_dft(i), o.engine.read_dft(i)
        assert np.linalg.norm(a - b) <= dtol * np.linalg.norm(b)


def test_conductor_damps_the_field():
    """Physical sanity on the device: the same scene with every sigma scaled by
    4 ends with less field energy."""
    def energy(scale):
        s = _cond_sim("f64", "fast", scene="pml")
        for obj in s.geometry:
            m = obj.material
            for k in ("σD", "σDx", "σB"):
                if getattr(m, k) is not None:
                    setattr(m, k, getattr(m, k) * scale)
        s.step(200)
        return sum(float(np.sum(s.read_field(C()) ** 2)) for C in COMPS)
    e1, e4 = energy(1.0), energy(4.0)
    assert e4 < e1
