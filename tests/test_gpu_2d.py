"""GPU parity of 2D simulations (cell_size z = 0: one z cell, dz = Inf, no z
PML profiles; src/DataStructures.jl:689-712, src/Boundaries.jl:130-162).

Bar as for 3D: EXACT mode equal to the oracle value for value, DFTs within
1e-13; FAST within rel-L2 1e-12 (float64) / 1e-5 (float32), DFTs within
1e-10 / 1e-4.  The scene drives both 2D polarizations (an Ez line source and
an Hz point source) through a high-index strip, a Lorentz + chi3 disc and a
conductor (sigma_D and sigma_B: in 2D the Ey component has a C stage but no U
stage, src/Kernels/Helpers.jl:80-93).  The 2D adjoint gradient from device
runs matches central finite differences.
"""
import math
import sys

import numpy as np
import pytest

import khronos_hip as kh
from oracle_engine import OracleEngine
from test_adjoint import slab2d_problem, directional_errors

pytestmark = pytest.mark.gpu

COMPS = [kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz]


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _sim2d(precision, mode, engine=None):
    geo = [kh.Object(kh.Box([0.0, 0.0, 0.0], [0.4, 3.0, 1.0]), kh.Material(ε=12.0)),
           kh.Object(kh.Sphere([0.5, 0.4, 0.0], 0.25),
                     kh.Material(ε=2.0, chi3=2.0, susceptibilities=[kh.LorentzianSusceptibility(1.0, 0.1, 0.5)])),
           kh.Object(kh.Box([-0.5, 0.5, 0.0], [0.3, 0.3, 1.0]), kh.Material(σD=1.2, σB=0.7))]
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.5), component=kh.Ez(),
                            center=[0.0, -0.8, 0.0], size=[0.4, 0.0, 0.0], amplitude=2.0),
           kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.1, fwidth=0.6), component=kh.Hz(),
                            center=[-0.3, -0.2, 0.0], size=[0.0, 0.0, 0.0], amplitude=3.0)]
    mons = [kh.FluxMonitor(center=[0.0, 0.8, 0.0], size=[1.2, 0.0, 0.0], frequencies=[0.9, 1.0, 1.1]),
            kh.DFTMonitor(component=kh.Ey(), center=[0.0, 0.0, 0.0], size=[1.0, 1.0, 0.0],
                          frequencies=[1.0], decimation=2)]
    return kh.Simulation(cell_size=[2.0, 2.4, 0.0], resolution=10, sources=src, geometry=geo, monitors=mons,
                         boundaries=[[0.3, 0.3], [0.3, 0.3]], precision=precision, mode=mode, engine=engine)


def _pair(precision, mode, steps):
    g = _sim2d(precision, mode)
    o = _sim2d(precision, "exact", engine=OracleEngine)
    _note(f"[2d] {precision} {mode} grid {g.Nx}x{g.Ny} steps {steps}")
    g.step(steps)
    o.step(steps)
    g.engine.sync()
    assert g.ndims == 2 and g.spec.N[2] == 1
    return g, o


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_2d_exact_mode_value_for_value(precision):
    g, o = _pair(precision, "exact", 150)
    for C in COMPS:
        a, b = g.read_field(C()), o.read_field(C())
        assert np.array_equal(a, b), (C.__name__, float(np.max(np.abs(a - b))))
    assert np.any(g.read_field(kh.Ey()) != 0) and np.any(g.read_field(kh.Ez()) != 0)
    for i in range(len(g.spec.dfts)):
        a, b = g.engine.read_dft(i), o.engine.read_dft(i)
        assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(b)


@pytest.mark.parametrize("precision,tol,dtol", [("f64", 1e-12, 1e-10), ("f32", 1e-5, 1e-4)])
def test_2d_fast_mode_matches_oracle(precision, tol, dtol):
    g, o = _pair(precision, "fast", 250)
    fa = [g.read_field(C()) for C in COMPS]
    fb = [o.read_field(C()) for C in COMPS]
    nb = math.sqrt(sum(float(np.sum(b.astype(np.float64) ** 2)) for b in fb))
    err = math.sqrt(sum(float(np.sum((a.astype(np.float64) - b) ** 2)) for a, b in zip(fa, fb))) / nb
    _note(f"[2d] fast {precision} rel {err:.3e}")
    assert err < tol, err
    for i in range(len(g.spec.dfts)):
        a, b = g.engine.read_dft(i), o.engine.read_dft(i)
        assert np.linalg.norm(a - b) <= dtol * np.linalg.norm(b)
    # device flux reduction (line monitor, per unit length in z) vs the host
    # formula on the oracle's DFTs
    from khronos_hip.monitors import flux_from_dft
    md = o.monitors[0].monitor_data
    mons = md.tangential_E_monitors + md.tangential_H_monitors
    oa = flux_from_dft(md, *[o.read_dft(m.monitor_data) for m in mons], [m.monitor_data.gv for m in mons])
    ga = kh.get_flux(g.monitors[0])
    assert np.max(np.abs(ga - oa)) <= dtol * np.max(np.abs(oa)), (ga, oa)


def test_2d_adjoint_gradient_matches_finite_differences():
    opt = slab2d_problem(engine=None, precision="f64", mode="fast")
    rho = 0.3 + 0.4 * np.random.default_rng(42).random(24)
    errs, pairs, _ = directional_errors(opt, rho)
    _note(f"[2d] adjoint directional errors {errs}")
    assert np.median(errs) < 0.03 and max(errs) < 0.05, (errs, pairs)
