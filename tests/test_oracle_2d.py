"""2D simulations (cell_size z = 0): host setup and oracle pins.

In 2D the reference keeps every 3D array with one z cell and dz = Inf, so the
z differences drop out through dt/dz = 0 (src/DataStructures.jl:689-712,
src/utils.jl:20-37, src/Fields.jl:236-247); the z PML profiles do not exist
(src/Boundaries.jl:130-162) and GridVolume z indices clamp to 1
(src/utils.jl:103-113).

Pins: the oracle's one-step map with dz = Inf has exactly the 2D cavity
spectrum (the 3D closed form with the z term 4 sin^2(k/2)/dz^2 = 0); a point
Ez source leaves Ex, Ey and Hz zero to the bit (the 2D TE/TM split); host
quantities (Nz, dz, dt, profile set, grid volumes) follow the reference's
2D rules.
"""
import math

import numpy as np

import khronos_hip as kh
from khronos_hip.simulation import SimSpec
from oracle_engine import OracleEngine
from test_oracle_pins import _analytic_lambdas, _one_step_map


def test_oracle_2d_cavity_spectrum():
    N = (3, 4, 1)
    delta = (0.1, 0.13, math.inf)
    dt = 0.5 * min(delta)
    eps_inv, mu_inv = 0.5, 0.8
    spec = SimSpec(N=N, real=np.float64, delta=delta, dt=dt, sigma=None, eps_inv=[None] * 3, mu_inv=[None] * 3,
                   eps_inv_scalar=eps_inv, mu_inv_scalar=mu_inv, sources=[], dfts=[], cutoff=-math.inf,
                   mode="exact", dft_accum_f64=True)
    M = _one_step_map(OracleEngine(spec), N, eps_inv, mu_inv)
    ev = np.linalg.eigvals(M)
    osc = ev[ev.imag > 1e-9]
    lam = np.sort((2 - 2 * osc.real) / dt ** 2) / (eps_inv * mu_inv)
    ref = _analytic_lambdas(N, delta)
    assert len(lam) == len(ref)
    assert np.max(np.abs(lam - ref) / ref) < 1e-9


def _sim2d(sources, engine=OracleEngine, **kw):
    return kh.Simulation(cell_size=[2.0, 2.4, 0.0], resolution=10, sources=sources,
                         boundaries=[[0.3, 0.3], [0.3, 0.3]], engine=engine, **kw)


def test_2d_host_setup_rules():
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.5), component=kh.Ez(),
                            center=[0.0, 0.0, 0.0], size=[0.0, 0.0, 0.0])]
    s = _sim2d(src, precision="f32")
    s.prepare()
    assert (s.Nx, s.Ny, s.Nz, s.ndims) == (20, 24, 0, 2)
    assert s.spec.N == (20, 24, 1) and math.isinf(s.delta_n[2])
    assert s.dt_n == float(np.float32(0.05))
    assert s.spec.sigma["B"][2] is None and s.spec.sigma["D"][2] is None
    assert all(s.spec.sigma[k][a] is not None for k in "BD" for a in range(2))
    for C in (kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz):
        gv = kh.grid.grid_volume(s, kh.Volume([0.3, -0.2, 5.0], [0.5, 0.0, 7.0]), C())
        assert gv.start_idx[2] == 1 and gv.end_idx[2] == 1 and gv.Nz == 1
        cnt = kh.component_voxel_count(s.Nx, s.Ny, s.Nz, C())
        assert cnt[2] == (2 if C in (kh.Ex, kh.Ey, kh.Hz) else 1)


def test_2d_te_tm_split_is_exact():
    geo = [kh.Object(kh.Box([0.0, 0.0, 0.0], [0.4, 3.0, 1.0]), kh.Material(ε=12.0))]
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.5), component=kh.Ez(),
                            center=[0.1, -0.3, 0.0], size=[0.0, 0.0, 0.0])]
    s = _sim2d(src, geometry=geo)
    s.step(120)
    assert np.max(np.abs(s.read_field(kh.Ez()))) > 0
    for C in (kh.Ex, kh.Ey, kh.Hz):
        assert not np.any(s.read_field(C()))
