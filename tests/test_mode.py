"""Mode solver, mode source and mode monitor (host side and oracle runs).

The reference takes its modes from the VectorModesolver package (not
vendored); the solver here is pinned by the mathematics:

* a slab (uniform along u between conducting walls that leave E_u free) has
  the analytic TE slab mode: the solver's n_eff converges to it at second
  order (error x1/4 per halving of the grid step, grids aligned with the
  slab faces);
* the mode carries P = 1/2 Re int (E x H*) . w dA = 1/2 and E, H satisfy the
  transverse curl relations used to build H (checked through the overlap);
* end to end on the oracle: a mode source in a strip waveguide launches the
  mode one way (the monitor behind the source sees |a-|^2 below 3% of the
  forward |a+|^2), and the forward mode monitor's |a+|^2 equals the Poynting
  flux through the same plane within 10% (the mode is normalized to P = 1/2,
  the flux has no 1/2).
"""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

import khronos_hip as kh
from khronos_hip.mode import solve_modes
from khronos_hip.monitors import flux_from_dft, mode_amplitudes_from_dft
from oracle_engine import OracleEngine


def _te_slab_neff(n1, n2, t, k0):
    a = t / 2
    f = lambda b: (math.sqrt(k0 * k0 * n1 * n1 - b * b) * math.tan(math.sqrt(k0 * k0 * n1 * n1 - b * b) * a)
                   - math.sqrt(b * b - k0 * k0 * n2 * n2))
    bmin = math.sqrt(max(k0 * k0 * n1 * n1 - (math.pi / (2 * a)) ** 2, (k0 * n2) ** 2)) * 1.0000001
    return brentq(f, bmin, k0 * n1 * 0.9999999) / k0


def test_slab_mode_converges_at_second_order():
    n1, n2, t, k0 = 3.5, 1.5, 0.3, 2 * math.pi / 1.55
    ref = _te_slab_neff(n1, n2, t, k0)
    eps = lambda comp, U, V: np.where(np.abs(V) <= t / 2, n1 * n1, n2 * n2)
    errs = []
    for res, v0 in ((50, -1.0), (100, -1.0 + 0.005)):     # slab faces between grid nodes
        m = solve_modes(eps, -0.5, 0.5, v0, v0 + 2.0, res, k0, num_modes=1)[0]
        errs.append(m.neff - ref)
        assert np.max(np.abs(m.Ev)) < 1e-12 * np.max(np.abs(m.Eu))      # pure TE
        du = m.uh[1] - m.uh[0]
        dv = m.vn[1] - m.vn[0]
        P = 0.5 * (np.sum(m.Eu * np.conj(m.Hv)) - np.sum(m.Ev * np.conj(m.Hu))).real * du * dv
        assert abs(P - 0.5) < 1e-12
    assert abs(errs[0]) < 3e-3
    assert 3.5 < errs[0] / errs[1] < 4.5, errs


def _waveguide(engine=OracleEngine, precision="f64", mode="exact"):
    fcen = 1 / 1.55
    geo = [kh.Object(kh.Box([0, 0, 0], [10, 0.45, 0.25]), kh.Material(ε=12.0)),
           kh.Object(kh.Box([0, 0, 0], [10, 10, 10]), kh.Material(ε=2.1))]
    src = [kh.ModeSource(frequency=fcen, time_profile=kh.GaussianPulseSource(fcen=fcen, fwidth=0.15 * fcen),
                         mode_solver_resolution=40, mode_index=1, center=[-0.8, 0, 0], size=[0, 1.4, 1.2],
                         geometry=geo)]
    freqs = [0.95 * fcen, fcen, 1.05 * fcen]
    spec = kh.ModeSpec(mode_solver_resolution=40)
    mons = [kh.ModeMonitor(center=[0.7, 0, 0], size=[0, 1.4, 1.2], frequencies=freqs, mode_spec=spec),
            kh.ModeMonitor(center=[-1.2, 0, 0], size=[0, 1.4, 1.2], frequencies=freqs, mode_spec=spec),
            kh.FluxMonitor(center=[0.7, 0, 0], size=[0, 1.4, 1.2], frequencies=freqs)]
    return kh.Simulation(cell_size=[3.2, 1.8, 1.6], resolution=20, sources=src, geometry=geo, monitors=mons,
                         boundaries=[[0.4, 0.4]] * 3, engine=engine, precision=precision, mode=mode)


def host_amplitudes(sim, m):
    md = m.monitor_data
    ms = md.tangential_E_monitors + md.tangential_H_monitors
    return mode_amplitudes_from_dft(md, *[sim.read_dft(x.monitor_data) for x in ms],
                                    [x.monitor_data.gv for x in ms])


def host_flux(sim, m):
    md = m.monitor_data
    ms = md.tangential_E_monitors + md.tangential_H_monitors
    return flux_from_dft(md, *[sim.read_dft(x.monitor_data) for x in ms], [x.monitor_data.gv for x in ms])


def check_mode_physics(ap, am, bp, bm, flux):
    k = 1   # centre frequency
    assert abs(bm[k]) ** 2 < 0.03 * abs(ap[k]) ** 2          # one-way launch
    assert np.all(np.abs(np.abs(ap) ** 2 / flux - 1) < 0.1)  # forward power is in the mode
    assert abs(am[k]) ** 2 < 0.05 * abs(ap[k]) ** 2


def test_mode_source_and_monitor_on_oracle():
    s = _waveguide()
    s.step(int(160 / s.dt_n))
    ap, am = host_amplitudes(s, s.monitors[0])
    bp, bm = host_amplitudes(s, s.monitors[1])
    check_mode_physics(ap, am, bp, bm, host_flux(s, s.monitors[2]))
