"""Mode monitors on the device: the overlap reduction (kh_mode_overlap)
equals the host restatement on the same DFTs within 1e-12; a FAST float64
run's mode amplitudes equal the oracle's within 1e-9; the physics bars of
tests/test_mode.py hold for the device run (one-way launch, forward power in
the mode)."""
import sys

import numpy as np
import pytest

import khronos_hip as kh
from oracle_engine import OracleEngine
from test_mode import _waveguide, check_mode_physics, host_amplitudes, host_flux

pytestmark = pytest.mark.gpu


def test_device_mode_amplitudes():
    g = _waveguide(engine=None, mode="fast")
    o = _waveguide(engine=OracleEngine, mode="exact")
    n = int(160 / g.dt_n)
    print(f"[mode] strip waveguide {g.Nx}x{g.Ny}x{g.Nz}, {n} steps", flush=True)
    g.step(n)
    o.step(n)
    for m in g.monitors[:2]:
        dp, dm = kh.compute_mode_amplitudes(m)
        hp, hm = host_amplitudes(g, m)
        assert np.allclose(dp, hp, rtol=1e-12, atol=1e-15 * np.max(np.abs(hp)))
        assert np.allclose(dm, hm, rtol=1e-12, atol=1e-15 * np.max(np.abs(hp)))
    for mg, mo in zip(g.monitors[:2], o.monitors[:2]):
        dp, dm = kh.compute_mode_amplitudes(mg)
        op, om = host_amplitudes(o, mo)
        scale = np.max(np.abs(op))
        assert np.max(np.abs(dp - op)) <= 1e-9 * scale and np.max(np.abs(dm - om)) <= 1e-9 * scale
    ap, am = kh.compute_mode_amplitudes(g.monitors[0])
    bp, bm = kh.compute_mode_amplitudes(g.monitors[1])
    print(f"[mode] |a+|^2 {np.abs(ap) ** 2} flux {kh.get_flux(g.monitors[2])} back |a-|^2 {np.abs(bm) ** 2}",
          flush=True)
    check_mode_physics(ap, am, bp, bm, kh.get_flux(g.monitors[2]))
    assert np.allclose(kh.get_mode_transmission(g.monitors[0]), np.abs(ap) ** 2)
