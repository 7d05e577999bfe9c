"""PlaneWaveSource (src/Sources/SpatialSources.jl:412-586, assembly
src/Sources/Sources.jl:142-233).

Pins: the polarization is normal to k and, rotated by the polarization angle,
stays in the plane normal to k; the E and H current sheets (n x (k x E) / Z0
and -n x E) make a one-way (Huygens) source, so in a periodic cell with the
matching Bloch phase the flux behind the sheet is below 1% of the flux in
front at the centre frequency, at normal and oblique incidence (a wrong sign
on either sheet sends half the power backwards); the transmitted power at
oblique incidence scales with |p|^2 cos(theta) for the reference's
unnormalized polarization |p| = sin(angle(n, k)).
"""
import math

import numpy as np

import khronos_hip as kh
from khronos_hip.monitors import flux_from_dft
from oracle_engine import OracleEngine


def test_polarization_geometry():
    for kv, ang in (([0.0, 0.0, 1.0], 0.0), ([0.3, 0.1, 1.0], 0.4), ([0.0, 0.5, 1.0], 1.2)):
        s = kh.PlaneWaveSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.2), center=[0, 0, 0],
                               size=[math.inf, math.inf, 0.0], k_vector=kv, polarization_angle=ang)
        assert abs(np.linalg.norm(s.k_vector) - 1) < 1e-15
        assert abs(float(np.dot(s.polarization, s.k_vector))) < 1e-15
        assert s.k == 2 * math.pi
        comps = [type(c).__name__ for c in s.components()]
        assert all(c in ("Ex", "Ey", "Hx", "Hy") for c in comps)
    s = kh.PlaneWaveSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.2), center=[0, 0, 0],
                           size=[math.inf, math.inf, 0.0], k_vector=[0.0, 0.0, 1.0])
    assert np.allclose(s.polarization, [1, 0, 0]) and [type(c).__name__ for c in s.components()] == ["Ex", "Hy"]


def _fluxes(kvec):
    k = 2 * math.pi
    kt = [k * kvec[0] / np.linalg.norm(kvec), k * kvec[1] / np.linalg.norm(kvec)]
    src = [kh.PlaneWaveSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.3), center=[0, 0, -1.0],
                              size=[math.inf, math.inf, 0.0], k_vector=kvec)]
    freqs = [0.9, 1.0, 1.1]
    mons = [kh.FluxMonitor(center=[0, 0, 1.0], size=[1.0, 1.0, 0.0], frequencies=freqs),
            kh.FluxMonitor(center=[0, 0, -1.6], size=[1.0, 1.0, 0.0], frequencies=freqs)]
    # the sheet's phase exp(-i k.r) repeats with exp(-i k_t L): Bloch k = -k_t
    bcs = [[kh.Bloch(k=-kt[0])] * 2, [kh.Bloch(k=-kt[1])] * 2, [kh.PML(), kh.PML()]]
    s = kh.Simulation(cell_size=[1.0, 1.0, 4.0], resolution=20, sources=src, monitors=mons,
                      boundaries=[[0, 0], [0, 0], [0.5, 0.5]], boundary_conditions=bcs, engine=OracleEngine)
    s.step(int(60 / s.dt_n))
    out = []
    for m in s.monitors:
        md = m.monitor_data
        ms = md.tangential_E_monitors + md.tangential_H_monitors
        out.append(flux_from_dft(md, *[s.read_dft(x.monitor_data) for x in ms], [x.monitor_data.gv for x in ms]))
    return out


def test_one_way_injection_normal_and_oblique():
    fn, bn = _fluxes([0.0, 0.0, 1.0])
    kz = math.sqrt(1 - 0.25 ** 2)
    fo, bo = _fluxes([0.25, 0.0, kz])
    assert fn[1] > 0 and fo[1] > 0
    assert abs(bn[1]) < 1e-2 * fn[1] and abs(bo[1]) < 1e-2 * fo[1], (bn / fn, bo / fo)
    # |p| = sin(angle(n, k)) = 0.25 for the oblique sheet: power ~ |p|^2 cos(theta)
    ratio = fo[1] / fn[1]
    assert abs(ratio / (0.25 ** 2 * kz) - 1) < 0.1, ratio
