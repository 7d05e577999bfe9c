"""Host setup pinned against the reference's own test fixtures and closed forms.

* GridVolume golden integers      -- test/test_grid_volume.jl:13-118
* interpolation-weight invariants -- test/test_interpolation.jl:39-94
* source footprint counts         -- test/test_sources.jl:40-83
* PML profile: integral closed form of the quadratic grading (R = 1e-15)
* Gaussian pulse: closed-form envelope, phase, cutoff, peak snapping
"""
import json
import math
import os

import numpy as np
import pytest

import khronos_hip as kh
from khronos_hip import grid as G
from khronos_hip.boundaries import compute_sigma
from khronos_hip.sources import assemble_source

HERE = os.path.dirname(os.path.abspath(__file__))


class _SimLike:
    def __init__(self, cell_size, cell_center, resolution, real=np.float64):
        self.cell_center = [float(v) for v in cell_center]
        self.cell_size_n = [float(real(v)) for v in cell_size]
        n = [int(math.floor(c * resolution)) for c in cell_size]
        self.delta_n = [float(real(cell_size[a] / n[a])) for a in range(3)]
        self.ndims = 3


def _inf(v):
    return float("inf") if v == "Inf" else float(v)


def test_grid_volume_golden():
    with open(os.path.join(HERE, "golden", "grid_volume_golden.json")) as f:
        gold = json.load(f)
    s = gold["sim"]
    sim = _SimLike(s["cell_size"], s["cell_center"], s["resolution"])
    comps = {"Ex": kh.Ex, "Ey": kh.Ey, "Ez": kh.Ez}
    for c in gold["cases"]:
        vol = G.Volume([_inf(v) for v in c["center"]], [_inf(v) for v in c["size"]])
        gv = G.grid_volume(sim, vol, comps[c["component"]]())
        assert [gv.Nx, gv.Ny, gv.Nz] == c["N"], c
        assert gv.start_idx == c["start"], c
        assert gv.end_idx == c["end"], c


def _restrict(volume, ndims, dx, dy, xs, ys):
    w = np.zeros((len(xs), len(ys)))
    for iy, y in enumerate(ys):
        for ix, x in enumerate(xs):
            w[ix, iy] = G.interpolation_weight([x, y, 0.0], volume, ndims, [dx, dy, 0.0])
    return w


def test_interpolation_weight_invariants():
    # test/test_interpolation.jl:14-94 (same grid, volumes and tolerances)
    dx = dy = 0.1
    xs = np.arange(-0.5, 0.5 + 1e-12, dx)
    ys = np.arange(-0.7, 0.7 + 1e-12, dy)
    w = _restrict(G.Volume([0, 0, 0], [0, 0, 0]), 2, dx, dy, xs, ys)
    assert abs(w.sum() - 1.0) <= 1e-15
    w = _restrict(G.Volume([-0.12, -0.26, 0], [0, 0, 0]), 2, dx, dy, xs, ys)
    assert abs(w.sum() - 1.0) <= 1e-15
    w = _restrict(G.Volume([0.14, -0.21, 0], [5 * dx, 0, 0]), 2, dx, dy, xs, ys)
    assert abs(w.sum() * dx - 5 * dx) <= 1e-15
    w = _restrict(G.Volume([0.14, -0.21, 0], [0, 5 * dy, 0]), 2, dx, dy, xs, ys)
    assert abs(w.sum() * dy - 5 * dy) <= 1e-15
    for sx, sy in ((0.4, 0.4), (0.4, 0.5)):
        for ix in np.arange(-1.6, 1.6 + 1e-9, 0.4):
            w = _restrict(G.Volume([ix * dx, ix * dx, 0], [sx, sy, 0]), 2, dx, dy, xs, ys)
            assert abs(w.sum() * dx * dy - sx * sy) <= 1e-15
    for ix in np.arange(-1.6, 1.6 + 1e-9, 0.4):
        w = _restrict(G.Volume([0.14, ix * dy, 0], [5.0 * dx, 0.5 * dy, 0]), 2, dx, dy, xs, ys)
        assert abs(w.sum() - 5.0 * 0.5) <= 1e-14
    for ix in np.arange(-0.8, 0.8 + 1e-9, 0.4):
        w = _restrict(G.Volume([ix * dx, 0.14, 0], [0.4 * dx, 6.0 * dy, 0]), 2, dx, dy, xs, ys)
        assert abs(w.sum() - 0.4 * 6.0) <= 1e-14


def _src_sim(component, size):
    # test/test_sources.jl:13-21: 10 x 10 x 10 cell, resolution 10, PML 1.0
    src = kh.UniformSource(time_profile=kh.ContinuousWaveSource(fcen=1.0), component=component,
                           center=[0.023, 0.784, 0.631], size=size)
    return kh.Simulation(cell_size=[10.0, 10.0, 10.0], resolution=10, sources=[src],
                         boundaries=[[1.0, 1.0]] * 3), src


@pytest.mark.parametrize("comp", [kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz])
def test_point_source_touches_8_nodes(comp):
    sim, src = _src_sim(comp(), [0.0, 0.0, 0.0])
    (sd,) = assemble_source(sim, src)
    assert np.count_nonzero(sd.amplitude) == 8
    assert abs(sd.amplitude.sum() - 1.0) < 1e-12  # point weights sum to one


@pytest.mark.parametrize("direction", [0, 1, 2])
@pytest.mark.parametrize("comp", [kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz])
def test_line_source_touches_4n_nodes(comp, direction):
    size = [0.0, 0.0, 0.0]
    size[direction] = float("inf")
    sim, src = _src_sim(comp(), size)
    (sd,) = assemble_source(sim, src)
    n = kh.component_voxel_count(sim.Nx, sim.Ny, sim.Nz, comp())[direction]
    assert np.count_nonzero(sd.amplitude) == 4 * n


@pytest.mark.parametrize("real", [np.float32, np.float64])
def test_pml_profile_integral(real):
    """The graded profile integrates to -ln(1e-15)/4 * dt/2 per layer
    (quadratic grading: sigma_max * L / 3 with sigma_max = -ln R / (4L/3))."""
    n_cells, delta, dt, L = 80, 0.05, 0.025, 0.5
    s = compute_sigma(2 * n_cells + 1, real(delta), real(dt), real(L), real(L), real).astype(np.float64)
    xs = np.arange(1, 2 * n_cells + 2) * delta / 2  # sample positions idx * delta / 2
    target = -math.log(1e-15) / 4 * (0.5 * dt)
    nl = int(round(L / (delta / 2)))  # samples 1..nl cover (0, L]
    # the layer is one quadratic on [0, L]: extrapolate the x = 0 value exactly
    # from three samples, then Simpson's rule (exact for quadratics)
    vals = np.r_[3 * s[0] - 3 * s[1] + s[2], s[:nl]]
    h = delta / 2
    integ = h / 3 * (vals[0] + vals[-1] + 4 * vals[1:-1:2].sum() + 2 * vals[2:-1:2].sum())
    assert abs(integ - target) / target < (1e-5 if real is np.float32 else 1e-12)
    # interior exactly zero, profile symmetric and monotone into the layer
    mid = (xs > L + 1e-9) & (xs < n_cells * delta - L - 1e-9)
    assert np.all(s[mid] == 0)
    nl = int(round(L / (delta / 2)))
    assert np.all(np.diff(s[:nl]) < 0)
    np.testing.assert_allclose(s[:nl], s[::-1][1:nl + 1], rtol=1e-4 if real is np.float32 else 1e-12)  # float32 position arithmetic


def test_pml_sampled_value_closed_form():
    """sigma at half-cell sample idx for a left layer of thickness L:
    (-ln R / (4L/3)) * (dt/2) * ((L - idx*delta/2)/L)^2."""
    n, delta, dt, L = 41, 0.05, 0.025, 0.5
    s = compute_sigma(n, delta, dt, L, 0.0, np.float64)
    for idx in (1, 2, 7, 19):
        x = idx * delta / 2
        expect = (math.log(1e15) / (4 * L / 3)) * (dt / 2) * ((L - x) / L) ** 2
        assert abs(s[idx - 1] - expect) <= 1e-15 * max(1.0, expect)


@pytest.mark.parametrize("real", [np.float32, np.float64])
def test_gaussian_pulse_closed_form(real):
    g = kh.GaussianPulseSource(fcen=1.0, fwidth=0.2)
    assert g.width == 5.0 and g.cutoff_t == 25.0
    assert g.peak_time == 12.0  # round(12.5 / 1.0) with ties to even
    assert abs(g.fwidth - math.sqrt(-2 * math.log(1e-7)) / (5.0 * math.pi)) < 1e-15
    tol = 2e-6 if real is np.float32 else 1e-12
    for t in (0.0, 3.3, 11.9, 12.0, 12.7, 30.0):
        v = g.eval(t, real)
        tt = t - 12.0
        env = math.exp(-tt * tt / (2 * 25.0))
        assert abs(abs(v) - env) <= tol * max(env, 1e-30) + 1e-30
        if env > 1e-3:
            # phase of exp(-i 2 pi fcen tt)
            d = (np.angle(v) + 2 * math.pi * tt) % (2 * math.pi)
            d = min(d, 2 * math.pi - d)
            assert d < (1e-4 if real is np.float32 else 1e-9)
    assert g.eval(12.0 + 25.0 + 0.5, real) == 0  # past cutoff
    g2 = kh.GaussianPulseSource(fcen=1 / 1.55, fwidth=0.2 / 1.55)
    # peak snapped to an integer number of periods
    k = g2.peak_time * (1 / 1.55)
    assert abs(k - round(k)) < 1e-9


def test_cw_source_value():
    c = kh.ContinuousWaveSource(fcen=0.5)
    for t in (0.0, 0.25, 1.3):
        v = c.eval(t, np.float64)
        assert abs(v - complex(math.cos(math.pi * t), -math.sin(math.pi * t))) < 1e-15


def test_auto_decimation_dipole_80():
    """The 80^3 dipole: fcen 1, fwidth 0.2 at dt 0.025 gives D = 16."""
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.2), component=kh.Ez())]
    mon = [kh.DFTMonitor(component=kh.Ez(), center=[0, 0, 0], size=[4, 4, 0], frequencies=[1.0])]
    sim = kh.Simulation(cell_size=[4, 4, 4], resolution=20, sources=src, monitors=mon,
                        boundaries=[[0.5, 0.5]] * 3)
    spec = sim.build_spec()
    assert spec.dfts[0].decimation == 16
    assert spec.dfts[0].gv.start_idx == [1, 1, 40] and spec.dfts[0].gv.dims == (81, 81, 2)
    assert spec.sources[0].gv.start_idx == [41, 41, 40] and spec.sources[0].gv.dims == (1, 1, 2)


def test_flux_from_dft_closed_form():
    from khronos_hip.monitors import FluxMonitorData, flux_from_dft
    from khronos_hip.grid import GridVolume
    nf, n1, n2 = 3, 4, 5
    md = FluxMonitorData(3, [], [], np.arange(nf, dtype=float), 0.1, 0.2, 0.3)
    a = np.array([1 + 2j, 0.5 - 1j, 2.0])
    b = np.array([0.3 - 0.1j, 1j, -1.0])
    e1 = np.broadcast_to(a, (n1, n2, 2, nf)).copy()
    h2 = np.broadcast_to(b, (n1, n2, 1, nf)).copy()
    e2 = np.zeros((n1, n2, 1, nf), complex)
    h1 = np.zeros((n1, n2, 2, nf), complex)
    gvs = [GridVolume(None, [1, 1, 1], [1, 1, 1], n1, n2, 2)] * 2 + [GridVolume(None, [1, 1, 1], [1, 1, 1], n1, n2, 1)] * 2
    fl = flux_from_dft(md, e1, e2, h1, h2, gvs)
    expect = np.real(a * np.conj(b)) * n1 * n2 * 0.1 * 0.2
    np.testing.assert_allclose(fl, expect, rtol=1e-14)
