"""Near-to-far projection on the device (kh_near2far).

Parity: the far fields of four surfaces (all three normal axes, both normal
signs, one- and two-plane DFTs across the normal, a non-vacuum medium,
far (r = 1e6 and 2e9) and near observation points, split and unsplit
sample sums)
equal the oracle's surface sum (oracle/khronos_oracle.c orc_near2far, pinned
in tests/test_near2far.py) on the same DFT arrays within 1e-12 of the
largest field component of each observation set and frequency.
"""
import numpy as np
import pytest

import khronos_hip as kh
import n2f_oracle as N
from khronos_hip import near2far as NF
from khronos_hip._lib import KhronosError

pytestmark = pytest.mark.gpu

FREQS = [0.8, 1.0, 1.2]


def _scene(precision="f64", dft_accum="f64", mode="fast"):
    rng = np.random.default_rng(11)
    near = rng.normal(size=(40, 3)) * 1.5
    near[:, 2] += 2.5
    srcs = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.5), component=kh.Ez(),
                             center=[0.1, -0.05, 0.02], size=[0, 0, 0]),
            kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.5), component=kh.Hx(),
                             center=[-0.1, 0.1, -0.05], size=[0, 0, 0], amplitude=0.7)]
    mons = [kh.Near2FarMonitor(center=[0, 0, 0.4375], size=[1.0, 0.9, 0], frequencies=FREQS,
                               theta=np.linspace(0.1, 1.5, 7), phi=np.linspace(0, 6, 5)),
            kh.Near2FarMonitor(center=[-0.45, 0.05, 0.0], size=[0, 0.8, 0.7], frequencies=FREQS,
                               observation_points=near, normal_dir="-"),
            kh.Near2FarMonitor(center=[0.02, 0.375, -0.03], size=[0.7, 0, 0.6], frequencies=FREQS,
                               theta=np.linspace(0.2, 2.9, 300), phi=np.linspace(0.0, 6.0, 9), r=50.0,
                               medium_eps=2.25, medium_mu=1.1),
            kh.Near2FarMonitor(center=[0, 0, -0.375], size=[0.2, 0.15, 0], frequencies=FREQS,
                               observation_points=near[:5], normal_dir="-"),
            # k r ~ 1e10: the FAST sincos hands these arguments to the library
            kh.Near2FarMonitor(center=[0, 0, 0.4375], size=[1.0, 0.9, 0], frequencies=FREQS,
                               observation_points=2e9 * np.array([[0.0, 0.0, 1.0], [0.6, 0.0, 0.8],
                                                                  [0.0, -0.6, 0.8], [0.48, 0.36, 0.8]]))]
    return kh.Simulation(cell_size=[2.0, 1.75, 1.5], resolution=8, boundaries=[[0.25, 0.25]] * 3, sources=srcs,
                         monitors=mons, precision=precision, dft_accum=dft_accum, mode=mode)


def _oracle_far_field(sim, m, f32_average=False):
    md = m.monitor_data
    mons = md.tangential_E_monitors + md.tangential_H_monitors
    planes = [sim.read_dft(x.monitor_data) for x in mons]
    base, n1, n2, d1, d2, dA = NF.surface_layout(md)
    ax = md.normal_axis - 1
    if f32_average:
        # float accumulators are averaged across the normal in float
        # (src/Monitors/Near2Far.jl:431-474); hand the oracle that plane
        out = []
        for p in planes:
            if p.shape[ax] >= 2:
                a = np.take(p, [0], axis=ax).astype(np.complex64)
                b = np.take(p, [1], axis=ax).astype(np.complex64)
                p = ((a + b) / np.float32(2)).astype(np.complex128)
            out.append(p)
        planes = out
    return N.near2far(planes, md.normal_axis, md.normal_sign, base, d1, d2, n1, n2, md.frequencies,
                      NF.observation_points(md), eps=md.medium_eps, mu=md.medium_mu, dA=dA)


def _check(dev, ref, tol):
    for k in range(ref.shape[2]):
        scale = np.abs(ref[:, :, k]).max()
        err = np.abs(dev[:, :, k] - ref[:, :, k]).max()
        assert err <= tol * scale, (k, err / scale)


@pytest.mark.parametrize("precision,dft_accum,mode", [("f64", "f64", "fast"), ("f64", "f64", "exact"),
                                                      ("f32", "f64", "fast"), ("f32", "f32", "fast")])
def test_near2far_matches_oracle(precision, dft_accum, mode):
    """FAST contexts project with green3d_fast, EXACT ones with the
    reference's operation order; both within 1e-12 of the oracle."""
    sim = _scene(precision, dft_accum, mode)
    n = int(12.0 / sim.dt_n)
    print(f"[near2far] {precision}/{dft_accum}/{mode} {sim.Nx}x{sim.Ny}x{sim.Nz}, {n} steps", flush=True)
    sim.step(n)
    for m in sim.monitors:
        md = m.monitor_data
        base, n1, n2, _, _, _ = NF.surface_layout(md)
        dev = kh.compute_far_field(m)
        ref = _oracle_far_field(sim, m, f32_average=(dft_accum == "f32"))
        assert dev.shape == ref.shape == (NF.observation_points(md).shape[0], 6, len(FREQS))
        assert np.abs(ref).max() > 0
        print(f"[near2far] axis {md.normal_axis} {n1}x{n2} samples, {dev.shape[0]} points: "
              f"max rel err {np.abs(dev - ref).max() / np.abs(ref).max():.2e}", flush=True)
        _check(dev, ref, 1e-12)


def test_near2far_repeat_and_shapes():
    """Two calls in a row (buffers reused, then grown) give the same fields;
    mismatched arguments are rejected with an error, not a launch."""
    sim = _scene()
    sim.step(int(8.0 / sim.dt_n))
    a = kh.compute_far_field(sim.monitors[3])
    big = kh.compute_far_field(sim.monitors[2])
    b = kh.compute_far_field(sim.monitors[3])
    np.testing.assert_array_equal(a, b)
    assert big.shape[0] == 300 * 9
    md = sim.monitors[0].monitor_data
    owner, ids = NF._owner(md)
    base, n1, n2, d1, d2, dA = NF.surface_layout(md)
    obs = NF.observation_points(md)
    with pytest.raises(KhronosError):
        owner.engine.near2far(ids, 3, 1.0, base, d1, d2, dA, n1 + 50, n2, md.frequencies, 1.0, 1.0, obs)
    with pytest.raises(KhronosError):
        owner.engine.near2far(ids, 3, 1.0, base, d1, d2, dA, n1, n2, md.frequencies[:2], 1.0, 1.0, obs)
    with pytest.raises(KhronosError):
        owner.engine.near2far(ids, 3, 1.0, base, d1, d2, dA, n1, n2, md.frequencies, -1.0, 1.0, obs)


def test_layered_spectrum_and_far_field():
    """Layered projection: the device's spectral bins equal numpy's FFT of
    the zero-padded planes within 1e-12, and the whole far field the numpy
    oracle (oracle/near2far_layered.py) within 1e-10, per frequency."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    import near2far_layered as LY
    stack = kh.LayerStack([kh.LayerSpec(-5.0, 0.05, 2.25), kh.LayerSpec(0.05, 0.2, 1.7, 1.1),
                           kh.LayerSpec(0.2, 5.0, 1.0)])
    mon = kh.Near2FarMonitor(center=[0.0, 0.0, -0.1875], size=[1.3, 1.1, 0.0], frequencies=FREQS,
                             theta=np.radians(np.linspace(0.0, 85.0, 35)), phi=np.linspace(0, 2 * np.pi, 13)[:-1],
                             r=1e4, layer_stack=stack)
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.5), component=kh.Ex(),
                            center=[0.05, -0.02, -0.5], size=[0, 0, 0])]
    sim = kh.Simulation(cell_size=[2.0, 1.75, 1.5], resolution=8, boundaries=[[0.25, 0.25]] * 3, sources=src,
                        monitors=[mon], precision="f32")
    sim.step(int(10.0 / sim.dt_n))
    md = mon.monitor_data
    mons = md.tangential_E_monitors + md.tangential_H_monitors
    planes = [sim.read_dft(x.monitor_data) for x in mons]
    base, n1, n2, d1, d2, dA = NF.surface_layout(md)
    owner, ids = NF._owner(md)
    p1, p2 = 4 * n1, 4 * n2
    rng = np.random.default_rng(1)
    bins = np.stack([rng.integers(-p1 // 2, p1 // 2, 300), rng.integers(-p2 // 2, p2 // 2, 300)], 1)
    for k in range(len(FREQS)):
        dev = owner.engine.near2far_spectrum(ids, 3, n1, n2, k, p1, p2, bins)
        for q, p in enumerate(planes):
            pl = (p[:, :, 0, k] + p[:, :, 1, k]) / 2 if p.shape[2] >= 2 else p[:, :, 0, k]
            pad = np.zeros((p1, p2), dtype=np.complex128)
            pad[:n1, :n2] = pl[:n1, :n2]
            F = np.fft.fft2(pad)
            want = F[np.mod(bins[:, 0], p1), np.mod(bins[:, 1], p2)]
            assert np.abs(dev[:, q] - want).max() <= 1e-12 * np.abs(F).max()
    got = kh.compute_far_field(mon)
    want = LY.far_field_layered(planes, base, md.dx, md.dy, md.frequencies, md.theta, md.phi, md.r,
                                md.normal_sign, [(L.z_min, L.z_max, L.eps, L.mu) for L in stack.layers])
    for k in range(len(FREQS)):
        scale = np.abs(want[:, :, k]).max()
        assert scale > 0
        err = np.abs(got[:, :, k] - want[:, :, k]).max() / scale
        print(f"[near2far] layered f={FREQS[k]}: max rel err {err:.2e}", flush=True)
        assert err <= 1e-10
