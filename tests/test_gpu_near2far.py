"""Near-to-far projection on the device (kh_near2far).

Parity: the far fields of four surfaces (all three normal axes, both normal
signs, one- and two-plane DFTs across the normal, a non-vacuum medium,
far (r = 1e6) and near observation points, split and unsplit sample sums)
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
                               observation_points=near[:5], normal_dir="-")]
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


@pytest.mark.parametrize("precision,dft_accum", [("f64", "f64"), ("f32", "f64"), ("f32", "f32")])
def test_near2far_matches_oracle(precision, dft_accum):
    sim = _scene(precision, dft_accum)
    n = int(12.0 / sim.dt_n)
    print(f"[near2far] {precision}/{dft_accum} {sim.Nx}x{sim.Ny}x{sim.Nz}, {n} steps", flush=True)
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
