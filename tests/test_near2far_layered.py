"""Layered near-to-far projection: the numpy oracle (oracle/near2far_layered.py)
pinned, and the product's host assembly checked against it.

Pins: transfer matrices against the Fresnel coefficients of one interface
(TE, and tangential TM), a quarter-wave anti-reflection layer and a stack of
identical layers (pure propagation); the whole projection against the
stationary-phase far field of an exactly synthesised upward Gaussian beam
(E_far = -2 pi i k cos(theta) a(k_t) exp(ikr) / r), sampled on staggered Yee
positions with E and H half a cell apart -- the reference's phase is that of
an aperture whose first E sample sits at the origin, which the test undoes.
The direct Green's-function sum on the same beam gives minus that field.
"""
import cmath
import math
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import near2far_layered as LY  # noqa: E402  (the oracle)
import n2f_oracle as N  # noqa: E402
from khronos_hip import near2far as NF  # noqa: E402


# ------------------------------------------------------------ TMM pins
def _kz(n, k0, kt):
    return cmath.sqrt(complex((n * k0) ** 2 - kt ** 2))


@pytest.mark.parametrize("theta_deg", [0.0, 20.0, 35.0])
def test_tmm_single_interface_fresnel(theta_deg):
    """Monitor on the interface (no propagation): t_TE = 2 kz1 / (kz1 + kz2)
    and, for tangential E of TM, 2 eps1 kz2 / (eps1 kz2 + eps2 kz1)."""
    f = 1.0
    k0 = 2 * math.pi * f
    n1, n2 = 1.5, 1.0
    kt = n2 * k0 * math.sin(math.radians(theta_deg))
    layers = [(-1.0, 0.0, n1 ** 2, 1.0), (0.0, 5.0, n2 ** 2, 1.0)]
    t_te, t_tm = LY.layer_transmission(layers, 0.0, f, kt, 0.0)
    k1, k2 = _kz(n1, k0, kt), _kz(n2, k0, kt)
    assert t_te == pytest.approx(2 * k1 / (k1 + k2), rel=1e-13)
    e1, e2 = n1 ** 2, n2 ** 2
    assert t_tm == pytest.approx(2 * e1 * k2 / (e1 * k2 + e2 * k1), rel=1e-13)


def test_tmm_quarter_wave_antireflection():
    """n1 | n2 = sqrt(n1 n3), lambda / (4 n2) thick | n3 at normal
    incidence: no reflection, so |t|^2 n3 / n1 = 1."""
    f = 1.0
    n1, n3 = 2.4, 1.0
    n2 = math.sqrt(n1 * n3)
    d = 1.0 / (4 * n2)
    layers = [(-1.0, 0.0, n1 ** 2, 1.0), (0.0, d, n2 ** 2, 1.0), (d, 3.0, n3 ** 2, 1.0)]
    t_te, t_tm = LY.layer_transmission(layers, 0.0, f, 0.0, 0.0)
    assert abs(t_te) ** 2 * n3 / n1 == pytest.approx(1.0, rel=1e-12)
    assert abs(t_tm) ** 2 * n3 / n1 == pytest.approx(1.0, rel=1e-12)
    # a half-wave layer is absent: the bare n1 | n3 interface
    layers2 = [(-1.0, 0.0, n1 ** 2, 1.0), (0.0, 2 * d, n2 ** 2, 1.0), (2 * d, 3.0, n3 ** 2, 1.0)]
    bare = [(-1.0, 0.0, n1 ** 2, 1.0), (0.0, 3.0, n3 ** 2, 1.0)]
    a = LY.layer_transmission(layers2, 0.0, f, 0.0, 0.0)
    b = LY.layer_transmission(bare, 0.0, f, 0.0, 0.0)
    assert abs(a[0]) == pytest.approx(abs(b[0]), rel=1e-12)


def test_tmm_identical_layers_propagate():
    """Identical layers: t = exp(i kz (z_top_interface - z_monitor))."""
    f, n = 0.9, 1.7
    k0 = 2 * math.pi * f
    kt = 0.4 * k0
    layers = [(-1.0, 0.3, n * n, 1.0), (0.3, 0.8, n * n, 1.0), (0.8, 2.0, n * n, 1.0)]
    zm = -0.2
    t_te, t_tm = LY.layer_transmission(layers, zm, f, kt * 0.6, kt * 0.8)
    want = cmath.exp(1j * _kz(n, k0, kt) * (0.8 - zm))
    assert t_te == pytest.approx(want, rel=1e-12) and t_tm == pytest.approx(want, rel=1e-12)
    assert LY.layer_transmission(layers, 1.5, f, kt, 0.0) == (1.0, 1.0)


def test_tmm_vectorised_matches_oracle():
    """The product's vectorised transfer matrices = the oracle's, direction
    by direction, including grazing and evanescent directions."""
    rng = np.random.default_rng(2)
    f = 1.1
    k0 = 2 * math.pi * f
    layers = [(-2.0, -0.4, 6.0, 1.0), (-0.4, 0.25, 2.1, 1.2), (0.25, 0.5, 3.3, 1.0), (0.5, 4.0, 1.0, 1.0)]
    kx = rng.uniform(-1.0, 1.0, 200) * k0
    ky = rng.uniform(-1.0, 1.0, 200) * k0
    kx[0], ky[0] = k0, 0.0                     # grazing in the top layer
    te, tm = NF.layer_transmission(layers, -1.0, f, kx, ky)
    for i in range(kx.size):
        a, b = LY.layer_transmission(layers, -1.0, f, kx[i], ky[i])
        assert te[i] == pytest.approx(a, rel=1e-12, abs=1e-300)
        assert tm[i] == pytest.approx(b, rel=1e-12, abs=1e-300)


# ------------------------------------------------------- beam pin
def _beam(f, eps, mu, w0, tilt, pol, pts_e, pts_h, nk=31):
    """Upward Gaussian beam as a plane-wave sum with spectrum
    a(k_t) = exp(-|k_t - k_c|^2 w0^2 / 4) (e0 - (e0 . k_hat) k_hat) dk^2,
    e0 = (pol[0], pol[1], 0) (smooth through normal incidence); returns E at
    pts_e, H at pts_h (each (n, 3) -> (n, 3)) and a(.)."""
    k = 2 * math.pi * f * math.sqrt(eps * mu)
    Z = math.sqrt(mu / eps)
    kc = np.array([k * math.sin(tilt[0]) * math.cos(tilt[1]), k * math.sin(tilt[0]) * math.sin(tilt[1])])
    span = 6.0 / w0
    g = np.linspace(-span, span, nk)
    dk = g[1] - g[0]
    KX, KY = np.meshgrid(kc[0] + g, kc[1] + g, indexing="ij")
    KX, KY = KX.ravel(), KY.ravel()

    def vec(kx, ky):
        kt = np.hypot(kx, ky)
        kz = np.sqrt(k * k - kt ** 2 + 0j).real
        kv = np.stack([kx, ky, kz], -1) / k
        e0 = np.array([pol[0], pol[1], 0.0], dtype=complex)
        amp = np.exp(-((kx - kc[0]) ** 2 + (ky - kc[1]) ** 2) * w0 * w0 / 4)
        e = amp[..., None] * (e0 - (kv @ e0)[..., None] * kv)
        h = np.cross(kv, e) / Z
        return e, h, kz

    ea, ha, kz = vec(KX, KY)
    kvec = np.stack([KX, KY, kz], 1)

    def field(pts, amps):
        ph = np.exp(1j * (np.asarray(pts) @ kvec.T))      # (n, m)
        return ph @ amps * dk * dk

    E = field(pts_e, ea)
    H = field(pts_h, ha)

    def a_of(kx, ky):
        e, _, _ = vec(np.atleast_1d(kx), np.atleast_1d(ky))
        return e
    return E, H, a_of, k


def _beam_planes(f, eps, mu, dx, n, z_e, pol, tilt, w0=2.0):
    """Yee-staggered planes of a z-normal surface: Ex at (x+dx/2, y), Ey at
    (x, y+dx/2), both at z_e; Hx at (x, y+dx/2), Hy at (x+dx/2, y), at
    z_e - dx/2.  Sample (0, 0) sits at x = y = -n dx / 2."""
    x0 = -n * dx / 2
    idx = np.arange(n) * dx + x0
    I, J = np.meshgrid(idx, idx, indexing="ij")
    zh = z_e - dx / 2
    bases = np.array([[x0 + dx / 2, x0, z_e], [x0, x0 + dx / 2, z_e], [x0, x0 + dx / 2, zh], [x0 + dx / 2, x0, zh]])
    planes = []
    for q, (axis, elec) in enumerate(((0, True), (1, True), (0, False), (1, False))):
        b = bases[q]
        pts = np.stack([(I - x0 + b[0]).ravel(), (J - x0 + b[1]).ravel(), np.full(I.size, b[2])], 1)
        E, H, a_of, k = _beam(f, eps, mu, w0, tilt, pol, pts, pts)
        v = (E if elec else H)[:, axis]
        planes.append(v.reshape(n, n)[:, :, None, None])
    return planes, bases, a_of, k


@pytest.mark.parametrize("pol,tilt", [((1.0, 0.0), (0.12, 0.6)), ((0.3, 0.8 + 0.2j), (0.1, 2.2))])
def test_layered_beam_stationary_phase(pol, tilt):
    """Directions on the padded FFT grid (phi on the axes, k sin(theta) a
    multiple of the bin spacing), where the bilinear weights vanish: the
    reference interpolates a spectrum whose phase turns by pi/4 per bin for
    any field centred in the aperture (4x padding), which off the grid costs
    up to ~8% per axis and is its method, not checked here."""
    f, eps, mu = 1.0, 1.0, 1.0
    dx, n = 0.1, 128
    planes, bases, a_of, k = _beam_planes(f, eps, mu, dx, n, 0.3, pol, tilt)
    dk = 2 * math.pi / (4 * n * dx)
    theta = np.arcsin(np.arange(0, 13, 2) * dk / k)
    phi = np.array([0.0, np.pi / 2, np.pi, 3 * np.pi / 2])
    r = 1e6
    L = LY.far_field_layered(planes, bases, dx, dx, [f], theta, phi, r, 1.0, [(-5.0, 5.0, eps, mu)])[:, :, 0]
    P, T = np.meshgrid(phi, theta, indexing="ij")
    T, P = T.ravel(), P.ravel()
    kx, ky, kz = k * np.sin(T) * np.cos(P), k * np.sin(T) * np.sin(P), k * np.cos(T)
    a = a_of(kx, ky)
    true = (-2j * math.pi * k * np.cos(T))[:, None] * a * (np.exp(1j * k * r) / r)
    # undo the reference's phase convention: spectrum referenced to the
    # first E sample (Ey's x, Ex's y) and the H plane's height
    corr = np.exp(-1j * (kx * bases[1][0] + ky * bases[0][1] + kz * bases[2][2]))
    got = L[:, :3] * corr[:, None]
    scale = np.abs(true).max()
    assert np.abs(got - true).max() < 2e-4 * scale, np.abs(got - true).max() / scale
    # H = r x E / Z in the far field
    rh = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], 1)
    np.testing.assert_allclose(L[:, 3:], np.cross(rh, L[:, :3]), atol=1e-12 * scale)
    # the direct Green's-function sum of the same surface: minus the field,
    # with the surface's true position
    obs = r * rh
    D = N.near2far(planes, 3, 1.0, bases, dx, dx, n, n, [f], obs)[:, :3, 0]
    assert np.abs(D + true).max() < 2e-4 * scale, np.abs(D + true).max() / scale


def test_bilinear_is_exact_on_bilinear_data():
    """The oracle's interpolation reproduces a bilinear function of (kx, ky)
    exactly, and is zero outside the grid (src/Monitors/Near2Far.jl:677-703)."""
    kxg = np.arange(-8, 8) * 0.37
    kyg = np.arange(-6, 6) * 0.21
    KX, KY = np.meshgrid(kxg, kyg, indexing="ij")
    c = (0.3 - 1j, 1.7 + 0.2j, -0.4j, 0.9)
    data = c[0] + c[1] * KX + c[2] * KY + c[3] * KX * KY
    rng = np.random.default_rng(4)
    for _ in range(50):
        x = rng.uniform(kxg[0], kxg[-1] - 1e-9)
        y = rng.uniform(kyg[0], kyg[-1] - 1e-9)
        want = c[0] + c[1] * x + c[2] * y + c[3] * x * y
        assert LY.bilinear(data, kxg, kyg, x, y) == pytest.approx(want, rel=1e-12, abs=1e-12)
    assert LY.bilinear(data, kxg, kyg, kxg[-1] + 0.01, 0.0) == 0
    assert LY.bilinear(data, kxg, kyg, 0.0, kyg[0] - 0.01) == 0


# ----------------------------------------------- product host vs oracle
def _layered_sim(stack_z):
    import khronos_hip as kh
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.4), component=kh.Ez(),
                            center=[0, 0, 0], size=[0, 0, 0])]
    mon = kh.Near2FarMonitor(center=[0.05, -0.03, stack_z], size=[1.2, 1.0, 0.0], frequencies=[0.9, 1.0, 1.15],
                             theta=np.radians(np.linspace(0.0, 80.0, 17)), phi=np.linspace(0, 2 * np.pi, 9)[:-1],
                             r=1e5, normal_dir="-")
    sim = kh.Simulation(cell_size=[2.0, 2.0, 1.6], resolution=10, boundaries=[[0.3, 0.3]] * 3, sources=src,
                        monitors=[mon])
    sim.build_spec()
    return sim, mon


@pytest.mark.parametrize("stack", [
    [(-2.0, 2.0, 1.0, 1.0)],
    [(-2.0, 0.1, 4.0, 1.0), (0.1, 0.35, 2.25, 1.1), (0.35, 3.0, 1.0, 1.0)],
])
def test_layered_host_matches_oracle(stack):
    """compute_far_field_layered (vectorised host assembly) on random DFT
    planes, with the spectral bins from numpy's FFT of the padded planes,
    equals the oracle within 1e-11 of the largest component."""
    sim, mon = _layered_sim(-0.0125)
    md = mon.monitor_data
    md.layer_stack = NF.LayerStack([NF.LayerSpec(*L) for L in stack])
    rng = np.random.default_rng(7)
    mons = md.tangential_E_monitors + md.tangential_H_monitors
    planes = []
    for m in mons:
        d = tuple(m.monitor_data.gv.dims) + (3,)
        planes.append(rng.normal(size=d) + 1j * rng.normal(size=d))
    base, n1, n2, d1, d2, dA = NF.surface_layout(md)

    def avg(p):
        return (p[:, :, 0] + p[:, :, 1]) / 2 if p.shape[2] >= 2 else p[:, :, 0]

    def spectrum(k, p1, p2, bins):
        out = np.zeros((len(bins), 4), dtype=np.complex128)
        for q, p in enumerate(planes):
            pad = np.zeros((p1, p2), dtype=np.complex128)
            pad[:n1, :n2] = avg(p)[:n1, :n2, k]
            F = np.fft.fft2(pad)
            out[:, q] = F[np.mod(bins[:, 0], p1), np.mod(bins[:, 1], p2)]
        return out

    got = NF.compute_far_field_layered(md, md.layer_stack, spectrum=spectrum)
    want = LY.far_field_layered(planes, base, md.dx, md.dy, md.frequencies, md.theta, md.phi, md.r,
                                md.normal_sign, [(L.z_min, L.z_max, L.eps, L.mu) for L in md.layer_stack.layers])
    for k in range(3):
        scale = np.abs(want[:, :, k]).max()
        assert scale > 0
        assert np.abs(got[:, :, k] - want[:, :, k]).max() <= 1e-11 * scale


def test_layered_requires_z_normal_and_angles():
    sim, mon = _layered_sim(0.0)
    md = mon.monitor_data
    st = NF.LayerStack([NF.LayerSpec(-2.0, 2.0, 1.0)])
    md.normal_axis = 1
    with pytest.raises(ValueError):
        NF.compute_far_field_layered(md, st, spectrum=lambda *a: None)
    md.normal_axis = 3
    md.theta = None
    with pytest.raises(ValueError):
        NF.compute_far_field_layered(md, st, spectrum=lambda *a: None)
