"""Near-to-far projection: the oracle pinned, and the host functions.

The oracle's green3d (oracle/khronos_oracle.c, src/Monitors/Near2Far.jl:40-94)
is pinned by the textbook Hertzian-dipole fields and by Maxwell's curl
equations; its surface sum (:103-246) by the surface equivalence theorem on a
closed box (the field outside, zero inside).  Planted mistakes (a wrong term2
coefficient, flipped magnetic currents) fail those pins.  The reference's own
tests (test/test_near2far.jl) are restated on the oracle and on the host
functions compute_far_field_power / compute_LEE.
"""
import math

import numpy as np
import pytest

import n2f_oracle as N
from khronos_hip import near2far as NF


# ------------------------------------------------------------ green3d pins
def test_green_nonzero():
    # test/test_near2far.jl:10-21
    eh = N.green3d([1.0, 0.0, 0.0], [0.0, 0.0, 0.0], 1, 1.0)
    assert np.any(np.abs(eh) > 0)


def test_green_reciprocity():
    # test/test_near2far.jl:23-41
    x1, x0 = [2.0, 1.0, 0.5], [0.0, 0.0, 0.0]
    fwd = N.green3d(x1, x0, 1, 1.0)
    bwd = N.green3d(x0, x1, 1, 1.0)
    assert fwd[0] == pytest.approx(bwd[0], rel=1e-10)


def test_green_far_field_decay():
    # test/test_near2far.jl:43-65
    d = np.ones(3) / math.sqrt(3.0)
    e1 = N.green3d(1e4 * d, np.zeros(3), 3, 1.0)[:3]
    e2 = N.green3d(2e4 * d, np.zeros(3), 3, 1.0)[:3]
    assert np.linalg.norm(e2) / np.linalg.norm(e1) == pytest.approx(0.5, rel=1e-3)


@pytest.mark.parametrize("r", [0.05, 0.3, 2.0, 50.0])
def test_green_hertzian_dipole_closed_form(r):
    """A unit z current element in a medium (eps, mu), time dependence
    exp(-i w t): with g = exp(ikr) / (4 pi r), on the equator (r_hat = x)
    E_z = i w mu g (1 + i/(kr) - 1/(kr)^2), H_y = -i k g (1 + i/(kr)); on the
    axis E_z = eta exp(ikr) / (2 pi r^2) (1 + i/(kr)).  By duality a z
    magnetic current gives H_z = i w eps g (1 + i/(kr) - 1/(kr)^2) and
    E_y = +i k g (1 + i/(kr)) on the equator."""
    f, eps, mu = 0.7, 2.0, 1.5
    w = 2 * math.pi * f
    k = w * math.sqrt(eps * mu)
    eta = math.sqrt(mu / eps)
    g = np.exp(1j * k * r) / (4 * math.pi * r)
    kr = k * r
    eq = N.green3d([r, 0.0, 0.0], [0.0, 0.0, 0.0], 3, 1.0, f, eps, mu)
    np.testing.assert_allclose(eq[2], 1j * w * mu * g * (1 + 1j / kr - 1 / kr ** 2), rtol=1e-12)
    np.testing.assert_allclose(eq[4], -1j * k * g * (1 + 1j / kr), rtol=1e-12)
    assert np.all(np.abs(eq[[0, 1, 3, 5]]) < 1e-14 * np.abs(eq[2]))
    ax = N.green3d([0.0, 0.0, r], [0.0, 0.0, 0.0], 3, 1.0, f, eps, mu)
    np.testing.assert_allclose(ax[2], eta * np.exp(1j * kr) / (2 * math.pi * r ** 2) * (1 + 1j / kr), rtol=1e-12)
    assert np.all(np.abs(ax[3:]) < 1e-14 * np.abs(ax[2]))
    m = N.green3d([r, 0.0, 0.0], [0.0, 0.0, 0.0], 6, 1.0, f, eps, mu)
    np.testing.assert_allclose(m[5], 1j * w * eps * g * (1 + 1j / kr - 1 / kr ** 2), rtol=1e-12)
    np.testing.assert_allclose(m[1], 1j * k * g * (1 + 1j / kr), rtol=1e-12)


def _curl_residuals(L, c0, x, eps=2.0, mu=1.5, f=1.0, h=1e-4):
    w = 2 * math.pi * f
    x0 = np.zeros(3)
    F = lambda p: N.green3d(p, x0, c0, 0.6 - 0.3j, f, eps, mu, L=L)
    d = np.zeros((3, 6), complex)
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        d[a] = (F(x + e) - F(x - e)) / (2 * h)
    v = F(x)
    cE = np.array([d[1][2] - d[2][1], d[2][0] - d[0][2], d[0][1] - d[1][0]])
    cH = np.array([d[1][5] - d[2][4], d[2][3] - d[0][5], d[0][4] - d[1][3]])
    rE = np.abs(cE - 1j * w * mu * v[3:]).max() / np.abs(cE).max()
    rH = np.abs(cH + 1j * w * eps * v[:3]).max() / np.abs(cH).max()
    return rE, rH


@pytest.mark.parametrize("c0", [1, 2, 3, 4, 5, 6])
def test_green_satisfies_maxwell(c0):
    """Away from the source: curl E = i w mu H and curl H = -i w eps E
    (central differences, step 1e-4: truncation ~1e-7)."""
    rE, rH = _curl_residuals(N.lib(), c0, np.array([0.7, -0.4, 0.9]))
    assert rE < 1e-6 and rH < 1e-6, (rE, rH)


def test_green_maxwell_pin_catches_planted_term():
    L = N.lib("liboracle_plant_green_f64.so")
    for c0 in (1, 4):
        rE, rH = _curl_residuals(L, c0, np.array([0.7, -0.4, 0.9]))
        assert max(rE, rH) > 1e-3


# ------------------------------------------------------ surface-sum pins
SOURCES = [((0.08, -0.05, 0.06), 3, 1.0 + 0.5j), ((-0.04, 0.07, -0.03), 4, 0.3 - 0.8j),
           ((0.02, 0.03, -0.08), 2, -0.4 + 0.2j)]


def _field(p, f, eps, mu):
    t = np.zeros(6, complex)
    for x0, c, a in SOURCES:
        t += N.green3d(p, x0, c, a, f, eps, mu)
    return t


def _box_projection(half, h, obs, f, eps, mu, L=None):
    """Sum of the six faces of the box [-half, half] (outward normals), each
    sampled at cell midpoints with spacing h[a] along axis a."""
    tot = np.zeros((len(obs), 6), complex)
    for a in range(3):
        t0, t1 = [b for b in range(3) if b != a]
        n1, n2 = int(round(2 * half[t0] / h[t0])), int(round(2 * half[t1] / h[t1]))
        for side in (-1.0, 1.0):
            dims = [1, 1, 1]
            dims[t0], dims[t1] = n1, n2
            planes = [np.zeros(dims + [1], complex) for _ in range(4)]
            for i1 in range(n1):
                for i2 in range(n2):
                    p = np.zeros(3)
                    p[a] = side * half[a]
                    p[t0] = -half[t0] + (i1 + 0.5) * h[t0]
                    p[t1] = -half[t1] + (i2 + 0.5) * h[t1]
                    F = _field(p, f, eps, mu)
                    idx = [0, 0, 0]
                    idx[t0], idx[t1] = i1, i2
                    ix = tuple(idx) + (0,)
                    planes[0][ix], planes[1][ix] = F[t0], F[t1]
                    planes[2][ix], planes[3][ix] = F[3 + t0], F[3 + t1]
            base = np.zeros((4, 3))
            base[:, a] = side * half[a]
            base[:, t0] = -half[t0] + h[t0] / 2
            base[:, t1] = -half[t1] + h[t1] / 2
            tot += N.near2far(planes, a + 1, side, base, h[t0], h[t1], n1, n2, [f], obs, eps=eps, mu=mu,
                              L=L)[:, :, 0]
    return tot


OUT_PTS = np.array([[1.3, 0.2, -0.4], [0.1, -2.0, 0.7], [-0.9, 0.8, 1.5], [0.0, 0.0, 3.0]])
IN_PTS = np.array([[0.2, 0.1, -0.2], [-0.3, -0.25, 0.3]])


@pytest.mark.parametrize("eps,mu", [(1.0, 1.0), (2.25, 1.0), (1.7, 1.3)])
def test_surface_equivalence_closed_box(eps, mu):
    """Surface equivalence on a closed rectangular box around J and M
    sources: outside, the projection reproduces the sources' field -- with
    the reference's overall sign, -1 (its currents are J = -n x H,
    M = n x E, src/Monitors/Near2Far.jl:147-150) -- and inside it vanishes
    (extinction).  Midpoint sampling: error O(h^2)."""
    f = 1.0
    half = np.array([0.5, 0.4, 0.45])
    h = np.array([0.05, 0.04, 0.045])
    P = _box_projection(half, h, np.vstack([OUT_PTS, IN_PTS]), f, eps, mu)
    D = np.array([_field(p, f, eps, mu) for p in OUT_PTS])
    Din = np.array([_field(p, f, eps, mu) for p in IN_PTS])
    ext = np.abs(P[: len(OUT_PTS)] + D).max() / np.abs(D).max()
    inside = np.abs(P[len(OUT_PTS):]).max() / np.abs(Din).max()
    assert ext < 5e-3, ext
    assert inside < 2e-3, inside


def test_surface_equivalence_converges_second_order():
    f, eps, mu = 1.0, 1.0, 1.0
    half = np.array([0.5, 0.5, 0.5])
    D = np.array([_field(p, f, eps, mu) for p in OUT_PTS])
    errs = []
    for h in (0.1, 0.05):
        P = _box_projection(half, np.array([h, h, h]), OUT_PTS, f, eps, mu)
        errs.append(np.abs(P + D).max() / np.abs(D).max())
    assert errs[0] / errs[1] == pytest.approx(4.0, rel=0.15), errs


def test_surface_equivalence_catches_planted_magnetic_sign():
    f = 1.0
    half = np.array([0.5, 0.4, 0.45])
    h = np.array([0.05, 0.04, 0.045])
    P = _box_projection(half, h, OUT_PTS, f, 1.0, 1.0, L=N.lib("liboracle_plant_n2f_f64.so"))
    D = np.array([_field(p, f, 1.0, 1.0) for p in OUT_PTS])
    assert np.abs(P + D).max() / np.abs(D).max() > 0.1


def test_normal_average_of_two_planes():
    """A DFT two planes deep across the normal is averaged
    (src/Monitors/Near2Far.jl:431-474): planes (v + d, v - d) project like v."""
    rng = np.random.default_rng(3)
    n1, n2, nf = 5, 4, 2
    for na in (1, 2, 3):
        ax = na - 1
        t0, t1 = [b for b in range(3) if b != ax]
        dims1 = [1, 1, 1]
        dims1[t0], dims1[t1] = n1, n2
        dims2 = list(dims1)
        dims2[ax] = 2
        one, two = [], []
        for q in range(4):
            v = rng.normal(size=dims1 + [nf]) + 1j * rng.normal(size=dims1 + [nf])
            d = 10 * (rng.normal(size=dims1 + [nf]) + 1j * rng.normal(size=dims1 + [nf]))
            one.append(v)
            two.append(np.concatenate([v + d, v - d], axis=ax))
        base = rng.normal(size=(4, 3)) * 0.1
        obs = rng.normal(size=(7, 3)) * 3
        a = N.near2far(one, na, -1.0, base, 0.1, 0.12, n1, n2, [0.8, 1.1], obs)
        b = N.near2far(two, na, -1.0, base, 0.1, 0.12, n1, n2, [0.8, 1.1], obs)
        np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_sample_positions_per_component():
    """One nonzero sample (i1, i2) = (2, 1) in one plane projects as a point
    current at that plane's base + 2 d1 (lower tangential axis) + d2."""
    table = {3: [(3, 1, 1.0), (2, 2, -1.0), (1, 4, -1.0), (0, 5, 1.0)],
             1: [(3, 2, 1.0), (2, 3, -1.0), (1, 5, -1.0), (0, 6, 1.0)],
             2: [(2, 3, 1.0), (3, 1, -1.0), (0, 6, -1.0), (1, 4, 1.0)]}
    rng = np.random.default_rng(5)
    d1, d2, ns, dA = 0.07, 0.05, -1.0, 0.07 * 0.05
    obs = np.array([[0.9, -1.2, 2.0]])
    for na, rows in table.items():
        ax = na - 1
        t0, t1 = [b for b in range(3) if b != ax]
        base = rng.normal(size=(4, 3))
        for q, c0, sg in rows:
            dims = [1, 1, 1]
            dims[t0], dims[t1] = 4, 3
            planes = [np.zeros(dims + [1], complex) for _ in range(4)]
            idx = [0, 0, 0]
            idx[t0], idx[t1] = 2, 1
            planes[q][tuple(idx) + (0,)] = 0.3 - 0.7j
            got = N.near2far(planes, na, ns, base, d1, d2, 4, 3, [0.9], obs)[0, :, 0]
            x0 = base[q].copy()
            x0[t0] += 2 * d1
            x0[t1] += 1 * d2
            want = N.green3d(obs[0], x0, c0, sg * ns * (0.3 - 0.7j) * dA, 0.9)
            np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-300)


def test_dipole_plane_roundtrip_sin2():
    """test/test_near2far.jl:67-163: the near field of a z dipole on a 20
    lambda plane 2 lambda above it, projected to r = 1e6, follows sin^2 theta
    for theta <= 45 deg (rms < 0.15); our power and the oracle's sum."""
    n, Lx, z0 = 81, 20.0, 2.0
    xs = np.linspace(-Lx / 2, Lx / 2, n)
    dx = xs[1] - xs[0]
    planes = [np.zeros((n, n, 1, 1), complex) for _ in range(4)]
    for i2 in range(n):
        for i1 in range(n):
            F = N.green3d([xs[i1], xs[i2], z0], [0.0, 0.0, 0.0], 3, 1.0)
            planes[0][i1, i2, 0, 0], planes[1][i1, i2, 0, 0] = F[0], F[1]
            planes[2][i1, i2, 0, 0], planes[3][i1, i2, 0, 0] = F[3], F[4]
    assert np.abs(planes[0]).max() > 0 and np.abs(planes[3]).max() > 0
    theta = np.radians(np.linspace(5.0, 45.0, 9))
    phi = np.linspace(0.0, 2 * np.pi - 2 * np.pi / 12, 12)
    md = NF.Near2FarMonitorData(3, 1.0, [], [], np.array([1.0]), theta=theta, phi=phi, r=1e6)
    obs = NF.observation_points(md)
    base = np.tile([xs[0], xs[0], z0], (4, 1))
    EH = N.near2far(planes, 3, 1.0, base, dx, dx, n, n, [1.0], obs)
    power = NF.compute_far_field_power(EH, theta, phi)
    pa = power.mean(axis=1)
    s2 = np.sin(theta) ** 2
    rms = math.sqrt(np.mean((pa / pa.max() - s2 / s2.max()) ** 2))
    assert rms < 0.15, rms


# ------------------------------------------------------------ host functions
def test_observation_points_order():
    """theta runs fastest, phi outer (src/Monitors/Near2Far.jl:1005-1016)."""
    md = NF.Near2FarMonitorData(3, 1.0, [], [], np.array([1.0]), theta=np.array([0.1, 0.2, 0.3]),
                                phi=np.array([0.0, 1.0]), r=5.0)
    o = NF.observation_points(md)
    assert o.shape == (6, 3)
    t, p = 0.2, 1.0
    np.testing.assert_allclose(o[4], [5 * math.sin(t) * math.cos(p), 5 * math.sin(t) * math.sin(p), 5 * math.cos(t)])
    md2 = NF.Near2FarMonitorData(3, 1.0, [], [], np.array([1.0]), observation_points=np.eye(3))
    np.testing.assert_array_equal(NF.observation_points(md2), np.eye(3))
    with pytest.raises(ValueError):
        NF.observation_points(NF.Near2FarMonitorData(3, 1.0, [], [], np.array([1.0])))


def test_far_field_power_closed_form():
    """0.5 (|E_theta|^2 + |E_phi|^2) / Z from Cartesian E; a radial part
    carries no power (src/Monitors/Near2Far.jl:1043-1072); only the first
    frequency is read."""
    rng = np.random.default_rng(0)
    theta = np.linspace(0.1, 1.4, 5)
    phi = np.linspace(0.0, 5.0, 4)
    eps, mu = 2.0, 1.2
    EH = np.zeros((20, 6, 2), complex)
    want = np.zeros((5, 4))
    idx = 0
    for ip, p in enumerate(phi):
        for it, t in enumerate(theta):
            et, ep, er = rng.normal(size=3) + 1j * rng.normal(size=3)
            th = np.array([math.cos(t) * math.cos(p), math.cos(t) * math.sin(p), -math.sin(t)])
            ph = np.array([-math.sin(p), math.cos(p), 0.0])
            rh = np.array([math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)])
            EH[idx, :3, 0] = et * th + ep * ph + er * rh
            EH[idx, :3, 1] = 100.0
            want[it, ip] = 0.5 * (abs(et) ** 2 + abs(ep) ** 2) / math.sqrt(mu / eps)
            idx += 1
    np.testing.assert_allclose(NF.compute_far_field_power(EH, theta, phi, eps=eps, mu=mu), want, rtol=1e-12)


def test_far_field_power_uniform_nonnegative():
    # test/test_near2far.jl:165-184
    theta = np.linspace(0, np.pi / 2, 10)
    phi = np.linspace(0, 2 * np.pi, 21)[:-1]
    EH = np.zeros((200, 6, 1), complex)
    EH[:, 0, 0] = 1.0
    p = NF.compute_far_field_power(EH, theta, phi)
    assert p.shape == (10, 20) and np.all(p >= 0)


def test_lee_closed_forms():
    # test/test_near2far.jl:186-207
    theta = np.linspace(0, np.pi / 2, 100)
    phi = np.linspace(0, 2 * np.pi, 101)[:-1]
    power = np.ones((100, 100))
    a = math.radians(30.0)
    assert NF.compute_LEE(power, theta, phi, cone_half_angle=a) == pytest.approx(1 - math.cos(a), rel=0.05)
    assert NF.compute_LEE(power, theta, phi, cone_half_angle=np.pi / 2) == pytest.approx(1.0, rel=0.01)


def test_lee_dipole_pattern():
    """P = sin^2 theta: LEE(a) = (2/3 - cos a + cos^3 a / 3) / (2/3) (the
    cone edge carries a full trapezoid weight: O(dtheta) error)."""
    theta = np.linspace(0, np.pi / 2, 1601)
    phi = np.linspace(0, 2 * np.pi, 33)
    power = np.tile((np.sin(theta) ** 2)[:, None], (1, phi.size))
    for a in (0.3, 0.8, 1.2):
        want = (2 / 3 - math.cos(a) + math.cos(a) ** 3 / 3) / (2 / 3)
        assert NF.compute_LEE(power, theta, phi, cone_half_angle=a) == pytest.approx(want, rel=3e-3)


def test_lee_single_angle_gives_zero():
    """A single phi (or theta) gets trapezoid weight 0 in the reference
    (its 2 pi / pi/2 fallbacks are unreachable), so LEE = 0."""
    theta = np.linspace(0, np.pi / 2, 10)
    assert NF.compute_LEE(np.ones((10, 1)), theta, np.array([0.3])) == 0.0
    assert NF.compute_LEE(np.ones((1, 8)), np.array([0.2]), np.linspace(0, 6, 8)) == 0.0


def test_monitor_construction_and_init():
    # test/test_near2far.jl:209-226, and the four tangential DFTs per axis
    import khronos_hip as kh
    n2f = kh.Near2FarMonitor(center=[0.0, 0.0, 1.0], size=[2.0, 2.0, 0.0], frequencies=[1.0],
                             theta=list(np.linspace(0, np.pi / 2, 10)), phi=list(np.linspace(0, 2 * np.pi, 20)),
                             normal_dir="+")
    assert list(n2f.center) == [0.0, 0.0, 1.0] and list(n2f.size) == [2.0, 2.0, 0.0]
    assert len(n2f.frequencies) == 1 and len(n2f.theta) == 10 and len(n2f.phi) == 20
    mons = []
    for size, comps in (([0.0, 0.6, 0.5], ("Ey", "Ez", "Hy", "Hz")), ([0.6, 0.0, 0.5], ("Ex", "Ez", "Hx", "Hz")),
                        ([0.6, 0.5, 0.0], ("Ex", "Ey", "Hx", "Hy"))):
        mons.append((kh.Near2FarMonitor(center=[0.1, -0.1, 0.05], size=size, frequencies=[1.0, 1.2],
                                        observation_points=np.ones((2, 3)), normal_dir="-"), comps))
    sim = kh.Simulation(cell_size=[2.0, 2.0, 2.0], resolution=10, boundaries=[[0.3, 0.3]] * 3,
                        sources=[kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.4),
                                                  component=kh.Ez(), center=[0, 0, 0], size=[0, 0, 0])],
                        monitors=[m for m, _ in mons])
    sim.build_spec()
    for m, comps in mons:
        md = m.monitor_data
        assert md.normal_sign == -1.0
        got = [type(x.component).__name__ for x in md.tangential_E_monitors + md.tangential_H_monitors]
        assert tuple(got) == comps
        base, n1, n2, d1, d2, dA = NF.surface_layout(md)
        assert dA == d1 * d2 and n1 >= 1 and n2 >= 1
        # tangential E and H sit half a cell apart along the normal axis
        ax = md.normal_axis - 1
        assert abs(abs(base[0, ax] - base[2, ax]) - 0.05) < 1e-12


def test_normal_dir_and_2d_rejected():
    import khronos_hip as kh
    with pytest.raises(ValueError):
        NF._normal_sign("up")
    assert NF._normal_sign(":-") == -1.0
    sim = kh.Simulation(cell_size=[2.0, 2.0, 0.0], resolution=10, boundaries=[[0.3, 0.3]] * 2,
                        sources=[kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.4),
                                                  component=kh.Ez(), center=[0, 0, 0], size=[0, 0, 0])],
                        monitors=[kh.Near2FarMonitor(center=[0, 0.5, 0], size=[1.0, 0.0, 0.0], frequencies=[1.0],
                                                     observation_points=np.ones((1, 3)))])
    with pytest.raises(ValueError):
        sim.build_spec()
