"""Near-to-far projection of planar DFT surfaces (host side).

* monitor structures             -- src/DataStructures.jl:359-456
* monitor = 4 tangential DFTs     -- src/Monitors/Near2Far.jl:1137-1200
* base positions of the samples   -- src/Monitors/Monitors.jl:398-421
* observation points (theta, phi) -- src/Monitors/Near2Far.jl:998-1033
* far-field power, LEE            -- src/Monitors/Near2Far.jl:1043-1129
* layered projection (TMM)        -- src/Monitors/Near2Far.jl:565-996

The surface sum itself (full dyadic Green's function of the equivalent
currents of every sample, at every observation point and frequency) runs in
the HIP library (kh_near2far).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from .grid import Ex, Ey, Ez, Hx, Hy, Hz, component_origin


@dataclass
class LayerSpec:
    """One dielectric layer of a planar stack, bottom to top (src/DataStructures.jl:364-381)."""

    z_min: float
    z_max: float
    eps: float
    mu: float = 1.0


@dataclass
class LayerStack:
    """Layers ordered bottom to top (src/DataStructures.jl:383-395)."""

    layers: List[LayerSpec]


@dataclass
class Near2FarMonitor:
    """Planar surface whose tangential E/H are projected to the far field
    (src/DataStructures.jl:436-456).  Give ``theta``/``phi`` (angles, far
    field at distance ``r``) or ``observation_points`` ((N, 3) positions).
    ``normal_dir`` is "+" or "-" (outward normal along the zero-size axis)."""

    center: Sequence[float]
    size: Sequence[float]
    frequencies: Sequence[float]
    observation_points: Optional[np.ndarray] = None
    theta: Optional[Sequence[float]] = None
    phi: Optional[Sequence[float]] = None
    r: float = 1e6
    normal_dir: str = "+"
    medium_eps: float = 1.0
    medium_mu: float = 1.0
    decimation: int = 1
    layer_stack: Optional[LayerStack] = None
    monitor_data: object = None


@dataclass(eq=False)
class Near2FarMonitorData:
    """src/DataStructures.jl:402-424."""

    normal_axis: int                 # 1-based like the reference
    normal_sign: float
    tangential_E_monitors: list
    tangential_H_monitors: list
    frequencies: np.ndarray
    medium_eps: float = 1.0
    medium_mu: float = 1.0
    observation_points: Optional[np.ndarray] = None
    theta: Optional[np.ndarray] = None
    phi: Optional[np.ndarray] = None
    r: float = 1e6
    dx: float = 0.0
    dy: float = 0.0
    dz: float = 0.0
    layer_stack: Optional[LayerStack] = None
    grid: tuple = None               # (cell_center, cell_size, delta) of the simulation


def _normal_sign(d) -> float:
    d = str(d).lstrip(":")
    if d not in ("+", "-"):
        raise ValueError(f"normal_dir must be '+' or '-', got {d!r}")
    return 1.0 if d == "+" else -1.0


def init_near2far_monitor(sim, m: Near2FarMonitor):
    """Build the monitor data and its four tangential DFT monitors
    (src/Monitors/Near2Far.jl:1137-1200); returns the DFT monitors (E1, E2,
    H1, H2) for the simulation to register."""
    from .monitors import DFTMonitor
    normal = [i for i, s in enumerate(m.size) if float(s) == 0.0]
    if not normal:
        raise ValueError("a near-to-far surface is a plane: exactly one size entry must be 0")
    if any(math.isinf(float(d)) for d in sim.delta_n):
        raise ValueError("near-to-far projection needs a 3D simulation")
    na = normal[0] + 1
    ecomps, hcomps = {1: ((Ey, Ez), (Hy, Hz)), 2: ((Ex, Ez), (Hx, Hz)), 3: ((Ex, Ey), (Hx, Hy))}[na]
    mk = lambda c: DFTMonitor(component=c(), center=list(m.center), size=list(m.size),
                              frequencies=list(m.frequencies), decimation=m.decimation)
    te = [mk(c) for c in ecomps]
    th = [mk(c) for c in hcomps]
    m.monitor_data = Near2FarMonitorData(
        normal_axis=na, normal_sign=_normal_sign(m.normal_dir),
        tangential_E_monitors=te, tangential_H_monitors=th,
        frequencies=np.asarray(m.frequencies, dtype=np.float64),
        medium_eps=float(m.medium_eps), medium_mu=float(m.medium_mu),
        observation_points=None if m.observation_points is None else np.asarray(m.observation_points, np.float64),
        theta=None if m.theta is None else np.asarray(m.theta, dtype=np.float64),
        phi=None if m.phi is None else np.asarray(m.phi, dtype=np.float64),
        r=float(m.r), dx=float(sim.delta_n[0]), dy=float(sim.delta_n[1]), dz=float(sim.delta_n[2]),
        layer_stack=m.layer_stack,
        grid=(list(sim.cell_center), list(sim.cell_size_n), list(sim.delta_n)))
    return te + th


def surface_layout(md: Near2FarMonitorData):
    """(base (4, 3), n1, n2, d1, d2, dA) of the surface: the position of
    DFT sample (1, 1, 1) of e1, e2, h1, h2 -- component origin + (start - 1)
    delta (src/Monitors/Monitors.jl:398-421) -- the common tangential sample
    counts and spacings (src/Monitors/Near2Far.jl:409-425, :477-482)."""
    cc, cs, dl = md.grid
    mons = md.tangential_E_monitors + md.tangential_H_monitors
    gvs = [m.monitor_data.gv for m in mons]
    base = np.zeros((4, 3))
    for q, m in enumerate(mons):
        org = component_origin(cc, cs, dl, m.component)
        st = m.monitor_data.gv.start_idx
        base[q] = [org[a] + (float(st[a]) - 1) * float(dl[a]) for a in range(3)]
    ax = md.normal_axis - 1
    t0, t1 = sorted(a for a in range(3) if a != ax)
    n1 = min(g.dims[t0] for g in gvs)
    n2 = min(g.dims[t1] for g in gvs)
    d = (md.dx, md.dy, md.dz)
    d1, d2 = d[t0], d[t1]
    return base, n1, n2, d1, d2, d1 * d2


def observation_points(md: Near2FarMonitorData) -> np.ndarray:
    """(n_obs, 3): r (sin t cos p, sin t sin p, cos t) with theta fastest, or
    the given points (src/Monitors/Near2Far.jl:1005-1024)."""
    if md.theta is not None and md.phi is not None:
        th = np.asarray(md.theta, dtype=np.float64)
        ph = np.asarray(md.phi, dtype=np.float64)
        P, T = np.meshgrid(ph, th, indexing="ij")      # (n_phi, n_theta): theta fastest
        T, P = T.ravel(), P.ravel()
        return np.stack([md.r * np.sin(T) * np.cos(P), md.r * np.sin(T) * np.sin(P), md.r * np.cos(T)], axis=1)
    if md.observation_points is not None:
        return np.asarray(md.observation_points, dtype=np.float64).reshape(-1, 3)
    raise ValueError("give the far-field directions (theta and phi) or the observation points")


def _owner(md: Near2FarMonitorData):
    mons = md.tangential_E_monitors + md.tangential_H_monitors
    owner = mons[0].monitor_data.owner
    ids = [owner.spec.dfts.index(m.monitor_data) for m in mons]
    return owner, ids


def compute_far_field(monitor) -> np.ndarray:
    """E and H (n_obs, 6, n_freq) complex (Ex, Ey, Ez, Hx, Hy, Hz) at the
    observation points (src/Monitors/Near2Far.jl:998-1033): the layered
    projection when a layer stack is set, otherwise the direct Green's
    function sum on the device."""
    md = monitor.monitor_data if isinstance(monitor, Near2FarMonitor) else monitor
    if md.layer_stack is not None:
        return compute_far_field_layered(md, md.layer_stack)
    obs = observation_points(md)
    owner, ids = _owner(md)
    base, n1, n2, d1, d2, dA = surface_layout(md)
    return owner.engine.near2far(ids, md.normal_axis, md.normal_sign, base, d1, d2, dA, n1, n2,
                                 md.frequencies, md.medium_eps, md.medium_mu, obs)


def compute_far_field_power(EH, theta, phi, eps: float = 1.0, mu: float = 1.0) -> np.ndarray:
    """Radiated power per solid angle 0.5 (|E_theta|^2 + |E_phi|^2) / Z,
    (n_theta, n_phi), from the first frequency of EH
    (src/Monitors/Near2Far.jl:1043-1072)."""
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    Z = math.sqrt(mu / eps)
    P, T = np.meshgrid(phi, theta, indexing="ij")       # observation order: theta fastest
    T, P = T.ravel(), P.ravel()
    n = T.size
    E = np.asarray(EH)[:n, 0:3, 0]
    th_hat = np.stack([np.cos(T) * np.cos(P), np.cos(T) * np.sin(P), -np.sin(T)], axis=1)
    ph_hat = np.stack([-np.sin(P), np.cos(P), np.zeros_like(P)], axis=1)
    e_t = E[:, 0] * th_hat[:, 0] + E[:, 1] * th_hat[:, 1] + E[:, 2] * th_hat[:, 2]
    e_p = E[:, 0] * ph_hat[:, 0] + E[:, 1] * ph_hat[:, 1] + E[:, 2] * ph_hat[:, 2]
    power = 0.5 * (np.abs(e_t) ** 2 + np.abs(e_p) ** 2) / Z
    return power.reshape(phi.size, theta.size).T.copy()


def compute_LEE(power, theta, phi, cone_half_angle: float = math.radians(90.0)) -> float:
    """Power inside a cone over power in the upper hemisphere, trapezoidal
    weights with the sin(theta) Jacobian (src/Monitors/Near2Far.jl:1082-1129)."""
    power = np.asarray(power, dtype=np.float64)
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)

    def weights(x):
        # a one-point axis gets weight diff = [0.0] / 2 = 0: the reference's
        # 2 pi and pi / 2 fallbacks sit behind `length(dphi) > 0`, which is
        # always true, so a single theta or phi gives LEE = 0
        n = x.size
        if n == 1:
            return np.zeros(1)
        d = np.diff(x)
        w = np.empty(n)
        w[0] = d[0] / 2
        w[-1] = d[-1] / 2
        w[1:-1] = (d[:-1] + d[1:]) / 2
        return w

    wt = weights(theta)
    wp = weights(phi)
    integrand = power * np.sin(theta)[:, None] * wt[:, None] * wp[None, :]
    total = float(np.sum(integrand[theta <= math.pi / 2, :]))
    cone = float(np.sum(integrand[theta <= cone_half_angle, :]))
    return cone / total if total > 0 else 0.0


def _kz(eps, mu, w, kt2):
    """kz with Im >= 0 (src/Monitors/Near2Far.jl:588-592), per direction."""
    kz = np.sqrt((eps * mu * w ** 2 - kt2).astype(np.complex128))
    return np.where(kz.imag < 0, -kz, kz)


def _admittance(pol, eps, mu, w, kt2):
    kz = _kz(eps, mu, w, kt2)
    if pol == 0:                                   # TE: kz / (w mu)
        return kz / (w * mu)
    with np.errstate(divide="ignore", invalid="ignore"):   # TM: w eps / kz, inf at grazing
        return np.where(np.abs(kz) < 1e-30, complex(math.inf, 0.0), (w * eps) / kz)


def layer_transmission(layers, z_monitor, freq, kx, ky):
    """Fresnel transmission (t_TE, t_TM) of tangential E from the monitor's
    layer to the top layer by 2x2 transfer matrices in the (A+, A-) basis,
    vectorised over directions (src/Monitors/Near2Far.jl:565-669).  layers:
    [(z_min, z_max, eps, mu)] bottom to top."""
    kx = np.atleast_1d(np.asarray(kx, dtype=np.float64))
    ky = np.atleast_1d(np.asarray(ky, dtype=np.float64))
    n = len(layers)
    w = 2 * math.pi * freq
    kt2 = kx ** 2 + ky ** 2
    mon = next((i for i, L in enumerate(layers) if L[0] <= z_monitor <= L[1]), None)
    if mon is None:
        raise ValueError(f"Monitor z={z_monitor} is not inside any layer in the stack")
    one = np.ones(kx.shape, dtype=np.complex128)
    if mon == n - 1:
        return one, one.copy()
    res = []
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        for pol in (0, 1):
            T = np.zeros(kx.shape + (2, 2), dtype=np.complex128)
            T[..., 0, 0] = 1.0
            T[..., 1, 1] = 1.0
            alive = np.ones(kx.shape, dtype=bool)
            for i in range(mon, n - 1):
                z0, z1, e, m = layers[i]
                kz_i = _kz(e, m, w, kt2)
                p_i = _admittance(pol, e, m, w, kt2)
                d = (z1 - z_monitor) if i == mon else (z1 - z0)
                if d > 0:
                    P = np.zeros_like(T)
                    P[..., 0, 0] = np.exp(1j * kz_i * d)
                    P[..., 1, 1] = np.exp(-1j * kz_i * d)
                    T = T @ P
                L1 = layers[i + 1]
                p_n = _admittance(pol, L1[2], L1[3], w, kt2)
                alive &= ~((np.abs(p_n) < 1e-30) | np.isinf(np.abs(p_n)) | np.isinf(np.abs(p_i)))
                a = (p_i + p_n) / (2 * p_n)
                b = (p_i - p_n) / (2 * p_n)
                D = np.empty_like(T)
                D[..., 0, 0] = a
                D[..., 0, 1] = b
                D[..., 1, 0] = b
                D[..., 1, 1] = a
                T = T @ D
            p_mon = _admittance(pol, layers[mon][2], layers[mon][3], w, kt2)
            p_top = _admittance(pol, layers[-1][2], layers[-1][3], w, kt2)
            t22 = T[..., 1, 1]
            bad = np.isinf(np.abs(p_top)) | (np.abs(p_top) < 1e-30) | (np.abs(t22) < 1e-30)
            res.append(np.where(alive & ~bad, (p_mon / p_top) / t22, 0j))
    return res[0], res[1]


def compute_far_field_layered(md, layer_stack, spectrum=None) -> np.ndarray:
    """Far field (n_obs, 6, n_freq) of a z-normal surface inside a planar
    layer stack, seen from the top layer (src/Monitors/Near2Far.jl:723-985):
    spectra of the four tangential fields at the zero-padded (4x) FFT bins
    around each direction (k_t = k_top sin(theta) (cos phi, sin phi)),
    bilinearly interpolated, corrected for the in-plane Yee offsets, split
    into TE/TM upward amplitudes (E and H half a cell apart along z), carried
    through the stack by transfer matrices and projected by stationary
    phase.  As in the reference, the phase is that of an aperture whose first
    sample sits at the origin (the exp(i k.x0) of the surface position is not
    applied).  The spectral sums run on the device (kh_near2far_spectrum);
    ``spectrum(k, pad1, pad2, bins) -> (n_bins, 4)`` replaces them when
    given."""
    md = md.monitor_data if isinstance(md, Near2FarMonitor) else md
    if md.normal_axis != 3:
        raise ValueError("the layered projection needs a surface normal to z (normal_axis 3)")
    if md.theta is None or md.phi is None:
        raise ValueError("the layered projection takes directions (theta and phi), not observation points")
    base, n1, n2, d1, d2, dA = surface_layout(md)
    if spectrum is None:
        owner, ids = _owner(md)
        spectrum = lambda k, p1, p2, bins: owner.engine.near2far_spectrum(ids, 3, n1, n2, k, p1, p2, bins)
    layers = [(float(L.z_min), float(L.z_max), float(L.eps), float(L.mu)) for L in layer_stack.layers]
    e1b, e2b, h1b, h2b = base
    z_mon = (e1b[2] + e2b[2] + h1b[2] + h2b[2]) / 4.0
    mon = next((i for i, L in enumerate(layers) if L[0] <= z_mon <= L[1]), None)
    if mon is None:
        raise ValueError(f"Monitor z={z_mon} is not inside any layer")
    eps_m, mu_m = layers[mon][2], layers[mon][3]
    eps_t, mu_t = layers[-1][2], layers[-1][3]
    Z_top = math.sqrt(mu_t / eps_t)
    dz = (e1b[2] + e2b[2]) / 2.0 - (h1b[2] + h2b[2]) / 2.0
    dx, dy = md.dx, md.dy
    ns = md.normal_sign
    th = np.asarray(md.theta, dtype=np.float64)
    ph = np.asarray(md.phi, dtype=np.float64)
    PH, TH = np.meshgrid(ph, th, indexing="ij")           # phi outer, theta fastest
    TH, PH = TH.ravel(), PH.ravel()
    st, ct, sp, cp = np.sin(TH), np.cos(TH), np.sin(PH), np.cos(PH)
    out = np.zeros((TH.size, 6, len(md.frequencies)), dtype=np.complex128)
    p1, p2 = 4 * n1, 4 * n2

    def grid(p, d):
        # sorted fftfreq(p, 2 pi / d): m * ((2 pi / d) / p), m = -p/2 .. p/2 - 1
        mult = (2 * math.pi / d) / p
        g0 = (-(p // 2)) * mult
        return g0, ((-(p // 2) + 1) * mult) - g0

    gx0, dkx = grid(p1, dx)
    gy0, dky = grid(p2, dy)
    for k, freq in enumerate(md.frequencies):
        w = 2 * math.pi * freq
        k_top = w * math.sqrt(eps_t * mu_t)
        kx = k_top * st * cp
        ky = k_top * st * sp
        kt2 = kx ** 2 + ky ** 2
        fx = (kx - gx0) / dkx + 1.0
        fy = (ky - gy0) / dky + 1.0
        ix = np.floor(fx).astype(np.int64)
        iy = np.floor(fy).astype(np.int64)
        ok = (ix >= 1) & (ix < p1) & (iy >= 1) & (iy < p2)
        wx, wy = fx - ix, fy - iy
        vals = np.zeros((TH.size, 4), dtype=np.complex128)
        if ok.any():
            cx = np.stack([ix[ok], ix[ok] + 1, ix[ok], ix[ok] + 1], axis=1) - 1 - p1 // 2
            cy = np.stack([iy[ok], iy[ok], iy[ok] + 1, iy[ok] + 1], axis=1) - 1 - p2 // 2
            pairs = np.stack([cx.ravel(), cy.ravel()], axis=1)
            uniq, inv = np.unique(pairs, axis=0, return_inverse=True)
            S = spectrum(k, p1, p2, uniq) * (dx * dy)
            C = S[inv.ravel()].reshape(-1, 4, 4)              # (obs, corner, field)
            a, b = wx[ok][:, None], wy[ok][:, None]
            vals[ok] = ((1 - a) * (1 - b) * C[:, 0] + a * (1 - b) * C[:, 1]
                        + (1 - a) * b * C[:, 2] + a * b * C[:, 3])
        ex = vals[:, 0] * np.exp(-1j * kx * (e1b[0] - e2b[0]))
        ey = vals[:, 1] * np.exp(-1j * ky * (e2b[1] - e1b[1]))
        hx = vals[:, 2] * np.exp(-1j * ky * (h1b[1] - h2b[1]))
        hy = vals[:, 3] * np.exp(-1j * kx * (h2b[0] - h1b[0]))
        e_te = -sp * ex + cp * ey
        e_tm = cp * ex + sp * ey
        h_p = cp * hx + sp * hy
        h_s = -sp * hx + cp * hy
        with np.errstate(divide="ignore", invalid="ignore"):
            kz = _kz(eps_m, mu_m, w, kt2)
            p_te = kz / (w * mu_m)
            p_tm = np.where(np.abs(kz) > 1e-30, (w * eps_m) / kz, complex(math.inf, 0.0))
            phz = np.exp(-1j * kz * dz)
            ckz = np.cos(kz * dz)
            ckz = np.where(np.abs(ckz) < 1e-30, 1e-30 + 0j, ckz)
            a_te = np.where((np.abs(p_te) > 1e-30) & ~np.isinf(np.abs(p_te)),
                            (e_te - ns * (h_p / p_te) * phz) / (2 * ckz), 0j)
            a_tm = np.where((np.abs(p_tm) > 1e-30) & ~np.isinf(np.abs(p_tm)),
                            (e_tm + ns * (h_s / p_tm) * phz) / (2 * ckz), 0j)
        t_te, t_tm = layer_transmission(layers, z_mon, freq, kx, ky)
        pref = -1j * k_top * np.exp(1j * k_top * md.r) / (2 * math.pi * md.r)
        e_th = pref * (t_tm * a_tm)
        e_ph = pref * ct * (t_te * a_te)
        Ex = e_th * ct * cp + e_ph * (-sp)
        Ey = e_th * ct * sp + e_ph * cp
        Ez = e_th * (-st)
        rx, ry, rz = st * cp, st * sp, ct
        out[:, 0, k], out[:, 1, k], out[:, 2, k] = Ex, Ey, Ez
        out[:, 3, k] = (ry * Ez - rz * Ey) / Z_top
        out[:, 4, k] = (rz * Ex - rx * Ez) / Z_top
        out[:, 5, k] = (rx * Ey - ry * Ex) / Z_top
    return out
