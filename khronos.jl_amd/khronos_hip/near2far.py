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
        raise ValueError("Near2FarMonitor size must have exactly one zero dimension")
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
    raise ValueError("Near2FarMonitorData must have either theta/phi or observation_points")


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


def compute_far_field_layered(md, layer_stack):
    raise NotImplementedError("layered near-to-far projection")
