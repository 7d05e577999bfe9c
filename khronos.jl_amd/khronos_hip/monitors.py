"""DFT and flux monitors (host setup and flux readout).

* auto-decimation from the source bandwidth -- src/Monitors/Monitors.jl:33-78
* DFT monitor grid volume                   -- src/Monitors/Monitors.jl:202-272
* flux monitor = 4 tangential DFT monitors  -- src/Monitors/FluxMonitor.jl:18-70
* Poynting flux readout                     -- src/Monitors/FluxMonitor.jl:92-172
* mode monitor = 4 tangential DFTs + modes  -- src/Monitors/ModeMonitor.jl:191-331
* mode amplitudes a+- (overlap on device)   -- src/Monitors/ModeMonitor.jl:345-520
* near-to-far surface = 4 tangential DFTs   -- src/Monitors/Near2Far.jl:1137-1200 (near2far.py)

The accumulation itself (``F += dt * exp(i 2 pi f t) * field`` every D-th
step, H at t and E at t + dt/2) runs in the HIP library.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from .grid import Ex, Ey, Ez, Hx, Hy, Hz, Field, Volume, grid_volume
from .near2far import Near2FarMonitor, init_near2far_monitor


@dataclass
class DFTMonitor:
    """src/DataStructures.jl:339-346."""

    component: Field
    center: Sequence[float]
    size: Sequence[float]
    frequencies: Sequence[float]
    decimation: int = 1
    monitor_data: object = None


@dataclass(eq=False)   # compared by identity (it is looked up in the simulation's list)
class DFTMonitorData:
    component: Field
    gv: object
    frequencies: np.ndarray
    decimation: int
    handle: int = -1
    owner: object = None

    @property
    def fields(self) -> np.ndarray:
        """Accumulated DFT, complex array of shape (Nx, Ny, Nz, n_freq)."""
        return self.owner.read_dft(self)


@dataclass
class FluxMonitor:
    """src/DataStructures.jl:556-563."""

    center: Sequence[float]
    size: Sequence[float]
    frequencies: Sequence[float]
    decimation: int = 1
    monitor_data: object = None


@dataclass
class FluxMonitorData:
    normal_axis: int  # 1-based like the reference
    tangential_E_monitors: List[DFTMonitor]
    tangential_H_monitors: List[DFTMonitor]
    frequencies: np.ndarray
    dx: float
    dy: float
    dz: float


@dataclass
class ModeMonitor:
    """Mode overlap on a planar cross-section (src/DataStructures.jl:504-519)."""

    center: Sequence[float]
    size: Sequence[float]
    frequencies: Sequence[float]
    mode_spec: object = None
    decimation: int = 1
    monitor_data: object = None


@dataclass
class ModeMonitorData(FluxMonitorData):
    modes: list = None            # one mode (mode.Mode) per frequency
    grid: tuple = None            # (cell_center, cell_size_n, delta_n) of the simulation


_MODE_CACHE = {}


def _solve_monitor_modes(sim, m: ModeMonitor):
    from .mode import ModeSpec, plane_modes
    spec = m.mode_spec or ModeSpec()
    geo = spec.geometry if spec.geometry else sim.geometry
    out = []
    for f in m.frequencies:
        key = (tuple(float(v) for v in m.center), tuple(float(v) for v in m.size), float(f),
               int(spec.mode_solver_resolution), int(spec.num_modes), float(spec.target_neff), id(geo))
        if key not in _MODE_CACHE:
            _MODE_CACHE[key] = plane_modes(m.center, m.size, geo, float(f), spec.mode_solver_resolution,
                                           num_modes=spec.num_modes, target_neff=spec.target_neff)
        modes = _MODE_CACHE[key]
        if len(modes) < spec.num_modes:
            raise ValueError(f"Mode solver found only {len(modes)} modes at frequency {f}, but mode index "
                             f"{spec.num_modes} was requested.")
        out.append(modes[spec.num_modes - 1])
    return out


def auto_decimate(sim) -> int:
    """Nyquist decimation from the source bandwidth (src/Monitors/Monitors.jl:33-78).

    Sets ``decimation`` on every monitor that still has 1; returns D.
    """
    dt = float(sim.dt_n)
    f_max = 0.0
    for src in sim.sources:
        tp = src.time_profile
        fm = getattr(tp, "f_max", None)
        if fm is None:
            fm = 1.0 / (2.0 * dt)
        f_max = max(f_max, fm)
    if f_max <= 0:
        return 1
    d_max = max(1, int(math.floor(1.0 / (2.0 * f_max * dt))))
    if d_max <= 1:
        return 1
    for m in sim.monitors or []:
        if m is not None and hasattr(m, "decimation") and m.decimation == 1:
            m.decimation = d_max
    return d_max


def expand_monitors(sim, monitors):
    """Flatten user monitors into DFT monitors (flux -> 4 tangential DFTs)."""
    dfts = []
    for m in monitors or []:
        if isinstance(m, DFTMonitor):
            dfts.append(m)
        elif isinstance(m, (FluxMonitor, ModeMonitor)):
            normal = [i for i, s in enumerate(m.size) if float(s) == 0.0]
            if not normal:
                raise ValueError(f"{type(m).__name__} is not planar: exactly one of its size entries must be 0")
            na = normal[0] + 1
            ecomps, hcomps = {1: ((Ey, Ez), (Hy, Hz)), 2: ((Ex, Ez), (Hx, Hz)), 3: ((Ex, Ey), (Hx, Hy))}[na]
            mk = lambda c: DFTMonitor(component=c(), center=list(m.center), size=list(m.size),
                                      frequencies=list(m.frequencies), decimation=m.decimation)
            te = [mk(c) for c in ecomps]
            th = [mk(c) for c in hcomps]
            args = (na, te, th, np.asarray(m.frequencies, dtype=np.float64),
                    float(sim.delta_n[0]), float(sim.delta_n[1]), float(sim.delta_n[2]))
            if isinstance(m, ModeMonitor):
                m.monitor_data = ModeMonitorData(*args, modes=_solve_monitor_modes(sim, m),
                                                 grid=(list(sim.cell_center), list(sim.cell_size_n),
                                                       list(sim.delta_n)))
            else:
                m.monitor_data = FluxMonitorData(*args)
            dfts.extend(te + th)
        elif isinstance(m, Near2FarMonitor):
            dfts.extend(init_near2far_monitor(sim, m))
        else:
            raise TypeError(f"unsupported monitor {type(m).__name__}")
    return dfts


def init_dft_monitor(sim, m: DFTMonitor) -> DFTMonitorData:
    gv = grid_volume(sim, Volume(m.center, m.size), m.component)
    md = DFTMonitorData(m.component, gv, np.asarray(m.frequencies, dtype=np.float64), int(m.decimation))
    m.monitor_data = md
    return md


def flux_from_dft(md: FluxMonitorData, e1, e2, h1, h2, gvs) -> np.ndarray:
    """Poynting flux per frequency from the four DFT arrays
    (src/Monitors/FluxMonitor.jl:92-156), accumulated in float64."""
    na = md.normal_axis
    ax = na - 1

    def avg(f):
        f = np.asarray(f, dtype=np.complex128)
        if f.shape[ax] >= 2:
            a = np.take(f, [0], axis=ax)
            b = np.take(f, [1], axis=ax)
            return (a + b) / 2
        return f

    e1, e2, h1, h2 = avg(e1), avg(e2), avg(h1), avg(h2)
    dx = 1.0 if math.isinf(md.dx) else md.dx
    dy = 1.0 if math.isinf(md.dy) else md.dy
    dz = 1.0 if math.isinf(md.dz) else md.dz
    dA = {1: dy * dz, 2: dx * dz, 3: dx * dy}[na]
    t = [a for a in range(3) if a != ax]
    n1 = min(g.dims[t[0]] for g in gvs)
    n2 = min(g.dims[t[1]] for g in gvs)
    nf = len(md.frequencies)
    flux = np.zeros(nf)

    def sl(f):
        idx = [slice(None)] * 4
        idx[ax] = 0
        idx[t[0]] = slice(0, n1)
        idx[t[1]] = slice(0, n2)
        return f[tuple(idx)]  # shape (n1, n2, nf)

    E1, E2, H1, H2 = sl(e1), sl(e2), sl(h1), sl(h2)
    for k in range(nf):
        # the reference sums in i1-fastest order, one term at a time
        term = (np.real(E1[:, :, k] * np.conj(H2[:, :, k]) - E2[:, :, k] * np.conj(H1[:, :, k])) * dA)
        flux[k] = math.fsum(term.ravel(order="F"))
    return flux


# ------------------------------------------------------------ mode overlaps
def mode_plane_samples(md: ModeMonitorData, gvs):
    """The mode's tangential (E1, E2, H1, H2) at the samples of the four
    tangential DFTs (each at its own Yee position, after the normal-axis
    average), complex (4, n_freq, n2, n1); plus n1, n2, dA."""
    from .grid import component_origin
    ax = md.normal_axis - 1
    t0, t1 = sorted(a for a in range(3) if a != ax)
    n1 = min(g.dims[t0] for g in gvs)
    n2 = min(g.dims[t1] for g in gvs)
    cc, cs, dl = md.grid
    comps = [m.component for m in md.tangential_E_monitors + md.tangential_H_monitors]
    nf = len(md.frequencies)
    out = np.zeros((4, nf, n2, n1), dtype=np.complex128)
    for q, (comp, gv) in enumerate(zip(comps, gvs)):
        org = component_origin(cc, cs, dl, comp)
        p0 = org[t0] + (np.arange(1, n1 + 1) + gv.start_idx[t0] - 2) * float(dl[t0])
        p1 = org[t1] + (np.arange(1, n2 + 1) + gv.start_idx[t1] - 2) * float(dl[t1])
        P0, P1 = np.meshgrid(p0, p1, indexing="xy")      # (n2, n1)
        for k, mode in enumerate(md.modes):
            u, v, w = mode.axes
            pos = {t0: P0, t1: P1}
            axis = comp.axis
            name = ("E" if comp.electric else "H") + ("u" if axis == u else ("v" if axis == v else "w"))
            out[q, k] = mode.field(name, pos[u], pos[v])
    dx, dy, dz = (1.0 if math.isinf(d) else d for d in (md.dx, md.dy, md.dz))
    dA = {1: dy * dz, 2: dx * dz, 3: dx * dy}[md.normal_axis]
    return out, n1, n2, dA


def _mode_power(samples, dA):
    # P_mode = 1/2 Re sum (m1 conj(n2) - m2 conj(n1)) dA, per frequency
    m1, m2, h1, h2 = samples
    return 0.5 * np.real(np.sum(m1 * np.conj(h2) - m2 * np.conj(h1), axis=(1, 2))) * dA


def compute_mode_amplitudes(monitor):
    """(a_plus, a_minus) per frequency: overlap sums on the device
    (kh_mode_overlap) over 4 P_mode (src/Monitors/ModeMonitor.jl:345-507)."""
    md = monitor.monitor_data if isinstance(monitor, ModeMonitor) else monitor
    mons = md.tangential_E_monitors + md.tangential_H_monitors
    owner = mons[0].monitor_data.owner
    ids = [owner.spec.dfts.index(m.monitor_data) for m in mons]
    samples, n1, n2, dA = mode_plane_samples(md, [m.monitor_data.gv for m in mons])
    ov = owner.engine.mode_overlap(ids, md.normal_axis, samples, dA)
    P = _mode_power(samples, dA)
    ok = np.abs(P) > 1e-30
    ap = np.where(ok, ov[:, 0] / (4 * np.where(ok, P, 1.0)), 0)
    am = np.where(ok, ov[:, 1] / (4 * np.where(ok, P, 1.0)), 0)
    return ap, am


def mode_amplitudes_from_dft(md: ModeMonitorData, e1, e2, h1, h2, gvs):
    """Host restatement of the overlap from the four DFT arrays."""
    ax = md.normal_axis - 1
    t0, t1 = sorted(a for a in range(3) if a != ax)
    samples, n1, n2, dA = mode_plane_samples(md, gvs)

    def plane(f):
        f = np.asarray(f, dtype=np.complex128)
        if f.shape[ax] >= 2:
            f = (np.take(f, [0], axis=ax) + np.take(f, [1], axis=ax)) / 2
        idx = [slice(None)] * 4
        idx[ax] = 0
        idx[t0] = slice(0, n1)
        idx[t1] = slice(0, n2)
        return np.transpose(f[tuple(idx)], (2, 1, 0))   # (nf, n2, n1)

    E1, E2, H1, H2 = plane(e1), plane(e2), plane(h1), plane(h2)
    m1, m2, q1, q2 = samples
    s = E1 * np.conj(q2) - E2 * np.conj(q1)
    t = np.conj(m1) * H2 - np.conj(m2) * H1
    P = _mode_power(samples, dA)
    ap = np.sum(s + t, axis=(1, 2)) * dA / (4 * P)
    am = np.sum(s - t, axis=(1, 2)) * dA / (4 * P)
    return ap, am


def get_mode_transmission(monitor) -> np.ndarray:
    """|a+|^2 per frequency (src/Monitors/ModeMonitor.jl:515-518)."""
    ap, _ = compute_mode_amplitudes(monitor)
    return np.abs(ap) ** 2
