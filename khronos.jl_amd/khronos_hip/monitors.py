"""DFT and flux monitors (host setup and flux readout).

* auto-decimation from the source bandwidth -- src/Monitors/Monitors.jl:33-78
* DFT monitor grid volume                   -- src/Monitors/Monitors.jl:202-272
* flux monitor = 4 tangential DFT monitors  -- src/Monitors/FluxMonitor.jl:18-70
* Poynting flux readout                     -- src/Monitors/FluxMonitor.jl:92-172

The accumulation itself (``F += dt * exp(i 2 pi f t) * field`` every D-th
step, H at t and E at t + dt/2) runs in the HIP library.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from .grid import Ex, Ey, Ez, Hx, Hy, Hz, Field, Volume, grid_volume


@dataclass
class DFTMonitor:
    """src/DataStructures.jl:339-346."""

    component: Field
    center: Sequence[float]
    size: Sequence[float]
    frequencies: Sequence[float]
    decimation: int = 1
    monitor_data: object = None


@dataclass
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
        elif isinstance(m, FluxMonitor):
            normal = [i for i, s in enumerate(m.size) if float(s) == 0.0]
            if not normal:
                raise ValueError("FluxMonitor size must have exactly one zero dimension")
            na = normal[0] + 1
            ecomps, hcomps = {1: ((Ey, Ez), (Hy, Hz)), 2: ((Ex, Ez), (Hx, Hz)), 3: ((Ex, Ey), (Hx, Hy))}[na]
            mk = lambda c: DFTMonitor(component=c(), center=list(m.center), size=list(m.size),
                                      frequencies=list(m.frequencies), decimation=m.decimation)
            te = [mk(c) for c in ecomps]
            th = [mk(c) for c in hcomps]
            m.monitor_data = FluxMonitorData(na, te, th, np.asarray(m.frequencies, dtype=np.float64),
                                             float(sim.delta_n[0]), float(sim.delta_n[1]), float(sim.delta_n[2]))
            dfts.extend(te + th)
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
