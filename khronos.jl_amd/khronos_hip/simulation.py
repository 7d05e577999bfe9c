"""Simulation object and run loop, mirroring the reference's user API.

* ``Simulation(...)`` keyword interface and derived grid quantities
  -- src/DataStructures.jl:714-801
* ``prepare_simulation!`` -- src/Simulation.jl:92-283 (host setup, then upload)
* ``run`` / ``run_until`` / ``stop_when_dft_decayed`` / ``run_benchmark``
  -- src/Simulation.jl:295-527
* ``reset_fields!`` -- src/Simulation.jl:571-637
* ``get_flux`` -- src/Monitors/FluxMonitor.jl:92-156

Precision is chosen per simulation (``precision="f32"|"f64"``), the role
``choose_backend(device, T)`` plays in the reference (src/load_deps.jl:23-64).
Every time step runs in libkhronos_hip on the GPU; this module only does
setup and control.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import grid as G
from .boundaries import apply_absorbers, has_absorbers, init_boundaries, needs_complex_fields, periodic_wrap
from .geometry import init_polarization, rasterize, rasterize_chi3, rasterize_conductivity
from .smoothing import NoSmoothing, apply_subpixel_smoothing
from .monitors import (DFTMonitor, FluxMonitor, FluxMonitorData, auto_decimate, expand_monitors,
                       init_dft_monitor, flux_from_dft)
from .sources import assemble_source


@dataclass
class SimSpec:
    """Everything the time step needs, in the reference's array layouts."""

    N: tuple
    real: type
    delta: tuple          # backend-precision spacings (as python floats)
    dt: float
    sigma: Optional[dict]
    eps_inv: list
    mu_inv: list
    eps_inv_scalar: float
    mu_inv_scalar: float
    sources: list         # SourceData
    dfts: list            # DFTMonitorData
    cutoff: float
    mode: str
    dft_accum_f64: bool
    chi3: Optional[np.ndarray] = None   # (Nx, Ny, Nz) or None
    poles: list = field(default_factory=list)  # geometry.PoleData
    complex_fields: bool = False        # any Bloch side
    periodic: tuple = (False, False, False)
    bloch_fwd: tuple = (1 + 0j, 1 + 0j, 1 + 0j)   # exp(+i k L): lower interior -> upper ghost
    bloch_rev: tuple = (1 + 0j, 1 + 0j, 1 + 0j)   # exp(-i k L): upper interior -> lower ghost
    field_real: Optional[type] = None   # field precision when it differs from the setup's
    # conductivity sigma_D of Ex,Ey,Ez then sigma_B of Hx,Hy,Hz (geometry layout, absorbers added)
    cond: list = field(default_factory=lambda: [None] * 6)


class Simulation:
    """Python mirror of ``Khronos.Simulation`` (src/DataStructures.jl:714-801)."""

    def __init__(self, *, cell_size, resolution, sources, cell_center=(0.0, 0.0, 0.0),
                 boundaries=None, boundary_conditions=None, geometry=None, monitors=None,
                 absorbers=None, subpixel_smoothing=None, Courant=0.5, precision="f64", mode="fast",
                 dft_accum="f64", num_chunks=1, engine=None, setup_precision=None, raster="host"):
        if num_chunks not in (1, None):
            raise NotImplementedError("only the single-domain layout (num_chunks=1) is provided")
        if precision not in ("f32", "f64"):
            raise ValueError("precision must be 'f32' or 'f64'")
        if setup_precision not in (None, "f32", "f64"):
            raise ValueError("setup_precision must be 'f32' or 'f64'")
        if raster not in ("host", "device"):
            raise ValueError("raster must be 'host' or 'device'")
        # "device": paint and smooth the material arrays with the HIP library
        self.raster = raster
        # the precision form: setup_precision="f64" with precision="f32" keeps
        # the grid snapping, weights, profiles and waveforms of the FP64 run
        # and runs the time step in FP32
        self.field_real = np.float32 if precision == "f32" else np.float64
        self.real = self.field_real if setup_precision is None else (
            np.float32 if setup_precision == "f32" else np.float64)
        R = self.real
        self.precision = precision
        self.mode = mode
        self.dft_accum_f64 = (dft_accum == "f64")
        self.cell_size = [float(v) for v in cell_size]
        self.cell_center = [float(v) for v in cell_center]
        self.resolution = resolution
        self.sources = list(sources)
        self.boundaries = boundaries
        self.boundary_conditions = boundary_conditions
        self.absorbers = absorbers
        self.subpixel_smoothing = subpixel_smoothing if subpixel_smoothing is not None else NoSmoothing()
        self.geometry = list(geometry) if geometry else []
        self.monitors = list(monitors) if monitors else []
        self.Courant = Courant
        # src/DataStructures.jl:730-741 (grid counts from the user's float64 values)
        self.Nx = int(math.floor(self.cell_size[0] * resolution))
        self.Ny = int(math.floor(self.cell_size[1] * resolution))
        # 2D (Nz = 0): dz = Inf, so dt/dz = 0, and every array keeps one z
        # cell (src/DataStructures.jl:689-712, src/utils.jl:20-37)
        if len(self.cell_size) < 3:
            self.cell_size = self.cell_size + [0.0] * (3 - len(self.cell_size))
        self.Nz = 0 if self.cell_size[2] == 0.0 else int(math.floor(self.cell_size[2] * resolution))
        self.ndims = 2 if self.Nz == 0 else 3
        self.cell_size_n = [float(R(v)) for v in self.cell_size]
        dz = R(math.inf) if self.Nz == 0 else R(self.cell_size[2] / self.Nz)
        self.delta_r = [R(self.cell_size[0] / self.Nx), R(self.cell_size[1] / self.Ny), dz]
        self.delta_n = [float(v) for v in self.delta_r]
        self.dt_r = R(min(self.delta_n) * Courant)
        self.dt_n = float(self.dt_r)
        self.timestep = 0
        self.is_prepared = False
        self.monitor_data = []
        self.source_data = []
        self.spec: Optional[SimSpec] = None
        self.engine = None
        self._engine_cls = engine

    # ------------------------------------------------------------ setup
    @property
    def Δt(self):
        return self.dt_n

    def last_source_time(self) -> float:
        """src/Sources/TimeSources.jl:46-47; +inf when a source has no cutoff."""
        ts = []
        for s in self.sources:
            try:
                ts.append(float(self.real(s.time_profile.cutoff())))
            except ValueError:
                return math.inf
        return max(ts) if ts else -math.inf

    def build_spec(self) -> SimSpec:
        """Host setup of ``prepare_simulation!`` (src/Simulation.jl:92-283)."""
        sig = init_boundaries(self)
        if self.raster == "device":
            from .raster import raster_device
            eps_inv, mu_inv = raster_device(self, self.geometry)
        else:
            eps_inv, mu_inv = rasterize(self, self.geometry)
            # subpixel smoothing of eps^-1 before anything reads it (src/Geometry.jl:602-603)
            eps_inv = apply_subpixel_smoothing(self, self.geometry, eps_inv)
        chi3 = rasterize_chi3(self, self.geometry)
        poles, eps_inv = init_polarization(self, self.geometry, eps_inv)
        # conductivity (src/Geometry.jl:467-511, :708-789)
        absorb = has_absorbers(self.absorbers)
        sd, sb = rasterize_conductivity(self, self.geometry, force=absorb)
        cond = list(sd) + list(sb)
        if absorb:
            apply_absorbers(self, self.absorbers, cond)
        if any(c is not None for c in cond) and sig is None:
            # without PML arrays the reference has no curl update for a
            # conductive component (src/Kernels/Helpers.jl:141-151 vs :225-236)
            raise ValueError("conductivity (material sigma or absorbers) needs `boundaries` "
                             "(PML thicknesses, zeros allowed)")
        wrap = periodic_wrap(self)
        srcs = []
        for s in self.sources:
            srcs.extend(assemble_source(self, s))
        auto_decimate(self)
        dft_mons = expand_monitors(self, self.monitors)
        dfts = [init_dft_monitor(self, m) for m in dft_mons]
        return SimSpec(
            N=(self.Nx, self.Ny, max(1, self.Nz)), real=self.real, field_real=self.field_real, delta=tuple(self.delta_n), dt=self.dt_n,
            sigma=sig, eps_inv=eps_inv, mu_inv=mu_inv, eps_inv_scalar=1.0, mu_inv_scalar=1.0,
            sources=srcs, dfts=dfts, cutoff=self.last_source_time(), mode=self.mode,
            dft_accum_f64=self.dft_accum_f64, chi3=chi3, poles=poles,
            complex_fields=needs_complex_fields(self.boundary_conditions),
            periodic=tuple(w[0] for w in wrap), bloch_fwd=tuple(w[1] for w in wrap),
            bloch_rev=tuple(w[2] for w in wrap), cond=cond)

    def prepare(self):
        if self.is_prepared:
            return
        self.spec = self.build_spec()
        self.source_data = self.spec.sources
        self.monitor_data = self.spec.dfts
        cls = self._engine_cls
        if cls is None:
            from .engine import HipEngine
            cls = HipEngine
        self.engine = cls(self.spec)
        for md in self.spec.dfts:
            md.owner = self
        self.is_prepared = True

    # ------------------------------------------------------------ stepping
    def round_time(self) -> float:
        """src/Simulation.jl:22: Float64(timestep * dt) with the product in R."""
        return float(self.real(self.timestep) * self.dt_r)

    def step(self, n: int = 1):
        self.prepare()
        self.engine.step(n)
        self.timestep += n

    def read_field(self, component) -> np.ndarray:
        self.prepare()
        return self.engine.read_field(G.component_code(component))

    def read_dft(self, md) -> np.ndarray:
        return self.engine.read_dft(self.spec.dfts.index(md))

    def sources_active(self) -> bool:
        return self.engine.sources_active()


# ---------------------------------------------------------------- run loop
def prepare_simulation(sim: Simulation):
    sim.prepare()


def run_until(sim: Simulation, until: float) -> Callable[[Simulation], bool]:
    """src/Simulation.jl:384-387."""
    return lambda s: s.round_time() > until


def stop_when_dft_decayed(*, tolerance=1e-6, minimum_runtime=0.0, maximum_runtime=math.inf):
    """src/Simulation.jl:411-485: per-monitor relative change of the DFT norm."""
    if minimum_runtime > maximum_runtime:
        raise ValueError(f"minimum_runtime {minimum_runtime} exceeds maximum_runtime {maximum_runtime}")
    state = {}

    def _stop(sim: Simulation) -> bool:
        if sim.round_time() < minimum_runtime:
            return False
        if sim.round_time() > maximum_runtime:
            return True
        all_conv = True
        n_active = 0
        norms = sim.engine.dft_norms()
        for i, cur in enumerate(norms):
            if cur == 0.0:
                continue
            if i not in state:
                state[i] = (cur, 0.0)
                all_conv = False
                continue
            prev, mx = state[i]
            change = abs(cur - prev)
            mx = max(mx, change)
            state[i] = (cur, mx)
            if mx == 0.0:
                all_conv = False
                continue
            n_active += 1
            if change / mx > tolerance:
                all_conv = False
        if n_active == 0:
            return False
        return all_conv

    return _stop


def run(sim: Simulation, *, until=None, until_after_sources=None, chunk_steps: int = 1):
    """src/Simulation.jl:295-382.

    With a numeric stop time the number of steps is known in advance and the
    steps are issued to the device in one call; with a callback the callback
    is evaluated after every step, as in the reference.
    """
    if until is None and until_after_sources is None:
        raise ValueError("run needs a stop condition: pass until or until_after_sources")
    if until is not None and until_after_sources is not None:
        raise ValueError("run takes one stop condition: pass until or until_after_sources, not both")
    sim.prepare()
    wait = until_after_sources is not None
    cond = until if until is not None else until_after_sources
    if wait:
        t_src = sim.last_source_time()
        n = _steps_while(sim, lambda t: t <= t_src)
        sim.step(n) if n else None
    if callable(cond):
        while not cond(sim):
            sim.step(1)
    else:
        n = _steps_while(sim, lambda t: not (t > cond))
        if n:
            sim.step(n)
    sim.engine.sync()


def _steps_while(sim: Simulation, pred) -> int:
    """Number of steps issued while ``pred(round_time)`` holds before each step."""
    n = 0
    ts = sim.timestep
    R = sim.real
    while pred(float(R(ts + n) * sim.dt_r)):
        n += 1
        if n > 10_000_000:
            raise RuntimeError("run length exceeds 1e7 steps")
    return n


def run_benchmark(sim: Simulation, until: int) -> float:
    """src/Simulation.jl:492-527: Mcells/s over steps 10..until."""
    sim.prepare()
    sim.step(9)
    sim.engine.sync()
    t0 = time.perf_counter()
    sim.step(until - 9)
    sim.engine.sync()
    dt = time.perf_counter() - t0
    return sim.Nx * sim.Ny * max(1, sim.Nz) * (until - 10) / dt / 1e6


def reset_fields(sim: Simulation):
    """src/Simulation.jl:571-637."""
    if not sim.is_prepared:
        return
    sim.engine.reset()
    sim.timestep = 0


def get_flux(monitor) -> np.ndarray:
    """Poynting flux per frequency of a FluxMonitor, reduced on the device
    (src/Monitors/FluxMonitor.jl:92-156)."""
    md: FluxMonitorData = monitor.monitor_data if isinstance(monitor, FluxMonitor) else monitor
    mons = md.tangential_E_monitors + md.tangential_H_monitors
    owner = mons[0].monitor_data.owner
    ids = [owner.spec.dfts.index(m.monitor_data) for m in mons]
    dx = 1.0 if math.isinf(md.dx) else md.dx
    dy = 1.0 if math.isinf(md.dy) else md.dy
    dz = 1.0 if math.isinf(md.dz) else md.dz
    dA = {1: dy * dz, 2: dx * dz, 3: dx * dy}[md.normal_axis]
    return owner.engine.flux(ids, md.normal_axis, dA)


def dft_fields_norm(sim: Simulation) -> float:
    """src/Simulation.jl:534-539."""
    return float(sum(sim.engine.dft_norms()))
