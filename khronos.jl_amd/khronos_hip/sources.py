"""Time profiles and spatial current sources (host setup).

* Gaussian pulse / CW / custom time profiles -- src/Sources/TimeSources.jl:53-190
* UniformSource, amplitude assembly        -- src/Sources/SpatialSources.jl:99-129,
                                              src/Sources/Sources.jl:44-131
* ArraySource: caller-supplied complex amplitude on the source grid volume
  (the role the reference's EquivalentSource/ModeSource profiles play:
  a host-computed array the time step only scales and adds).

The time step sees a source as (component, grid volume, complex amplitude
array, waveform f(t) per half step); ``S += Re(f(t) * A)`` on the volume
(src/Sources/Sources.jl:346-357).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

from .grid import Field, Volume, grid_volume, component_origin, weight_1d


# --------------------------------------------------------------------------
# Time profiles
# --------------------------------------------------------------------------
class TimeSource:
    def eval(self, t: float, R) -> complex:  # pragma: no cover - interface
        raise NotImplementedError

    def cutoff(self) -> float:
        raise NotImplementedError


def _f32_exp(x):
    return np.float32(math.exp(float(x)))


@dataclass
class ContinuousWaveData(TimeSource):
    """src/Sources/TimeSources.jl:53-72."""

    fcen: float

    def eval(self, t: float, R) -> complex:
        two_pi = R(2) * R(math.pi)
        ph = (two_pi * R(self.fcen)) * R(t)
        # exp(-i * ph): cos/sin evaluated in float64 and rounded to R
        return complex(R(math.cos(float(ph))), -R(math.sin(float(ph))))

    def cutoff(self) -> float:
        raise ValueError("a continuous-wave source has no cutoff time")

    @property
    def f_max(self):
        return float(self.fcen)


@dataclass
class GaussianPulseData(TimeSource):
    """src/Sources/TimeSources.jl:78-138 (fields as stored by the reference)."""

    fcen: float
    fwidth: float
    width: float
    peak_time: float
    cutoff_t: float

    def eval(self, t: float, R) -> complex:
        fcen, w, peak, cut = R(self.fcen), R(self.width), R(self.peak_time), R(self.cutoff_t)
        tt = R(t) - peak
        if tt > cut:
            return complex(0.0, 0.0)
        two_pi = R(2) * R(math.pi)
        g = R(math.exp(float(((-tt) * tt) / ((R(2) * w) * w))))
        b = ((-two_pi) * fcen) * tt  # imaginary part of -2*pi*i*fcen*tt
        c = R(math.cos(float(b)))
        s = R(math.sin(float(b)))
        return complex(g * c, g * s)

    def cutoff(self) -> float:
        return float(self.cutoff_t)

    @property
    def f_max(self):
        return float(self.fcen) + float(self.fwidth) / 2


def _gaussian_bandwidth(width: float) -> float:
    # src/Sources/TimeSources.jl:118-121
    tol = 1e-7
    return math.sqrt(-2.0 * math.log(tol)) / (width * math.pi)


def GaussianPulseSource(*, fcen, fwidth, start_time=0.0, cutoff_scale=5.0, real=np.float64):
    """src/Sources/TimeSources.jl:91-109 (peak snapped to a multiple of 1/fcen)."""
    width = 1.0 / fwidth
    cutoff = width * cutoff_scale + start_time
    fw = _gaussian_bandwidth(width)
    while math.exp(-cutoff * cutoff / (2 * width * width)) < 1e-100:
        cutoff *= 0.9
    peak_raw = cutoff / 2
    period = 1.0 / fcen
    peak = round(peak_raw / period) * period  # ties to even, like Julia's round
    R = real
    return GaussianPulseData(float(R(fcen)), float(R(fw)), float(R(width)), float(R(peak)), float(R(cutoff)))


def ContinuousWaveSource(*, fcen, real=np.float64):
    return ContinuousWaveData(float(real(fcen)))


@dataclass
class CustomSourceData(TimeSource):
    """src/Sources/TimeSources.jl:157-190: arbitrary ``src_func(t)``."""

    src_func: Callable[[float], complex]
    fcen: float
    fwidth: float
    end_time: float

    def eval(self, t: float, R) -> complex:
        if R(t) > R(self.end_time):
            return complex(0.0, 0.0)
        v = complex(self.src_func(float(R(t))))
        return complex(R(v.real), R(v.imag))

    def cutoff(self) -> float:
        return float(self.end_time)

    @property
    def f_max(self):
        return float(self.fcen) + float(self.fwidth) / 2


def CustomSource(*, src_func, fcen, fwidth, end_time):
    return CustomSourceData(src_func, float(fcen), float(fwidth), float(end_time))


# --------------------------------------------------------------------------
# Spatial sources
# --------------------------------------------------------------------------
class Source:
    time_profile: TimeSource
    center: Sequence[float]
    size: Sequence[float]
    amplitude: complex

    def components(self):
        raise NotImplementedError

    def profile(self, x, y, z, component) -> np.ndarray:
        """Complex spatial profile evaluated on the broadcast grid (x, y, z)."""
        return np.ones(np.broadcast(x, y, z).shape, dtype=np.complex128)


@dataclass
class UniformSource(Source):
    """src/Sources/SpatialSources.jl:99-129."""

    time_profile: TimeSource
    component: Field
    amplitude: complex = 1.0
    center: Sequence[float] = (0.0, 0.0, 0.0)
    size: Sequence[float] = (0.0, 0.0, 0.0)

    def components(self):
        return [self.component]


@dataclass
class ArraySource(Source):
    """A source whose complex profile is a host function of position.

    ``profile_fn(x, y, z, component)`` receives broadcastable coordinate arrays
    and returns the complex profile; the amplitude assembly multiplies it by
    the interpolation weight and ``amplitude`` exactly as the reference does
    for its non-uniform spatial sources (src/Sources/Sources.jl:101-131).
    """

    time_profile: TimeSource
    component: Field
    profile_fn: Callable = None
    amplitude: complex = 1.0
    center: Sequence[float] = (0.0, 0.0, 0.0)
    size: Sequence[float] = (0.0, 0.0, 0.0)

    def components(self):
        return [self.component]

    def profile(self, x, y, z, component):
        return np.asarray(self.profile_fn(x, y, z, component), dtype=np.complex128)


def GaussianProfileSource(*, time_profile, component, center, size, waist, amplitude=1.0):
    """Planar source with a Gaussian transverse envelope of 1/e field radius ``waist``."""
    c = [float(v) for v in center]

    def fn(x, y, z, comp):
        r2 = np.zeros(np.broadcast(x, y, z).shape)
        for a, coord in enumerate((x, y, z)):
            if float(size[a]) != 0.0:
                r2 = r2 + (coord - c[a]) ** 2
        return np.exp(-r2 / (waist * waist)).astype(np.complex128)

    return ArraySource(time_profile=time_profile, component=component, profile_fn=fn,
                       amplitude=amplitude, center=center, size=size)


class EquivalentSource(Source):
    """Currents equivalent to given E and H on a plane
    (src/Sources/SpatialSources.jl:155-257): J = n x H on the transverse E
    components, M = -n x E on the transverse H components.  ``fields(name,
    X, Y, Z)`` returns the complex field ``name`` ('Ex' ... 'Hz') at the
    simulation points; n is the plane's +normal, the injection direction."""

    def __init__(self, *, time_profile, fields: Callable, center=(0.0, 0.0, 0.0), size=(0.0, 0.0, 0.0),
                 amplitude=1.0):
        self.time_profile = time_profile
        self.fields = fields
        self.center = [float(v) for v in center]
        self.size = [float(v) for v in size]
        self.amplitude = amplitude
        zeros = [a for a in range(3) if self.size[a] == 0.0]
        if len(zeros) != 1:
            raise ValueError(f"expected a planar volume (one size entry equal to 0), got size={self.size}")
        self.normal_axis = zeros[0]

    def components(self):
        from .grid import Ex, Ey, Ez, Hx, Hy, Hz
        trans = {0: [Ey(), Ez(), Hy(), Hz()], 1: [Ex(), Ez(), Hx(), Hz()], 2: [Ex(), Ey(), Hx(), Hy()]}
        return trans[self.normal_axis]

    def profile(self, x, y, z, component):
        n = self.normal_axis
        a = component.axis
        # (n x F)_a = eps_{a n b} F_b with b the remaining axis
        b = 3 - n - a
        sign = 1.0 if (a, n, b) in ((0, 1, 2), (1, 2, 0), (2, 0, 1)) else -1.0
        X, Y, Z = np.broadcast_arrays(x, y, z)
        if component.electric:
            return sign * np.asarray(self.fields("H" + "xyz"[b], X, Y, Z), dtype=np.complex128)
        return -sign * np.asarray(self.fields("E" + "xyz"[b], X, Y, Z), dtype=np.complex128)


def ModeSource(*, frequency, time_profile, mode_solver_resolution, mode_index, center, size,
               solver_tolerance=1e-6, amplitude=1.0, geometry, target_neff=0.0):
    """The mode_index-th mode of the plane (center, size) at ``frequency``,
    launched along +normal as an equivalent source
    (src/Sources/SpatialSources.jl:588-646)."""
    from .mode import plane_modes
    modes = plane_modes(center, size, geometry, float(frequency), mode_solver_resolution,
                        num_modes=mode_index, target_neff=target_neff)
    if len(modes) < mode_index:
        raise ValueError(f"Mode solver found only {len(modes)} modes, mode index {mode_index} was requested")
    mode = modes[mode_index - 1]
    u, v, w = mode.axes

    def fields(name, X, Y, Z):
        P = (X, Y, Z)
        ax = "xyz".index(name[1])
        local = {u: "u", v: "v", w: "w"}[ax]
        return mode.field(name[0] + local, P[u], P[v])

    src = EquivalentSource(time_profile=time_profile, fields=fields, center=center, size=size,
                           amplitude=amplitude)
    src.mode = mode
    return src


class PlaneWaveSource(Source):
    """Plane-wave current sheet (src/Sources/SpatialSources.jl:412-586).

    ``k_vector`` or the angles (theta azimuth, phi polar) give the direction;
    k = 2 pi fcen sqrt(eps mu).  The polarization is the normal of the plane
    spanned by the sheet normal and k (fixed axes when they coincide) rotated
    by ``polarization_angle`` about k.  Electric currents carry
    n x (k x E / Z0), magnetic ones -n x E, on the sheet's transverse
    components whose factor is nonzero."""

    def __init__(self, *, time_profile, center=(0.0, 0.0, 0.0), size=(0.0, 0.0, 0.0), amplitude=1.0,
                 ε=1.0, μ=1.0, polarization_angle=0.0, angle_theta=None, angle_phi=None, k_vector=None):
        if k_vector is None and angle_theta is None and angle_phi is None:
            raise ValueError("give the propagation direction: k_vector, or angle_theta together with angle_phi")
        if k_vector is None:
            k_vector = [math.sin(angle_phi) * math.cos(angle_theta), math.sin(angle_phi) * math.sin(angle_theta),
                        math.cos(angle_phi)]
        self.time_profile = time_profile
        self.center = [float(v) for v in center]
        self.size = [float(v) for v in size]
        self.amplitude = amplitude
        self.ε, self.μ = ε, μ
        self.k = 2 * math.pi * float(time_profile.fcen) * math.sqrt(ε * μ)
        kv = np.asarray(k_vector, dtype=float)
        self.k_vector = kv / np.linalg.norm(kv)
        zeros = [a for a in range(3) if self.size[a] == 0.0]
        if len(zeros) != 1:
            raise ValueError(f"expected a planar volume (one size entry equal to 0), got size={self.size}")
        self.normal_axis = int(np.argmin(self.size))
        n = np.zeros(3)
        n[self.normal_axis] = 1.0
        self.normal_direction = n
        if float(np.dot(n, self.k_vector)) == 1.0:
            first = int(np.nonzero(self.k_vector != 0.0)[0][0])
            ppn = {0: [0.0, 1.0, 0.0], 1: [0.0, 0.0, 1.0], 2: [1.0, 0.0, 0.0]}[first]
            ppn = np.asarray(ppn)
        else:
            ppn = np.cross(n, self.k_vector)
        # Rodrigues' rotation about k (the k (k . p) term vanishes: p is normal to k)
        self.polarization = ppn * math.cos(polarization_angle) + np.cross(self.k_vector, ppn) * math.sin(
            polarization_angle)

    def pol_factor(self, component) -> float:
        a = component.axis
        if component.electric:
            z0 = math.sqrt(self.μ / self.ε)
            return float(np.cross(self.normal_direction, np.cross(self.k_vector, (1 / z0) * self.polarization))[a])
        return float(-np.cross(self.normal_direction, self.polarization)[a])

    def components(self):
        from .grid import Ex, Ey, Ez, Hx, Hy, Hz
        trans = {0: [Ey(), Ez(), Hy(), Hz()], 1: [Ex(), Ez(), Hx(), Hz()], 2: [Ex(), Ey(), Hx(), Hy()]}
        return [c for c in trans[self.normal_axis] if self.pol_factor(c) != 0.0]


def _planewave_amplitude(sim, src: PlaneWaveSource, comp, gv, org, dims, mn, mx, d):
    """The reference's separable assembly (src/Sources/Sources.jl:142-233):
    per axis weight(p) exp(-i k k_a p), weight 1 along an infinite extent;
    the z factor is dropped when the sheet is one cell thick in z; times
    (pol_factor * amplitude)."""
    n1 = dims[0]
    n2 = dims[1] if sim.ndims >= 2 else 1
    n3 = dims[2] if sim.ndims >= 3 else 1
    k = float(src.k)
    fac = []
    for a, n in enumerate((n1, n2, n3)):
        out = np.empty(n, dtype=np.complex128)
        for i in range(1, n + 1):
            p = org[a] + (i + gv.start_idx[a] - 2) * d[a]
            if a == 2 and (math.isinf(p) or math.isnan(p)):
                p = float(sim.cell_center[2])
            w = 1.0 if math.isinf(float(src.size[a])) else weight_1d(p, mn[a], mx[a], float(src.size[a]), d[a])
            out[i - 1] = w * np.exp(-1j * k * float(src.k_vector[a]) * p)
        fac.append(out)
    scale = complex(src.pol_factor(comp) * src.amplitude)
    amp = np.zeros(dims, dtype=np.complex128)
    cxy = fac[0][:, None] * fac[1][None, :]
    cz = fac[2] if n3 > 1 else np.ones(1, dtype=np.complex128)
    amp[:n1, :n2, :n3] = (cxy[:, :, None] * cz[None, None, :]) * scale
    return amp


@dataclass
class SourceData:
    """What the time step needs for one (source, component) pair."""

    component: Field
    gv: object
    amplitude: np.ndarray  # complex, shape (Nx, Ny, Nz) of the grid volume
    time_profile: TimeSource


def assemble_source(sim, src: Source):
    """Complex amplitude arrays for every component of ``src``
    (src/Sources/Sources.jl:44-131): weight(point) * amplitude * profile(point).
    """
    vol = Volume(src.center, src.size)
    mn, mx = vol.min_corner(), vol.max_corner()
    out = []
    d = [float(v) for v in sim.delta_n]
    for comp in src.components():
        gv = grid_volume(sim, vol, comp)
        org = component_origin(sim.cell_center, sim.cell_size_n, sim.delta_n, comp)
        dims = (gv.Nx, gv.Ny, max(1, gv.Nz))
        if isinstance(src, PlaneWaveSource):
            amp = _planewave_amplitude(sim, src, comp, gv, org, dims, mn, mx, d)
            out.append(SourceData(comp, gv, np.ascontiguousarray(amp), src.time_profile))
            continue
        coords = []
        weights = []
        for a in range(3):
            pts = []
            for i in range(1, dims[a] + 1):
                p = org[a] + (i + gv.start_idx[a] - 2) * d[a]
                if a == 2 and (math.isinf(p) or math.isnan(p)):
                    p = float(sim.cell_center[2])
                pts.append(p)
            pts = np.array(pts)
            coords.append(pts)
            if a < sim.ndims:
                weights.append(np.array([weight_1d(p, mn[a], mx[a], float(src.size[a]), d[a]) for p in pts]))
            else:
                weights.append(np.ones(len(pts)))
        X = coords[0][:, None, None]
        Y = coords[1][None, :, None]
        Z = coords[2][None, None, :]
        w = weights[0][:, None, None] * weights[1][None, :, None] * weights[2][None, None, :]
        prof = src.profile(X, Y, Z, comp)
        amp = (w * complex(src.amplitude)) * prof
        amp = np.broadcast_to(amp, dims).astype(np.complex128)
        out.append(SourceData(comp, gv, np.ascontiguousarray(amp), src.time_profile))
    return out
