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


def _interp_weights(x, p):
    """Per-point neighbour indices and weights of the reference's linear
    interpolation (src/utils.jl:587-621) on nodes ``x``: the nearest node
    alone when it equals the point or is an end node, else the two nodes
    around the point, the left one weighted by (p - x_left) / dx and the
    right one by (x_right - p) / dx, as the reference computes them."""
    x = np.asarray(x, dtype=np.float64)
    p = np.asarray(p, dtype=np.float64).ravel()
    c = np.argmin(np.abs(x[None, :] - p[:, None]), axis=1)
    xc = x[c]
    single = (xc == p) | (c == 0) | (c == len(x) - 1)
    left = np.where(single, c, np.where(xc > p, c - 1, c))
    right = np.where(single, c, np.where(xc > p, c, c + 1))
    dx = x[right] - x[left]
    safe = np.where(single, 1.0, dx)
    wl = np.where(single, 1.0, (p - x[left]) / safe)
    wr = np.where(single, 0.0, (x[right] - p) / safe)
    return left, right, wl, wr


def interpolate_array(data, center, size, X, Y, Z):
    """The reference's trilinear interpolator of a 3D array spanning a volume
    (gen_interpolator_from_array, src/utils.jl:659-728), evaluated on the
    broadcastable coordinate arrays X, Y, Z (one-dimensional along their own
    axis).  A singleton axis of ``data`` sits at the volume's centre."""
    data = np.asarray(data)
    if data.ndim != 3:
        raise ValueError(f"the field array must be 3D (got {data.ndim} dimensions)")
    axes = []
    for a, P in enumerate((X, Y, Z)):
        n = data.shape[a]
        nodes = [float(center[a])] if n == 1 else np.linspace(center[a] - size[a] / 2, center[a] + size[a] / 2, n)
        pts = np.asarray(P, dtype=np.float64).reshape(-1)
        axes.append(_interp_weights(nodes, pts))
    (l1, r1, a1, b1), (l2, r2, a2, b2), (l3, r3, a3, b3) = axes
    out = np.zeros((len(l1), len(l2), len(l3)), dtype=np.complex128)
    for i1, w1 in ((l1, a1), (r1, b1)):
        for i2, w2 in ((l2, a2), (r2, b2)):
            for i3, w3 in ((l3, a3), (r3, b3)):
                w = w1[:, None, None] * w2[None, :, None] * w3[None, None, :]
                out += w * data[i1[:, None, None], i2[None, :, None], i3[None, None, :]]
    return out


_CURRENT_OF = {
    # J = n x H on E components, M = -n x E on H components
    # (src/Sources/SpatialSources.jl:186-213): (n index, field, n index, field)
    "Ex": ((1, "Hz"), (2, "Hy")), "Ey": ((2, "Hx"), (0, "Hz")), "Ez": ((0, "Hy"), (1, "Hx")),
    "Hx": ((2, "Ey"), (1, "Ez")), "Hy": ((0, "Ez"), (2, "Ex")), "Hz": ((1, "Ex"), (0, "Ey")),
}


class EquivalentSource(Source):
    """Currents equivalent to given E and H on a plane
    (src/Sources/SpatialSources.jl:155-257): J = n x H on the transverse E
    components, M = -n x E on the transverse H components; n is the plane's
    +normal, the injection direction.  ``fields`` is either a callable
    ``fields(name, X, Y, Z)`` returning the complex field ``name`` ('Ex' ...
    'Hz') at the simulation points, or, as in the reference, a dict of the six
    components (Field instances or names) to 3D arrays sampled on the plane,
    interpolated the reference's way (``interpolate_array``)."""

    def __init__(self, *, time_profile, fields, center=(0.0, 0.0, 0.0), size=(0.0, 0.0, 0.0),
                 amplitude=1.0):
        self.time_profile = time_profile
        self.center = [float(v) for v in center]
        self.size = [float(v) for v in size]
        self.amplitude = amplitude
        zeros = [a for a in range(3) if self.size[a] == 0.0]
        if len(zeros) != 1:
            raise ValueError(f"expected a planar volume (one size entry equal to 0), got size={self.size}")
        self.normal_axis = zeros[0]
        self.currents = None
        if isinstance(fields, dict):
            arrays = {(k if isinstance(k, str) else type(k).__name__): np.asarray(v) for k, v in fields.items()}
            missing = [c for c in ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz") if c not in arrays]
            if missing:
                raise ValueError(f"an equivalent source needs all six field components; missing {missing}")
            for c, v in arrays.items():
                if v.ndim != 3:
                    raise ValueError(f"field {c} must be a 3D array (got {v.ndim} dimensions)")
            nv = [0.0, 0.0, 0.0]
            nv[self.normal_axis] = 1.0
            self.currents = {c: nv[i] * arrays[f] - nv[j] * arrays[g]
                             for c, ((i, f), (j, g)) in _CURRENT_OF.items()}
            fields = None
        self.fields = fields

    def components(self):
        from .grid import Ex, Ey, Ez, Hx, Hy, Hz
        trans = {0: [Ey(), Ez(), Hy(), Hz()], 1: [Ex(), Ez(), Hx(), Hz()], 2: [Ex(), Ey(), Hx(), Hy()]}
        return trans[self.normal_axis]

    def profile(self, x, y, z, component):
        if self.currents is not None:
            return interpolate_array(self.currents[type(component).__name__], self.center, self.size, x, y, z)
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


def TFSFSource(*, time_profile, size, center=(0.0, 0.0, 0.0), injection_axis=3, direction=-1,
               polarization_angle=0.0, amplitude=1.0, ε=1.0, μ=1.0, n_grid=50):
    """Total-field / scattered-field box (src/Sources/SpatialSources.jl:297-425):
    six equivalent-source faces, each injecting the incident plane wave
    E = amplitude e_pol exp(-i k k_hat . r), H = k_hat x E / Z0, sampled on an
    n_grid-scaled grid of the face and interpolated the reference's way;
    k_hat = direction along ``injection_axis`` (1-based), e_pol = (cos a,
    sin a) on the two other axes in order.  The upper face of each axis
    carries amplitude -1, the lower +1.  Returns the list of six sources."""
    if injection_axis not in (1, 2, 3):
        raise ValueError("injection_axis must be 1, 2 or 3")
    if direction not in (-1, 1):
        raise ValueError("direction must be -1 or +1")
    ax = injection_axis - 1
    k_hat = np.zeros(3)
    k_hat[ax] = float(direction)
    k_mag = 2 * math.pi * float(time_profile.fcen) * math.sqrt(ε * μ)
    z0 = math.sqrt(μ / ε)
    trans = [a for a in range(3) if a != ax]
    e_pol = np.zeros(3)
    e_pol[trans[0]] = math.cos(polarization_angle)
    e_pol[trans[1]] = math.sin(polarization_angle)
    h_pol = np.cross(k_hat, e_pol) / z0
    size = [float(v) for v in size]
    half = [v / 2 for v in size]
    smax = max(size)
    out = []
    for axis in range(3):
        for side in (-1, 1):
            fc = [float(v) for v in center]
            fc[axis] += side * half[axis]
            fs = list(size)
            fs[axis] = 0.0
            o1, o2 = [a for a in range(3) if a != axis]
            n1 = max(2, int(round(fs[o1] * n_grid / smax)))
            n2 = max(2, int(round(fs[o2] * n_grid / smax)))
            r1 = np.linspace(fc[o1] - fs[o1] / 2, fc[o1] + fs[o1] / 2, n1)
            r2 = np.linspace(fc[o2] - fs[o2] / 2, fc[o2] + fs[o2] / 2, n2)
            shape = [1, 1, 1]
            shape[o1], shape[o2] = n1, n2
            pos = [None, None, None]
            pos[axis] = np.full((n1, n2), fc[axis])
            pos[o1], pos[o2] = np.meshgrid(r1, r2, indexing="ij")
            phase = np.exp(-1j * k_mag * (k_hat[0] * pos[0] + k_hat[1] * pos[1] + k_hat[2] * pos[2]))
            fields = {}
            for name, vec in (("E", e_pol), ("H", h_pol)):
                for a in range(3):
                    fields[name + "xyz"[a]] = (amplitude * vec[a] * phase).reshape(shape)
            out.append(EquivalentSource(time_profile=time_profile, fields=fields, center=fc, size=fs,
                                        amplitude=-1.0 if side == 1 else 1.0))
    return out


class GaussianBeamSource(Source):
    """Gaussian beam as the complex-source-point field
    (src/Sources/SpatialSources.jl:650-857): E and H of a point dipole at the
    complex position beam_center + i z0 k_hat (z0 = k w0^2 / 2), normalised
    to the field at the waist centre, injected as J = n x H (E components)
    and M = -n x E (H components) on the plane's transverse components.  A
    polarization with an imaginary part fails, as the reference's
    imaginary-part branch does (it names an undefined variable)."""

    def __init__(self, *, time_profile, beam_center, beam_waist, k_vector, polarization, center=(0.0, 0.0, 0.0),
                 size=(0.0, 0.0, 0.0), amplitude=1.0, ε=1.0, μ=1.0):
        self.time_profile = time_profile
        self.center = [float(v) for v in center]
        self.size = [float(v) for v in size]
        self.amplitude = amplitude
        self.ε, self.μ = ε, μ
        self.beam_center = np.asarray(beam_center, dtype=np.complex128)
        self.beam_waist = float(beam_waist)
        self.k_vector = np.asarray(k_vector, dtype=np.float64)
        self.polarization = np.asarray(polarization, dtype=np.complex128)
        zeros = [a for a in range(3) if self.size[a] == 0.0]
        if len(zeros) != 1:
            raise ValueError(f"expected a planar volume (one size entry equal to 0), got size={self.size}")
        self.normal_axis = zeros[0]
        if np.linalg.norm(self.polarization.imag) > 1e-6:
            raise NotImplementedError("a polarization with an imaginary part is not supported: the reference's "
                                      "branch for it cannot run")

    def components(self):
        from .grid import Ex, Ey, Ez, Hx, Hy, Hz
        trans = {0: [Ey(), Ez(), Hy(), Hz()], 1: [Ex(), Ez(), Hx(), Hz()], 2: [Ex(), Ey(), Hx(), Hy()]}
        return trans[self.normal_axis]

    def fields(self, X, Y, Z):
        """(E, H): complex arrays of shape (3, *broadcast shape) at the points
        (get_gaussianbeam_field_profiles, :715-831, per point)."""
        X, Y, Z = np.broadcast_arrays(np.asarray(X, float), np.asarray(Y, float), np.asarray(Z, float))
        eps_r, mu_r = float(np.real(self.ε)), float(np.real(self.μ))
        k = complex(2 * math.pi * float(self.time_profile.fcen) * math.sqrt(eps_r * mu_r))
        zr = math.sqrt(mu_r / eps_r)
        z0 = k * self.beam_waist ** 2 / 2.0
        kz0 = k * z0
        xrel = np.stack([X - self.beam_center[0], Y - self.beam_center[1], Z - self.beam_center[2]])
        zhat = self.k_vector / np.linalg.norm(self.k_vector)
        cr = np.cross(zhat[:, None], xrel.reshape(3, -1), axis=0).reshape(xrel.shape)
        rho = np.sqrt(np.sum(np.abs(cr) ** 2, axis=0))
        z = np.tensordot(zhat, xrel, axes=1)
        zc = z - 1j * z0
        rsq = rho * rho + zc * zc
        R = np.sqrt(rsq)
        kR = k * R
        kRsq = kR * kR
        kR3 = kRsq * kR
        big = np.abs(kR) > 1e-15
        resc = big & (np.abs(np.imag(kR)) > 30.0)
        with np.errstate(all="ignore"):
            expi = np.exp(1j * np.real(kR))
            ep = np.exp(np.imag(kR) - kz0)
            em = np.exp(-np.imag(kR) - kz0)
            cos_r = 0.5 * (expi * em + np.conj(expi) * ep)
            sin_r = -1j * 0.5 * (expi * em - np.conj(expi) * ep)
            coskR = np.where(resc, cos_r, np.cos(kR))
            sinkR = np.where(resc, sin_r, np.sin(kR))
            f = -3 * (coskR / kRsq - sinkR / kR3)
            g = 1.5 * (sinkR / kR + coskR / kRsq - sinkR / kR3)
            fmg = (f - g) / rsq
        kR4 = kRsq * kRsq
        f = np.where(big, f, kR4 / 280.0 - kRsq / 10.0 + 1.0)
        g = np.where(big, g, kR4 * 3.0 / 280.0 - kRsq / 5.0 + 1.0)
        fmg = np.where(big, fmg, (kR4 / 5040.0 - kRsq / 140.0 + 0.1) * (k * k))
        i2fk = 0.5j * f * k
        E = np.zeros((3,) + X.shape, dtype=np.complex128)
        H = np.zeros((3,) + X.shape, dtype=np.complex128)
        rnorm = float(np.linalg.norm(self.polarization.real))
        if rnorm > 1e-6:
            xhat = self.polarization.real / rnorm
            yhat = np.cross(zhat, xhat)
            x = np.tensordot(xhat, xrel, axes=1)
            y = np.tensordot(yhat, xrel, axes=1)
            e_loc = [g + fmg * x * x + i2fk * zc, fmg * x * y, fmg * x * zc - i2fk * x]
            h_loc = [fmg * x * y, g + fmg * y * y + i2fk * zc, fmg * y * zc - i2fk * y]
            for a in range(3):
                E[a] += rnorm * (xhat[a] * e_loc[0] + yhat[a] * e_loc[1] + zhat[a] * e_loc[2])
                H[a] += rnorm * (xhat[a] * h_loc[0] + yhat[a] * h_loc[1] + zhat[a] * h_loc[2])
        # the normalisation of each point's branch (the unscaled one overflows
        # for a wide beam, as it does in the reference, at points that take it)
        with np.errstate(all="ignore"):
            e_resc = 3.0 / (2 * kz0 ** 3) * (kz0 * (kz0 - 1) + 0.5 * (1.0 - np.exp(-2.0 * kz0)))
            e_plain = 3.0 / (2 * kz0 ** 3) * (np.exp(kz0) * kz0 * (kz0 - 1) + np.sinh(kz0))
        e_orig = np.where(resc, e_resc, e_plain)
        return E / e_orig, H / (e_orig * zr)

    def profile(self, x, y, z, component):
        E, H = self.fields(x, y, z)
        n = self.normal_axis
        a = component.axis
        b = 3 - n - a
        sign = 1.0 if (a, n, b) in ((0, 1, 2), (1, 2, 0), (2, 0, 1)) else -1.0
        # (n x F)_a = eps_{a n b} F_b with n the unit normal
        if component.electric:
            return sign * H[b]
        return -sign * E[b]


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
