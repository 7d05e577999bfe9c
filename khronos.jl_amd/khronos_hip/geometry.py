"""Materials, shapes and per-voxel permittivity rasterization (host setup).

Mirrors src/Geometry.jl:449-663: every component's inverse-permittivity array
(component voxel count, no ghosts) starts at 1 and objects are painted in
reverse order, so the first object in ``geometry`` wins where shapes overlap
(src/Geometry.jl:233-249).  A voxel belongs to a shape when its Yee-staggered
sample point lies inside it (boundaries inclusive).

The reference tests membership with the GeometryPrimitives package (not
vendored); the box / sphere / cylinder membership tests here are the closed
forms of those shapes.  Subpixel smoothing is applied afterwards
(``smoothing.apply_subpixel_smoothing``).

Also here: the per-voxel conductivity arrays sigma_D / sigma_B
(src/Geometry.jl:467-476, painted as :117-118 with every object, zero where an
object sets none), the chi3 array (src/Geometry.jl:610-636) and the dispersive-pole
setup of ``init_polarization!`` (src/Geometry.jl:1049-1355): unique poles,
per-pole oscillator-strength arrays, their zeroing inside the PML, and the
chi1 correction of the inverse permittivity.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Tuple, List, Optional, Sequence

import numpy as np

from .grid import (Center, Ex, Ey, Ez, Hx, Hy, Hz, build_coords, component_origin, full_grid_volume, gv_coords,
                   component_voxel_count)
from .susceptibility import (LorentzianSusceptibility, chi1_coefficient, compute_ade_coefficients,
                             has_susceptibilities)


@dataclass
class Material:
    """Subset of src/DataStructures.jl:244-273 used on the time-step path."""

    ε: Optional[float] = None
    εx: Optional[float] = None
    εy: Optional[float] = None
    εz: Optional[float] = None
    σD: Optional[float] = None
    σDx: Optional[float] = None
    σDy: Optional[float] = None
    σDz: Optional[float] = None
    μ: Optional[float] = None
    μx: Optional[float] = None
    μy: Optional[float] = None
    μz: Optional[float] = None
    σB: Optional[float] = None
    σBx: Optional[float] = None
    σBy: Optional[float] = None
    σBz: Optional[float] = None
    chi3: Optional[float] = None
    susceptibilities: List[LorentzianSusceptibility] = field(default_factory=list)

    def perm(self, electric: bool, axis: int) -> Optional[float]:
        """get_material_perm_comp, else get_material_perm (src/Geometry.jl:305-312, :64-70)."""
        comp = (self.εx, self.εy, self.εz)[axis] if electric else (self.μx, self.μy, self.μz)[axis]
        return comp if comp is not None else (self.ε if electric else self.μ)

    def conductivity(self, electric: bool, axis: int) -> Optional[float]:
        """get_material_conductivity_comp, else get_material_conductivity
        (src/Geometry.jl:327-334, :72-80)."""
        comp = (self.σDx, self.σDy, self.σDz)[axis] if electric else (self.σBx, self.σBy, self.σBz)[axis]
        return comp if comp is not None else (self.σD if electric else self.σB)


class Shape:
    def bounds(self):
        raise NotImplementedError

    def contains(self, X, Y, Z) -> np.ndarray:
        raise NotImplementedError


@dataclass
class Box(Shape):
    center: Sequence[float]
    size: Sequence[float]

    def bounds(self):
        c, s = np.asarray(self.center, float), np.asarray(self.size, float)
        return c - s / 2, c + s / 2

    def contains(self, X, Y, Z):
        c, s = self.center, self.size
        return ((np.abs(X - c[0]) <= s[0] / 2) & (np.abs(Y - c[1]) <= s[1] / 2)
                & (np.abs(Z - c[2]) <= s[2] / 2))


@dataclass
class Sphere(Shape):
    center: Sequence[float]
    radius: float

    def bounds(self):
        c = np.asarray(self.center, float)
        return c - self.radius, c + self.radius

    def contains(self, X, Y, Z):
        c = self.center
        return (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 <= self.radius ** 2


@dataclass
class Cylinder(Shape):
    """Circular cylinder with its axis along ``axis`` (0/1/2)."""

    center: Sequence[float]
    radius: float
    height: float
    axis: int = 2

    def bounds(self):
        c = np.asarray(self.center, float)
        half = np.full(3, self.radius)
        half[self.axis] = self.height / 2
        return c - half, c + half

    def contains(self, X, Y, Z):
        P = (X - self.center[0], Y - self.center[1], Z - self.center[2])
        r2 = sum(P[a] ** 2 for a in range(3) if a != self.axis)
        return (r2 <= self.radius ** 2) & (np.abs(P[self.axis]) <= self.height / 2)


@dataclass
class Prism(Shape):
    """Right prism: a convex polygon base (vertices relative to ``center``,
    counter-clockwise) in the x-y plane, extruded over ``height`` along z --
    GeometryPrimitives' Prism with identity axes, as src/GDS.jl:984-997
    builds them.  Boundaries inclusive."""

    center: Sequence[float]
    vertices: Sequence[Tuple[float, float]]
    height: float

    def bounds(self):
        v = np.asarray(self.vertices, float)
        c = np.asarray(self.center, float)
        return (c + np.array([v[:, 0].min(), v[:, 1].min(), -self.height / 2]),
                c + np.array([v[:, 0].max(), v[:, 1].max(), self.height / 2]))

    def contains(self, X, Y, Z):
        c = self.center
        x, y = X - c[0], Y - c[1]
        inside = np.abs(Z - c[2]) <= self.height / 2
        v = [(float(a), float(b)) for a, b in self.vertices]
        for i in range(len(v)):
            ax, ay = v[i]
            bx, by = v[(i + 1) % len(v)]
            # left of (or on) every edge of the counter-clockwise boundary
            inside = inside & ((bx - ax) * (y - ay) - (by - ay) * (x - ax) >= 0)
        return inside


@dataclass
class Object:
    shape: Shape
    material: Material


def rasterize(sim, geometry: List[Object]):
    """Per-component inverse permittivity/permeability arrays.

    Returns ``(eps_inv[3], mu_inv[3])``; an entry is ``None`` when no object
    sets that material property (the scalar 1 is then used, as in
    src/Geometry.jl:652-660).
    """
    R = sim.real
    need_eps = [any(o.material.perm(True, a) is not None for o in geometry) for a in range(3)]
    need_mu = [any(o.material.perm(False, a) is not None for o in geometry) for a in range(3)]
    out_eps, out_mu = [None] * 3, [None] * 3
    for electric, comps, need, out in ((True, (Ex, Ey, Ez), need_eps, out_eps),
                                       (False, (Hx, Hy, Hz), need_mu, out_mu)):
        for a, C in enumerate(comps):
            if not need[a]:
                continue
            c = C()
            cnt = component_voxel_count(sim.Nx, sim.Ny, sim.Nz, c)
            arr = np.ones(cnt, dtype=R)
            gv = full_grid_volume(sim, c)
            xs, ys, zs = gv_coords(sim, gv)
            xs, ys, zs = xs[:cnt[0]], ys[:cnt[1]], zs[:cnt[2]]
            for obj in reversed(geometry):
                p = obj.material.perm(electric, a)
                if p is None:
                    continue
                lo, hi = obj.shape.bounds()
                ix = np.nonzero((xs >= lo[0]) & (xs <= hi[0]))[0]
                iy = np.nonzero((ys >= lo[1]) & (ys <= hi[1]))[0]
                iz = np.nonzero((zs >= lo[2]) & (zs <= hi[2]))[0]
                if len(ix) == 0 or len(iy) == 0 or len(iz) == 0:
                    continue
                X = xs[ix][:, None, None]
                Y = ys[iy][None, :, None]
                Z = zs[iz][None, None, :]
                inside = obj.shape.contains(X, Y, Z)
                sub = arr[np.ix_(ix, iy, iz)]
                sub[inside] = R(1) / R(p)
                arr[np.ix_(ix, iy, iz)] = sub
            out[a] = arr
    return out_eps, out_mu


def rasterize_conductivity(sim, geometry: List[Object], force: bool = False):
    """Per-component conductivity arrays ``(sigma_D[3], sigma_B[3])``
    (src/Geometry.jl:467-476, :233-249, :117-118).

    A component's array exists when some object sets a conductivity for it
    (``needs_conductivities``, :336-347), or for all six when ``force`` (the
    reference allocates every array, zero-filled, when absorbers are given,
    :479-511).  Objects are painted in reverse order with ``get_sigma`` (0 for
    an object that sets none), so the first object wins where shapes overlap.
    """
    R = sim.real
    out_d, out_b = [None] * 3, [None] * 3
    for electric, comps, out in ((True, (Ex, Ey, Ez), out_d), (False, (Hx, Hy, Hz), out_b)):
        for a, C in enumerate(comps):
            if not (force or any(o.material.conductivity(electric, a) is not None for o in geometry)):
                continue
            c = C()
            cnt = component_voxel_count(sim.Nx, sim.Ny, sim.Nz, c)
            arr = np.zeros(cnt, dtype=R)
            gv = full_grid_volume(sim, c)
            xs, ys, zs = gv_coords(sim, gv)
            xs, ys, zs = xs[:cnt[0]], ys[:cnt[1]], zs[:cnt[2]]
            for obj in reversed(geometry):
                v = obj.material.conductivity(electric, a)
                _paint(arr, xs, ys, zs, obj.shape, R(0) if v is None else R(v))
            out[a] = arr
    return out_d, out_b


def _index_range(coords, lo, hi):
    """searchsortedfirst(coords, lo) .. searchsortedlast(coords, hi), 0-based."""
    return np.nonzero((coords >= lo) & (coords <= hi))[0]


def _paint(arr, xs, ys, zs, shape, value):
    lo, hi = shape.bounds()
    ix = _index_range(xs, lo[0], hi[0])
    iy = _index_range(ys, lo[1], hi[1])
    iz = _index_range(zs, lo[2], hi[2])
    if len(ix) == 0 or len(iy) == 0 or len(iz) == 0:
        return
    inside = shape.contains(xs[ix][:, None, None], ys[iy][None, :, None], zs[iz][None, None, :])
    sub = arr[np.ix_(ix, iy, iz)]
    sub[inside] = value
    arr[np.ix_(ix, iy, iz)] = sub


def rasterize_chi3(sim, geometry: List[Object]) -> Optional[np.ndarray]:
    """chi3 array (Nx, Ny, Nz) on the cell centres, or None (src/Geometry.jl:610-636).

    Objects are painted last to first and only where their chi3 is non-zero,
    so an object without chi3 does not clear one painted below it."""
    R = sim.real
    if not any(o.material.chi3 is not None for o in geometry):
        return None
    nz = max(1, sim.Nz)
    arr = np.zeros((sim.Nx, sim.Ny, nz), dtype=R)
    xs, ys, zs = gv_coords(sim, full_grid_volume(sim, Center()))
    xs, ys, zs = xs[:sim.Nx], ys[:sim.Ny], zs[:nz]
    for obj in reversed(geometry):
        c3 = R(0) if obj.material.chi3 is None else R(obj.material.chi3)
        if c3 != 0:
            _paint(arr, xs, ys, zs, obj.shape, c3)
    return arr


@dataclass
class PoleData:
    """One dispersive pole as the time step consumes it: ADE coefficients and
    the oscillator strength per cell, (Nx, Ny, Nz) in the backend type."""

    pole: LorentzianSusceptibility
    coeffs: object
    sigma: np.ndarray


def collect_unique_poles(geometry: List[Object]) -> List[LorentzianSusceptibility]:
    """Poles keyed by (omega_0, gamma), in first-seen order (src/Geometry.jl:1053-1073)."""
    poles, seen = [], set()
    for obj in geometry:
        if not has_susceptibilities(obj.material):
            continue
        for s in obj.material.susceptibilities:
            key = (s.omega_0, s.gamma)
            if key not in seen:
                seen.add(key)
                poles.append(s)
    return poles


def init_polarization(sim, geometry: List[Object], eps_inv: list):
    """Host part of ``init_polarization!`` (src/Geometry.jl:1136-1355) for one chunk.

    Returns ``(poles, eps_inv)``: the per-pole data and the inverse
    permittivity arrays after the chi1 correction eps_inv / (1 + eps_inv chi1).
    The reference applies that correction with the linear index of the
    (Nx, Ny, Nz) chi1 array into each component's array
    (src/Geometry.jl:1334-1342).  With a scalar permittivity the arrays it
    corrects are (Nx, Ny, Nz) too and the linear index is the cell; with
    per-voxel arrays (component voxel count) the linear index lands on other
    cells, and that is reproduced here as the reference executes it.
    """
    R = sim.real
    poles = collect_unique_poles(geometry)
    if not poles:
        return [], eps_inv
    # Ex sample points (src/Geometry.jl:1146-1167)
    gv = full_grid_volume(sim, Ex())
    org = component_origin(sim.cell_center, sim.cell_size_n, sim.delta_n, Ex())
    cnt = component_voxel_count(sim.Nx, sim.Ny, sim.Nz, Ex())
    xs = build_coords(sim.delta_n[0], org[0], gv.start_idx[0], gv.Nx)[:cnt[0]]
    ys = build_coords(sim.delta_n[1], org[1], gv.start_idx[1], gv.Ny)[:cnt[1]]
    if sim.ndims == 3:
        zs = build_coords(sim.delta_n[2], org[2], gv.start_idx[2], gv.Nz)[:cnt[2]]
    else:
        zs = np.array([0.0])   # 2D: one z sample at 0 (src/Geometry.jl:1153-1162)
    sigmas = []
    for pole in poles:
        arr = np.zeros((len(xs), len(ys), len(zs)), dtype=R)
        key = (pole.omega_0, pole.gamma)
        for obj in reversed(geometry):
            if not has_susceptibilities(obj.material):
                continue
            val = None
            for s in obj.material.susceptibilities:
                if (s.omega_0, s.gamma) == key:
                    val = R(s.sigma)
                    break
            if val is not None:
                _paint(arr, xs, ys, zs, obj.shape, val)
        sigmas.append(arr)
    # zero the oscillator strength inside the PML (src/Geometry.jl:1180-1222)
    if sim.boundaries is not None:
        # cell_size and boundaries are stored in the backend type
        # (src/DataStructures.jl:719, :723); the comparisons run in float64
        half = [float(R(v) / R(2)) for v in sim.cell_size_n]
        cc = [float(v) for v in sim.cell_center]
        inpml = np.zeros((len(xs), len(ys), len(zs)), dtype=bool)
        for a, co in enumerate((xs, ys, zs)):
            if len(sim.boundaries) <= a:
                continue
            lo, hi = cc[a] - half[a], cc[a] + half[a]
            left, right = float(R(sim.boundaries[a][0])), float(R(sim.boundaries[a][1]))
            m = np.zeros(len(co), dtype=bool)
            if left > 0:
                m |= co < lo + left
            if right > 0:
                m |= co > hi - right
            shp = [1, 1, 1]
            shp[a] = len(co)
            inpml |= m.reshape(shp)
        for arr in sigmas:
            arr[inpml] = R(0)
    # chi1 = sum_k sigma_k c_k in the backend type (src/Geometry.jl:1224-1237)
    chi1 = np.zeros((len(xs), len(ys), len(zs)), dtype=R)
    coeffs = []
    for pole, arr in zip(poles, sigmas):
        c = compute_ade_coefficients(pole, sim.dt_n)
        coeffs.append(c)
        chi1 = chi1 + arr * R(chi1_coefficient(c))
    N = (sim.Nx, sim.Ny, max(1, sim.Nz))
    sl = (slice(0, N[0]), slice(0, N[1]), slice(0, N[2]))
    out = [PoleData(p, c, np.ascontiguousarray(arr[sl])) for p, c, arr in zip(poles, coeffs, sigmas)]
    chunk_chi1 = chi1[sl]
    if np.max(np.abs(chunk_chi1)) > 0:
        new_eps = []
        flat_chi1 = chunk_chi1.ravel(order="F")
        nz = np.nonzero(flat_chi1 != 0)[0]
        for a in range(3):
            if eps_inv[a] is None:
                # scalar: the corrected array is per cell; cells 1..N hold the
                # corrected values, the component's extra plane is never read
                cnt_a = component_voxel_count(*N, (Ex(), Ey(), Ez())[a])
                arr = np.ones(cnt_a, dtype=R)
                e = np.ones(N, dtype=R)
                f = e.ravel(order="F")
                f[nz] = f[nz] / (R(1) + f[nz] * flat_chi1[nz])
                arr[sl] = f.reshape(N, order="F")
            else:
                arr = np.array(eps_inv[a], dtype=R, order="F", copy=True)
                f = arr.ravel(order="F")
                f[nz] = f[nz] / (R(1) + f[nz] * flat_chi1[nz])
                arr = f.reshape(arr.shape, order="F")
            new_eps.append(arr)
        eps_inv = new_eps
    return out, eps_inv
