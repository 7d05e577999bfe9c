"""Materials, shapes and per-voxel permittivity rasterization (host setup).

Mirrors src/Geometry.jl:449-663: every component's inverse-permittivity array
(component voxel count, no ghosts) starts at 1 and objects are painted in
reverse order, so the first object in ``geometry`` wins where shapes overlap
(src/Geometry.jl:233-249).  A voxel belongs to a shape when its Yee-staggered
sample point lies inside it (boundaries inclusive).

The reference tests membership with the GeometryPrimitives package (not
vendored); the box / sphere / cylinder membership tests here are the closed
forms of those shapes.  Subpixel smoothing is not applied here.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from .grid import Ex, Ey, Ez, Hx, Hy, Hz, full_grid_volume, gv_coords, component_voxel_count


@dataclass
class Material:
    """Subset of src/DataStructures.jl:244-273 used on the time-step path."""

    ε: Optional[float] = None
    εx: Optional[float] = None
    εy: Optional[float] = None
    εz: Optional[float] = None
    μ: Optional[float] = None
    chi3: Optional[float] = None

    def perm(self, electric: bool, axis: int) -> Optional[float]:
        if electric:
            comp = (self.εx, self.εy, self.εz)[axis]
            return comp if comp is not None else self.ε
        return self.μ


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
