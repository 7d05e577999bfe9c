"""Subpixel smoothing of the inverse permittivity at material interfaces
(host setup; src/Geometry.jl:792-955, types src/DataStructures.jl:62-76).

After rasterization, every E-component voxel whose value differs from one of
its six face neighbours (relative tolerance 1e-6, first differing neighbour in
the order +x, -x, +y, -y, +z, -z) is an interface voxel.  For it:

* the nearest shape surface over all objects (first object wins ties) gives
  a surface point r0 and an outward normal n;
* f = fraction of the voxel [p - d/2, p + d/2] on the inner side of the plane
  through r0 normal to n;
* with eps_shape the value on the side of the voxel centre inside the nearest
  shape and eps_bg the neighbour's (or the other way round when the centre is
  outside), <eps> = f eps_shape + (1 - f) eps_bg and
  <1/eps> = f / eps_shape + (1 - f) / eps_bg;
* VolumeAveraging: eps_inv = 1 / <eps>;
  AnisotropicSmoothing (Farjadpour et al. 2006): for component k,
  eps_inv = (1 - n_k^2) <1/eps> + n_k^2 / <eps>.

The reference takes the surface point, the normal and the volume fraction from
the GeometryPrimitives package (``surfpt_nearby``, ``volfrac``, ``level``),
which is not vendored.  They are restated here for the shapes this package
provides: the sphere's radial projection; for boxes and cylinders the
projection onto the nearest face inside, and outside the clamp onto the solid
with the normal along x - clamp(x) (a face normal off a face, a diagonal one
off an edge or corner); the volume fraction is the exact volume of the voxel
cut by the plane.
"""
from __future__ import annotations

import math
from typing import List

import numpy as np

from .grid import Ex, Ey, Ez, full_grid_volume, gv_coords


class SubpixelSmoothing:
    """Smoothing mode of ``Simulation(subpixel_smoothing=...)``."""


class NoSmoothing(SubpixelSmoothing):
    """Point sampling (the default)."""


class VolumeAveraging(SubpixelSmoothing):
    """Isotropic eps_bar = f eps_1 + (1 - f) eps_2."""


class AnisotropicSmoothing(SubpixelSmoothing):
    """Direction-dependent effective eps from the interface normal."""


# ------------------------------------------------------------ geometry kernels
def _normalize(v):
    n = np.linalg.norm(v, axis=1)
    out = np.tile(np.array([0.0, 0.0, 1.0]), (len(v), 1))
    ok = n > 0
    out[ok] = v[ok] / n[ok, None]
    return out


def surfpt_box(P, center, size):
    """Nearest surface point and outward normal of an axis-aligned box, (n, 3) each."""
    c = np.asarray(center, float)
    h = np.asarray(size, float) / 2
    L = P - c
    s = np.where(L >= 0, 1.0, -1.0)
    d = np.abs(L) - h
    inside = np.all(d <= 0, axis=1)
    sp = np.empty_like(P)
    nout = np.zeros_like(P)
    # inside: nearest face
    if np.any(inside):
        k = np.argmax(d[inside], axis=1)
        rows = np.nonzero(inside)[0]
        sp[rows] = P[rows]
        sp[rows, k] = c[k] + s[rows, k] * h[k]
        nout[rows, k] = s[rows, k]
    out = ~inside
    if np.any(out):
        clamp = np.clip(P[out], c - h, c + h)
        sp[out] = clamp
        delta = P[out] - clamp
        nout[out] = _normalize(delta)
    return sp, nout


def surfpt_sphere(P, center, radius):
    c = np.asarray(center, float)
    v = P - c
    n = np.linalg.norm(v, axis=1)
    nout = np.tile(np.array([1.0, 0.0, 0.0]), (len(P), 1))
    ok = n > 0
    nout[ok] = v[ok] / n[ok, None]
    return c + radius * nout, nout


def surfpt_cylinder(P, center, radius, height, axis):
    c = np.asarray(center, float)
    a = np.zeros(3)
    a[axis] = 1.0
    v = P - c
    ell = v[:, axis]
    rho = v.copy()
    rho[:, axis] = 0.0
    rn = np.linalg.norm(rho, axis=1)
    rhat = np.zeros_like(P)
    ok = rn > 0
    rhat[ok] = rho[ok] / rn[ok, None]
    # a point on the axis: any radial direction; take the first axis after `axis`
    rhat[~ok, (axis + 1) % 3] = 1.0
    sgn = np.where(ell >= 0, 1.0, -1.0)
    d_side = rn - radius
    d_cap = np.abs(ell) - height / 2
    n_side = rhat
    n_cap = sgn[:, None] * a[None, :]
    inside = (d_side <= 0) & (d_cap <= 0)
    use_side = inside & (d_side >= d_cap)
    use_cap = inside & ~use_side
    sp = np.empty_like(P)
    nout = np.empty_like(P)
    sp[use_side] = P[use_side] - d_side[use_side, None] * n_side[use_side]
    nout[use_side] = n_side[use_side]
    sp[use_cap] = P[use_cap] - d_cap[use_cap, None] * n_cap[use_cap]
    nout[use_cap] = n_cap[use_cap]
    out = ~inside
    if np.any(out):
        ds = np.maximum(d_side[out], 0.0)
        dc = np.maximum(d_cap[out], 0.0)
        delta = ds[:, None] * n_side[out] + dc[:, None] * n_cap[out]
        sp[out] = P[out] - delta
        nout[out] = _normalize(delta)
    return sp, nout


def surfpt_prism(P, center, vertices, height):
    """Nearest surface point and outward normal of a z-extruded convex
    polygon (vertices about ``center``, counter-clockwise), by the cylinder's
    rule with the polygon in place of the circle: inside, the nearer of the
    closest side (largest edge distance) and the cap; outside, the clamped
    offset from the nearest boundary point of the polygon and the cap."""
    c = np.asarray(center, float)
    v = np.asarray(vertices, float)
    q = P - c
    x, y, ell = q[:, 0], q[:, 1], q[:, 2]
    k = len(v)
    best = np.full(len(P), -np.inf)
    bn = np.zeros((len(P), 2))
    dmin = np.full(len(P), np.inf)
    dn = np.zeros((len(P), 2))
    for i in range(k):
        ax, ay = v[i]
        bx, by = v[(i + 1) % k]
        ex, ey = bx - ax, by - ay
        ln = np.sqrt(ex * ex + ey * ey)
        nx, ny = ey / ln, -ex / ln
        sd = (x - ax) * nx + (y - ay) * ny
        up = sd > best
        best = np.where(up, sd, best)
        bn[up] = (nx, ny)
        t = np.clip(((x - ax) * ex + (y - ay) * ey) / (ln * ln), 0.0, 1.0)
        px, py = x - (ax + t * ex), y - (ay + t * ey)
        d = np.sqrt(px * px + py * py)
        closer = d < dmin
        dmin = np.where(closer, d, dmin)
        dn[closer, 0] = px[closer]
        dn[closer, 1] = py[closer]
    inpoly = best <= 0
    d_side = np.where(inpoly, best, dmin)
    n_side = np.zeros_like(P)
    # a point on the boundary that rounding puts outside (best > 0 with
    # dmin == 0, e.g. on a vertex) takes the best edge's normal
    edge_n = inpoly | ~(dmin > 0)
    n_side[edge_n, :2] = bn[edge_n]
    out_p = ~edge_n
    n_side[out_p, 0] = dn[out_p, 0] / dmin[out_p]
    n_side[out_p, 1] = dn[out_p, 1] / dmin[out_p]
    sgn = np.where(ell >= 0, 1.0, -1.0)
    d_cap = np.abs(ell) - height / 2
    n_cap = np.zeros_like(P)
    n_cap[:, 2] = sgn
    inside = (d_side <= 0) & (d_cap <= 0)
    use_side = inside & (d_side >= d_cap)
    use_cap = inside & ~use_side
    sp = np.empty_like(P)
    nout = np.empty_like(P)
    sp[use_side] = P[use_side] - d_side[use_side, None] * n_side[use_side]
    nout[use_side] = n_side[use_side]
    sp[use_cap] = P[use_cap] - d_cap[use_cap, None] * n_cap[use_cap]
    nout[use_cap] = n_cap[use_cap]
    out = ~inside
    if np.any(out):
        ds = np.maximum(d_side[out], 0.0)
        dc = np.maximum(d_cap[out], 0.0)
        delta = ds[:, None] * n_side[out] + dc[:, None] * n_cap[out]
        sp[out] = P[out] - delta
        nout[out] = _normalize(delta)
    return sp, nout


def surfpt_nearby(P, shape):
    from .geometry import Box, Cylinder, Prism, Sphere
    if isinstance(shape, Prism):
        return surfpt_prism(P, shape.center, shape.vertices, shape.height)
    if isinstance(shape, Box):
        return surfpt_box(P, shape.center, shape.size)
    if isinstance(shape, Sphere):
        return surfpt_sphere(P, shape.center, shape.radius)
    if isinstance(shape, Cylinder):
        return surfpt_cylinder(P, shape.center, shape.radius, shape.height, shape.axis)
    raise NotImplementedError(f"subpixel smoothing for {type(shape).__name__}")


def volfrac(lo, hi, n, r0):
    """Fraction of each box [lo, hi] (rows) with (x - r0) . n <= 0.

    Exact: with the box mapped to [0, L] and the axes flipped so that every
    nonzero n_i > 0, the volume of {n . x <= t} is
    sum_v (-1)^|v| max(0, t - n . (v L))^k / (k! prod n_i) over the corners v
    of the k axes with nonzero n_i (the other axes contribute their length).
    """
    lo = np.atleast_2d(np.asarray(lo, float))
    hi = np.atleast_2d(np.asarray(hi, float))
    n = np.atleast_2d(np.asarray(n, float))
    r0 = np.atleast_2d(np.asarray(r0, float))
    L = hi - lo
    # corner of the box with the lowest n . x; the plane's offset from it
    base = np.where(n >= 0, lo, hi)
    t = np.sum(n * (r0 - base), axis=1)
    an = np.abs(n)
    # an axis whose n_i L_i is negligible against the largest is dropped (its
    # effect on the fraction is below that ratio; keeping it would divide a
    # cancelling corner sum by a tiny product)
    span = an * np.abs(L)
    nz = span > 1e-9 * np.maximum(span.max(axis=1, keepdims=True), 1e-300)
    out = np.where(t >= 0, 1.0, 0.0)
    codes = nz[:, 0] * 1 + nz[:, 1] * 2 + nz[:, 2] * 4
    for code in range(1, 8):
        rows = np.nonzero(codes == code)[0]
        if len(rows) == 0:
            continue
        ks = [a for a in range(3) if code >> a & 1]
        k = len(ks)
        nk = an[np.ix_(rows, ks)]
        Lk = L[np.ix_(rows, ks)]
        acc = np.zeros(len(rows))
        for bits in range(1 << k):
            w = t[rows].copy()
            sign = 1.0
            for j in range(k):
                if bits >> j & 1:
                    w = w - nk[:, j] * Lk[:, j]
                    sign = -sign
            acc += sign * np.where(w > 0, w, 0.0) ** k
        frac = acc / (math.factorial(k) * np.prod(nk, axis=1) * np.prod(Lk, axis=1))
        out[rows] = np.clip(frac, 0.0, 1.0)
    return out


# ------------------------------------------------------------------ smoothing
def apply_subpixel_smoothing(sim, geometry, eps_inv: List) -> List:
    """Smoothed copies of the three E-component inverse-permittivity arrays
    (src/Geometry.jl:805-955); ``None`` entries and 2D simulations are left
    as they are (the reference warns and returns for 2D)."""
    mode = getattr(sim, "subpixel_smoothing", None)
    if mode is None or isinstance(mode, NoSmoothing) or not geometry:
        return eps_inv
    if sim.ndims != 3:
        return eps_inv
    aniso = isinstance(mode, AnisotropicSmoothing)
    R = sim.real
    half = np.array([float(sim.delta_n[0]) / 2, float(sim.delta_n[1]) / 2, float(sim.delta_n[2]) / 2])
    rtol = R(1e-6)
    out = list(eps_inv)
    for k, C in enumerate((Ex, Ey, Ez)):
        arr = eps_inv[k]
        if arr is None:
            continue
        orig = np.asarray(arr, dtype=R)
        nx, ny, nz = orig.shape
        xs, ys, zs = gv_coords(sim, full_grid_volume(sim, C()))
        xs, ys, zs = xs[:nx], ys[:ny], zs[:nz]
        # interface detection: first differing face neighbour, +x -x +y -y +z -z
        nbr = np.full(orig.shape, np.nan, dtype=R)
        found = np.zeros(orig.shape, dtype=bool)
        for ax, step in ((0, 1), (0, -1), (1, 1), (1, -1), (2, 1), (2, -1)):
            sl_dst = [slice(None)] * 3
            sl_src = [slice(None)] * 3
            if step == 1:
                sl_dst[ax], sl_src[ax] = slice(0, -1), slice(1, None)
            else:
                sl_dst[ax], sl_src[ax] = slice(1, None), slice(0, -1)
            c = orig[tuple(sl_dst)]
            b = orig[tuple(sl_src)]
            diff = np.abs(b - c) > rtol * np.maximum(np.abs(b), np.abs(c))
            fd = found[tuple(sl_dst)]
            new = diff & ~fd
            nbr_v = nbr[tuple(sl_dst)]
            nbr_v[new] = b[new]
            nbr[tuple(sl_dst)] = nbr_v
            fd |= new
            found[tuple(sl_dst)] = fd
        idx = np.nonzero(found)
        if len(idx[0]) == 0:
            continue
        P = np.stack([xs[idx[0]], ys[idx[1]], zs[idx[2]]], axis=1)
        eps_c = (R(1) / orig[idx]).astype(np.float64)
        eps_n = (R(1) / nbr[idx]).astype(np.float64)
        # nearest surface over the objects, first object on ties
        best = np.full(len(P), np.inf)
        bsp = P.copy()
        bn = np.tile(np.array([0.0, 0.0, 1.0]), (len(P), 1))
        bidx = np.full(len(P), -1)
        for gi, obj in enumerate(geometry):
            sp, nout = surfpt_nearby(P, obj.shape)
            d2 = np.sum((sp - P) ** 2, axis=1)
            better = d2 < best
            best[better] = d2[better]
            bsp[better] = sp[better]
            bn[better] = nout[better]
            bidx[better] = gi
        ok = bidx >= 0
        nhat = _normalize(bn)
        f = volfrac(P - half, P + half, nhat, bsp)
        inside = np.zeros(len(P), dtype=bool)
        for gi, obj in enumerate(geometry):
            sel = bidx == gi
            if np.any(sel):
                Q = P[sel]
                inside[sel] = obj.shape.contains(Q[:, 0], Q[:, 1], Q[:, 2])
        e_shape = np.where(inside, eps_c, eps_n)
        e_bg = np.where(inside, eps_n, eps_c)
        e_avg = f * e_shape + (1 - f) * e_bg
        e_harm = f / e_shape + (1 - f) / e_bg
        if aniso:
            n2 = nhat[:, k] ** 2
            val = (1 - n2) * e_harm + n2 / e_avg
        else:
            val = 1.0 / e_avg
        res = np.array(orig, dtype=R, copy=True)
        sel = tuple(i[ok] for i in idx)
        res[sel] = val[ok].astype(R)
        out[k] = res
    return out
