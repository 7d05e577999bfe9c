"""PML conductivity profiles (host setup).

Mirrors src/Boundaries.jl:23-72 (``sigma_helper`` / ``compute_sigma``) and the
per-axis assembly of ``init_boundaries`` (src/Boundaries.jl:99-164): quadratic
grading with R = 1e-15, pre-multiplied by dt/2, sampled on a half-cell grid of
2N+1 points.  The time step reads entry ``2*i - 1`` (1-based) for cell ``i``
of every component (src/Kernels/Helpers.jl:277).

The arithmetic follows the reference's type promotion: thicknesses, grid
spacing and dt are backend-precision numbers, so in single precision the
position/length terms are evaluated in float32 and the profile itself in
float64 before the final rounding to the backend type.
"""
from __future__ import annotations

import math

import numpy as np


class BoundaryCondition:
    """Boundary condition of one side of one axis (src/DataStructures.jl:154-162).
    ``zero_pml``: the side's PML thickness is overridden to 0
    (src/Boundaries.jl:100-109)."""

    zero_pml = False

    def __repr__(self):
        return type(self).__name__ + "()"


class PML(BoundaryCondition):
    pass


class Periodic(BoundaryCondition):
    zero_pml = True


class Bloch(BoundaryCondition):
    """Bloch-periodic side with wave number ``k`` (the field at x = L is the
    field at x = 0 times exp(i k L)); any Bloch side makes the fields complex
    (src/Fields.jl:145-159)."""

    zero_pml = True

    def __init__(self, k: float = 0.0):
        self.k = float(k)

    def __repr__(self):
        return f"Bloch(k={self.k})"


class PECBoundary(BoundaryCondition):
    zero_pml = True


class PMCBoundary(BoundaryCondition):
    zero_pml = True


class Absorber:
    """Adiabatic absorber on one side of one axis (src/Boundaries.jl:10-21):
    conductivity sigma_max * (d / L)^sigma_order added to every sigma_D and
    sigma_B array over the outermost ``num_layers`` cells; ``sigma_max = 0``
    picks the value for a reflection of about 1e-6."""

    def __init__(self, num_layers: int = 40, sigma_order: int = 3, sigma_max: float = 0.0):
        self.num_layers = int(num_layers)
        self.sigma_order = int(sigma_order)
        self.sigma_max = float(sigma_max)

    def __repr__(self):
        return f"Absorber(num_layers={self.num_layers}, sigma_order={self.sigma_order}, sigma_max={self.sigma_max})"


def has_absorbers(absorbers) -> bool:
    return absorbers is not None and any(
        side is not None for ax in absorbers if ax is not None for side in ax)


def apply_absorbers(sim, absorbers, arrays) -> None:
    """Add the absorber profiles to the conductivity arrays in place
    (src/Geometry.jl:708-789).  ``absorbers`` is per axis: None or
    ``[left, right]`` with None or an :class:`Absorber` per side; ``arrays``
    are the six sigma_D / sigma_B arrays (None entries skipped).

    Arithmetic as the reference: L = num_layers * delta in the backend type;
    the automatic sigma_max = -(p + 1) ln(1e-6) / (2 L) in float64, rounded to
    the backend type; per layer (1 = the outermost cell)
    sigma = R(sigma_max * ((num_layers - layer + 1) / num_layers)^p), added to
    the whole plane of cells at that index of the array's own extent.
    """
    R = sim.real
    if absorbers is None:
        return
    for axis, ax in enumerate(absorbers):
        if ax is None:
            continue
        for side, spec in enumerate(ax):
            if spec is None:
                continue
            n, p = spec.num_layers, spec.sigma_order
            delta = sim.delta_r[axis]
            L = R(R(n) * delta)
            if spec.sigma_max > 0:
                smax = float(spec.sigma_max)
            else:
                smax = float(R(-(p + 1) * math.log(1e-6) / (2.0 * float(L))))
            for arr in arrays:
                if arr is None:
                    continue
                N = arr.shape[axis]
                for layer in range(1, min(n, N) + 1):
                    d = (n - layer + 1) / n
                    val = R(smax * d ** p)
                    i = (N - layer) if side == 1 else (layer - 1)
                    sl = [slice(None)] * 3
                    sl[axis] = i
                    arr[tuple(sl)] += val


def needs_complex_fields(boundary_conditions) -> bool:
    """Any Bloch side (src/Fields.jl:149-159)."""
    if boundary_conditions is None:
        return False
    return any(isinstance(bc, Bloch) for bcs in boundary_conditions for bc in bcs)


def periodic_wrap(sim):
    """Per axis: (wraps, exp(+i k L), exp(-i k L)) of the single-domain
    wraparound (src/Chunking.jl:1725-1768).  An axis wraps when both sides are
    Periodic or Bloch; k is the first Bloch side's; L is the backend-precision
    cell size; the phases are float64 complex."""
    out = []
    bcs_all = sim.boundary_conditions
    for axis in range(3):
        bcs = bcs_all[axis] if bcs_all is not None and axis < len(bcs_all) else None
        if bcs is None or len(bcs) < 2 or not all(isinstance(b, (Periodic, Bloch)) for b in bcs):
            out.append((False, 1 + 0j, 1 + 0j))
            continue
        k = 0.0
        if isinstance(bcs[0], Bloch):
            k = bcs[0].k
        elif isinstance(bcs[1], Bloch):
            k = bcs[1].k
        L = float(sim.real(sim.cell_size_n[axis]))
        x = k * L
        out.append((True, complex(math.cos(x), math.sin(x)), complex(math.cos(-x), math.sin(-x))))
    return out


def sigma_value(idx: int, n: int, delta, dt, left, right, R=np.float64) -> float:
    """Sample ``idx`` (1-based) of a profile of length ``n``, as float64.

    ``delta``, ``dt``, ``left``, ``right`` are backend-precision scalars of type
    ``R`` (src/Boundaries.jl:23-38, :41-42).
    """
    delta, dt, left, right = R(delta), R(dt), R(left), R(right)
    two = R(2)

    def u0(length):
        # -log(1e-15) / (4 * L * 1 / 3) * (0.5 * dt): the denominator stays in R
        den = ((R(4) * length) * R(1)) / R(3)
        return -math.log(1e-15) / float(den) * (0.5 * float(dt))

    def u(x):
        # x^2 * 0.5 * (sign(x) + 1): x^2 and sign(x)+1 in R, the rest in float64
        sgn = R(0) if x == 0 else R(math.copysign(1.0, float(x)))
        return float(x * x) * 0.5 * float(sgn + R(1))

    total = (R(n) * delta) / two
    pos = (R(idx) * delta) / two
    if pos < left:
        return u0(left) * u((left - pos) / left)
    if (total - pos) < right:
        return u0(right) * u((right - (total - pos)) / right)
    return 0.0


def compute_sigma(n: int, delta, dt, left, right, R) -> np.ndarray:
    """Length-``n`` profile in the backend precision (src/Boundaries.jl:64-72)."""
    s = np.zeros(n, dtype=R)
    if float(left) != 0.0 or float(right) != 0.0:
        for idx in range(1, n + 1):
            s[idx - 1] = R(sigma_value(idx, n, delta, dt, left, right, R))
    return s


def init_boundaries(sim) -> dict:
    """Return ``{"B": [sx, sy, sz], "D": [sx, sy, sz]}``, or ``None`` when the
    simulation has no PML (src/Boundaries.jl:94-164).

    Periodic/Bloch/PEC/PMC axes get zero thickness, except that the z profile
    of the D field keeps the raw ``boundaries`` thickness, as the reference
    does (src/Boundaries.jl:154-162).
    """
    if sim.boundaries is None:
        return None
    R = sim.real
    nb = [[R(v) for v in b] for b in sim.boundaries]
    eff = [list(b) for b in nb]
    if sim.boundary_conditions is not None:
        for axis, bcs in enumerate(sim.boundary_conditions):
            for side, bc in enumerate(bcs):
                if bc is not None and getattr(bc, "zero_pml", False):
                    eff[axis][side] = R(0)
    N = [sim.Nx, sim.Ny, sim.Nz]
    out = {"B": [], "D": []}
    for a in range(3):
        if a == 2 and sim.ndims <= 2:
            out["B"].append(None)
            out["D"].append(None)
            continue
        n = 2 * N[a] + 1
        d = sim.delta_r[a]
        out["B"].append(compute_sigma(n, d, sim.dt_r, eff[a][0], eff[a][1], R))
        zb = nb[a] if a == 2 else eff[a]
        out["D"].append(compute_sigma(n, d, sim.dt_r, zb[0], zb[1], R))
    return out
