"""Yee-grid index math: field components, volumes, grid volumes, weights.

Host-side setup that mirrors the reference's grid conventions so that source
and monitor footprints land on exactly the same grid indices:

* component voxel counts        -- src/utils.jl:26-37
* Yee shifts / component origin -- src/utils.jl:139-171
* point -> grid index snapping  -- src/utils.jl:176-214
* GridVolume from a Volume      -- src/utils.jl:103-113
* interpolation weights         -- src/utils.jl:486-540

Grid index convention (same as the reference): index g = 1..N is an updated
cell, 0 and N+1 are the always-zero boundary planes.  Position of index g of
component c along axis a is ``origin_c[a] + (g - 1) * delta[a]``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np


# --------------------------------------------------------------------------
# Field components (src/DataStructures.jl:12-49)
# --------------------------------------------------------------------------
class Field:
    """Base of every field component; ``axis`` is 0/1/2 for x/y/z."""

    axis: int = -1
    electric: bool = False
    name: str = "Field"

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self).__name__)

    def __repr__(self):
        return f"{self.name}()"


def _mk(name, axis, electric):
    return type(name, (Field,), {"axis": axis, "electric": electric, "name": name})


Ex = _mk("Ex", 0, True)
Ey = _mk("Ey", 1, True)
Ez = _mk("Ez", 2, True)
Hx = _mk("Hx", 0, False)
Hy = _mk("Hy", 1, False)
Hz = _mk("Hz", 2, False)
# displacement / induction components: the Yee positions of E / H
# (src/DataStructures.jl:12-49); DFT monitors may record them
Dx = _mk("Dx", 0, True)
Dy = _mk("Dy", 1, True)
Dz = _mk("Dz", 2, True)
Bx = _mk("Bx", 0, False)
By = _mk("By", 1, False)
Bz = _mk("Bz", 2, False)


class Center(Field):
    axis = -1
    electric = False
    name = "Center"


E_COMPONENTS = (Ex, Ey, Ez)
H_COMPONENTS = (Hx, Hy, Hz)

# C-ABI component codes (include/khronos_hip.h: KH_EX..KH_BZ)
COMPONENT_CODE = {"Ex": 0, "Ey": 1, "Ez": 2, "Hx": 3, "Hy": 4, "Hz": 5,
                  "Dx": 6, "Dy": 7, "Dz": 8, "Bx": 9, "By": 10, "Bz": 11}


def component_code(c: Field) -> int:
    return COMPONENT_CODE[c.name]


def component_from_code(code: int) -> Field:
    return (Ex, Ey, Ez, Hx, Hy, Hz, Dx, Dy, Dz, Bx, By, Bz)[code]()


def is_electric(c: Field) -> bool:
    return bool(c.electric)


def is_magnetic(c: Field) -> bool:
    return not c.electric and c.axis >= 0


# --------------------------------------------------------------------------
# Volumes
# --------------------------------------------------------------------------
@dataclass
class Volume:
    """Axis-aligned box (src/DataStructures.jl:82-85)."""

    center: Sequence[float] = (0.0, 0.0, 0.0)
    size: Sequence[float] = (0.0, 0.0, 0.0)

    def min_corner(self):
        # src/utils.jl:281-283
        return [float(c) - float(s) / 2 for c, s in zip(self.center, self.size)]

    def max_corner(self):
        # src/utils.jl:285-287
        return [float(c) + float(s) / 2 for c, s in zip(self.center, self.size)]


@dataclass
class GridVolume:
    """Integer box of grid indices of one component (src/DataStructures.jl:87-94)."""

    component: Field
    start_idx: List[int]
    end_idx: List[int]
    Nx: int
    Ny: int
    Nz: int

    @property
    def dims(self):
        return (self.Nx, self.Ny, self.Nz)


def component_voxel_count(nx: int, ny: int, nz: int, c: Field):
    """src/utils.jl:26-37 (the staggered extent of each component array)."""
    nz1 = max(1, nz)
    if isinstance(c, Center):
        return [nx, ny, nz1]
    a = c.axis
    if c.electric:
        cnt = [nx + 1, ny + 1, nz1 + 1]
        cnt[a] = [nx, ny, nz1][a]
    else:
        cnt = [nx, ny, nz1]
        cnt[a] = cnt[a] + 1
    return cnt


def yee_shift(delta, c: Field):
    """src/utils.jl:139-148."""
    dx, dy, dz = delta
    if isinstance(c, Center):
        return [0.0, 0.0, 0.0]
    if c.electric:
        s = [-dx / 2.0, -dy / 2.0, -dz / 2.0]
        s[c.axis] = 0.0
    else:
        s = [0.0, 0.0, 0.0]
        s[c.axis] = [-dx / 2.0, -dy / 2.0, -dz / 2.0][c.axis]
    return s


def component_origin(cell_center, cell_size, delta, c: Field):
    """src/utils.jl:160-171: position of grid index 1 of component ``c``."""
    base = [
        float(cell_center[a]) - float(cell_size[a]) / 2 + float(delta[a]) / 2
        for a in range(3)
    ]
    sh = yee_shift(delta, c)
    return [base[a] + sh[a] for a in range(3)]


def grid_idx(cell_center, cell_size, delta, point, c: Field):
    """Fractional 1-based grid index of ``point`` for ``c`` (src/utils.jl:176-200)."""
    sim_vol = Volume(cell_center, cell_size)
    mn = sim_vol.min_corner()
    mx = sim_vol.max_corner()
    sh = yee_shift(delta, c)
    out = []
    for a in range(3):
        d = float(delta[a])
        c2c = d / 2
        lo = mn[a] + c2c + sh[a]
        hi = mx[a] - c2c - sh[a]
        p = min(float(point[a]), hi)
        p = max(p, lo)
        with np.errstate(invalid="ignore", divide="ignore"):
            v = (p - lo) / d + 1 if d != math.inf else float("nan")
        if math.isnan(v):
            v = 0.0
        out.append(v)
    return out


def grid_volume(sim_like, volume: Volume, c: Field) -> GridVolume:
    """GridVolume snapping of a continuous Volume (src/utils.jl:103-113, :205-214)."""
    cc, cs, dl = sim_like.cell_center, sim_like.cell_size_n, sim_like.delta_n
    lo = grid_idx(cc, cs, dl, volume.min_corner(), c)
    hi = grid_idx(cc, cs, dl, volume.max_corner(), c)
    start = [int(math.floor(v)) for v in lo]
    end = [int(math.ceil(v)) for v in hi]
    if sim_like.ndims == 2:
        start[2] = max(1, start[2])
        end[2] = max(1, end[2])
    n = [end[a] - start[a] + 1 for a in range(3)]
    return GridVolume(c, start, end, n[0], n[1], n[2])


def full_grid_volume(sim_like, c: Field) -> GridVolume:
    return grid_volume(sim_like, Volume(sim_like.cell_center, sim_like.cell_size_n), c)


# --------------------------------------------------------------------------
# Interpolation weights (src/utils.jl:486-540), one axis at a time
# --------------------------------------------------------------------------
def weight_1d(p: float, lo: float, hi: float, sz: float, d: float) -> float:
    if (p <= (lo - d)) or (p >= (hi + d)):
        return 0.0
    if sz == 0.0:
        return 1 - min(abs(p - (lo + hi) * 0.5) / d, 1.0)
    if sz < d:
        if (p >= lo) and (p <= hi):
            return 1 - 0.5 * (1.0 - (p - lo) / d) ** 2 - 0.5 * (1.0 - (hi - p) / d) ** 2
        if (p <= lo) and (abs(p - lo) < d):
            if hi < (p + d):
                return 0.5 * (1.0 - (lo - p) / d) ** 2 - 0.5 * (1.0 - (hi - p) / d) ** 2
            return 0.5 * (1.0 - (lo - p) / d) ** 2
        if (p >= hi) and (abs(p - hi) < d):
            if lo > (p - d):
                return 0.5 * (1.0 - (p - hi) / d) ** 2 - 0.5 * (1.0 - (p - lo) / d) ** 2
            return 0.5 * (1.0 - (p - hi) / d) ** 2
        return 1.0
    if (p < lo) and (abs(p - lo) < d):
        return 0.5 * (1.0 - (lo - p) / d) ** 2
    if (p >= lo) and (abs(p - lo) < d):
        return 1 - 0.5 * (1.0 - (p - lo) / d) ** 2
    if (p <= hi) and (abs(p - hi) < d):
        return 1 - 0.5 * (1.0 - (hi - p) / d) ** 2
    if (p > hi) and (abs(p - hi) < d):
        return 0.5 * (1.0 - (p - hi) / d) ** 2
    return 1.0


def interpolation_weight(point, volume: Volume, ndims: int, delta) -> float:
    """Separable product of :func:`weight_1d` over the first ``ndims`` axes."""
    mn, mx = volume.min_corner(), volume.max_corner()
    w = 1.0
    for a in range(ndims):
        wa = weight_1d(float(point[a]), mn[a], mx[a], float(volume.size[a]), float(delta[a]))
        if wa == 0.0:
            return 0.0
        w *= wa
    return w


def gv_coords(sim_like, gv: GridVolume):
    """Physical coordinates of the grid volume's points (src/Geometry.jl:353-366)."""
    org = component_origin(sim_like.cell_center, sim_like.cell_size_n, sim_like.delta_n, gv.component)
    d = sim_like.delta_n
    xs = np.array([org[0] + (i + gv.start_idx[0] - 2) * d[0] for i in range(1, gv.Nx + 1)])
    ys = np.array([org[1] + (i + gv.start_idx[1] - 2) * d[1] for i in range(1, gv.Ny + 1)])
    # z is accumulated from a base point, in that order (src/Geometry.jl:359-364)
    pz_base = org[2] + (gv.start_idx[2] - 2) * d[2]
    zs = []
    for i in range(1, gv.Nz + 1):
        v = pz_base + i * d[2]
        zs.append(float(sim_like.cell_center[2]) if (math.isinf(v) or math.isnan(v)) else v)
    return xs, ys, np.array(zs)


def build_coords(delta: float, origin: float, gv_offset: int, n: int) -> np.ndarray:
    """Uniform-grid coordinates origin + (i + gv_offset - 2) delta, i = 1..n
    (src/Geometry.jl:369-376)."""
    return np.array([origin + (i + gv_offset - 2) * delta for i in range(1, n + 1)])
