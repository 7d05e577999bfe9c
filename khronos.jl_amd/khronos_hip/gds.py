"""GDSII layout import: polygons of a cell, extruded into prism objects.

The on-disk geometry format feeding rasterisation (src/GDS.jl):

* record stream, UNITS, BOUNDARY/BOX, PATH, SREF/AREF, STRANS/MAG/ANGLE,
  COLROW, PATHTYPE/BGNEXTN/ENDEXTN       -- src/GDS.jl:343-534
* excess-64 base-16 reals                 -- src/GDS.jl:105-131
* PATH outline with miter joins, clamped at 4 half-widths; pathtypes 0/1
  flush, 2 half-width extension, 4 explicit extensions -- src/GDS.jl:147-302
* SREF/AREF transform (magnify, reflect about x, rotate, translate) and
  recursive flattening with layer filters -- src/GDS.jl:308-330, :620-727
* polygon -> Prism objects: convex polygons whole, others by ear clipping
  into triangles, vertices about their centroid, CCW -- src/GDS.jl:745-1002

Coordinates are in the library's user units (database integers times the
first UNITS real).  As in the reference, ``axis`` only places the prism's
centre (axis 3: (cx, cy, cz), 2: (cx, cz, cy), 1: (cz, cx, cy)); the base
polygon stays in the x-y plane and the height along z (its prisms are built
with identity axes, src/GDS.jl:984-997).
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

from .geometry import Object, Prism

HEADER, BGNLIB, LIBNAME, UNITS, ENDLIB, BGNSTR, STRNAME, ENDSTR = 0x00, 0x01, 0x02, 0x03, 0x04, 0x05, 0x06, 0x07
BOUNDARY, PATH, SREF, AREF, TEXT, LAYER, DATATYPE, WIDTH, XY, ENDEL = (0x08, 0x09, 0x0A, 0x0B, 0x0C, 0x0D, 0x0E,
                                                                       0x0F, 0x10, 0x11)
SNAME, COLROW, NODE, TEXTTYPE, STRANS, MAG, ANGLE = 0x12, 0x13, 0x15, 0x16, 0x1A, 0x1B, 0x1C
PATHTYPE, BOX, BOXTYPE, BGNEXTN, ENDEXTN = 0x21, 0x2D, 0x2E, 0x30, 0x31

Point = Tuple[float, float]


@dataclass
class GDSPolygon:
    vertices: List[Point]
    layer: int
    datatype: int


@dataclass
class GDSReference:
    cell_name: str
    origin: Point
    rotation: float = 0.0          # radians
    magnification: float = 1.0
    x_reflection: bool = False
    is_array: bool = False
    columns: int = 1
    rows: int = 1
    col_vec: Point = (0.0, 0.0)
    row_vec: Point = (0.0, 0.0)


@dataclass
class GDSCell:
    name: str
    polygons: List[GDSPolygon] = field(default_factory=list)
    references: List[GDSReference] = field(default_factory=list)


@dataclass
class GDSLibrary:
    name: str
    unit: float        # user unit in meters
    precision: float   # database unit in meters
    cells: Dict[str, GDSCell]


def decode_gdsii_real(b: bytes) -> float:
    """8 big-endian bytes: sign bit, 7-bit excess-64 base-16 exponent,
    56-bit fraction."""
    raw = int.from_bytes(bytes(b[:8]), "big")
    exponent = ((raw >> 56) & 0x7F) - 64
    value = (raw & 0x00FFFFFFFFFFFFFF) / 2.0 ** 56 * 2.0 ** (4 * exponent)
    return -value if (raw >> 63) & 1 else value


def _miter(center, n0, t0, n1, t1, offset):
    """Intersection of the two offset lines at an interior vertex (their
    mean when parallel), the step along the incoming line clamped to
    4 |offset|."""
    p0 = (center[0] + n0[0] * offset, center[1] + n0[1] * offset)
    p1 = (center[0] + n1[0] * offset, center[1] + n1[1] * offset)
    den = t0[0] * t1[1] - t0[1] * t1[0]
    if abs(den) < 1e-15:
        return ((p0[0] + p1[0]) / 2, (p0[1] + p1[1]) / 2)
    u = ((p1[0] - p0[0]) * t1[1] - (p1[1] - p0[1]) * t1[0]) / den
    lim = abs(offset) * 4.0
    u = min(max(u, -lim), lim)
    return (p0[0] + u * t0[0], p0[1] + u * t0[1])


def _unit(dx, dy):
    n = math.sqrt(dx * dx + dy * dy)
    return None if n < 1e-30 else (dx / n, dy / n)


def expand_path(centerline: List[Point], half_width: float, pathtype: int = 0, bgnextn: float = 0.0,
                endextn: float = 0.0) -> List[Point]:
    """Closed outline of a constant-width path: start cap, left edge
    forward, end cap, right edge backward."""
    n = len(centerline)
    if n < 2:
        return []
    tang = []
    for i in range(n):
        a, b = (0, 1) if i == 0 else ((n - 2, n - 1) if i == n - 1 else (i - 1, i + 1))
        t = _unit(centerline[b][0] - centerline[a][0], centerline[b][1] - centerline[a][1])
        tang.append(t if t is not None else (1.0, 0.0))
    right = [(c[0] + t[1] * half_width, c[1] - t[0] * half_width) for c, t in zip(centerline, tang)]
    left = [(c[0] - t[1] * half_width, c[1] + t[0] * half_width) for c, t in zip(centerline, tang)]
    for i in range(1, n - 1):
        t0 = _unit(centerline[i][0] - centerline[i - 1][0], centerline[i][1] - centerline[i - 1][1])
        t1 = _unit(centerline[i + 1][0] - centerline[i][0], centerline[i + 1][1] - centerline[i][1])
        if t0 is None or t1 is None:
            continue
        n0, n1 = (-t0[1], t0[0]), (-t1[1], t1[0])
        right[i] = _miter(centerline[i], n0, t0, n1, t1, -half_width)
        left[i] = _miter(centerline[i], n0, t0, n1, t1, half_width)
    ext0 = half_width if pathtype == 2 else (bgnextn if pathtype == 4 else 0.0)
    ext1 = half_width if pathtype == 2 else (endextn if pathtype == 4 else 0.0)
    ts, te = tang[0], tang[-1]
    out = [(right[0][0] - ts[0] * ext0, right[0][1] - ts[1] * ext0),
           (left[0][0] - ts[0] * ext0, left[0][1] - ts[1] * ext0)]
    out += left[1:n - 1]
    out += [(left[-1][0] + te[0] * ext1, left[-1][1] + te[1] * ext1),
            (right[-1][0] + te[0] * ext1, right[-1][1] + te[1] * ext1)]
    out += right[n - 2:0:-1]
    return out


def transform_gds_point(x: float, y: float, mag: float, x_reflect: bool, rotation: float,
                        origin: Point) -> Point:
    """Magnify, reflect about the x axis, rotate, translate."""
    x, y = x * mag, y * mag
    if x_reflect:
        y = -y
    ca, sa = math.cos(rotation), math.sin(rotation)
    return (x * ca - y * sa + origin[0], x * sa + y * ca + origin[1])


def _i16(p, o=0):
    return struct.unpack(">h", bytes(p[o:o + 2]))[0]


def _i32(p, o=0):
    return struct.unpack(">i", bytes(p[o:o + 4]))[0]


def _string(p):
    return bytes(p).decode("latin-1").rstrip("\x00 ")


def read_gds(source) -> GDSLibrary:
    """Parse a GDSII stream (a path or the file's bytes)."""
    if isinstance(source, (bytes, bytearray, memoryview)):
        data = bytes(source)
    else:
        with open(source, "rb") as f:
            data = f.read()
    lib_name, db_user, db_meter, factor = "", 1.0, 1.0, 1.0
    cells: Dict[str, GDSCell] = {}
    cur: Optional[GDSCell] = None
    e: dict = {"kind": None}
    pos = 0
    while pos + 4 <= len(data):
        rlen = (data[pos] << 8) | data[pos + 1]
        rtype = data[pos + 2]
        if rlen < 4 or pos + rlen > len(data):
            break
        p = data[pos + 4:pos + rlen]
        pos += rlen
        if rtype == LIBNAME:
            lib_name = _string(p)
        elif rtype == UNITS:
            db_user, db_meter = decode_gdsii_real(p[0:8]), decode_gdsii_real(p[8:16])
            factor = db_user
        elif rtype == ENDLIB:
            break
        elif rtype == BGNSTR:
            cur = GDSCell("")
        elif rtype == STRNAME:
            if cur is not None:
                cur.name = _string(p)
        elif rtype == ENDSTR:
            if cur is not None and cur.name:
                cells[cur.name] = cur
            cur = None
        elif rtype in (BOUNDARY, BOX):
            e = {"kind": "boundary", "layer": 0, "datatype": 0, "xy": []}
        elif rtype == PATH:
            e = {"kind": "path", "layer": 0, "datatype": 0, "xy": [], "width": 0.0, "pathtype": 0,
                 "bgnextn": 0.0, "endextn": 0.0}
        elif rtype in (SREF, AREF):
            e = {"kind": "sref" if rtype == SREF else "aref", "sname": "", "xy": [], "reflect": False, "mag": 1.0,
                 "angle": 0.0, "colrow": (1, 1)}
        elif rtype in (TEXT, NODE):
            e = {"kind": "other"}
        elif rtype == LAYER:
            e["layer"] = _i16(p)
        elif rtype in (DATATYPE, BOXTYPE, TEXTTYPE):
            e["datatype"] = _i16(p)
        elif rtype == WIDTH:
            e["width"] = factor * abs(float(_i32(p)))
        elif rtype == XY:
            e["xy"] = [(factor * float(_i32(p, 8 * i)), factor * float(_i32(p, 8 * i + 4)))
                       for i in range(len(p) // 8)]
        elif rtype == PATHTYPE:
            e["pathtype"] = _i16(p)
        elif rtype == BGNEXTN:
            e["bgnextn"] = factor * float(_i32(p))
        elif rtype == ENDEXTN:
            e["endextn"] = factor * float(_i32(p))
        elif rtype == SNAME:
            e["sname"] = _string(p)
        elif rtype == COLROW:
            e["colrow"] = (_i16(p, 0), _i16(p, 2))
        elif rtype == STRANS:
            e["reflect"] = ((p[0] << 8 | p[1]) >> 15) & 1 == 1
        elif rtype == MAG:
            e["mag"] = decode_gdsii_real(p[0:8])
        elif rtype == ANGLE:
            e["angle"] = math.radians(decode_gdsii_real(p[0:8]))
        elif rtype == ENDEL:
            if cur is not None:
                _finalize(e, cur)
            e = {"kind": None}
    # the reference stores db_in_user * db_in_meters / db_in_user here, i.e.
    # the database unit in meters (src/GDS.jl:530-531); kept as it is
    return GDSLibrary(lib_name, db_user * db_meter / db_user if db_user else db_meter, db_meter, cells)


def _finalize(e, cell: GDSCell):
    kind, pts = e.get("kind"), e.get("xy", [])
    if kind == "boundary" and len(pts) >= 3:
        v = list(pts)
        if abs(v[0][0] - v[-1][0]) < 1e-15 and abs(v[0][1] - v[-1][1]) < 1e-15:
            v.pop()
        cell.polygons.append(GDSPolygon(v, e["layer"], e["datatype"]))
    elif kind == "path" and len(pts) >= 2 and e["width"] > 0.0:
        out = expand_path(pts, e["width"] / 2.0, e["pathtype"], e["bgnextn"], e["endextn"])
        if len(out) >= 3:
            cell.polygons.append(GDSPolygon(out, e["layer"], e["datatype"]))
    elif kind == "sref" and e["sname"] and pts:
        cell.references.append(GDSReference(e["sname"], pts[0], e["angle"], e["mag"], e["reflect"]))
    elif kind == "aref" and e["sname"] and len(pts) >= 3:
        o = pts[0]
        cols, rows = e["colrow"]
        cv = ((pts[1][0] - o[0]) / max(cols, 1), (pts[1][1] - o[1]) / max(cols, 1))
        rv = ((pts[2][0] - o[0]) / max(rows, 1), (pts[2][1] - o[1]) / max(rows, 1))
        cell.references.append(GDSReference(e["sname"], o, e["angle"], e["mag"], e["reflect"], True, cols, rows,
                                            cv, rv))


def get_layers(cell: GDSCell):
    return {(p.layer, p.datatype) for p in cell.polygons}


def get_cell_names(lib: GDSLibrary):
    return list(lib.cells.keys())


def _layer_set(layers):
    if isinstance(layers, tuple) and len(layers) == 2 and all(isinstance(v, int) for v in layers):
        return {layers}
    return set(tuple(v) for v in layers)


def flatten_cell(lib: GDSLibrary, cell_name: str, layer: Optional[int] = None, datatype: int = 0,
                 layers=None) -> List[GDSPolygon]:
    """All polygons of a cell with its references resolved (transforms
    composed), optionally filtered to (layer, datatype) pairs."""
    out: List[GDSPolygon] = []
    _flatten(out, lib, cell_name, 1.0, False, 0.0, (0.0, 0.0), 0)
    if layer is not None:
        layers = {(layer, datatype)}
    if layers is not None:
        keep = _layer_set(layers)
        out = [p for p in out if (p.layer, p.datatype) in keep]
    return out


def _flatten(out, lib, name, mag, refl, rot, origin, depth):
    if depth > 128:
        raise RuntimeError("GDS cell hierarchy exceeds 128 levels (circular reference?)")
    cell = lib.cells.get(name)
    if cell is None:
        return
    for poly in cell.polygons:
        out.append(GDSPolygon([transform_gds_point(x, y, mag, refl, rot, origin) for x, y in poly.vertices],
                              poly.layer, poly.datatype))
    for ref in cell.references:
        new_mag = mag * ref.magnification
        new_refl = refl != ref.x_reflection
        new_rot = rot + (-ref.rotation if refl else ref.rotation)
        if ref.is_array:
            for ci in range(ref.columns):
                for ri in range(ref.rows):
                    ox = ci * ref.col_vec[0] + ri * ref.row_vec[0] + ref.origin[0]
                    oy = ci * ref.col_vec[1] + ri * ref.row_vec[1] + ref.origin[1]
                    _flatten(out, lib, ref.cell_name, new_mag, new_refl, new_rot,
                             transform_gds_point(ox, oy, mag, refl, rot, origin), depth + 1)
        else:
            _flatten(out, lib, ref.cell_name, new_mag, new_refl, new_rot,
                     transform_gds_point(ref.origin[0], ref.origin[1], mag, refl, rot, origin), depth + 1)


def signed_area(v: List[Point]) -> float:
    n = len(v)
    return sum(v[i][0] * v[(i + 1) % n][1] - v[(i + 1) % n][0] * v[i][1] for i in range(n)) / 2.0


def is_convex(v: List[Point]) -> bool:
    """Every non-degenerate turn has the same sign."""
    n = len(v)
    if n < 3:
        return False
    sign = 0
    for i in range(n):
        a, b, c = v[i], v[(i + 1) % n], v[(i + 2) % n]
        cr = (b[0] - a[0]) * (c[1] - b[1]) - (b[1] - a[1]) * (c[0] - b[0])
        if abs(cr) > 1e-15:
            s = 1 if cr > 0 else -1
            if sign == 0:
                sign = s
            elif s != sign:
                return False
    return True


def _in_triangle(p, a, b, c):
    d1 = (p[0] - b[0]) * (a[1] - b[1]) - (a[0] - b[0]) * (p[1] - b[1])
    d2 = (p[0] - c[0]) * (b[1] - c[1]) - (b[0] - c[0]) * (p[1] - c[1])
    d3 = (p[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (p[1] - a[1])
    neg = d1 < 0 or d2 < 0 or d3 < 0
    pos = d1 > 0 or d2 > 0 or d3 > 0
    return not (neg and pos)


def ear_clip_triangulate(vertices: List[Point]) -> List[List[Point]]:
    """Triangles of a simple polygon: repeatedly cut the first convex vertex
    (CCW order) whose triangle holds no other vertex; if none qualifies, the
    first three vertices are cut (degenerate input); at most n^2 passes."""
    n = len(vertices)
    tris: List[List[Point]] = []
    if n < 3:
        return tris
    idx = list(range(n))
    if signed_area(vertices) < 0:
        idx.reverse()
    it = 0
    while len(idx) > 3:
        it += 1
        if it > n * n:
            break
        m = len(idx)
        cut = False
        for i in range(m):
            ip, inx = (i - 1) % m, (i + 1) % m
            a, b, c = vertices[idx[ip]], vertices[idx[i]], vertices[idx[inx]]
            if (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0]) <= 1e-15:
                continue
            if any(_in_triangle(vertices[idx[j]], a, b, c) for j in range(m) if j not in (ip, i, inx)):
                continue
            tris.append([a, b, c])
            del idx[i]
            cut = True
            break
        if not cut:
            tris.append([vertices[idx[0]], vertices[idx[1]], vertices[idx[2]]])
            del idx[1]
    if len(idx) == 3:
        tris.append([vertices[idx[0]], vertices[idx[1]], vertices[idx[2]]])
    return tris


def _make_prism(verts: List[Point], z_min: float, z_max: float, material, axis: int) -> Optional[Object]:
    k = len(verts)
    if k < 3:
        return None
    area = signed_area(verts)
    if abs(area) < 1e-20:
        return None
    cx = sum(v[0] for v in verts) / k
    cy = sum(v[1] for v in verts) / k
    cz = (z_min + z_max) / 2.0
    ordered = verts if area > 0 else list(reversed(verts))
    rel = [(x - cx, y - cy) for x, y in ordered]
    if axis == 3:
        center = (cx, cy, cz)
    elif axis == 2:
        center = (cx, cz, cy)
    elif axis == 1:
        center = (cz, cx, cy)
    else:
        raise ValueError("axis must be 1, 2 or 3")
    return Object(Prism(center, rel, float(z_max - z_min)), material)


def polygon_to_objects(vertices: List[Point], z_min: float, z_max: float, material, axis: int = 3) -> List[Object]:
    if len(vertices) < 3 or z_max - z_min <= 0:
        return []
    pieces = [vertices] if is_convex(vertices) else ear_clip_triangulate(vertices)
    return [o for o in (_make_prism(p, z_min, z_max, material, axis) for p in pieces) if o is not None]


def gds_polygons_to_objects(polygons: List[GDSPolygon], z_min: float, z_max: float, material,
                            axis: int = 3) -> List[Object]:
    out: List[Object] = []
    for p in polygons:
        out += polygon_to_objects(p.vertices, z_min, z_max, material, axis)
    return out


def gds_to_objects(lib: GDSLibrary, cell_name: str, layer_map: dict, axis: int = 3) -> List[Object]:
    """layer_map: (layer, datatype) -> (z_min, z_max, material)."""
    out: List[Object] = []
    for (layer, datatype), (z0, z1, mat) in layer_map.items():
        out += gds_polygons_to_objects(flatten_cell(lib, cell_name, layer, datatype), z0, z1, mat, axis)
    return out
