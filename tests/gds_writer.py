"""Minimal GDSII writer for the reader's tests (records of the GDSII stream
format: 2-byte length, record type, data type, big-endian payload)."""
import math
import struct


def real8(v: float) -> bytes:
    """Excess-64 base-16 real of v (exact for the values the tests use)."""
    if v == 0:
        return bytes(8)
    sign = 0x80 if v < 0 else 0
    v = abs(v)
    e = 0
    while v >= 1:
        v /= 16.0
        e += 1
    while v < 1 / 16.0:
        v *= 16.0
        e -= 1
    mant = int(round(v * 2 ** 56))
    if mant >= 2 ** 56:
        mant //= 16
        e += 1
    return bytes([sign | (e + 64)]) + mant.to_bytes(7, "big")


def rec(rtype: int, dtype: int, payload: bytes = b"") -> bytes:
    if len(payload) % 2:
        payload += b"\x00"
    return struct.pack(">HBB", 4 + len(payload), rtype, dtype) + payload


def i16(*v):
    return struct.pack(">" + "h" * len(v), *v)


def i32(*v):
    return struct.pack(">" + "i" * len(v), *v)


def string(s):
    return s.encode("latin-1")


class Writer:
    def __init__(self, name="library", user=1e-3, meter=1e-9):
        self.user = user
        self.out = [rec(0x00, 0x02, i16(600)), rec(0x01, 0x02, i16(*([0] * 12))), rec(0x02, 0x06, string(name)),
                    rec(0x03, 0x05, real8(user) + real8(meter))]

    def db(self, v):
        return int(round(v / self.user))

    def xy(self, pts):
        flat = []
        for x, y in pts:
            flat += [self.db(x), self.db(y)]
        return rec(0x10, 0x03, i32(*flat))

    def begin(self, name):
        self.out += [rec(0x05, 0x02, i16(*([0] * 12))), rec(0x06, 0x06, string(name))]

    def end(self):
        self.out.append(rec(0x07, 0x00))

    def boundary(self, pts, layer, datatype=0):
        closed = list(pts) + [pts[0]]
        self.out += [rec(0x08, 0x00), rec(0x0D, 0x02, i16(layer)), rec(0x0E, 0x02, i16(datatype)), self.xy(closed),
                     rec(0x11, 0x00)]

    def path(self, pts, width, layer, pathtype=0, datatype=0):
        self.out += [rec(0x09, 0x00), rec(0x0D, 0x02, i16(layer)), rec(0x0E, 0x02, i16(datatype)),
                     rec(0x21, 0x02, i16(pathtype)), rec(0x0F, 0x03, i32(self.db(width))), self.xy(pts),
                     rec(0x11, 0x00)]

    def sref(self, name, origin, angle=0.0, mag=1.0, reflect=False):
        r = [rec(0x0A, 0x00), rec(0x12, 0x06, string(name))]
        if reflect or angle or mag != 1.0:
            r.append(rec(0x1A, 0x01, i16(-32768 if reflect else 0)))
            if mag != 1.0:
                r.append(rec(0x1B, 0x05, real8(mag)))
            if angle:
                r.append(rec(0x1C, 0x05, real8(angle)))
        r += [self.xy([origin]), rec(0x11, 0x00)]
        self.out += r

    def aref(self, name, origin, cols, rows, col_pitch, row_pitch):
        o = origin
        pts = [o, (o[0] + cols * col_pitch[0], o[1] + cols * col_pitch[1]),
               (o[0] + rows * row_pitch[0], o[1] + rows * row_pitch[1])]
        self.out += [rec(0x0B, 0x00), rec(0x12, 0x06, string(name)), rec(0x13, 0x02, i16(cols, rows)),
                     self.xy(pts), rec(0x11, 0x00)]

    def bytes(self):
        return b"".join(self.out) + rec(0x04, 0x00)
