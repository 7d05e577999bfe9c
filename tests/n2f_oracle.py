"""ctypes wrappers of the oracle's near-to-far functions (oracle/khronos_oracle.c:
orc_green3d = green3d!, orc_near2far = the surface sum) for the tests.

The oracle is the checker: tests compare the HIP library's kh_near2far with
it, and pin it against closed forms (Maxwell's curl equations, the surface
equivalence theorem, far-field decay).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
_LIBS = {}


def lib(name: str = "liboracle_f64.so"):
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(ORACLE_DIR, "build", name)
    if not os.path.exists(path):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR, os.path.join("build", name)], check=True)
    L = C.CDLL(path)
    P = C.c_void_p
    L.orc_green3d.restype = None
    L.orc_green3d.argtypes = [P, P, C.c_double, C.c_double, C.c_double, P, C.c_int, C.c_double, C.c_double]
    L.orc_near2far.restype = C.c_int
    L.orc_near2far.argtypes = [P, P, P, P, P, C.c_int, C.c_int, C.c_double, P, C.c_double, C.c_double, C.c_double,
                               C.c_int, C.c_int, P, C.c_double, C.c_double, P, C.c_int, P]
    _LIBS[name] = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def green3d(x, x0, c0: int, f0: complex = 1.0, freq: float = 1.0, eps: float = 1.0, mu: float = 1.0,
            L=None) -> np.ndarray:
    """(Ex, Ey, Ez, Hx, Hy, Hz) complex at x of a unit current component c0
    (1-3 J, 4-6 M) with amplitude f0 at x0."""
    L = L or lib()
    eh = np.zeros(12)
    xa = np.ascontiguousarray(x, dtype=np.float64)
    x0a = np.ascontiguousarray(x0, dtype=np.float64)
    L.orc_green3d(_p(eh), _p(xa), float(freq), float(eps), float(mu), _p(x0a), int(c0), float(np.real(f0)),
                  float(np.imag(f0)))
    return eh[0::2] + 1j * eh[1::2]


def near2far(planes, normal_axis: int, ns: float, base, d1: float, d2: float, n1: int, n2: int, freqs, obs,
             eps: float = 1.0, mu: float = 1.0, dA=None, L=None) -> np.ndarray:
    """Oracle surface sum; planes = 4 complex arrays (d0, d1, d2, nf) (e1, e2,
    h1, h2).  Returns (n_obs, 6, nf) complex like kh_near2far."""
    L = L or lib()
    bufs = [np.asfortranarray(np.asarray(p, dtype=np.complex128)) for p in planes]
    nf = bufs[0].shape[3]
    flat = [np.ascontiguousarray(b.reshape(-1, order="F")).view(np.float64) for b in bufs]
    dims = np.array([b.shape[:3] for b in bufs], dtype=np.int32).ravel()
    b = np.ascontiguousarray(base, dtype=np.float64).ravel()
    f = np.ascontiguousarray(freqs, dtype=np.float64).ravel()
    o = np.ascontiguousarray(obs, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(2 * nf * 6 * o.shape[0])
    dA = d1 * d2 if dA is None else dA
    rc = L.orc_near2far(_p(flat[0]), _p(flat[1]), _p(flat[2]), _p(flat[3]), _p(dims), nf, int(normal_axis),
                        float(ns), _p(b), float(d1), float(d2), float(dA), int(n1), int(n2), _p(f), float(eps),
                        float(mu), _p(o), o.shape[0], _p(out))
    assert rc == 0, rc
    z = (out[0::2] + 1j * out[1::2]).reshape(nf, 6, o.shape[0])
    return np.transpose(z, (2, 1, 0))


def dipole_field(points, x0, c0: int, f0: complex = 1.0, freq: float = 1.0, eps: float = 1.0, mu: float = 1.0,
                 L=None) -> np.ndarray:
    """green3d at many points: (n, 6)."""
    return np.array([green3d(x, x0, c0, f0, freq, eps, mu, L=L) for x in np.asarray(points, dtype=np.float64)])
