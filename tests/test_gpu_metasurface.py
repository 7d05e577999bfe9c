"""The periodic dispersive metasurface (Bloch x/y, PML z, 8 x 8
Lorentz + chi3 cylinders, plane-wave pulse, complex FP32 fields, FP64 DFT at
100 frequencies).

Small lattice (2 x 2 cells, 64 x 64 x 48): EXACT mode equals the oracle value
for value, FAST within rel-L2 1e-5 and DFTs within 1e-4.  Full size
(256 x 256 x 384): FAST within rel-L2 1e-5 of EXACT after 200 steps (EXACT is
the oracle's operation order), DFTs within 1e-4; the FAST step rate is
reported.
"""
import math
import sys
import time

import numpy as np
import pytest

import khronos_hip as kh
from khronos_hip.workloads import metasurface
from oracle_engine import OracleEngine

pytestmark = pytest.mark.gpu

COMPS = [kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz]


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(fa, fb):
    nb = math.sqrt(sum(float(np.sum(np.abs(b.astype(np.complex128)) ** 2)) for b in fb))
    d = math.sqrt(sum(float(np.sum(np.abs(a.astype(np.complex128) - b) ** 2)) for a, b in zip(fa, fb)))
    return d / nb


def test_small_lattice_against_oracle():
    kw = dict(nz=48, lattice=2, n_freq=10)
    o = metasurface(engine=OracleEngine, mode="exact", **kw)
    e = metasurface(mode="exact", **kw)
    f = metasurface(mode="fast", **kw)
    _note(f"[metasurface] small lattice {e.Nx}x{e.Ny}x{e.Nz}, 120 steps")
    for s in (o, e, f):
        s.step(120)
    fo = [o.read_field(C()) for C in COMPS]
    for C, a, b in zip(COMPS, [e.read_field(C()) for C in COMPS], fo):
        assert np.array_equal(a, b), C.__name__
    assert np.any(fo[0].imag != 0)
    assert _rel([f.read_field(C()) for C in COMPS], fo) < 1e-5
    for i in range(len(o.spec.dfts)):
        b = o.engine.read_dft(i)
        assert np.linalg.norm(e.engine.read_dft(i) - b) <= 1e-13 * np.linalg.norm(b)
        assert np.linalg.norm(f.engine.read_dft(i) - b) <= 1e-4 * np.linalg.norm(b)


def test_full_size_fast_vs_exact():
    steps = 200
    f = metasurface(mode="fast")
    e = metasurface(mode="exact")
    _note(f"[metasurface] full size {f.Nx}x{f.Ny}x{f.Nz}, {steps} steps FAST and EXACT")
    f.prepare()
    f.step(8)
    f.engine.sync()
    t0 = time.perf_counter()
    f.step(steps - 8)
    f.engine.sync()
    dt = time.perf_counter() - t0
    _note(f"[metasurface] FAST {f.Nx * f.Ny * f.Nz * (steps - 8) / dt / 1e6:.0f} Mcell-updates/s "
          f"({dt / (steps - 8) * 1e3:.3f} ms/step, complex FP32, 2 poles + chi3, 100-frequency flux)")
    e.step(steps)
    fa = [f.read_field(C()) for C in COMPS]
    fb = [e.read_field(C()) for C in COMPS]
    assert all(np.all(np.isfinite(x)) for x in fa)
    err = _rel(fa, fb)
    _note(f"[metasurface] FAST vs EXACT rel-L2 {err:.3e}")
    assert err < 1e-5
    for i in range(len(f.spec.dfts)):
        b = e.engine.read_dft(i)
        assert np.linalg.norm(f.engine.read_dft(i) - b) <= 1e-4 * np.linalg.norm(b)
