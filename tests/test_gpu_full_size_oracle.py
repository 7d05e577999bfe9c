"""Full-size configurations against the CPU oracle itself (not against the
device's EXACT mode):

* strip waveguide (384 x 192 x 160, 12-cell PML, anisotropic smoothing,
  4000 steps): both flux spectra against the FP64 oracle's, committed as
  tests/golden/strip_waveguide_flux_fp64_oracle.json (written by
  tests/golden/make_strip_waveguide_flux.py).  FP32 fields on the FP64 setup
  with FP64 DFT accumulators: within 1e-4 of each spectrum's peak (the
  precision form's bar); FP64 FAST: within 1e-9.
* strip waveguide, 300 steps, FP32: fields and DFT accumulators against a
  live FP32 oracle run, rel-L2 within 1e-5 (fields) and 1e-4 (DFTs); the
  device's EXACT mode equal to the oracle value for value (DFTs within
  1e-13).
* metasurface (256 x 256 x 384, Bloch x/y, 16-cell PML z, complex FP32, 2
  Lorentz poles + chi3 cylinders, 100-frequency flux plane), 80 steps:
  fields and DFTs against a live FP32 oracle run at the same tolerances, and
  EXACT equal to it value for value.  In
  80 steps the pulse has not reached the cylinders; their ADE and chi3
  updates are compared with the oracle on the 2 x 2 lattice
  (test_gpu_metasurface.py).
"""
import json
import math
import os
import sys
import time

import numpy as np
import pytest

import khronos_hip as kh
from khronos_hip.workloads import metasurface, strip_waveguide
from oracle_engine import OracleEngine

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
COMPS = [kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz]


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(fa, fb):
    nb = math.sqrt(sum(float(np.sum(np.abs(np.asarray(b, np.complex128)) ** 2)) for b in fb))
    d = math.sqrt(sum(float(np.sum(np.abs(np.asarray(a, np.complex128) - b) ** 2)) for a, b in zip(fa, fb)))
    return d / nb


def _fluxes(precision, setup=None, steps=4000):
    s = strip_waveguide(precision=precision, mode="fast", setup_precision=setup)
    s.prepare()
    s.step(steps)
    s.engine.sync()
    out = np.stack([kh.get_flux(m) for m in s.monitors])
    s.engine.close()
    return out


def test_strip_waveguide_flux_against_fp64_oracle_fixture():
    fix = json.load(open(os.path.join(HERE, "golden", "strip_waveguide_flux_fp64_oracle.json")))
    ref = np.asarray(fix["flux"])
    peak = np.max(np.abs(ref), axis=1, keepdims=True)
    assert fix["steps"] == 4000 and ref.shape == (2, 50) and np.all(peak > 0)
    for precision, setup, bar in (("f64", None, 1e-9), ("f32", "f64", 1e-4)):
        _note(f"[full-size] strip waveguide 4000 steps {precision} (setup {setup or precision}) vs FP64 oracle flux")
        got = _fluxes(precision, setup)
        assert np.all(np.isfinite(got))
        err = float(np.max(np.abs(got - ref) / peak))
        _note(f"[full-size] max |flux - oracle FP64| / peak = {err:.3e} (bar {bar:g})")
        assert err <= bar, (precision, err)


def _exact_equals_oracle(x, o, tag):
    """The device's EXACT mode against the oracle at full size: every field
    value equal, DFT accumulators within 1e-13 (relative L2)."""
    for C in COMPS:
        assert np.array_equal(x.read_field(C()), o.read_field(C())), (tag, C.__name__)
    for i in range(len(o.spec.dfts)):
        a, b = x.engine.read_dft(i), o.engine.read_dft(i)
        nb = float(np.linalg.norm(b))
        assert (not np.any(a)) if nb == 0 else float(np.linalg.norm(a - b) / nb) <= 1e-13, (tag, i)
    _note(f"[full-size] {tag}: device EXACT equals the oracle value for value")


def test_strip_waveguide_fp32_fields_against_live_oracle():
    steps = 300
    g = strip_waveguide(precision="f32", mode="fast")
    x = strip_waveguide(precision="f32", mode="exact")
    o = strip_waveguide(precision="f32", mode="exact", engine=OracleEngine)
    _note(f"[full-size] strip waveguide FP32 {steps} steps: device FAST and EXACT, CPU oracle")
    g.step(steps)
    x.step(steps)
    t0 = time.perf_counter()
    o.step(steps)
    _note(f"[full-size] oracle {steps} steps in {time.perf_counter() - t0:.1f} s")
    _exact_equals_oracle(x, o, "strip waveguide")
    x.engine.close()
    fa = [g.read_field(C()) for C in COMPS]
    fb = [o.read_field(C()) for C in COMPS]
    err = _rel(fa, fb)
    _note(f"[full-size] fields rel-L2 {err:.3e}")
    assert np.any(fb[2] != 0) and err <= 1e-5
    seen = 0
    for i in range(len(o.spec.dfts)):
        a, b = g.engine.read_dft(i), o.engine.read_dft(i)
        nb = float(np.linalg.norm(b))
        if nb == 0:  # the pulse has not reached the far plane: zero on both sides
            assert not np.any(a), i
            continue
        seen += 1
        d = float(np.linalg.norm(a - b) / nb)
        assert d <= 1e-4, (i, d)
    assert seen >= 4  # the near plane's four tangential DFTs
    g.engine.close()
    o.engine.close()


def test_metasurface_full_size_against_live_oracle():
    steps = 80
    g = metasurface(mode="fast")
    x = metasurface(mode="exact")
    o = metasurface(mode="exact", engine=OracleEngine)
    _note(f"[full-size] metasurface {g.Nx}x{g.Ny}x{g.Nz} complex FP32 {steps} steps: device FAST and EXACT, "
          "CPU oracle")
    g.step(steps)
    x.step(steps)
    t0 = time.perf_counter()
    o.step(steps)
    _note(f"[full-size] oracle {steps} steps in {time.perf_counter() - t0:.1f} s")
    _exact_equals_oracle(x, o, "metasurface")
    x.engine.close()
    fa = [g.read_field(C()) for C in COMPS]
    fb = [o.read_field(C()) for C in COMPS]
    err = _rel(fa, fb)
    _note(f"[full-size] fields rel-L2 {err:.3e}")
    assert any(np.any(np.asarray(b).imag != 0) for b in fb) and err <= 1e-5
    for i in range(len(o.spec.dfts)):
        a, b = g.engine.read_dft(i), o.engine.read_dft(i)
        nb = float(np.linalg.norm(b))
        if nb == 0:  # the pulse has not reached the flux plane: zero on both sides
            assert not np.any(a), i
            continue
        d = float(np.linalg.norm(a - b) / nb)
        assert d <= 1e-4, (i, d)
    g.engine.close()
    o.engine.close()
