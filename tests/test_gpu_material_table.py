"""The byte-per-cell material table of the PML kernels.

When the per-voxel inverse permittivity holds at most 256 distinct values the
update reads a byte per cell into a table of them instead of the float array;
the numbers are the same, so the fields must equal the float-array run value
for value.  Past 256 values (and after kh_set_eps_inv pushes it past) the
float arrays serve, still equal to the oracle's bar.
"""
import os
import sys

import numpy as np
import pytest

import khronos_hip as kh
from oracle_engine import OracleEngine

pytestmark = pytest.mark.gpu

COMPS = [kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz]


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _scene(precision, engine=None, smoothing=True):
    geo = [kh.Object(kh.Box([0.13, -0.11, 0.04], [0.5, 0.45, 0.37]), kh.Material(ε=6.0)),
           kh.Object(kh.Box([0.0, 0.0, 0.0], [3.0, 0.3, 0.2]), kh.Material(ε=12.0)),
           kh.Object(kh.Box([0.0, 0.0, 0.0], [10.0, 10.0, 10.0]), kh.Material(ε=2.0))]
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.5), component=kh.Ez(),
                            center=[-0.4, 0.0, 0.0], size=[0.0, 0.4, 0.0])]
    sm = kh.AnisotropicSmoothing() if smoothing else kh.NoSmoothing()
    return kh.Simulation(cell_size=[1.6, 1.2, 1.0], resolution=10, sources=src, geometry=geo,
                         boundaries=[[0.3, 0.3]] * 3, precision=precision, mode="fast", engine=engine,
                         subpixel_smoothing=sm)


def _run(precision, lut, steps=120):
    old = os.environ.get("KH_LUT")
    os.environ["KH_LUT"] = "1" if lut else "0"
    try:
        s = _scene(precision)
        s.prepare()
    finally:
        if old is None:
            os.environ.pop("KH_LUT")
        else:
            os.environ["KH_LUT"] = old
    s.step(steps)
    return s


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_table_equals_float_arrays(precision):
    a = _run(precision, True)
    b = _run(precision, False)
    vals = np.unique(np.concatenate([np.ravel(e) for e in a.spec.eps_inv]))
    _note(f"[table] {precision} {len(vals)} distinct eps^-1 values")
    assert 3 < len(vals) <= 255
    for C in COMPS:
        assert np.array_equal(a.read_field(C()), b.read_field(C())), C.__name__


def test_set_eps_inv_past_256_values_falls_back():
    s = _run("f64", True, steps=0)
    o = _scene("f64", engine=OracleEngine)
    o.prepare()
    rng = np.random.default_rng(0)
    for a in range(3):
        arr = np.array(s.spec.eps_inv[a], copy=True)
        arr[2:-2, 2:-2, 2:-2] = 1.0 / (2.0 + rng.random(arr[2:-2, 2:-2, 2:-2].shape))
        s.engine.set_eps_inv(a, arr)
        o.engine.set_eps_inv(a, arr)
    s.engine.set_graph(8)
    s.step(60)
    o.step(60)
    fa = [s.read_field(C()) for C in COMPS]
    fb = [o.read_field(C()) for C in COMPS]
    nb = np.sqrt(sum(np.sum(b ** 2) for b in fb))
    err = np.sqrt(sum(np.sum((x - y) ** 2) for x, y in zip(fa, fb))) / nb
    assert err < 1e-12, err
