"""Edge shapes: odd and one-cell extents, sources and monitors on the first
and last cells, against the oracle (EXACT value for value, FAST within
1e-12 in float64)."""
import math
import sys

import numpy as np
import pytest

import khronos_hip as kh
from oracle_engine import OracleEngine

pytestmark = pytest.mark.gpu

COMPS = [kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz]


def _sim(n, pml, mode, engine=None, res=10):
    cell = [v / res for v in n]
    lo = [-c / 2 + 0.5 / res for c in cell]   # first cell centre
    hi = [c / 2 - 0.5 / res for c in cell]    # last cell centre
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.5), component=kh.Ez(),
                            center=[lo[0], lo[1], 0.0], size=[0.0, 0.0, 0.0]),
           kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.2, fwidth=0.5), component=kh.Hx(),
                            center=[hi[0], hi[1], hi[2]], size=[0.0, 0.0, 0.0])]
    mons = [kh.DFTMonitor(component=kh.Ez(), center=[0.0, 0.0, 0.0], size=cell, frequencies=[1.0, 1.1]),
            kh.DFTMonitor(component=kh.Hy(), center=[hi[0], 0.0, 0.0], size=[0.0, cell[1], cell[2]],
                          frequencies=[0.9])]
    geo = [kh.Object(kh.Box([0.0, 0.0, 0.0], [cell[0] / 3, cell[1] / 2, cell[2]]), kh.Material(ε=3.0))]
    return kh.Simulation(cell_size=cell, resolution=res, sources=src, monitors=mons, geometry=geo,
                         boundaries=pml, precision="f64", mode=mode, engine=engine)


CASES = [((11, 7, 5), [[0.2, 0.2]] * 3), ((9, 6, 1), [[0.2, 0.2], [0.2, 0.2], [0.0, 0.0]]),
         ((1, 5, 8), [[0.0, 0.0], [0.2, 0.2], [0.3, 0.3]]), ((131, 3, 4), [[0.5, 0.3], [0.0, 0.0], [0.1, 0.1]])]


@pytest.mark.parametrize("n,pml", CASES)
@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_edge_shapes_match_oracle(n, pml, mode):
    g = _sim(n, pml, mode)
    o = _sim(n, pml, "exact", engine=OracleEngine)
    print(f"[edge] {mode} grid {g.Nx}x{g.Ny}x{g.Nz} pml {pml}", flush=True)
    g.step(45)
    o.step(45)
    fa = [g.read_field(C()) for C in COMPS]
    fb = [o.read_field(C()) for C in COMPS]
    nb = math.sqrt(sum(float(np.sum(b * b)) for b in fb))
    assert nb > 0
    if mode == "exact":
        for C, a, b in zip(COMPS, fa, fb):
            assert np.array_equal(a, b), C.__name__
    else:
        tot = math.sqrt(sum(float(np.sum((a - b) ** 2)) for a, b in zip(fa, fb))) / nb
        assert tot < 1e-12, tot
    for i in range(len(g.spec.dfts)):
        a, b = g.engine.read_dft(i), o.engine.read_dft(i)
        assert np.linalg.norm(a - b) <= 1e-10 * max(np.linalg.norm(b), 1e-300)


def _overlap_sim(pml, mode, engine=None):
    """Three Ez sources on overlapping cells (a plane, a line in it, a point on
    both) plus an Hy line: every overlapped cell takes each source in turn."""
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.5), component=kh.Ez(),
                            center=[0.0, 0.0, 0.05], size=[0.6, 0.4, 0.0], amplitude=1.0),
           kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.3, fwidth=0.7), component=kh.Ez(),
                            center=[0.1, 0.0, 0.05], size=[0.0, 0.6, 0.0], amplitude=-2.0),
           kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=0.9, fwidth=0.4), component=kh.Ez(),
                            center=[0.1, 0.1, 0.05], size=[0.0, 0.0, 0.0], amplitude=3.0),
           kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.1, fwidth=0.5), component=kh.Hy(),
                            center=[-0.2, 0.0, -0.1], size=[0.5, 0.0, 0.0], amplitude=1.5)]
    geo = [kh.Object(kh.Box([0.0, 0.0, 0.0], [0.4, 0.4, 0.4]), kh.Material(ε=2.0))]
    mons = [kh.DFTMonitor(component=kh.Ez(), center=[0.0, 0.0, 0.2], size=[1.0, 1.0, 0.0],
                          frequencies=[1.0])]
    return kh.Simulation(cell_size=[1.2, 1.0, 0.8], resolution=10, sources=src, monitors=mons, geometry=geo,
                         boundaries=pml, precision="f64", mode=mode, engine=engine)


@pytest.mark.parametrize("pml", [[[0.3, 0.3]] * 3, [[0.0, 0.0]] * 3])
def test_overlapping_sources_fast_mode(pml):
    """With PML the FAST source correction is applied inside the whole-grid
    update kernel; without PML, after the inner-box kernel, one launch per
    source.  Overlapping sources of one component must all land."""
    g = _overlap_sim(pml, "fast")
    o = _overlap_sim(pml, "exact", engine=OracleEngine)
    print(f"[edge] overlapping sources pml {pml[0]}", flush=True)
    n = int(g.last_source_time() / g.dt_n) + 15
    for chunk in (n // 2, n - n // 2):
        g.step(chunk)
        o.step(chunk)
        fa = [g.read_field(C()) for C in COMPS]
        fb = [o.read_field(C()) for C in COMPS]
        nb = math.sqrt(sum(float(np.sum(b * b)) for b in fb))
        assert nb > 0
        tot = math.sqrt(sum(float(np.sum((a - b) ** 2)) for a, b in zip(fa, fb))) / nb
        assert tot < 1e-12, tot
    assert not g.sources_active()
