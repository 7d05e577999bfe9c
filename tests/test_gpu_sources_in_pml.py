"""FAST-mode sources applied inside the PML launch.

Three forms of the same correction, chosen per half step:
* the sources (real fields) are all on one component with pairwise disjoint
  boxes: their corrections for a block of steps are computed by one launch at
  the block's start and the PML kernel adds them before its stores;
* otherwise small sources (points and lines: few wave-wide row segments of
  the launch) are applied by the PML kernel after its lanes store their
  values, source by source in registration order;
* otherwise (and with KH_SRC_LAUNCH=1, read when a context plans its first
  step) source_fix_kernel launches after the update.
All perform the same operations on the same values, so a run with every
source in separate launches must give the same fields and DFTs bit for bit:
point sources on E and H components, two of them on one cell, a line source
crossing a PML slab and a point inside one (own-axis sigma non-zero: the W
stage takes the correction), sources that switch off mid-run inside a graph
block, real and complex (Bloch) fields, FP32 and FP64.
"""
import sys

import numpy as np
import pytest

import khronos_hip as kh

pytestmark = pytest.mark.gpu

COMPS = [kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz]


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _single_component_scene(precision):
    """E sources all on Ez with disjoint boxes, one a plane too large for the
    end-of-lane form and one a point inside the z PML; H sources all on Hy, a
    plane: both half steps take the precomputed form."""
    g1 = kh.GaussianPulseSource(fcen=1.0, fwidth=1.2)
    g2 = kh.GaussianPulseSource(fcen=0.8, fwidth=1.0)
    src = [kh.UniformSource(time_profile=g1, component=kh.Ez(), center=[-1.0, 0.0, 0.0], size=[0.0, 3.6, 3.6],
                            amplitude=2.0),
           kh.UniformSource(time_profile=g2, component=kh.Ez(), center=[0.5, 0.2, 0.3], size=[0.4, 0.3, 0.0],
                            amplitude=-0.7 + 0.4j),
           kh.UniformSource(time_profile=g1, component=kh.Ez(), center=[0.6, -0.5, -1.85], size=[0.0, 0.0, 0.0]),
           kh.UniformSource(time_profile=g2, component=kh.Hy(), center=[1.2, 0.1, -0.2], size=[0.0, 3.4, 3.4])]
    geo = [kh.Object(kh.Box([0.0, 0.0, 0.0], [0.5, 0.4, 0.4]), kh.Material(ε=2.5))]
    mons = [kh.DFTMonitor(component=kh.Ez(), center=[0.0, 0.0, 0.1], size=[0.8, 0.8, 0.0],
                          frequencies=[0.9, 1.1])]
    return kh.Simulation(cell_size=[4.0, 4.0, 4.0], resolution=10, sources=src, geometry=geo, monitors=mons,
                         boundaries=[[0.3, 0.3]] * 3, precision=precision, mode="fast")


def _scene(precision, bloch):
    bcs = None
    bnd = [[0.3, 0.3]] * 3
    if bloch:
        bcs = [[kh.Bloch(k=0.9), kh.Bloch(k=0.9)], [kh.Periodic(), kh.Periodic()], [kh.PML(), kh.PML()]]
        bnd = [[0.0, 0.0], [0.0, 0.0], [0.3, 0.3]]
    g1 = kh.GaussianPulseSource(fcen=1.0, fwidth=1.2)
    g2 = kh.GaussianPulseSource(fcen=0.8, fwidth=1.0)
    src = [kh.UniformSource(time_profile=g1, component=kh.Ez(), center=[0.1, 0.0, 0.0], size=[0.0, 0.0, 0.0],
                            amplitude=2.0),
           kh.UniformSource(time_profile=g2, component=kh.Ez(), center=[0.1, 0.0, 0.0], size=[0.0, 0.0, 0.0],
                            amplitude=-0.7 + 0.4j),
           kh.UniformSource(time_profile=g2, component=kh.Hx(), center=[-0.3, 0.1, 0.2], size=[0.0, 0.3, 0.0]),
           # along x through the x PML slabs: Ex's own-axis sigma is non-zero there
           kh.UniformSource(time_profile=g1, component=kh.Ex(), center=[0.0, -0.2, -0.1], size=[1.2, 0.0, 0.0])]
    geo = [kh.Object(kh.Box([0.0, 0.0, 0.0], [0.5, 0.4, 0.4]), kh.Material(ε=2.5))]
    mons = [kh.DFTMonitor(component=kh.Ez(), center=[0.0, 0.0, 0.1], size=[0.8, 0.8, 0.0],
                          frequencies=[0.9, 1.1]),
            kh.DFTMonitor(component=kh.Hx(), center=[0.0, 0.0, -0.2], size=[0.8, 0.8, 0.0],
                          frequencies=[0.9, 1.1], decimation=3)]
    return kh.Simulation(cell_size=[1.2, 1.0, 1.6], resolution=10, sources=src, geometry=geo, monitors=mons,
                         boundaries=bnd, boundary_conditions=bcs, precision=precision, mode="fast")


def _state(sim):
    return [sim.read_field(C()) for C in COMPS] + [sim.engine.read_dft(i) for i in range(len(sim.spec.dfts))]


@pytest.mark.parametrize("precision,bloch,single", [("f32", False, False), ("f64", False, False),
                                                    ("f32", True, False), ("f32", False, True),
                                                    ("f64", False, True)])
def test_sources_in_pml_launch_equal_separate_launches(precision, bloch, single, monkeypatch):
    make = (lambda: _single_component_scene(precision)) if single else (lambda: _scene(precision, bloch))
    fused = make()
    fused.prepare()
    fused.step(1)  # plans with sources inside the PML launch
    monkeypatch.setenv("KH_SRC_LAUNCH", "1")
    sep = make()
    sep.prepare()
    sep.step(1)
    monkeypatch.delenv("KH_SRC_LAUNCH")
    n_cut = int(fused.last_source_time() / fused.dt_n)
    _note(f"[sources in pml] {precision} bloch={bloch} single-component={single} grid "
          f"{fused.Nx}x{fused.Ny}x{fused.Nz} cutoff step ~{n_cut}")
    assert 40 < n_cut < 260
    done = 1
    for chunk in (20, 57, 3, 150, 64):
        fused.step(chunk)
        sep.step(chunk)
        done += chunk
        for a, b in zip(_state(fused), _state(sep)):
            assert np.array_equal(a, b), done
    assert done > n_cut + 20 and not fused.sources_active()
    assert any(np.any(f != 0) for f in _state(fused))
