"""FAST mode: small E monitors accumulated by the E update itself.

The E launch applies the DFT update of up to four small E monitors (at most
8192 cell-frequencies each, inside cells 1..N, real fields, float64
accumulators, no material pass or periodic axis, no E source writing a
monitor cell of its component) from the values it stores, with weights
formed once per block of steps; dft_kernel runs only for the other monitors
(KH_DFT_LAUNCH=1, read when a context plans its steps: every monitor in
dft_kernel).  Both perform the same operations on the same values, so the
fields and every DFT agree bit for bit: 3D and 2D, FP32 and FP64, monitors
with and without decimation, a fifth small E monitor (left to dft_kernel),
an H monitor, a monitor under a source of its component, and a plane source
with block-precomputed corrections, through graph replays and eager steps.
"""
import os
import sys

import numpy as np
import pytest

import khronos_hip as kh

pytestmark = pytest.mark.gpu

COMPS3 = [kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz]


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mons(two_d):
    z = 0.0 if two_d else 0.1
    sz = 0.0
    mons = [kh.DFTMonitor(component=kh.Ez(), center=[0.3, 0.0, z], size=[0.0, 0.5, sz], frequencies=[0.9, 1.1]),
            kh.DFTMonitor(component=kh.Ez(), center=[-0.2, 0.1, z], size=[0.3, 0.3, sz],
                          frequencies=[0.8, 1.0, 1.2], decimation=3),
            kh.DFTMonitor(component=kh.Ex(), center=[0.0, -0.2, z], size=[0.4, 0.0, sz], frequencies=[1.0]),
            kh.DFTMonitor(component=kh.Ey(), center=[0.1, 0.0, z], size=[0.2, 0.2, sz],
                          frequencies=[0.95, 1.05], decimation=2),
            # a fifth small E monitor: past four, dft_kernel takes it
            kh.DFTMonitor(component=kh.Ez(), center=[0.0, 0.3, z], size=[0.2, 0.0, sz], frequencies=[1.0]),
            kh.DFTMonitor(component=kh.Hz(), center=[0.0, 0.0, z], size=[0.3, 0.3, sz], frequencies=[1.0])]
    return mons


def _scene(precision, two_d=False, plane=False, src_on_mon=False):
    g1 = kh.GaussianPulseSource(fcen=1.0, fwidth=1.2)
    cell = [1.6, 1.4, 0.0 if two_d else 1.4]
    if plane:
        # 36 x 36 cells: too large for the end-of-lane form, so its
        # corrections are computed per block of steps
        cell = [4.0, 4.0, 4.0]
        src = [kh.UniformSource(time_profile=g1, component=kh.Ez(), center=[-1.0, 0.0, 0.0], size=[0.0, 3.6, 3.6])]
    else:
        src = [kh.UniformSource(time_profile=g1, component=kh.Ez(), center=[-0.35, 0.05, 0.0],
                                size=[0.0, 0.0, 0.0], amplitude=2.0)]
    if two_d:
        # TE fields too (Ex, Ey, Hz)
        src.append(kh.UniformSource(time_profile=g1, component=kh.Ey(), center=[-0.3, -0.1, 0.0],
                                    size=[0.0, 0.0, 0.0]))
    if src_on_mon:
        # writes cells of the first Ez monitor: that monitor stays with dft_kernel
        src.append(kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=0.8, fwidth=1.0), component=kh.Ez(),
                                    center=[0.3, 0.0, 0.0 if two_d else 0.1], size=[0.0, 0.2, 0.0]))
    geo = [kh.Object(kh.Box([0.0, 0.0, 0.0], [0.5, 0.4, 0.4]), kh.Material(ε=2.5))]
    bnd = [[0.3, 0.3]] * (2 if two_d else 3)
    return kh.Simulation(cell_size=cell, resolution=10, sources=src, geometry=geo, monitors=_mons(two_d),
                         boundaries=bnd, precision=precision, mode="fast")


def _state(sim):
    return [sim.read_field(C()) for C in COMPS3] + [sim.engine.read_dft(i) for i in range(len(sim.spec.dfts))]


def _step(sim, n, separate):
    if separate:
        os.environ["KH_DFT_LAUNCH"] = "1"
    try:
        sim.step(n)
    finally:
        os.environ.pop("KH_DFT_LAUNCH", None)


@pytest.mark.parametrize("precision,two_d,plane,src_on_mon", [("f32", False, False, False),
                                                              ("f64", False, False, True),
                                                              ("f32", True, False, False),
                                                              ("f64", False, True, False)])
def test_dft_in_e_update_equals_dft_launches(precision, two_d, plane, src_on_mon):
    fused = _scene(precision, two_d, plane, src_on_mon)
    sep = _scene(precision, two_d, plane, src_on_mon)
    fused.prepare()
    sep.prepare()
    _note(f"[dft in update] {precision} 2d={two_d} plane={plane} source-on-monitor={src_on_mon} "
          f"grid {fused.Nx}x{fused.Ny}x{fused.Nz}")
    fused.engine.launch_counts()
    sep.engine.launch_counts()
    nf = ns = 0
    done = 0
    for chunk in (1, 20, 57, 3, 64, 16):
        _step(fused, chunk, False)
        _step(sep, chunk, True)
        done += chunk
        nf += fused.engine.launch_counts()["dft_launches"]
        ns += sep.engine.launch_counts()["dft_launches"]
        for a, b in zip(_state(fused), _state(sep)):
            assert np.array_equal(a, b), done
    _note(f"[dft in update] dft_kernel launches {nf} (E update on) vs {ns} (off) in {done} steps")
    assert nf < ns
    assert all(np.any(sep.engine.read_dft(i) != 0) for i in range(5))
