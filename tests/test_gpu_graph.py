"""Graph replay of the time step (hipGraph of a block of steps).

The same scene is stepped with every kernel launched eagerly (graph steps 0)
and with captured graphs; the fields, polarization and DFTs must be equal
value for value after every chunk of steps, with reads (device syncs and
host copies) between the replays.  The scene has sources that switch off
mid-run (the FAST correction drops out and the graph is captured again), DFT
monitors with decimation, a Lorentz + chi3 block and Bloch boundaries, so
every per-step quantity the device timestep drives is exercised.
"""
import sys
import time

import numpy as np
import pytest

import khronos_hip as kh

pytestmark = pytest.mark.gpu

COMPS = [kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz]


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _scene(precision, mode, bloch):
    bcs = None
    bnd = [[0.3, 0.3]] * 3
    if bloch:
        bcs = [[kh.Bloch(k=1.1), kh.Bloch(k=1.1)], [kh.Periodic(), kh.Periodic()], [kh.PML(), kh.PML()]]
        bnd = [[0.0, 0.0], [0.0, 0.0], [0.3, 0.3]]
    geo = [kh.Object(kh.Box([0.0, 0.0, 0.0], [0.5, 0.4, 0.4]),
                     kh.Material(ε=2.5, chi3=10.0, susceptibilities=[kh.LorentzianSusceptibility(1.0, 0.1, 0.6)]))]
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=1.2), component=kh.Ez(),
                            center=[0.1, 0.0, 0.0], size=[0.0, 0.0, 0.0], amplitude=2.0),
           kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=0.8, fwidth=1.0), component=kh.Hx(),
                            center=[-0.3, 0.1, 0.2], size=[0.0, 0.3, 0.0])]
    mons = [kh.FluxMonitor(center=[0.0, 0.0, 0.5], size=[1.2, 1.0, 0.0], frequencies=[0.8, 1.0, 1.2]),
            kh.DFTMonitor(component=kh.Ez(), center=[0.0, 0.0, 0.1], size=[0.8, 0.8, 0.0],
                          frequencies=[0.9, 1.1], decimation=3)]
    return kh.Simulation(cell_size=[1.2, 1.0, 1.6], resolution=10, sources=src, geometry=geo, monitors=mons,
                         boundaries=bnd, boundary_conditions=bcs, precision=precision, mode=mode)


def _state(sim):
    out = [sim.read_field(C()) for C in COMPS]
    out += [sim.engine.read_pole(0, w, a) for w in (0, 1) for a in range(3)]
    out += [sim.engine.read_dft(i) for i in range(len(sim.spec.dfts))]
    return out


@pytest.mark.parametrize("precision,mode,bloch", [("f32", "fast", False), ("f64", "exact", False),
                                                   ("f32", "fast", True)])
def test_graph_replay_equals_eager(precision, mode, bloch):
    eager = _scene(precision, mode, bloch)
    graph = _scene(precision, mode, bloch)
    eager.prepare()
    graph.prepare()
    eager.engine.set_graph(0)
    graph.engine.set_graph(8)
    cut = eager.last_source_time()
    n_cut = int(cut / eager.dt_n)
    _note(f"[graph] {precision} {mode} bloch={bloch} grid {eager.Nx}x{eager.Ny}x{eager.Nz} cutoff step ~{n_cut}")
    assert 40 < n_cut < 260
    done = 0
    for chunk in (37, 8, 100, 5, 64, 129):
        eager.step(chunk)
        graph.step(chunk)
        done += chunk
        for a, b in zip(_state(eager), _state(graph)):
            assert np.array_equal(a, b), done
    assert done > n_cut + 20 and not graph.sources_active()


def test_graph_speeds_up_small_grid():
    """Config-1 sized dipole (80^3): graphs remove launch gaps (reported, not asserted)."""
    def mk():
        src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.2), component=kh.Ez(),
                                center=[0.0, 0.0, 0.0], size=[0.0, 0.0, 0.0])]
        mon = [kh.DFTMonitor(component=kh.Ez(), center=[0.0, 0.0, 0.0], size=[4.0, 4.0, 0.0],
                             frequencies=list(np.linspace(0.9, 1.1, 5)))]
        return kh.Simulation(cell_size=[4.0, 4.0, 4.0], resolution=20, sources=src, monitors=mon,
                             boundaries=[[0.5, 0.5]] * 3, precision="f32", mode="fast")
    res = {}
    for steps_per_graph in (0, 32):
        s = mk()
        s.prepare()
        s.engine.set_graph(steps_per_graph)
        s.step(64)
        s.engine.sync()
        t0 = time.perf_counter()
        s.step(512)
        s.engine.sync()
        res[steps_per_graph] = (time.perf_counter() - t0) / 512 * 1e3
    _note(f"[graph] 80^3 f32 ms/step eager {res[0]:.4f}  graph {res[32]:.4f}")
    assert res[32] > 0


def test_launch_counts_and_prepared_graphs():
    """kh_get_launch_counts reports what kh_step ran: every step either inside
    a replay or eager, and one DFT launch per half step with a due monitor
    (the monitors of a half step share a launch), counted from the monitors'
    decimations; kh_prepare_graphs captures without stepping (timestep and
    fields unchanged), and runs of every length and phase keep replay equal
    to eager (short graphs at several phases of the DFT period)."""
    eager = _scene("f32", "fast", False)
    graph = _scene("f32", "fast", False)
    eager.prepare()
    graph.prepare()
    eager.engine.set_graph(0)
    graph.engine.set_graph(8)
    decs = [(md.component.electric, md.decimation) for md in graph.spec.dfts]
    assert any(d > 1 for _, d in decs)

    def due_launches(t0, n):
        k = 0
        for t in range(t0, t0 + n):
            for electric in (False, True):
                k += any(e == electric and t % d == 0 for e, d in decs)
        return k

    graph.engine.launch_counts()
    before = [graph.read_field(C()) for C in COMPS]
    graph.engine.prepare_graphs(50)
    assert graph.engine.timestep() == 0
    for a, b in zip(before, [graph.read_field(C()) for C in COMPS]):
        assert np.array_equal(a, b)
    assert graph.engine.launch_counts() == {"graph_replays": 0, "graph_steps": 0, "eager_steps": 0,
                                            "dft_launches": 0}
    t = 0
    for chunk in (50, 3, 8, 20, 9, 31, 17, 64, 11):
        eager.step(chunk)
        graph.step(chunk)
        c = graph.engine.launch_counts()
        _note(f"[graph] steps {t}..{t + chunk - 1}: {c}")
        assert c["graph_steps"] + c["eager_steps"] == chunk
        assert c["dft_launches"] == due_launches(t, chunk)
        if chunk >= 8:
            assert c["graph_replays"] >= 1
        t += chunk
        for a, b in zip(_state(eager), _state(graph)):
            assert np.array_equal(a, b), t


def test_kernel_timing_counts_and_results():
    """kh_set_timing: every update, material and DFT launch stamps an event
    pair at its kernel's start and end (hipExtLaunchKernel).  The counts per
    kernel class are the launches made, the durations are positive and add up
    to less than the wall time, and the fields equal an untimed eager run's
    value for value."""
    import time as _time
    timed = _scene("f32", "fast", False)
    plain = _scene("f32", "fast", False)
    timed.prepare()
    plain.prepare()
    plain.engine.set_graph(0)
    decs = [(md.component.electric, md.decimation) for md in timed.spec.dfts]
    due = sum(any(e == el and t % d == 0 for e, d in decs) for t in range(40) for el in (False, True))
    timed.engine.set_timing(True)
    timed.engine.get_timing()
    t0 = _time.perf_counter()
    timed.step(40)
    timed.engine.sync()
    wall_ms = (_time.perf_counter() - t0) * 1e3
    tm = timed.engine.get_timing()
    timed.engine.set_timing(False)
    plain.step(40)
    _note(f"[timing] 40 steps, wall {wall_ms:.2f} ms: {tm}")
    for a, b in zip(_state(timed), _state(plain)):
        assert np.array_equal(a, b)
    assert tm["h_full"][1] == 40 and tm["e_full"][1] == 40
    assert tm["material"][1] == 40
    assert tm["dft"][1] == due
    assert tm["h_inner"][1] == 0 and tm["e_inner"][1] == 0
    for k in ("h_full", "e_full", "material", "dft"):
        assert tm[k][0] > 0, k
    assert sum(v[0] for v in tm.values()) < wall_ms
