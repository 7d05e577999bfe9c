"""The run loop and its stop rules (src/Simulation.jl:295-485) on the CPU
oracle engine.

Pins:
* run(until=T) stops at the first timestep n with round_time(n) > T
  (round_time = Float64(n * dt) with the product in the backend type,
  src/Simulation.jl:22, :384-387), which the test computes in closed form;
  the fields equal those of n plain steps;
* run(until_after_sources=T) first steps while round_time <= the last
  source's cutoff time, then to round_time > T (an absolute time, not one
  counted from the sources' end);
* stop_when_dft_decayed with decimation D > 1 (quirk Q7): after the sources,
  the first check records each monitor's norm, the next check after a
  sampled step sees a change (ratio 1), and the first check after a step
  that sampled nothing sees change 0 with a non-zero largest change, i.e.
  "converged" -- so the run stops at n1 + 1, n1 the first timestep after the
  first check whose previous step sampled (n1 - 1 = 0 mod D); the test
  derives n1 from the timestep arithmetic alone;
* with D = 1 (a source band reaching the Nyquist limit), the stop step is
  the first check at which every monitor's step-to-step change is at most
  tolerance times its largest change so far, evaluated here from norms
  recorded step by step in a second run;
* minimum_runtime delays every check; maximum_runtime stops at the first
  check past it.
"""
import math

import numpy as np
import pytest

import khronos_hip as kh
from khronos_hip import simulation as S
from oracle_engine import OracleEngine


def _sim(fcen=1.0, fwidth=0.4, decimations=(None, None), precision="f64", engine=OracleEngine, mode="exact"):
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=fcen, fwidth=fwidth), component=kh.Ez(),
                            center=[0.05, -0.1, 0.0], size=[0.0, 0.0, 0.0])]
    mons = []
    for k, d in enumerate(decimations):
        kw = {} if d is None else {"decimation": d}
        mons.append(kh.DFTMonitor(component=kh.Ez() if k == 0 else kh.Hx(), center=[0.0, 0.0, 0.2 * k],
                                  size=[1.0, 0.8, 0.0], frequencies=[0.9 * fcen, fcen, 1.1 * fcen], **kw))
    return kh.Simulation(cell_size=[1.6, 1.4, 1.2], resolution=10, sources=src, monitors=mons,
                         boundaries=[[0.3, 0.3]] * 3, precision=precision, mode=mode, engine=engine)


def _first_after(sim, t, start=0):
    """First timestep n >= start with round_time(n) > t, by the backend
    arithmetic of src/Simulation.jl:22."""
    R = sim.real
    n = start
    while not float(R(n) * sim.dt_r) > t:
        n += 1
    return n


def _fields(sim):
    return [sim.read_field(C()) for C in (kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz)]


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_run_until_stops_at_first_time_past_T(precision):
    sim = _sim(precision=precision)
    T = 3.01
    S.run(sim, until=T)
    sim.prepare()
    n = _first_after(sim, T)
    assert sim.timestep == n and sim.engine.timestep() == n
    assert sim.round_time() > T >= float(sim.real(n - 1) * sim.dt_r)
    ref = _sim(precision=precision)
    ref.step(n)
    for a, b in zip(_fields(sim), _fields(ref)):
        assert np.array_equal(a, b)
    # running again with a time already passed takes no step
    S.run(sim, until=T)
    assert sim.timestep == n


def test_run_until_after_sources_uses_an_absolute_time():
    sim = _sim()
    sim.prepare()
    t_src = sim.last_source_time()
    assert math.isfinite(t_src)
    n_src = _first_after(sim, t_src)
    # T before the sources' end: the run stops as soon as the sources are done
    S.run(sim, until_after_sources=0.5 * t_src)
    assert sim.timestep == n_src
    # T after it: on to the first step past T
    sim2 = _sim()
    T = t_src + 1.37
    S.run(sim2, until_after_sources=T)
    assert sim2.timestep == _first_after(sim2, T)


def test_run_needs_exactly_one_stop_condition():
    sim = _sim()
    with pytest.raises(ValueError):
        S.run(sim)
    with pytest.raises(ValueError):
        S.run(sim, until=1.0, until_after_sources=1.0)
    with pytest.raises(ValueError):
        S.stop_when_dft_decayed(minimum_runtime=2.0, maximum_runtime=1.0)


@pytest.mark.parametrize("decimations", [(None, None), (3, 3)])
def test_dft_decay_stop_with_decimation_fires_on_a_non_sampling_step(decimations):
    """Quirk Q7 (src/Simulation.jl:437-481 checks every step; DFTs change
    every D-th step, src/Kernels/Kernels.jl:470)."""
    sim = _sim(decimations=decimations)
    sim.prepare()
    D = sim.spec.dfts[0].decimation
    assert D > 1 and all(md.decimation == D for md in sim.spec.dfts)
    n0 = _first_after(sim, sim.last_source_time())  # first check
    n1 = n0 + 1
    while (n1 - 1) % D != 0:                          # the step before sampled
        n1 += 1
    S.run(sim, until_after_sources=S.stop_when_dft_decayed(tolerance=1e-6))
    assert sim.timestep == n1 + 1, (sim.timestep, n0, n1, D)


def _norm_trace(sim, n_end):
    """Norms after every step up to timestep n_end (one step at a time)."""
    sim.prepare()
    out = {}
    while sim.timestep < n_end:
        sim.step(1)
        out[sim.timestep] = np.array(sim.engine.dft_norms())
    return out


def _expected_stop(sim, trace, n0, tol, min_rt=0.0, max_rt=math.inf):
    """First check n >= n0 that stops, from the recorded norms: a monitor
    counts once it has signal; the check after it first shows signal only
    records it; a counted monitor whose norm never changed yet blocks the
    stop; the run stops when at least one monitor has changed and every
    changed one moved by at most tol times its largest move so far.  Checks
    before min_rt do nothing; the first past max_rt stops."""
    R = sim.real
    state = {}
    for n in range(n0, max(trace) + 1):
        t = float(R(n) * sim.dt_r)
        if t < min_rt:
            continue
        if t > max_rt:
            return n
        ok, active = True, 0
        for m, cur in enumerate(trace[n]):
            if cur == 0.0:
                continue
            if m not in state:
                state[m] = (cur, 0.0)
                ok = False
                continue
            prev, big = state[m]
            big = max(big, abs(cur - prev))
            state[m] = (cur, big)
            if big == 0.0:
                ok = False
                continue
            active += 1
            ok = ok and abs(cur - prev) <= tol * big
        if active and ok:
            return n
    return None


def test_dft_decay_stop_without_decimation():
    # f_max = 5.5 + 0.5 > 1/(4 dt) at resolution 10: the auto decimation is 1
    kw = dict(fcen=5.5, fwidth=1.0)
    sim = _sim(**kw)
    sim.prepare()
    assert all(md.decimation == 1 for md in sim.spec.dfts)
    tol = 1e-3
    n0 = _first_after(sim, sim.last_source_time())
    trace = _norm_trace(_sim(**kw), n0 + 400)
    expect = _expected_stop(sim, trace, n0, tol)
    assert expect is not None and expect > n0 + 2
    S.run(sim, until_after_sources=S.stop_when_dft_decayed(tolerance=tol))
    assert sim.timestep == expect


def test_dft_decay_stop_minimum_and_maximum_runtime():
    sim = _sim()
    sim.prepare()
    t_src = sim.last_source_time()
    n0 = _first_after(sim, t_src)
    # maximum_runtime: the first check past it stops, before the decay rule
    # (which needs at least two more checks) could
    t_max = t_src + 0.5 * float(sim.dt_r)
    S.run(sim, until_after_sources=S.stop_when_dft_decayed(maximum_runtime=t_max, minimum_runtime=0.0))
    assert sim.timestep == max(n0, _first_after(sim, t_max)) <= n0 + 1
    # minimum_runtime: the Q7 pattern starts at the first check at or past it
    sim2 = _sim()
    sim2.prepare()
    D = sim2.spec.dfts[0].decimation
    t_min = t_src + 2.0
    R = sim2.real
    m0 = n0
    while float(R(m0) * sim2.dt_r) < t_min:
        m0 += 1
    m1 = m0 + 1
    while (m1 - 1) % D != 0:
        m1 += 1
    S.run(sim2, until_after_sources=S.stop_when_dft_decayed(minimum_runtime=t_min))
    assert sim2.timestep == m1 + 1


def test_run_benchmark_steps_until():
    sim = _sim()
    rate = S.run_benchmark(sim, 30)
    assert sim.timestep == 30 and rate > 0
