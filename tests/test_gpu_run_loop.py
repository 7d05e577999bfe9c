"""The run loop through the HIP library (src/Simulation.jl:295-485): run with
a stop time, with until_after_sources and with stop_when_dft_decayed stops on
the same timestep as the same simulation on the CPU oracle, with the fields
within the parity tolerances; the device DFT norms the decay rule reads
(kh_dft_norms, reused between sampled steps) equal the oracle's at every
step."""
import numpy as np
import pytest

import khronos_hip as kh
from khronos_hip import simulation as S
from oracle_engine import OracleEngine
from test_run_loop import _expected_stop, _fields, _first_after, _norm_trace, _sim

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = np.concatenate([x.ravel() for x in a])
    b = np.concatenate([x.ravel() for x in b])
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("mode,tol", [("exact", 0.0), ("fast", 1e-12)])
def test_run_until_matches_oracle(mode, tol):
    print(f"[run-loop] until, {mode}", flush=True)
    sim = _sim(engine=None, mode=mode)
    ref = _sim()
    S.run(sim, until=4.2)
    S.run(ref, until=4.2)
    assert sim.timestep == ref.timestep == _first_after(ref, 4.2)
    assert sim.engine.timestep() == sim.timestep
    assert _rel(_fields(sim), _fields(ref)) <= tol


@pytest.mark.parametrize("mode", ["exact", "fast"])
def test_stop_when_dft_decayed_matches_oracle(mode):
    for kw in (dict(), dict(decimations=(3, 5)), dict(fcen=5.5, fwidth=1.0)):
        print(f"[run-loop] decay rule, {mode}, {kw}", flush=True)
        sim = _sim(engine=None, mode=mode, **kw)
        ref = _sim(**kw)
        rule = dict(tolerance=1e-3, minimum_runtime=0.0)
        S.run(sim, until_after_sources=S.stop_when_dft_decayed(**rule))
        S.run(ref, until_after_sources=S.stop_when_dft_decayed(**rule))
        print(f"[run-loop]   stopped at {sim.timestep} (oracle {ref.timestep})", flush=True)
        assert sim.timestep == ref.timestep
        assert _rel(_fields(sim), _fields(ref)) <= (0.0 if mode == "exact" else 1e-12)
        a = np.array(sim.engine.dft_norms())
        b = np.array(ref.engine.dft_norms())
        assert np.allclose(a, b, rtol=1e-12, atol=0)


def test_device_norms_track_the_oracle_step_by_step():
    """Mixed decimations: the engine answers from its last device query on
    steps where no monitor sampled, and queries again after a sampled step."""
    sim = _sim(engine=None, mode="fast", decimations=(3, 4))
    ref = _sim(decimations=(3, 4))
    tr = _norm_trace(ref, 60)
    sim.prepare()
    for n in range(1, 61):
        sim.step(1)
        a = np.array(sim.engine.dft_norms())
        assert np.allclose(a, tr[n], rtol=1e-12, atol=0), (n, a, tr[n])
    # the decay rule evaluated on the device norms, the oracle's stop step
    sim2 = _sim(engine=None, mode="fast", fcen=5.5, fwidth=1.0)
    sim2.prepare()
    n0 = _first_after(sim2, sim2.last_source_time())
    tr2 = _norm_trace(_sim(fcen=5.5, fwidth=1.0), n0 + 400)
    expect = _expected_stop(sim2, tr2, n0, 1e-3)
    S.run(sim2, until_after_sources=S.stop_when_dft_decayed(tolerance=1e-3))
    assert sim2.timestep == expect
