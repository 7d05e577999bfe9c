"""Serving form: continuous batching of independent simulations.

Every served request must give the flux it gives when run alone (the
contexts are independent; concurrency changes only the timing), and the
scheduler must respect its memory budget and slot limit.
"""
import sys

import numpy as np
import pytest

from khronos_hip import serving

pytestmark = pytest.mark.gpu

GRIDS = [(32, 32, 32), (40, 32, 36), (48, 40, 32)]


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_served_fluxes_equal_solo_runs():
    reqs = serving.serving_trace(7, seed=3, grids=GRIDS, steps=(60, 180), n_freq=6)
    _note(f"[serving] {len(reqs)} requests, steps {[r.steps for r in reqs]}")
    solo = serving.run_one_at_a_time(reqs)
    served = serving.Server(budget_bytes=2e9, max_active=3, chunk=40, prep_threads=2).run(reqs)
    assert [r.rid for r in served.results] == [r.rid for r in solo.results]
    for a, b in zip(served.results, solo.results):
        assert a.flux.shape == (6,)
        assert np.array_equal(a.flux, b.flux), a.rid
        assert np.any(a.flux != 0)
    _note(f"[serving] throughput served {served.throughput:.1f} vs solo {solo.throughput:.1f} Mcell/s, "
          f"p50 {served.latency_pct(50):.3f}s p99 {served.latency_pct(99):.3f}s")


def test_budget_limits_concurrency():
    reqs = serving.serving_trace(4, seed=5, grids=GRIDS, steps=(40, 60), n_freq=4)
    one = max(r.footprint for r in reqs)
    rep = serving.Server(budget_bytes=one * 1.5, max_active=4, chunk=20).run(reqs)
    # room for one request at a time: a request is admitted (prepared) only
    # after the previous one retired, so the served intervals do not overlap
    spans = sorted((r.start, r.done) for r in rep.results)
    for (s0, d0), (s1, d1) in zip(spans, spans[1:]):
        assert s1 >= d0, (spans,)
    assert len(rep.results) == 4
