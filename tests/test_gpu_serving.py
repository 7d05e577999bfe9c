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


def _config5_trace(seed=11, rate=4.0):
    """Eight requests of the serving workload (64^3 ... 256^3, 10-cell PML,
    1-3 dielectric boxes/spheres, Ez dipole, 1000-3000 steps, 20-frequency
    flux plane), one forced to 256^3 and one to 64^3, arriving as a Poisson
    process at ``rate`` requests/s; each request's draws come from its own
    seeded generator, so its FP64 twin is the same problem."""
    import math
    grids = [None, None, None, None, None, None, [(256, 256, 256)], [(64, 64, 64)]]
    reqs = []
    t = 0.0
    arrivals = np.random.default_rng(seed)
    for i, g in enumerate(grids):
        kw = {} if g is None else {"grids": g}
        r = serving.serving_request(i, np.random.default_rng([seed, i]), **kw)
        t += float(arrivals.exponential(1.0 / rate))
        r.arrival = t
        r.footprint = serving.request_footprint(r.make())
        reqs.append(r)
    assert math.isfinite(reqs[-1].arrival)
    return reqs


def test_config5_stream_served_equals_solo_and_fp64_and_oracle():
    """The serving configuration at its real sizes, Poisson arrivals: served
    fluxes equal the one-at-a-time fluxes bit for bit, each is within 1e-4
    (of the spectrum's peak) of the same request in FP64, and the 64^3
    request's FP64 flux is within 1e-4 of the CPU oracle's FP64 flux."""
    import torch

    from khronos_hip.monitors import flux_from_dft
    from oracle_engine import OracleEngine
    reqs = _config5_trace()
    _note("[serving config5] " + ", ".join(f"{r.rid}: {round(r.cells ** (1 / 3))}^3-class {r.cells} cells "
                                           f"{r.steps} steps at {r.arrival:.2f} s" for r in reqs))
    free = torch.cuda.mem_get_info()[0]
    served = serving.Server(budget_bytes=0.8 * free, max_active=6, chunk=256).run(reqs)
    solo = serving.run_one_at_a_time(reqs)
    _note(f"[serving config5] served {served.throughput:.0f} vs one at a time {solo.throughput:.0f} Mcell/s; "
          f"p50 {served.latency_pct(50):.2f} s p99 {served.latency_pct(99):.2f} s (one at a time "
          f"{solo.latency_pct(50):.2f} / {solo.latency_pct(99):.2f} s)")
    for a, b in zip(served.results, solo.results):
        assert a.rid == b.rid and a.flux.shape == (20,)
        assert np.array_equal(a.flux, b.flux), a.rid
    worst = 0.0
    small = None
    for r, a in zip(reqs, served.results):
        s64 = r.make(precision="f64")
        s64.prepare()
        s64.step(r.steps)
        f64 = serving._flux_of(s64)
        s64.engine.close()
        err = float(np.max(np.abs(a.flux - f64)) / np.max(np.abs(f64)))
        worst = max(worst, err)
        _note(f"[serving config5] request {r.rid} ({r.cells} cells): |flux32 - flux64| / peak = {err:.2e}")
        assert err <= 1e-4, (r.rid, err)
        if r.cells == 64 ** 3:
            small = (r, f64)
    r, f64 = small
    o = r.make(precision="f64", engine=OracleEngine)
    o.step(r.steps)
    fm = o.monitors[0]
    md = fm.monitor_data
    mons = md.tangential_E_monitors + md.tangential_H_monitors
    ref = flux_from_dft(md, *[o.read_dft(m.monitor_data) for m in mons], [m.monitor_data.gv for m in mons])
    err = float(np.max(np.abs(f64 - ref)) / np.max(np.abs(ref)))
    err32 = float(np.max(np.abs(served.results[r.rid].flux - ref)) / np.max(np.abs(ref)))
    _note(f"[serving config5] 64^3 request, {r.steps} steps: device FP64 vs oracle FP64 {err:.2e}, "
          f"served FP32 vs oracle FP64 {err32:.2e}; worst served FP32 vs FP64 {worst:.2e}")
    assert err <= 1e-9 and err32 <= 1e-4
