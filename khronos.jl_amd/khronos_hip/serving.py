"""Serving form: many independent simulations on one GPU with continuous batching.

The reference batches simulations by running them one after another
(``run_batch``, src/Batch.jl:82-140) or a wave of them at a time on separate
streams (``run_batch_concurrent``, :142-263), sizing waves with
``plan_batch`` (:11-68) from an HBM estimate (src/Memory.jl).  Here requests
are admitted continuously instead of in waves:

* admission -- a request is admitted while the device bytes of the admitted
  ones (``kh_mem_footprint`` plus monitor arrays) stay within the budget and
  fewer than ``max_active`` are in flight;
* preparation -- host setup and the device context of an admitted request
  are built on worker threads, so the GPU keeps stepping the others;
* stepping -- every active request owns a context (its own stream and its
  own captured step graphs); the library's scheduler (``kh_server_poll``)
  gives each idle context its next ``chunk`` steps (``kh_step`` is
  asynchronous) instead of synchronizing, so up to ``max_active`` streams run
  concurrently;
* retirement -- a finished request's flux is reduced on the device, its
  context freed, and its slot and bytes go to the next request in arrival
  order.

``run_one_at_a_time`` is the baseline: the same trace, one request at a time,
set up, run to completion and read before the next starts.

Latency of a request = completion time - arrival time; throughput = total
cell-updates / (last completion - first arrival).
"""
from __future__ import annotations

import math
import queue
import threading
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

from .simulation import Simulation, get_flux


@dataclass
class Request:
    """One simulation to serve: ``make()`` builds the Simulation (not yet
    prepared), ``steps`` time steps, ``arrival`` seconds after the trace starts."""

    rid: int
    make: Callable[[], Simulation]
    steps: int
    arrival: float = 0.0
    cells: int = 0
    footprint: int = 0


@dataclass
class Result:
    rid: int
    flux: np.ndarray
    arrival: float
    start: float
    done: float
    cells: int
    steps: int

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class ServeReport:
    results: List[Result]
    makespan: float
    cell_updates: float

    @property
    def throughput(self) -> float:
        """Mcell-updates/s over the whole trace."""
        return self.cell_updates / self.makespan / 1e6

    def latency_pct(self, q: float) -> float:
        return float(np.percentile([r.latency for r in self.results], q))


def request_footprint(sim: Simulation) -> int:
    """Device bytes of a simulation before it is set up: kh_mem_footprint of
    its descriptor (fields, auxiliaries, per-voxel materials when it has
    geometry) plus DFT accumulators sized from the monitors' extents."""
    import ctypes as C
    from . import _lib
    from .monitors import FluxMonitor
    lib = _lib.load()
    d = _lib.KhSimDesc()
    d.nx, d.ny, d.nz = sim.Nx, sim.Ny, sim.Nz
    d.precision = _lib.KH_F32 if sim.real is np.float32 else _lib.KH_F64
    d.mode = _lib.KH_MODE_EXACT if sim.mode == "exact" else _lib.KH_MODE_FAST
    d.dx, d.dy, d.dz = sim.delta_n
    d.dt = sim.dt_n
    keep = np.zeros(1)
    if sim.geometry:
        for a in range(3):
            d.eps_inv[a] = keep.ctypes.data_as(C.c_void_p)
    n = C.c_size_t()
    _lib.check(lib.kh_mem_footprint(C.byref(d), C.byref(n)))
    mon = 0
    for m in sim.monitors:
        cells = 1
        for a in range(3):
            cells *= int(math.floor(float(m.size[a]) * sim.resolution)) + 2
        comps = 4 if isinstance(m, FluxMonitor) else 1
        mon += comps * cells * len(m.frequencies) * 16
    return int(n.value) + mon


def _flux_of(sim: Simulation) -> np.ndarray:
    from .monitors import FluxMonitor
    out = [get_flux(m) for m in sim.monitors if isinstance(m, FluxMonitor)]
    return np.concatenate(out) if out else np.zeros(0)


class Server:
    """Continuous-batching scheduler over libkhronos_hip contexts (one GPU)."""

    def __init__(self, budget_bytes: float, max_active: int = 6, chunk: int = 256, prep_threads: int = 2,
                 graph_steps: int = 16, progress=None):
        # progress: optional callable(done, total, seconds) after each retired request
        self.progress = progress
        self.budget = float(budget_bytes)
        self.max_active = max_active
        self.chunk = chunk
        self.prep_threads = prep_threads
        self.graph_steps = graph_steps

    def run(self, requests: List[Request]) -> ServeReport:
        """Serve ``requests`` through the library's scheduler
        (kh_server_admit / kh_server_submit / kh_server_poll): admission and
        stepping run natively, host setup on worker threads."""
        import ctypes as C
        from . import _lib
        lib = _lib.load()  # load the library (and the HIP runtime) before the worker threads start
        srv = C.c_void_p()
        _lib.check(lib.kh_server_create(float(self.budget), int(self.max_active), int(self.chunk), C.byref(srv)))
        reqs = sorted(requests, key=lambda r: r.arrival)
        pending = list(reqs)
        jobs: "queue.Queue" = queue.Queue()
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        sims = {}             # request id -> (request, simulation, start time)
        lock = threading.Lock()
        results: List[Result] = []
        errors = []

        def worker():
            while True:
                r = jobs.get()
                if r is None:
                    return
                try:
                    sim = r.make()
                    sim.prepare()
                    sim.engine.set_graph(self.graph_steps)
                    sim.engine._ensure_waveforms(r.steps)  # the scheduler steps the context natively
                    with lock:
                        sims[r.rid] = (r, sim, now())
                    _lib.check(lib.kh_server_submit(srv, sim.engine.ctx, int(r.steps), float(r.footprint),
                                                    int(r.rid)))
                except Exception as e:  # surfaced on the scheduler thread
                    errors.append(e)
                    _lib.check(lib.kh_server_cancel(srv, float(r.footprint)))

        threads = [threading.Thread(target=worker, daemon=True) for _ in range(self.prep_threads)]
        for th in threads:
            th.start()
        done_buf = (C.c_int64 * 64)()
        n_done = C.c_int32()
        inflight = C.c_int32(0)
        admitted = C.c_int32()
        try:
            while True:
                if errors:
                    raise errors[0]
                t = now()
                # admit in arrival order while the budget and the slots allow
                while pending and pending[0].arrival <= t:
                    _lib.check(lib.kh_server_admit(srv, float(pending[0].footprint), C.byref(admitted)))
                    if not admitted.value:
                        break
                    jobs.put(pending.pop(0))
                _lib.check(lib.kh_server_poll(srv, done_buf, 64, C.byref(n_done), C.byref(inflight)))
                for k in range(n_done.value):
                    with lock:
                        r, sim, ts = sims.pop(int(done_buf[k]))
                    sim.engine.ts = r.steps
                    sim.timestep = r.steps
                    flux = _flux_of(sim)
                    done = now()
                    sim.engine.close()
                    results.append(Result(r.rid, flux, r.arrival, ts, done, r.cells, r.steps))
                    if self.progress is not None:
                        self.progress(len(results), len(reqs), done)
                if not pending and inflight.value == 0:
                    break
                if n_done.value == 0:
                    time.sleep(0.0002)
        finally:
            for _ in threads:
                jobs.put(None)
            lib.kh_server_destroy(srv)
        makespan = max(r.done for r in results) - min(r.arrival for r in results)
        return ServeReport(sorted(results, key=lambda r: r.rid), makespan,
                           float(sum(r.cells * r.steps for r in results)))


def run_one_at_a_time(requests: List[Request], graph_steps: int = 16, progress=None) -> ServeReport:
    """Baseline: the same trace served one request at a time."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    t0 = time.perf_counter()
    results = []
    for r in reqs:
        wait = r.arrival - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        start = time.perf_counter() - t0
        sim = r.make()
        sim.prepare()
        sim.engine.set_graph(graph_steps)
        sim.step(r.steps)
        sim.engine.sync()
        flux = _flux_of(sim)
        done = time.perf_counter() - t0
        sim.engine.close()
        results.append(Result(r.rid, flux, r.arrival, start, done, r.cells, r.steps))
        if progress is not None:
            progress(len(results), len(reqs), done)
    makespan = max(x.done for x in results) - min(x.arrival for x in results)
    return ServeReport(sorted(results, key=lambda x: x.rid), makespan,
                       float(sum(x.cells * x.steps for x in results)))


# ------------------------------------------------------------------ workload
SERVING_GRIDS = [(64, 64, 64), (96, 96, 96), (128, 128, 128), (192, 192, 128), (256, 256, 256)]


def serving_request(rid: int, rng: np.random.Generator, grids=SERVING_GRIDS, steps=(1000, 3000),
                    resolution: int = 20, precision: str = "f32", n_freq: int = 20,
                    setup_precision: Optional[str] = "f64") -> Request:
    """A request of the serving workload: a grid drawn from ``grids``,
    10-cell PML, 1-3 boxes/spheres with eps ~ U[1.5, 12], an Ez Gaussian
    dipole and one flux plane at ``n_freq`` frequencies.  FP32 requests take
    the FP64 setup (``setup_precision``: grid snapping, weights, waveforms)
    with FP64 DFT accumulators, so that their fluxes are those of the FP64
    problem up to the FP32 stencil's rounding."""
    from . import geometry as geo_mod
    from .grid import Ez
    from .monitors import FluxMonitor
    from .sources import GaussianPulseSource, UniformSource
    n = grids[int(rng.integers(len(grids)))]
    nsteps = int(rng.integers(steps[0], steps[1] + 1))
    cell = [v / resolution for v in n]
    pml = 10 / resolution
    objs = []
    for _ in range(int(rng.integers(1, 4))):
        eps = float(rng.uniform(1.5, 12.0))
        c = [float(rng.uniform(-0.25, 0.25)) * cell[a] for a in range(3)]
        if rng.random() < 0.5:
            size = [float(rng.uniform(0.1, 0.3)) * cell[a] for a in range(3)]
            shape = geo_mod.Box(c, size)
        else:
            shape = geo_mod.Sphere(c, float(rng.uniform(0.05, 0.15)) * min(cell))
        objs.append(geo_mod.Object(shape, geo_mod.Material(ε=eps)))
    freqs = list(np.linspace(0.8, 1.2, n_freq))
    plane_x = 0.5 * (cell[0] / 2 - pml)

    def make(precision=precision, engine=None):
        src = [UniformSource(time_profile=GaussianPulseSource(fcen=1.0, fwidth=0.2), component=Ez(),
                             center=[0.0, 0.0, 0.0], size=[0.0, 0.0, 0.0])]
        mon = [FluxMonitor(center=[plane_x, 0.0, 0.0], size=[0.0, cell[1] - 2 * pml, cell[2] - 2 * pml],
                           frequencies=freqs)]
        return Simulation(cell_size=cell, resolution=resolution, sources=src, geometry=objs, monitors=mon,
                          boundaries=[[pml, pml]] * 3, precision=precision, mode="fast",
                          setup_precision=setup_precision, engine=engine)

    return Request(rid, make, nsteps, cells=n[0] * n[1] * n[2])


def serving_trace(n_requests: int, seed: int = 0, rate: float = math.inf, **kw) -> List[Request]:
    """Poisson arrivals at ``rate`` requests/s (inf: all at t = 0)."""
    rng = np.random.default_rng(seed)
    reqs = []
    t = 0.0
    for i in range(n_requests):
        r = serving_request(i, rng, **kw)
        if math.isfinite(rate):
            t += float(rng.exponential(1.0 / rate))
        r.arrival = t
        reqs.append(r)
    for r in reqs:
        r.footprint = request_footprint(r.make())
    return reqs
