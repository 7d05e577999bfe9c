#!/usr/bin/env python3
"""Benchmark: FP32 3D Yee + CPML time step on one MI355X.

Workload: the strip-waveguide configuration (384 x 192 x 160 cells, 12-cell
CPML on all faces, per-voxel eps, Ez planar Gaussian-pulse source, two flux
planes at 50 frequencies with float64 DFT accumulators), FP32 fields, FAST
mode.  A step is one full time step (H half step, E half step, sources, due
DFT updates).  Fields, materials and monitors are resident in HBM before the
timed region starts.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
    python bench.py --near2far | --serving | --adjoint   (the other forms)

Prints one JSON line.  Mcell-updates/s = Nx*Ny*Nz*K / wall_s / 1e6
(src/Simulation.jl:517-518).  The warm-up runs on to a step where the flux
planes sample, the graphs the window replays are captured before it, and
``config.window`` says what the window ran (graph replays, eager steps, DFT
launches).  The roofline entry uses the E-update kernel (the dominant one):
algorithmic bytes per launch / mean launch time from the HIP event pair each
launch stamps at its kernel's start and end on the library's stream
(hipExtLaunchKernel), beside the measured HBM traffic
per launch (committed counter summary) over the same time.  Also in the
line: the 80^3 FP64 dipole's rate, a short Poisson serving stream (served vs
one at a time, p50/p99, ideal-overlap bound), the complex metasurface's rate
(200 steps), the 400 x 400 adjoint gradient's latency and the CPU oracle on
all the process's cores (at most 16) and on one core.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "khronos.jl_amd"))

HBM_PEAK_GBS = 8000.0


def _omp_threads(n: int):
    """Thread count of the oracle's OpenMP regions from here on (the oracle
    runs its parallel loops on the calling thread's team)."""
    import ctypes
    ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def cpu_baseline(spec, cores: int, budget_s: float = 12.0, warmup: int = 10, label: str = "strip-waveguide"):
    """Time the CPU oracle (C restatement of the reference CPU path, OpenMP on
    ``cores`` threads) on the same workload: ``warmup`` untimed steps, as
    run_benchmark does (src/Simulation.jl:492-527), then a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_engine import OracleEngine  # the checker, used here only as the CPU baseline
    _omp_threads(cores)
    eng = OracleEngine(spec)
    t0 = time.perf_counter()
    eng.step(warmup)
    t1 = (time.perf_counter() - t0) / max(1, warmup)
    n = max(3, min(200, int(budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    eng.step(n)
    dt = time.perf_counter() - t0
    cells = spec.N[0] * spec.N[1] * spec.N[2]
    del eng
    return {
        "value": round(cells * n / dt / 1e6, 3),
        "unit": "Mcell-updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{label} {spec.N[0]}x{spec.N[1]}x{spec.N[2]} "
                  f"{'FP32' if np.dtype(spec.real).itemsize == 4 else 'FP64'}, {n} steps after {warmup} warm-up "
                  f"steps (C oracle, -O3 -fopenmp, {cores} thread{'s' if cores > 1 else ''})",
    }


def dft_period(spec) -> int:
    """lcm of the DFT monitors' decimations (the library's graph period)."""
    p = 1
    for md in spec.dfts:
        p = p * max(1, md.decimation) // math.gcd(p, max(1, md.decimation))
    return p


def config1_leg(steps: int = 500):
    """The 80^3 FP64 dipole (vacuum, 10-cell PML, Ez Gaussian dipole, 5-frequency
    DFT plane): Mcell-updates/s over steps 10..steps, as run_benchmark times it
    (src/Simulation.jl:492-527)."""
    from khronos_hip.workloads import dipole_80
    sim = dipole_80(precision="f64", mode="fast")
    sim.prepare()
    sim.step(10)
    sim.engine.prepare_graphs(steps - 10)
    sim.engine.sync()
    t0 = time.perf_counter()
    sim.step(steps - 10)
    sim.engine.sync()
    dt = time.perf_counter() - t0
    rate = sim.Nx * sim.Ny * sim.Nz * (steps - 10) / dt / 1e6
    sim.engine.close()
    return {"workload": "dipole 80^3 FP64, 10-cell PML, Ez Gaussian dipole, 5-frequency DFT plane, "
                        f"steps 10..{steps}", "value": round(rate, 1), "unit": "Mcell-updates/s",
            "reference_gh200": 640.1, "reference_source": "benchmark/dipole.yml:380 (Khronos.jl, FP64, 80^3)"}


def dipole_bench(args):
    """The 80^3 FP64 dipole (vacuum, 10-cell PML, Ez Gaussian dipole,
    5-frequency DFT plane) with its kernels: the timed window starts after
    ``--warmup`` steps (10 in run_benchmark, src/Simulation.jl:492-527) and
    runs ``--steps`` steps (490: to step 500, the reference's run, sources on
    throughout); then 100 eager steps give the per-kernel times.  The roofline
    entry is the E update against its algorithmic bytes."""
    import torch
    if not torch.cuda.is_available():
        sys.stderr.write("no GPU visible\n")
        return 1
    from khronos_hip.workloads import algorithmic_bytes, dipole_80
    sim = dipole_80(precision="f64", mode="fast")
    sim.prepare()
    eng = sim.engine
    eng.set_graph(args.graph_steps)
    cells = sim.Nx * sim.Ny * sim.Nz
    ab = algorithmic_bytes(sim.spec)
    sim.step(args.warmup)
    eng.prepare_graphs(args.steps)
    eng.sync()
    eng.launch_counts()
    t_start = sim.timestep
    t0 = time.perf_counter()
    sim.step(args.steps)
    eng.sync()
    wall = time.perf_counter() - t0
    counts = eng.launch_counts()
    k_steps = 100
    eng.set_timing(k_steps * 8 + 64)
    eng.get_timing()
    sim.step(k_steps)
    eng.sync()
    tm = eng.get_timing()
    eng.set_timing(0)
    avg = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in tm.items()}
    e_ms = avg["e_full"]
    b_alg = (ab["H"] + ab["E"]) / cells
    line = {
        "metric": "Mcell-updates/s (FP64 3D Yee + CPML time step, 80^3 dipole)",
        "value": round(cells * args.steps / wall / 1e6, 2), "unit": "Mcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "dipole 80^3 FP64, 10-cell PML, Ez Gaussian dipole (fcen 1, fwidth 0.2), "
                               "5-frequency DFT plane", "cells": cells, "graph_steps": args.graph_steps,
                   "window": {"timesteps": [t_start, t_start + args.steps - 1], **counts},
                   "B_alg_bytes_per_cell_update": round(b_alg, 2),
                   "kernel_ms": {k: round(v, 5) for k, v in avg.items()},
                   "hbm_fraction_whole_step": round(cells * args.steps / wall * b_alg / 1e9 / HBM_PEAK_GBS, 4),
                   "reference_gh200": 640.1, "reference_source": "benchmark/dipole.yml:380 (Khronos.jl, FP64, 80^3)"},
        "roofline": {"bound": "hbm", "achieved": round(ab["E"] / (e_ms * 1e-3) / 1e9, 1) if e_ms > 0 else 0.0,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ab["E"] / (e_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if e_ms > 0 else 0.0,
                     "traffic": None, "kernel": "pml_e_kernel<double,false,false,false>",
                     "alg_bytes_per_launch": ab["E"], "kernel_ms": round(e_ms, 5)},
    }
    eng.close()
    print(json.dumps(line), flush=True)
    return 0


def concurrency_bench(args):
    """Serving form, small grids: ``--batch`` same-shape requests of the
    serving workload stepped concurrently (each context on its own stream,
    all their graphs queued from one host thread) against one
    context holding the same cells in one grid (n x n x batch*n: what one
    launch batching the requests would stream) and against one request
    alone, by algorithmic bytes moved per second (the PML share differs
    between the shapes, so cell rates alone would not compare)."""
    import torch
    if not torch.cuda.is_available():
        sys.stderr.write("no GPU visible\n")
        return 1
    from khronos_hip import serving
    from khronos_hip.workloads import algorithmic_bytes
    K, steps, warm = args.batch, args.steps, 64

    def ready(s):
        s.prepare()
        s.step(warm)
        s.engine.prepare_graphs(steps)
        s.engine.sync()
        ab = algorithmic_bytes(s.spec)
        return ab["H"] + ab["E"]

    def timed(sims):
        per_step = [ready(s) for s in sims]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in sims:  # queued asynchronously, each on its own stream
            s.step(steps)
        for s in sims:
            s.engine.sync()
        wall = time.perf_counter() - t0
        cells = sum(s.Nx * s.Ny * s.Nz for s in sims)
        for s in sims:
            s.engine.close()
        return {"GBps": round(sum(per_step) * steps / wall / 1e9, 1),
                "Mcell_updates_per_s": round(cells * steps / wall / 1e6, 1), "ms_per_step": round(wall / steps * 1e3, 4)}

    rows = []
    for n in args.grids:
        mk = lambda i, shape: serving.serving_request(i, np.random.default_rng([n, i]), grids=[shape]).make()
        alone = timed([mk(0, (n, n, n))])
        conc = timed([mk(i, (n, n, n)) for i in range(K)])
        one = timed([mk(0, (n, n, K * n))])
        rows.append({"grid": f"{n}^3", "alone": alone, f"{K}_concurrent": conc, f"one_grid_{n}x{n}x{K * n}": one,
                     "concurrent_over_one_grid": round(conc["GBps"] / one["GBps"], 3)})
        sys.stderr.write(f"{n}^3: alone {alone['GBps']} GB/s, {K} concurrent {conc['GBps']} GB/s, "
                         f"one {n}x{n}x{K * n} grid {one['GBps']} GB/s\n")
    print(json.dumps({"metric": "algorithmic GB/s of same-shape serving requests: concurrent contexts vs one grid",
                      "unit": "GB/s", "n_gpus": 1, "steps": steps, "batch": K, "dtype": "f32", "data": "synthetic",
                      "rows": rows}), flush=True)
    return 0


def stream_ceiling():
    """kh_stream_rate for the updates' read/write mixes (FP32 arrays of 48 Mi
    values, 192 MiB each; 20 timed launches per mix after two untimed)."""
    import ctypes as C
    import torch
    if not torch.cuda.is_available():
        sys.stderr.write("no GPU visible\n")
        return 1
    from khronos_hip import _lib
    lib = _lib.load()
    elements, iters = 48 * 1024 * 1024, 20
    rates = {}
    for nr, nw in ((6, 3), (3, 3), (2, 1), (1, 1)):
        out = C.c_double()
        _lib.check(lib.kh_stream_rate(nr, nw, elements, iters, C.byref(out)))
        rates[f"{nr}r{nw}w"] = round(out.value, 1)
        sys.stderr.write(f"{nr} reads, {nw} writes: {out.value:.1f} GB/s\n")
    print(json.dumps({"metric": "streaming rate of a plain one-pass kernel (GB/s)", "unit": "GB/s",
                      "peak": HBM_PEAK_GBS, "rates": rates,
                      "fraction_of_peak": {k: round(v / HBM_PEAK_GBS, 4) for k, v in rates.items()},
                      "config": {"elements_per_array": elements, "bytes_per_array": elements * 4,
                                 "timed_launches": iters, "dtype": "f32, two values per lane, nontemporal"}}),
          flush=True)
    return 0


def serving_warmup(reqs, budget):
    """Untimed: one short request per grid size of the trace, served and then
    run one at a time, so that both timed runs start with the device pool
    holding those sizes and every kernel variant loaded (the first use of a
    size otherwise pays its allocations inside whichever run comes first)."""
    import dataclasses
    from khronos_hip import serving
    seen, warm = set(), []
    for r in reqs:
        if r.cells not in seen:
            seen.add(r.cells)
            warm.append(dataclasses.replace(r, steps=64, arrival=0.0))
    serving.Server(budget_bytes=budget, max_active=6, chunk=256).run(warm)
    serving.run_one_at_a_time(warm)


def serving_leg(n_requests: int, rate: float, seed: int = 0):
    """A short Poisson stream of the serving workload, continuous batching vs
    one at a time, with the ideal-overlap bound: every request's algorithmic
    bytes (its time steps' compulsory traffic, which grows with the PML share
    of a small grid) moved at the bandwidth the stream's largest grid reaches
    alone, no earlier than the request arrives.  Also printed: the same bound
    in cell updates, every request at the largest grid's Mcell-updates/s."""
    import torch
    from khronos_hip import serving
    reqs = serving.serving_trace(n_requests, seed=seed, rate=rate)
    free = torch.cuda.mem_get_info()[0]
    serving_warmup(reqs, 0.85 * free)
    served = serving.Server(budget_bytes=0.85 * free, max_active=6, chunk=256).run(reqs)
    solo = serving.run_one_at_a_time(reqs)
    bound, bound_cells, r_max, bw = overlap_bound(reqs)
    same = all(np.array_equal(a.flux, b.flux) for a, b in zip(served.results, solo.results))
    return {
        "workload": f"{len(reqs)} requests (grids 64^3..256^3, 1000-3000 steps, FP32 on the FP64 setup, "
                    f"20-frequency flux), Poisson arrivals at {rate} requests/s (seed {seed})",
        "served": round(served.throughput, 1), "one_at_a_time": round(solo.throughput, 1),
        "unit": "Mcell-updates/s", "speedup": round(served.throughput / solo.throughput, 3),
        "ideal_overlap_bound": round(bound, 1), "served_over_bound": round(served.throughput / bound, 3),
        "largest_grid_alone": round(r_max / 1e6, 1), "largest_grid_GBps": round(bw / 1e9, 1),
        "bound_at_largest_grid_rate": round(bound_cells, 1),
        "latency_s": {"served_p50": round(served.latency_pct(50), 3), "served_p99": round(served.latency_pct(99), 3),
                      "solo_p50": round(solo.latency_pct(50), 3), "solo_p99": round(solo.latency_pct(99), 3)},
        "fluxes_identical": bool(same),
    }


def overlap_bound(reqs):
    """(bound, bound in cell updates, largest grid's cell rate, its bytes/s):
    the ideal-overlap throughput of a request stream in Mcell-updates/s.
    The largest grid of the stream is timed alone (320 steps from graphs);
    every request's algorithmic bytes per step (algorithmic_bytes of its
    setup) then move at that grid's bandwidth, no earlier than it arrives."""
    from khronos_hip.workloads import algorithmic_bytes
    big = max(reqs, key=lambda r: r.cells)
    sim = big.make()
    sim.prepare()
    sim.step(32)
    sim.engine.prepare_graphs(320)
    sim.engine.sync()
    t0 = time.perf_counter()
    sim.step(320)
    sim.engine.sync()
    dt_big = time.perf_counter() - t0
    r_max = big.cells * 320 / dt_big
    sim.engine.close()
    step_bytes = {}
    for r in reqs:
        ab = algorithmic_bytes(r.make().build_spec())
        step_bytes[r.rid] = ab["E"] + ab["H"]
    bw = step_bytes[big.rid] * 320 / dt_big  # bytes/s the largest grid moves alone
    t_first = min(r.arrival for r in reqs)
    work = sum(r.cells * r.steps for r in reqs)

    def bound_of(cost):
        finish = max(t_first + sum(cost(r) for r in reqs), max(r.arrival + cost(r) for r in reqs))
        return work / (finish - t_first) / 1e6
    bound = bound_of(lambda r: step_bytes[r.rid] * r.steps / bw)
    bound_cells = bound_of(lambda r: r.cells * r.steps / r_max)
    return bound, bound_cells, r_max, bw


def measured_traffic(kernel_prefix: str, summary: str = "pmc_traffic_strip_waveguide_fp32_fast.md"):
    """HBM bytes per dispatch of a kernel from a committed counter summary
    under profiles/ (FETCH_SIZE and WRITE_SIZE runs scaled by a 1 GiB copy's
    counter readings), or None."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[0].startswith(kernel_prefix):
            try:
                return int(float(cells[5]))
            except ValueError:
                return None
    return None


def serving_bench(args):
    """Serving form: a trace of independent FP32 requests (grids drawn from
    64^3, 96^3, 128^3, 192x192x128, 256^3; 10-cell PML; 1-3 random boxes or
    spheres; Ez dipole; one flux plane at 20 frequencies; 1000-3000 steps),
    served with continuous batching and, as the baseline, one at a time."""
    import torch
    if not torch.cuda.is_available():
        sys.stderr.write("no GPU visible\n")
        return 1
    from khronos_hip import serving
    reqs = serving.serving_trace(args.requests, seed=0, rate=args.rate)
    free, total = torch.cuda.mem_get_info()
    budget = 0.85 * free
    sys.stderr.write(f"serving {len(reqs)} requests ({sum(r.cells * r.steps for r in reqs) / 1e9:.1f} G cell-updates)\n")
    serving_warmup(reqs, budget)

    def progress(label):
        every = max(1, len(reqs) // 16)

        def report(done, total, t):
            if done % every == 0 or done == total:
                sys.stderr.write(f"{label}: {done}/{total} requests done at {t:.1f} s\n")
                sys.stderr.flush()
        return report

    served = serving.Server(budget_bytes=budget, max_active=args.max_active, chunk=args.chunk,
                            progress=progress("served")).run(reqs)
    sys.stderr.write(f"served: makespan {served.makespan:.2f} s; one at a time next\n")
    solo = serving.run_one_at_a_time(reqs, progress=progress("one at a time"))
    sys.stderr.write(f"one at a time: makespan {solo.makespan:.2f} s\n")
    same = all(np.array_equal(a.flux, b.flux) for a, b in zip(served.results, solo.results))
    bound, bound_cells, r_max, bw = overlap_bound(reqs)
    line = {
        "metric": "serving throughput (Mcell-updates/s), continuous batching vs one at a time",
        "value": round(served.throughput, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{len(reqs)} requests, grids 64^3..256^3, 1000-3000 steps, "
                               + ("all arrived at t=0" if math.isinf(args.rate) else
                                  f"Poisson arrivals at {args.rate} requests/s (seed 0)"),
                   "max_active": args.max_active, "chunk_steps": args.chunk,
                   "budget_GB": round(budget / 1e9, 1), "cell_updates": served.cell_updates},
        "one_at_a_time": round(solo.throughput, 2),
        "speedup": round(served.throughput / solo.throughput, 3),
        "latency_s": {"served_p50": round(served.latency_pct(50), 3), "served_p99": round(served.latency_pct(99), 3),
                      "solo_p50": round(solo.latency_pct(50), 3), "solo_p99": round(solo.latency_pct(99), 3)},
        "makespan_s": {"served": round(served.makespan, 3), "solo": round(solo.makespan, 3)},
        "fluxes_identical": bool(same),
        "ideal_overlap_bound": round(bound, 1), "served_over_bound": round(served.throughput / bound, 3),
        "largest_grid_alone": round(r_max / 1e6, 1), "largest_grid_GBps": round(bw / 1e9, 1),
        "bound_at_largest_grid_rate": round(bound_cells, 1),
    }
    print(json.dumps(line), flush=True)
    return 0


def near2far_bench(args):
    """Near-to-far projection: the far field of a 256 x 256-sample z-normal
    surface (a metasurface-scale aperture at resolution 32) at 91 x 73
    (theta, phi) directions and 5 frequencies, from FP32 fields with float64
    DFT accumulators; FP64 Green's-function sums on the device."""
    import torch
    if not torch.cuda.is_available():
        sys.stderr.write("no GPU visible\n")
        return 1
    import khronos_hip as kh
    from khronos_hip import near2far as NF
    n, res = 256, 32
    W = n / res
    theta = np.radians(np.linspace(0.0, 90.0, 91))
    phi = np.radians(np.linspace(0.0, 360.0, 73))
    freqs = [0.9, 0.95, 1.0, 1.05, 1.1]
    mon = kh.Near2FarMonitor(center=[0.0, 0.0, 0.25], size=[W - 1.0 / res, W - 1.0 / res, 0.0], frequencies=freqs,
                             theta=theta, phi=phi, r=1e6)
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.5), component=kh.Ex(),
                            center=[0.0, 0.0, 0.0], size=[W, W, 0.0])]
    sim = kh.Simulation(cell_size=[W + 0.75, W + 0.75, 1.5], resolution=res, boundaries=[[0.375, 0.375]] * 3,
                        sources=src, monitors=[mon], precision="f32")
    sim.step(60)
    sim.engine.sync()
    md = mon.monitor_data
    base, n1, n2, d1, d2, dA = NF.surface_layout(md)
    obs = NF.observation_points(md)
    owner, ids = NF._owner(md)
    call = lambda: owner.engine.near2far(ids, 3, md.normal_sign, base, d1, d2, dA, n1, n2, md.frequencies,
                                         md.medium_eps, md.medium_mu, obs)
    for _ in range(max(1, args.warmup)):
        call()
    k = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(k):
        EH = call()
    wall = (time.perf_counter() - t0) / k
    evals = obs.shape[0] * n1 * n2 * len(freqs)
    line = {
        "metric": "near-to-far surface-sample evaluations/s (full dyadic Green's function, FP64)",
        "value": round(evals / wall / 1e9, 3),
        "unit": "G sample-evals/s",
        "n_gpus": 1, "steps": k, "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"z-normal surface {n1}x{n2} samples, {obs.shape[0]} directions (91 theta x 73 phi) "
                               f"at r = 1e6, {len(freqs)} frequencies; a step is one kh_near2far call "
                               "(host copies of 0.1 MB in, 1.6 MB out included)",
                   "green3d_per_s": round(4 * evals / wall / 1e9, 3)},
    }
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import n2f_oracle as N  # the checker, used here only as the CPU baseline
        mons = md.tangential_E_monitors + md.tangential_H_monitors
        planes = [sim.read_dft(x.monitor_data)[..., :1] for x in mons]
        sub = obs[::2]
        t0 = time.perf_counter()
        N.near2far(planes, 3, md.normal_sign, base, d1, d2, n1, n2, freqs[:1], sub, dA=dA)
        dt = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
        line["cpu_baseline"] = {"value": round(sub.shape[0] * n1 * n2 / dt / 1e9, 4), "unit": "G sample-evals/s",
                                "cores": cores, "kind": "port",
                                "sample": f"{sub.shape[0]} directions x 1 frequency of the same surface "
                                          f"(C oracle orc_near2far, -O3 -fopenmp, {cores} threads)"}
    print(json.dumps(line), flush=True)
    return 0


def adjoint_bend_bench(args):
    """Differentiation form: the gradient of the 2D bend's
    objective (400 x 400 cells, 120 x 120 design parameters, 2001 steps, 3
    frequencies) with respect to all 14400 parameters by the time-reversed
    adjoint (checkpoints every 100 steps), FP32 fields on the FP64 setup.  A
    step is one gradient evaluation: the forward run with checkpoints, then
    the sweep (re-run of each interval and the transposed steps)."""
    import torch
    if not torch.cuda.is_available():
        sys.stderr.write("no GPU visible\n")
        return 1
    print(json.dumps(adjoint_line(args.steps, args.warmup, cpu=not args.no_cpu_baseline)), flush=True)
    return 0


def adjoint_line(steps: int, warmup: int, cpu: bool = True) -> dict:
    """The adjoint bench line (see adjoint_bend_bench); cpu: time the
    frequency-domain gradient on the CPU oracle beside it."""
    from khronos_hip import workloads as W
    opt, rho = W.adjoint_bend(precision="f32", setup_precision="f64")
    for _ in range(max(1, warmup)):
        opt(rho)
    k = max(1, steps)
    tf = tg = 0.0
    for _ in range(k):
        opt.update_design(rho)
        t0 = time.perf_counter()
        opt.forward_run()
        t1 = time.perf_counter()
        opt.adjoint_run()
        opt.calculate_gradient()
        t2 = time.perf_counter()
        tf += t1 - t0
        tg += t2 - t1
    n_steps = opt._fwd_sim.engine.timestep()
    wall = (tf + tg) / k
    line = {
        "metric": "adjoint gradient latency (forward run + time-reversed sweep, all design parameters)",
        "value": round(wall * 1e3, 2),
        "unit": "ms per gradient",
        "n_gpus": 1, "steps": k, "warmup": warmup,
        "ms_per_step": round(wall * 1e3, 2),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"2D waveguide bend 400x400 cells, 120x120 design parameters rho ~ U[0,1] (seed 0), "
                               f"{n_steps} steps, 3 frequencies, checkpoints every 100 steps; FP32 fields, "
                               "FP64 setup and gradient accumulation",
                   "forward_ms": round(tf / k * 1e3, 2), "sweep_ms": round(tg / k * 1e3, 2)},
    }
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_engine import OracleEngine  # the checker, used here only as the CPU baseline
        cores = min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "16") or 16), 16)
        _omp_threads(cores)
        ref, _ = W.adjoint_bend(precision="f32", setup_precision="f64", method="frequency", engine=OracleEngine)
        ref.update_design(rho)
        sys.stderr.write(f"CPU baseline: frequency-domain gradient on the oracle, {cores} threads\n")
        sys.stderr.flush()
        t0 = time.perf_counter()
        ref.forward_run()
        t1 = time.perf_counter()
        sys.stderr.write(f"  forward run {t1 - t0:.1f} s; adjoint runs next\n")
        sys.stderr.flush()
        ref.adjoint_run()
        g_ref = ref.calculate_gradient()
        dt = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": round(dt * 1e3, 1), "unit": "ms per gradient", "cores": cores,
                                "kind": "port",
                                "sample": "one whole gradient by the reference's frequency-domain method "
                                          "(forward run with the design-region D monitors, one adjoint run per "
                                          "frequency, the overlap sum) on the C oracle (-O3 -fopenmp, "
                                          f"{cores} threads), timed once: forward {(t1 - t0) * 1e3:.0f} ms",
                                "gradient_shape": list(np.shape(g_ref))}
    opt._close_forward()
    return line


def metasurface_bench(args):
    """The periodic dispersive metasurface at full size: 256 x 256 x 384
    cells, Bloch x/y, 16-cell PML z, 8 x 8 cylinders with two Lorentz poles
    and chi3, plane-wave Gaussian pulse, complex FP32 fields (real and
    imaginary parts as separate arrays), 100-frequency flux plane with FP64
    accumulators.  A step is one full time step.  The roofline entry is the
    E half step's update: both parts' launches against the half step's
    algorithmic bytes (complex field accesses, real eps^-1 read once)."""
    import torch
    if not torch.cuda.is_available():
        sys.stderr.write("no GPU visible\n")
        return 1
    print(json.dumps(metasurface_line(args.steps, args.warmup, args.graph_steps)), flush=True)
    return 0


def metasurface_line(steps: int, warmup: int, graph_steps: int = 16) -> dict:
    """The metasurface bench line (see metasurface_bench)."""
    import torch
    from khronos_hip.workloads import metasurface, algorithmic_bytes
    sim = metasurface(precision="f32", mode="fast")
    sim.prepare()
    eng = sim.engine
    cells = sim.Nx * sim.Ny * sim.Nz
    ab = algorithmic_bytes(sim.spec)
    eng.set_graph(graph_steps)
    period = dft_period(sim.spec)
    sim.step(warmup)
    extra = (-sim.timestep) % period
    sim.step(extra)
    eng.prepare_graphs(steps)
    eng.sync()
    torch.cuda.synchronize()
    eng.launch_counts()
    t_start = sim.timestep
    t0 = time.perf_counter()
    sim.step(steps)
    eng.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    counts = eng.launch_counts()
    k_steps = min(steps, 100)
    eng.set_timing(k_steps * 24 + 64)
    eng.get_timing()
    sim.step(k_steps)
    eng.sync()
    tm = eng.get_timing()
    eng.set_timing(0)
    # per half step: total ms of each kernel class over the timed steps
    per_step = {k: round(v[0] / k_steps, 5) for k, v in tm.items()}
    e_ms = tm["e_full"][0] / k_steps
    achieved = ab["E"] / (e_ms * 1e-3) / 1e9 if e_ms > 0 else 0.0
    # (the summary may name the kernel with or without its trailing flags)
    traffic = measured_traffic("kh::pml_e_kernel<float, true, false, true, true",
                               "pmc_traffic_metasurface_256x256x384_fp32.md")
    b_alg = (ab["H"] + ab["E"]) / cells
    value = cells * steps / wall / 1e6
    line = {
        "metric": "Mcell-updates/s (complex FP32 3D Yee + CPML + Lorentz ADE + chi3, Bloch x/y)",
        "value": round(value, 2), "unit": "Mcell-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(wall / steps * 1e3, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "metasurface 256x256x384, Bloch x/y (seeded kx, ky), 16-cell CPML z, 8x8 cylinders "
                               "(radii U[0.2,0.4]) with 2 Lorentz poles + chi3 1e-3, plane-wave Gaussian pulse, "
                               "complex FP32 fields, 100-frequency flux plane (f64 accumulators)",
                   "cells": cells, "warmup_steps_run": warmup + extra,
                   "window": {"timesteps": [t_start, t_start + steps - 1], "dft_period": period, **counts},
                   "B_alg_bytes_per_cell_update": round(b_alg, 2),
                   "kernel_ms_per_step": per_step,
                   "launches_per_step": {k: round(v[1] / k_steps, 2) for k, v in tm.items()},
                   "hbm_fraction_whole_step": round(cells * steps / wall * b_alg / 1e9 / HBM_PEAK_GBS, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_achieved": round(traffic / (e_ms * 1e-3) / 1e9, 1) if traffic and e_ms > 0 else None,
                     "traffic_frac": (round(traffic / (e_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                      if traffic and e_ms > 0 else None),
                     "kernel": "E update of one half step (pml_e_kernel, real and imaginary parts)",
                     "alg_bytes_per_launch": ab["E"], "kernel_ms": round(e_ms, 5)},
    }
    sim.engine.close()
    return line


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mode", default="fast", choices=["fast", "exact"])
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"],
                    help="field precision of the strip waveguide (the headline is f32; f64 for the comparison)")
    ap.add_argument("--graph-steps", type=int, default=16, help="steps per captured hipGraph (0: eager launches)")
    ap.add_argument("--stream-ceiling", action="store_true",
                    help="streaming rate of a plain one-pass kernel for several read/write mixes (the practical "
                         "HBM ceiling the update kernels are set against)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of 1 GiB five times (counter calibration)")
    ap.add_argument("--serving", action="store_true", help="serving form instead of the headline benchmark")
    ap.add_argument("--near2far", action="store_true", help="near-to-far projection instead of the headline benchmark")
    ap.add_argument("--metasurface", action="store_true",
                    help="the periodic dispersive metasurface (complex fields) instead of the headline benchmark")
    ap.add_argument("--dipole", action="store_true",
                    help="the 80^3 FP64 dipole with its kernel times (pass --steps 490 --warmup 10 for its run)")
    ap.add_argument("--concurrency", action="store_true",
                    help="serving: same-shape requests concurrently vs one grid of the same cells")
    ap.add_argument("--batch", type=int, default=6, help="requests per group (--concurrency)")
    ap.add_argument("--grids", type=int, nargs="+", default=[64, 96, 128], help="cube edges (--concurrency)")
    ap.add_argument("--adjoint", action="store_true",
                    help="time-reversed adjoint gradient of the 2D bend instead of the headline benchmark")
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--max-active", type=int, default=6)
    ap.add_argument("--chunk", type=int, default=256)
    ap.add_argument("--rate", type=float, default=math.inf, help="serving: Poisson arrival rate (requests/s)")
    ap.add_argument("--no-legs", action="store_true",
                    help="headline only: skip the legs (80^3 dipole, serving stream, metasurface, adjoint)")
    ap.add_argument("--leg-requests", type=int, default=24, help="requests of the serving leg")
    ap.add_argument("--leg-rate", type=float, default=8.0, help="Poisson arrival rate of the serving leg")
    args = ap.parse_args(argv)
    if args.serving:
        return serving_bench(args)
    if args.near2far:
        return near2far_bench(args)
    if args.adjoint:
        return adjoint_bend_bench(args)
    if args.dipole:
        return dipole_bench(args)
    if args.concurrency:
        return concurrency_bench(args)
    if args.metasurface:
        return metasurface_bench(args)
    if args.stream_ceiling:
        return stream_ceiling()
    if args.calibrate_copy:
        import torch
        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": src.numel() * 4, "bytes_written": dst.numel() * 4}))
        return 0
    if args.gpus != 1:
        sys.stderr.write("bench.py measures one GPU: the time step runs on a single MI355X "
                         "(multi-GPU runs are not part of this build); use --gpus 1\n")
        return 2

    # the host cores this process may use (at most 16, the GPU box's share)
    cores = min(len(os.sched_getaffinity(0)), int(os.environ.get("OMP_NUM_THREADS", "16") or 16), 16)
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))

    import torch
    if not torch.cuda.is_available():
        sys.stderr.write("no GPU visible\n")
        return 1

    from khronos_hip.workloads import strip_waveguide, algorithmic_bytes

    sim = strip_waveguide(precision=args.precision, mode=args.mode)
    ctype = "float" if args.precision == "f32" else "double"
    pname = "FP32" if args.precision == "f32" else "FP64"
    sim.prepare()
    eng = sim.engine
    cells = sim.Nx * sim.Ny * sim.Nz
    ab = algorithmic_bytes(sim.spec)

    eng.set_graph(args.graph_steps)
    # warm-up: W steps, then on to the next multiple of the DFT period, so
    # that the timed window starts on a step where the flux planes sample
    # (the window then holds its share of DFT launches)
    period = dft_period(sim.spec)
    sim.step(args.warmup)
    extra = (-sim.timestep) % period
    sim.step(extra)
    eng.prepare_graphs(args.steps)  # capture outside the timed window
    eng.sync()
    torch.cuda.synchronize()
    eng.launch_counts()  # clear
    t_start = sim.timestep

    # timed region: K steps as the library runs them (captured step graphs)
    t0 = time.perf_counter()
    sim.step(args.steps)
    eng.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    counts = eng.launch_counts()

    # per-kernel durations: the start/end events every launch stamps on the
    # library's stream (events need eager launches, so this is a separate run)
    k_steps = min(args.steps, 200)
    eng.set_timing(k_steps * 16 + 64)
    eng.get_timing()  # clear
    sim.step(k_steps)
    eng.sync()
    tm = eng.get_timing()
    eng.set_timing(0)

    value = cells * args.steps / wall / 1e6
    avg = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in tm.items()}
    # dominant kernel: the E update over the whole grid (FAST: the PML kernel,
    # whose sigma = 0 branch is the inner update; EXACT: the cascade kernel)
    if args.mode == "fast" and avg["e_inner"] > 0:
        dom, dom_bytes, dom_name = avg["e_inner"], ab["E_inner"], f"inner_e_kernel<{ctype},true,true>"
        trace_name = f"kh::inner_e_kernel<{ctype}, true, true>"
    elif args.mode == "fast":
        # the library reads a byte per cell into a value table when the
        # per-voxel eps^-1 (with the ghost value 1) holds at most 256 values
        vals = np.unique(np.concatenate([np.ravel(a) for a in sim.spec.eps_inv if a is not None] + [np.ones(1)]))
        lut = os.environ.get("KH_LUT", "1") != "0" and len(vals) <= 256
        tmpl = f"{ctype}, true, false, " + ("true" if lut else "false")
        dom, dom_bytes, dom_name = avg["e_full"], ab["E"], "pml_e_kernel<" + tmpl.replace(" ", "") + ">"
        # (the summary may name the kernel with or without its trailing
        # single-part flag)
        trace_name = "kh::pml_e_kernel<" + tmpl
    else:
        dom, dom_bytes, dom_name = avg["e_full"], ab["E"], f"cascade_e_kernel<{ctype}>"
        trace_name = f"kh::cascade_e_kernel<{ctype}>"
    achieved = dom_bytes / (dom * 1e-3) / 1e9 if dom > 0 else 0.0
    traffic = measured_traffic(trace_name) if args.precision == "f32" else None
    b_alg = (ab["H"] + ab["E"]) / cells
    step_kernel_ms = sum(v[0] for v in tm.values()) / k_steps
    # the H update beside it (FAST whole-grid PML kernel, scalar mu)
    h_ms = avg["h_full"]
    roofline_h = None
    if args.mode == "fast" and h_ms > 0 and avg["h_inner"] == 0:
        h_traffic = measured_traffic(f"kh::pml_h_kernel<{ctype}, false, false, false") if args.precision == "f32" \
            else None
        roofline_h = {"kernel": f"pml_h_kernel<{ctype},false,false,false>", "alg_bytes_per_launch": ab["H"],
                      "kernel_ms": round(h_ms, 5), "achieved": round(ab["H"] / (h_ms * 1e-3) / 1e9, 1),
                      "frac": round(ab["H"] / (h_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": h_traffic,
                      "traffic_frac": (round(h_traffic / (h_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                       if h_traffic else None)}

    line = {
        "metric": f"Mcell-updates/s ({pname} 3D Yee + CPML time step)",
        "value": round(value, 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {
            "workload": "strip waveguide 384x192x160, 12-cell CPML, anisotropically smoothed diagonal "
                        "per-voxel eps, Ez Gaussian-pulse planar source with Gaussian transverse profile, "
                        f"2 flux planes x 50 freqs (f64 DFT accumulators), {pname} fields",
            "cells": cells,
            "mode": args.mode,
            "graph_steps": args.graph_steps,
            "warmup_steps_run": args.warmup + extra,
            "window": {"timesteps": [t_start, t_start + args.steps - 1], "dft_period": period, **counts},
            "B_alg_bytes_per_cell_update": round(b_alg, 2),
            "kernel_ms": {k: round(v, 5) for k, v in avg.items()},
            "kernel_ms_per_step": round(step_kernel_ms, 5),
            "hbm_fraction_whole_step": round(cells * args.steps / wall * b_alg / 1e9 / HBM_PEAK_GBS, 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_achieved": round(traffic / (dom * 1e-3) / 1e9, 1) if traffic and dom > 0 else None,
            "traffic_frac": round(traffic / (dom * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic and dom > 0 else None,
            "kernel": dom_name,
            "alg_bytes_per_launch": dom_bytes,
            "kernel_ms": round(dom, 5),
        },
    }
    if roofline_h:
        line["roofline_h_update"] = roofline_h
    if not args.no_legs:
        sys.stderr.write("config-1 leg (80^3 FP64 dipole)\n")
        line["config1_dipole_80_f64"] = config1_leg()
        sys.stderr.write(f"serving leg ({args.leg_requests} Poisson requests)\n")
        line["serving"] = serving_leg(args.leg_requests, args.leg_rate)
        sys.stderr.write("metasurface leg (256x256x384 complex FP32)\n")
        m = metasurface_line(200, 20)
        line["metasurface"] = {"workload": m["config"]["workload"], "value": m["value"], "unit": m["unit"],
                               "ms_per_step": m["ms_per_step"], "steps": m["steps"],
                               "kernel_ms_per_step": m["config"]["kernel_ms_per_step"],
                               "roofline_frac": m["roofline"]["frac"], "traffic_frac": m["roofline"]["traffic_frac"],
                               "full_line": "python bench.py --metasurface"}
        sys.stderr.write("adjoint leg (400x400 bend, all 14400 design parameters)\n")
        a = adjoint_line(3, 1, cpu=False)
        line["adjoint_bend"] = {"workload": a["config"]["workload"], "value": a["value"], "unit": a["unit"],
                                "forward_ms": a["config"]["forward_ms"], "sweep_ms": a["config"]["sweep_ms"],
                                "full_line": "python bench.py --adjoint (with its timed CPU baseline)"}
    if not args.no_cpu_baseline:
        spec = sim.spec
        eng.close()
        del sim, eng
        sys.stderr.write(f"CPU baseline ({cores} threads, then 1)\n")
        line["cpu_baseline"] = cpu_baseline(spec, cores)
        line["cpu_baseline_1core"] = cpu_baseline(spec, 1, budget_s=10.0)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
