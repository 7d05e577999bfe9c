"""Writes strip_waveguide_flux_fp64_oracle.json: both flux spectra of the
full-size strip waveguide (384 x 192 x 160 cells, 12-cell PML, anisotropically
smoothed eps 12.1 strip in eps 2.085, Ez Gaussian-pulse planar source with a
Gaussian transverse envelope, two x-normal flux planes at 50 frequencies,
4000 steps) computed by the CPU oracle in FP64.

The oracle (oracle/khronos_oracle.c, the reference CPU path's operation order)
steps the fields; the host setup (grid snapping, smoothing, waveforms) is the
FP64 setup every engine receives; the flux is the host FP64 reduction of the
oracle's DFT accumulators (src/Monitors/FluxMonitor.jl:92-156).  Runs in a
few minutes on 8 cores:

    OMP_NUM_THREADS=8 python tests/golden/make_strip_waveguide_flux.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "khronos.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

from khronos_hip.monitors import flux_from_dft  # noqa: E402
from khronos_hip.workloads import strip_waveguide  # noqa: E402
from oracle_engine import OracleEngine  # noqa: E402

STEPS = 4000
OUT = os.path.join(HERE, "strip_waveguide_flux_fp64_oracle.json")


def main():
    sim = strip_waveguide(precision="f64", mode="exact", engine=OracleEngine)
    sim.prepare()
    t0 = time.perf_counter()
    done = 0
    while done < STEPS:
        n = min(250, STEPS - done)
        sim.step(n)
        done += n
        print(f"{done} steps, {time.perf_counter() - t0:.0f} s", flush=True)
    fluxes = []
    for fm in sim.monitors:
        md = fm.monitor_data
        mons = md.tangential_E_monitors + md.tangential_H_monitors
        arrs = [sim.read_dft(m.monitor_data) for m in mons]
        fluxes.append(flux_from_dft(md, *arrs, [m.monitor_data.gv for m in mons]).tolist())
    out = {
        "workload": "strip waveguide 384x192x160, 12-cell PML, anisotropic smoothing, FP64, CPU oracle",
        "steps": STEPS,
        "frequencies": [float(f) for f in sim.monitors[0].frequencies],
        "monitor_x": [-6.0, 6.0],
        "flux": fluxes,
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {OUT} ({time.perf_counter() - t0:.0f} s)")


if __name__ == "__main__":
    main()
