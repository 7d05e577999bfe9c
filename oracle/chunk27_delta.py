"""Delta between the reference's default 27-chunk plan arithmetic and the
single-chunk semantics the library implements, on the CPU oracle (quirk Q1:
src/Chunking.jl:907-923, :1090-1122; src/Kernels/Helpers.jl:200-223).

    OMP_NUM_THREADS=8 python oracle/chunk27_delta.py [dipole] [strip STEPS]

dipole: the 80^3 FP64 dipole, 500 steps (fields and the 5-frequency DFT
plane).  strip: the 384 x 192 x 160 strip waveguide in FP64 for STEPS steps
(default 1000; fields and both flux planes every 250 steps).  Prints the
relative L2 difference of all six field components and of the monitors,
and the field energy of both runs.
"""
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "khronos.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

import khronos_hip as kh  # noqa: E402
from oracle_engine import OracleEngine, chunk27_intervals  # noqa: E402

COMPS = [kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz]


def _rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float("nan")


def _report(tag, a, b):
    fa = np.concatenate([a.read_field(C()).ravel() for C in COMPS])
    fb = np.concatenate([b.read_field(C()).ravel() for C in COMPS])
    with np.errstate(all="ignore"):
        ea, eb = float(np.sum(fa * fa)), float(np.sum(fb * fb))
        mon = [_rel(b.engine.read_dft(i), a.engine.read_dft(i)) for i in range(len(a.spec.dfts))]
        frel = _rel(fb, fa)
    print(f"{tag}: fields rel-L2 {frel:.3e}, energy single {ea:.4e} chunk27 {eb:.4e}, "
          f"monitors rel-L2 max {max(mon) if mon else float('nan'):.3e}", flush=True)


def pair(make):
    a, b = make(), make()
    a.prepare()
    b.prepare()
    e, s = chunk27_intervals(b)
    b.engine.set_chunk27(e, s)
    return a, b


def dipole():
    from khronos_hip.workloads import dipole_80
    a, b = pair(lambda: dipole_80(precision="f64", mode="exact", engine=OracleEngine))
    print("dipole 80^3 FP64, PML intervals", chunk27_intervals(b), flush=True)
    for n in (100, 200, 300, 500):
        a.step(n - a.timestep)
        b.step(n - b.timestep)
        _report(f"  step {n}", a, b)


def strip(steps):
    from khronos_hip.workloads import strip_waveguide
    a, b = pair(lambda: strip_waveguide(precision="f64", mode="exact", engine=OracleEngine))
    print("strip waveguide 384x192x160 FP64, PML intervals", chunk27_intervals(b), flush=True)
    t0 = time.perf_counter()
    for n in range(250, steps + 1, 250):
        a.step(n - a.timestep)
        b.step(n - b.timestep)
        _report(f"  step {n} ({time.perf_counter() - t0:.0f} s)", a, b)


if __name__ == "__main__":
    args = sys.argv[1:] or ["dipole"]
    if "dipole" in args:
        dipole()
    if "strip" in args:
        i = args.index("strip")
        strip(int(args[i + 1]) if len(args) > i + 1 else 1000)
