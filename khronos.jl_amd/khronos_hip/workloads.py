"""Named simulation set-ups used by the benchmark, the smoke run and the tests.

* ``dipole_80`` -- 4 x 4 x 4 vacuum cell at resolution 20 (80^3), 10-cell PML,
  one Ez point current with a Gaussian pulse (fcen 1.0, fwidth 0.2), Courant
  0.5, one z-normal DFT plane at 5 frequencies (the reference's dipole
  benchmark geometry, benchmark/dipole.jl:33-52, at size 4 / resolution 20).
* ``strip_waveguide`` -- 19.2 x 9.6 x 8.0 um at 20 cells/um (384 x 192 x 160),
  12-cell PML, eps 12.1 strip of 0.5 x 0.22 um along x in eps 2.085 cladding
  with anisotropic subpixel smoothing (diagonal per-component eps^-1),
  Ez planar Gaussian-pulse source (fcen 1/1.55, fwidth 0.2 fcen) with a
  Gaussian transverse envelope at x = -7 um, two x-normal flux planes at 50
  frequencies.
"""
from __future__ import annotations

import numpy as np

from . import grid as G
from .geometry import Box, Material, Object
from .monitors import DFTMonitor, FluxMonitor
from .simulation import Simulation
from .smoothing import AnisotropicSmoothing, NoSmoothing
from .sources import GaussianProfileSource, GaussianPulseSource, UniformSource


def dipole_80(precision="f64", mode="fast", engine=None, dft_accum="f64"):
    src = [UniformSource(time_profile=GaussianPulseSource(fcen=1.0, fwidth=0.2), component=G.Ez(),
                         center=[0.0, 0.0, 0.0], size=[0.0, 0.0, 0.0])]
    mon = [DFTMonitor(component=G.Ez(), center=[0.0, 0.0, 0.0], size=[4.0, 4.0, 0.0],
                      frequencies=list(np.linspace(0.9, 1.1, 5)))]
    return Simulation(cell_size=[4.0, 4.0, 4.0], resolution=20, sources=src, monitors=mon,
                      boundaries=[[0.5, 0.5]] * 3, precision=precision, mode=mode, engine=engine,
                      dft_accum=dft_accum)


def strip_waveguide(precision="f32", mode="fast", engine=None, amplitude=1.0, n_freq=50,
                    flux=True, waist=1.0, dft_accum="f64", smoothing=True):
    fcen = 1 / 1.55
    src = [GaussianProfileSource(time_profile=GaussianPulseSource(fcen=fcen, fwidth=0.2 * fcen),
                                 component=G.Ez(), center=[-7.0, 0.0, 0.0], size=[0.0, 8.4, 6.8],
                                 waist=waist, amplitude=amplitude)]
    geo = [Object(Box([0.0, 0.0, 0.0], [30.0, 0.5, 0.22]), Material(ε=12.1)),
           Object(Box([0.0, 0.0, 0.0], [30.0, 30.0, 30.0]), Material(ε=2.085))]
    freqs = list(np.linspace(0.8 * fcen, 1.2 * fcen, n_freq))
    mons = []
    if flux:
        mons = [FluxMonitor(center=[-6.0, 0.0, 0.0], size=[0.0, 8.4, 6.8], frequencies=freqs),
                FluxMonitor(center=[6.0, 0.0, 0.0], size=[0.0, 8.4, 6.8], frequencies=freqs)]
    return Simulation(cell_size=[19.2, 9.6, 8.0], resolution=20, sources=src, geometry=geo,
                      monitors=mons, boundaries=[[0.6, 0.6]] * 3, precision=precision, mode=mode,
                      engine=engine, dft_accum=dft_accum,
                      subpixel_smoothing=AnisotropicSmoothing() if smoothing else NoSmoothing())


def algorithmic_bytes(spec) -> dict:
    """Compulsory HBM bytes per launch of the H and E update kernels.

    Per cell and half step, with p the number of axes on which the cell has a
    non-zero PML sigma: interior (p = 0) cells read the differentiated field
    and the updated field and write the updated field (9 accesses); PML cells
    read the 3 differentiated components, read+write D/B (6), write E/H (3),
    read E/H for the p own-axis components, read+write W and U for p
    components each (12 + 5p).  A per-voxel inverse permittivity (or
    permeability) adds 3 reads.  s = bytes of the backend real.
    """
    s = np.dtype(spec.real).itemsize
    N = spec.N
    out = {}
    for kind, key, minv in (("H", "B", spec.mu_inv), ("E", "D", spec.eps_inv)):
        if spec.sigma is None:
            nz = [np.zeros(N[a], bool) for a in range(3)]
        else:
            nz = [np.asarray(spec.sigma[key][a])[0:2 * N[a]:2] != 0 for a in range(3)]
        p = (nz[0][:, None, None].astype(int) + nz[1][None, :, None] + nz[2][None, None, :])
        counts = [int(np.count_nonzero(p == k)) for k in range(4)]
        acc = 9 * counts[0] + sum((12 + 5 * k) * counts[k] for k in (1, 2, 3))
        if any(m is not None for m in minv):
            acc += 3 * sum(counts)
        out[kind] = acc * s
        out[kind + "_counts"] = counts
        # the inner-box kernel alone: p = 0 cells, 9 accesses (+3 per-voxel)
        out[kind + "_inner"] = (9 + (3 if any(m is not None for m in minv) else 0)) * counts[0] * s
    return out
