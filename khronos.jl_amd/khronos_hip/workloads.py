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
* ``metasurface`` -- 8 x 8 lattice (a = 1 um) at 32 cells/um, 256 x 256 x 384
  cells, Bloch-periodic x/y with (kx, ky) ~ U(-pi/a, pi/a), 16-cell PML in z;
  cylinders of radius U[0.2, 0.4] a, height 0.5 um, eps_inf 2 with two
  Lorentz poles (omega_0 ~ U[0.8, 1.5], gamma 0.05, per-cylinder sigma ~
  U[0.5, 2]) and chi3 = 1e-3; a plane-wave Gaussian pulse whose transverse k
  matches the Bloch phase; one z-normal flux plane at 100 frequencies;
  complex FP32 fields, FP64 DFT accumulators (all draws from ``seed``).
* ``adjoint_bend`` -- the 2D 90-degree waveguide bend of the reference's
  inverse-design example (examples/waveguide_bend_inverse_design.jl:16-109)
  at resolution 40 with 2 um padding: 400 x 400 cells, 40-cell PML, 0.5 um
  eps 12.1 guides in eps 2.085 (the strip waveguide's silicon and
  cladding), a 3 x 3 um design region of 120 x 120 parameters rho ~ U[0, 1]
  (``seed``), eps in [2.085, 12.1], Ez line
  source (Gaussian pulse, fcen 0.625, fwidth 0.5), objective sum |Ez|^2 on
  the output guide's line at 3 frequencies; 2000 steps (until = 25),
  checkpoints every 100 steps for the time-reversed adjoint.
"""
from __future__ import annotations

import numpy as np

from . import grid as G
from .boundaries import PML, Bloch
from .geometry import Box, Cylinder, Material, Object
from .monitors import DFTMonitor, FluxMonitor
from .simulation import Simulation
from .smoothing import AnisotropicSmoothing, NoSmoothing
from .sources import GaussianProfileSource, GaussianPulseSource, PlaneWaveSource, UniformSource
from .susceptibility import LorentzianSusceptibility


def dipole_80(precision="f64", mode="fast", engine=None, dft_accum="f64"):
    src = [UniformSource(time_profile=GaussianPulseSource(fcen=1.0, fwidth=0.2), component=G.Ez(),
                         center=[0.0, 0.0, 0.0], size=[0.0, 0.0, 0.0])]
    mon = [DFTMonitor(component=G.Ez(), center=[0.0, 0.0, 0.0], size=[4.0, 4.0, 0.0],
                      frequencies=list(np.linspace(0.9, 1.1, 5)))]
    return Simulation(cell_size=[4.0, 4.0, 4.0], resolution=20, sources=src, monitors=mon,
                      boundaries=[[0.5, 0.5]] * 3, precision=precision, mode=mode, engine=engine,
                      dft_accum=dft_accum)


def strip_waveguide(precision="f32", mode="fast", engine=None, amplitude=1.0, n_freq=50,
                    flux=True, waist=1.0, dft_accum="f64", smoothing=True, setup_precision=None):
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
                      engine=engine, dft_accum=dft_accum, setup_precision=setup_precision,
                      subpixel_smoothing=AnisotropicSmoothing() if smoothing else NoSmoothing())


def metasurface(precision="f32", mode="fast", engine=None, seed=0, n_freq=100, dft_accum="f64", nz=384,
                lattice=8, res=32):
    import math
    rng = np.random.default_rng(seed)
    a = 1.0
    kx, ky = (float(v) for v in rng.uniform(-math.pi / a, math.pi / a, 2))
    w0 = [float(v) for v in rng.uniform(0.8, 1.5, 2)]
    L = lattice * a
    geo = []
    for i in range(lattice):
        for j in range(lattice):
            r = float(rng.uniform(0.2, 0.4)) * a
            s1, s2 = (float(v) for v in rng.uniform(0.5, 2.0, 2))
            poles = [LorentzianSusceptibility(w0[0], 0.05, s1), LorentzianSusceptibility(w0[1], 0.05, s2)]
            c = [-L / 2 + (i + 0.5) * a, -L / 2 + (j + 0.5) * a, 0.0]
            geo.append(Object(Cylinder(c, r, 0.5, axis=2), Material(ε=2.0, chi3=1e-3, susceptibilities=poles)))
    fcen, fwidth = 1.0, 0.5
    k = 2 * math.pi * fcen
    # the sheet's phase exp(-i k.r) repeats with exp(-i k_t L): transverse k = -Bloch k
    kv = [-kx / k, -ky / k, math.sqrt(max(0.0, 1 - (kx * kx + ky * ky) / (k * k)))]
    cz = nz / res
    src = [PlaneWaveSource(time_profile=GaussianPulseSource(fcen=fcen, fwidth=fwidth), center=[0.0, 0.0, -cz / 4],
                           size=[math.inf, math.inf, 0.0], k_vector=kv)]
    freqs = list(np.linspace(fcen - fwidth / 2, fcen + fwidth / 2, n_freq))
    mons = [FluxMonitor(center=[0.0, 0.0, cz / 4], size=[L, L, 0.0], frequencies=freqs)]
    pml = 16 / res
    bcs = [[Bloch(k=kx), Bloch(k=kx)], [Bloch(k=ky), Bloch(k=ky)], [PML(), PML()]]
    return Simulation(cell_size=[L, L, cz], resolution=res, sources=src, geometry=geo, monitors=mons,
                      boundaries=[[0.0, 0.0], [0.0, 0.0], [pml, pml]], boundary_conditions=bcs,
                      precision=precision, mode=mode, engine=engine, dft_accum=dft_accum)


def algorithmic_bytes(spec) -> dict:
    """Compulsory HBM bytes per launch of the H and E update kernels.

    Per cell and half step, with p the number of axes on which the cell has a
    non-zero PML sigma: interior (p = 0) cells read the differentiated field
    and the updated field and write the updated field (9 accesses); PML cells
    read the 3 differentiated components, read+write D/B (6), write E/H (3),
    read E/H for the p own-axis components, read+write W and U for p
    components each (12 + 5p).  A per-voxel inverse permittivity (or
    permeability) adds 3 reads.  Field accesses are of the field precision,
    twice that for complex (Bloch) fields; material reads are real.
    """
    s_mat = np.dtype(spec.field_real or spec.real).itemsize
    s = s_mat * (2 if spec.complex_fields else 1)
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
        mat = 3 * sum(counts) if any(m is not None for m in minv) else 0
        out[kind] = acc * s + mat * s_mat
        out[kind + "_counts"] = counts
        # the inner-box kernel alone: p = 0 cells, 9 accesses (+3 per-voxel)
        out[kind + "_inner"] = (9 * s + (3 * s_mat if mat else 0)) * counts[0]
    return out


def adjoint_bend(precision="f32", mode="fast", seed=0, setup_precision=None, method="time_reversed",
                 checkpoint_every=100, n_design=120, resolution=40, pad=2.0, steps=2000, engine=None):
    """(OptimizationProblem, rho) of the 2D bend (see the module docstring)."""
    import torch

    from . import adjoint as A
    eps_si, eps_sio2 = 12.1, 2.085
    wg, pml = 0.5, 1.0
    design = n_design / resolution
    fcen, fwidth = 0.625, 0.5
    dz = 1.0 / resolution
    cell = design + 2 * pad + 2 * pml + 2 * wg
    geo = [Object(Box([-(design / 2 + pad + wg / 2), 0.0, 0.0], [pad + wg + pml, wg, dz]), Material(ε=eps_si)),
           Object(Box([0.0, design / 2 + pad + wg / 2, 0.0], [wg, pad + wg + pml, dz]), Material(ε=eps_si)),
           Object(Box([0.0, 0.0, 0.0], [cell + 1, cell + 1, dz + 1]), Material(ε=eps_sio2))]
    src = [UniformSource(time_profile=GaussianPulseSource(fcen=fcen, fwidth=fwidth), component=G.Ez(),
                         center=[-(design / 2 + pad), 0.0, 0.0], size=[0.0, wg, 0.0])]
    sim = Simulation(cell_size=[cell, cell, 0.0], resolution=resolution, sources=src, geometry=geo,
                     boundaries=[[pml, pml], [pml, pml]], precision=precision, mode=mode,
                     setup_precision=setup_precision, engine=engine)
    dr = A.DesignRegion(volume=G.Volume([0.0, 0.0, 0.0], [design, design, 0.0]),
                        design_parameters=np.full(n_design * n_design, 0.5), grid_size=(n_design, n_design),
                        ε_min=eps_sio2, ε_max=eps_si)
    obj = A.FourierFieldsObjective(volume=G.Volume([0.0, design / 2 + pad, 0.0], [2 * wg, 0.0, 0.0]),
                                   component=G.Ez())
    until = steps * float(sim.dt_n)
    opt = A.OptimizationProblem(sim, [lambda ez: torch.sum(torch.abs(ez) ** 2)], [obj], [dr],
                                [0.55, 0.625, 0.7], until=until, method=method,
                                checkpoint_every=checkpoint_every)
    rho = np.random.default_rng(seed).random(n_design * n_design)
    return opt, rho
