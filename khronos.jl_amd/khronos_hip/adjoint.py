"""Adjoint gradients of DFT-field objectives (differentiation form).

Mirrors the reference's adjoint module (src/Adjoint/): a forward run with
design-region D-field DFT monitors, an adjoint run driven by sources built
from the objective's derivative, and the overlap integral

    df/drho_k = sum_voxels w_ki Re[ deps_inv/drho * D_adj . D_fwd ],
    deps_inv/drho = -(eps_max - eps_min) / eps(rho)^2

contracted onto the design grid with the transpose of the bilinear
interpolation (src/Adjoint/GradientKernel.jl:35-109).

* DesignRegion, bilinear weights, update_design!  -- src/Adjoint/DesignRegion.jl:14-288
* D monitors on the design region                 -- src/Adjoint/DesignRegion.jl:290-312
* FourierFieldsObjective                          -- src/Adjoint/ObjectiveQuantities.jl:250-345
* adjoint source scale                            -- src/Adjoint/AdjointSourceScale.jl:61-131
* forward / adjoint runs, Wirtinger derivative    -- src/Adjoint/OptimizationProblem.jl:104-374
* central finite-difference check                 -- src/Adjoint/GradientKernel.jl:117-152
* reverse-mode rule (ChainRules rrule)            -- src/Adjoint/ChainRulesIntegration.jl:20-60;
  here a torch.autograd.Function (``simulate_torch``)

Both runs step on the GPU through libkhronos_hip (the engine of the
simulation); the objective's derivative uses torch autograd on the real and
imaginary parts, as the reference uses ForwardDiff.

``method="time_reversed"`` replaces the second run: the forward run keeps
field checkpoints every ``checkpoint_every`` steps on the device, and the
library's discrete transpose of those steps (``kh_adjoint_gradient``) returns
dJ/d(eps^-1) per voxel of the design region, exact for the discrete objective
(multi-frequency objectives included), which is then contracted onto the
design grid with the same interpolation weights.
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from . import grid as G
from .monitors import DFTMonitor, init_dft_monitor
from .simulation import Simulation, run, stop_when_dft_decayed
from .sources import SourceData


# ------------------------------------------------------------ design region
@dataclass
class DesignRegion:
    """src/Adjoint/DesignRegion.jl:14-35: a 2-D design grid (extruded along z)
    whose parameters rho in [0, 1] set eps = eps_min + rho (eps_max - eps_min)."""

    volume: G.Volume
    design_parameters: np.ndarray
    grid_size: tuple
    ε_min: float = 1.0
    ε_max: float = 12.0
    weights: dict = field(default_factory=dict)   # comp name -> (yee_idx, design_idx, w) arrays, 0-based
    gv: dict = field(default_factory=dict)        # comp name -> GridVolume


def design_grid_coords(dr: DesignRegion):
    """src/Adjoint/DesignRegion.jl:41-55."""
    c, s = dr.volume.center, dr.volume.size
    nx, ny = dr.grid_size
    x0, x1 = c[0] - s[0] / 2, c[0] + s[0] / 2
    y0, y1 = c[1] - s[1] / 2, c[1] + s[1] / 2
    xs = np.linspace(x0, x1, nx) if nx > 1 else np.array([c[0]])
    ys = np.linspace(y0, y1, ny) if ny > 1 else np.array([c[1]])
    return xs, ys


def bilinear_weights(xs, ys, sim, comp: G.Field, dr: DesignRegion):
    """Design-grid -> Yee-point weights of one component
    (src/Adjoint/DesignRegion.jl:63-133), 0-based (yee, design, w)."""
    gv = G.grid_volume(sim, dr.volume, comp)
    org = G.component_origin(sim.cell_center, sim.cell_size_n, sim.delta_n, comp)
    dx, dy = sim.delta_n[0], sim.delta_n[1]
    nxd, nyd = len(xs), len(ys)
    yee, des, wts = [], [], []
    for iy in range(1, gv.Ny + 1):
        for ix in range(1, gv.Nx + 1):
            px = org[0] + (ix + gv.start_idx[0] - 2) * dx
            py = org[1] + (iy + gv.start_idx[1] - 2) * dy
            yi = (iy - 1) * gv.Nx + (ix - 1)
            if nxd == 1:
                ixl = ixh = 1
                wx = 1.0
            else:
                fx = min(max((px - xs[0]) / (xs[1] - xs[0]) + 1.0, 1.0), float(nxd))
                ixl = min(int(math.floor(fx)), nxd - 1)
                ixh = ixl + 1
                wx = fx - ixl
            if nyd == 1:
                iyl = iyh = 1
                wy = 1.0
            else:
                fy = min(max((py - ys[0]) / (ys[1] - ys[0]) + 1.0, 1.0), float(nyd))
                iyl = min(int(math.floor(fy)), nyd - 1)
                iyh = iyl + 1
                wy = fy - iyl
            for di, w in (((iyl - 1) * nxd + ixl, (1 - wx) * (1 - wy)), ((iyl - 1) * nxd + ixh, wx * (1 - wy)),
                          ((iyh - 1) * nxd + ixl, (1 - wx) * wy), ((iyh - 1) * nxd + ixh, wx * wy)):
                if abs(w) > 1e-15:
                    yee.append(yi)
                    des.append(di - 1)
                    wts.append(w)
    return gv, (np.array(yee, dtype=np.int64), np.array(des, dtype=np.int64), np.array(wts))


def init_design_region(sim, dr: DesignRegion) -> DesignRegion:
    """src/Adjoint/DesignRegion.jl:135-149."""
    xs, ys = design_grid_coords(dr)
    for C in (G.Ex, G.Ey, G.Ez):
        gv, w = bilinear_weights(xs, ys, sim, C(), dr)
        dr.gv[C.name] = gv
        dr.weights[C.name] = w
    return dr


def interpolate_to_yee(rho, w, n_yee):
    """src/Adjoint/DesignRegion.jl:157-168."""
    yee, des, wts = w
    out = np.zeros(n_yee)
    np.add.at(out, yee, rho[des] * wts)
    return out


def update_design(spec, sim, dr: DesignRegion, rho):
    """Write eps_inv of the design region into the spec's per-voxel arrays
    (src/Adjoint/DesignRegion.jl:186-288): the 2-D map is extruded over the
    region's z range, clamped to the update domain and the array."""
    R = spec.real
    dr.design_parameters = np.asarray(rho, dtype=np.float64).copy()
    N = spec.N
    for a, C in enumerate((G.Ex, G.Ey, G.Ez)):
        gv = dr.gv[C.name]
        r = interpolate_to_yee(dr.design_parameters, dr.weights[C.name], gv.Nx * gv.Ny)
        eps_inv = (1.0 / (dr.ε_min + r * (dr.ε_max - dr.ε_min))).reshape((gv.Nx, gv.Ny), order="F")
        arr = spec.eps_inv[a]
        st = gv.start_idx
        en = [st[0] + gv.Nx - 1, st[1] + gv.Ny - 1, st[2] + gv.Nz - 1]
        lo = [max(st[k], 1) for k in range(3)]
        hi = [min(en[k], N[k], arr.shape[k]) for k in range(3)]
        if any(hi[k] < lo[k] for k in range(3)):
            continue
        sub = eps_inv[lo[0] - st[0]:hi[0] - st[0] + 1, lo[1] - st[1]:hi[1] - st[1] + 1].astype(R)
        arr[lo[0] - 1:hi[0], lo[1] - 1:hi[1], lo[2] - 1:hi[2]] = sub[:, :, None]


# ------------------------------------------------------------ objectives
@dataclass
class FourierFieldsObjective:
    """DFT of one component over a volume (src/Adjoint/ObjectiveQuantities.jl:250-290)."""

    volume: G.Volume
    component: G.Field
    monitor: Optional[DFTMonitor] = None
    md: object = None
    eval_data: Optional[np.ndarray] = None


def corrected_fdtd_frequency(w: float, dt: float) -> complex:
    """src/Adjoint/AdjointSourceScale.jl:27-29."""
    return (1 - np.exp(-1j * w * dt)) / (-1j * dt)


def adj_src_scale(sim_R, dt: float, t_end: float, frequencies, time_profile, ndims=3, resolution=1.0,
                  include_resolution=True, fwidth_frac=0.1) -> np.ndarray:
    """src/Adjoint/AdjointSourceScale.jl:61-114 (t_end = the forward run's final time)."""
    dV = 1.0 / float(resolution) ** ndims if include_resolution else 1.0
    n_steps = int(math.floor(t_end / dt))
    t = np.arange(n_steps) * dt
    y = np.array([complex(time_profile.eval(float(tv), sim_R)) for tv in t])
    f = np.asarray(frequencies, dtype=np.float64)
    dtft = np.array([np.sum(dt / math.sqrt(2 * math.pi) * np.exp(1j * 2 * math.pi * fi * t) * y) for fi in f])
    if len(f) == 1:
        return 2.0 * dV / (math.pi * math.sqrt(2 * math.pi) * dtft)
    iomega = (1.0 - np.exp(-1j * (2 * math.pi * f) * dt)) / dt
    fcen = time_profile.fcen
    center = np.sum(dt / math.sqrt(2 * math.pi) * np.exp(1j * 2 * math.pi * fcen * t) * y)
    fwidth_scale = np.exp(-2j * math.pi * 5.0 / fwidth_frac)
    phase = np.exp(1j * np.angle(center)) * fwidth_scale
    return dV * iomega / phase * 2


def wirtinger_derivative(func: Callable, args: List[np.ndarray], idx: int) -> np.ndarray:
    """(df/dx + i df/dy) / 2 of a real objective w.r.t. complex argument idx
    (src/Adjoint/OptimizationProblem.jl:336-374), through torch autograd."""
    import torch
    x = np.asarray(args[idx])
    # may run inside an autograd backward, where grad mode is off
    with torch.enable_grad():
        xr = torch.tensor(x.real, dtype=torch.float64, requires_grad=True)
        xi = torch.tensor(x.imag, dtype=torch.float64, requires_grad=True)
        targs = [torch.tensor(np.asarray(a)) for a in args]
        targs[idx] = torch.complex(xr, xi)
        out = func(*targs)
        out = out if torch.is_tensor(out) else torch.tensor(out)
        gr, gi = torch.autograd.grad(out.real.sum(), [xr, xi], allow_unused=True)
    gr = np.zeros(x.shape) if gr is None else gr.numpy()
    gi = np.zeros(x.shape) if gi is None else gi.numpy()
    return (gr + 1j * gi).ravel(order="C") / 2


def _eval_objective(func, results):
    import torch
    out = func(*[torch.tensor(np.asarray(r)) for r in results])
    return float(out.real.item()) if torch.is_tensor(out) else float(np.real(out))


# ------------------------------------------------------------ problem
class OptimizationProblem:
    """src/Adjoint/OptimizationProblem.jl:21-374.

    ``sim`` is an unprepared Simulation (geometry, sources, boundaries); the
    design region's permittivity replaces whatever the geometry puts there.
    ``until`` (not in the reference) runs both simulations a fixed time
    instead of until the DFTs decay.
    """

    def __init__(self, sim: Simulation, objective_functions: Sequence[Callable],
                 objective_arguments: Sequence[FourierFieldsObjective], design_regions: Sequence[DesignRegion],
                 frequencies: Sequence[float], decay_by: float = 1e-6, minimum_run_time: float = 0.0,
                 maximum_run_time: float = math.inf, until: Optional[float] = None,
                 reference_scale: bool = False, method: str = "frequency", checkpoint_every: int = 100):
        self.sim = sim
        self.objective_functions = list(objective_functions)
        self.objective_arguments = list(objective_arguments)
        self.design_regions = list(design_regions)
        self.frequencies = [float(f) for f in frequencies]
        self.decay_by = decay_by
        self.minimum_run_time = minimum_run_time
        self.maximum_run_time = maximum_run_time
        self.until = until
        self.reference_scale = reference_scale
        if method not in ("frequency", "time_reversed"):
            raise ValueError("method must be 'frequency' or 'time_reversed'")
        if method == "time_reversed" and int(checkpoint_every) < 1:
            raise ValueError("checkpoint_every must be >= 1")
        self.method = method
        self.checkpoint_every = int(checkpoint_every)
        self._fwd_sim = None
        self.base = sim.build_spec()
        R = self.base.real
        for a, C in enumerate((G.Ex, G.Ey, G.Ez)):
            if self.base.eps_inv[a] is None:
                cnt = G.component_voxel_count(sim.Nx, sim.Ny, sim.Nz, C())
                self.base.eps_inv[a] = np.full(cnt, R(self.base.eps_inv_scalar), dtype=R, order="F")
        for dr in self.design_regions:
            init_design_region(sim, dr)
        self.f0 = None
        self.gradient = None
        self.results_list = None
        self.fwd_design = None
        self.f_bank = []
        self.current_state = "INIT"
        self._t_end = None

    # -- runs
    def _design_monitors(self):
        out = []
        for dr in self.design_regions:
            for C in (G.Dx, G.Dy, G.Dz):
                out.append(DFTMonitor(component=C(), center=list(dr.volume.center), size=list(dr.volume.size),
                                      frequencies=self.frequencies))
        return out

    def _run(self, spec, checkpoints: int = 0):
        """Step a copy of the simulation carrying ``spec`` on its engine
        (with field checkpoints every ``checkpoints`` steps if > 0)."""
        s = copy.copy(self.sim)
        s.spec = spec
        s.source_data = spec.sources
        s.monitor_data = spec.dfts
        cls = self.sim._engine_cls
        if cls is None:
            from .engine import HipEngine
            cls = HipEngine
        s.engine = cls(spec)
        s.timestep = 0
        s.is_prepared = True
        for md in spec.dfts:
            md.owner = s
        if checkpoints > 0:
            s.engine.set_checkpoints(checkpoints)
        if self.until is not None:
            run(s, until=self.until)
        else:
            run(s, until_after_sources=stop_when_dft_decayed(tolerance=self.decay_by,
                                                             minimum_runtime=self.minimum_run_time,
                                                             maximum_runtime=self.maximum_run_time))
        return s

    def _spec_with(self, sources, monitors, rho=None, negate_bloch=False):
        spec = copy.copy(self.base)
        spec.eps_inv = [None if e is None else e.copy(order="F") for e in self.base.eps_inv]
        if rho is not None:
            for dr in self.design_regions:
                update_design(spec, self.sim, dr, rho)
        spec.sources = list(sources)
        spec.dfts = [init_dft_monitor(self.sim, m) for m in monitors]
        for m, md in zip(monitors, spec.dfts):
            m.monitor_data = md
        if negate_bloch and spec.complex_fields:
            spec.bloch_fwd, spec.bloch_rev = spec.bloch_rev, spec.bloch_fwd
        return spec

    def update_design(self, rho):
        """src/Adjoint/OptimizationProblem.jl:84-99."""
        self.rho = np.asarray(rho, dtype=np.float64).copy()
        self.current_state = "INIT"

    def forward_run(self):
        """src/Adjoint/OptimizationProblem.jl:101-167."""
        obj_mons = []
        for oq in self.objective_arguments:
            oq.monitor = DFTMonitor(component=oq.component, center=list(oq.volume.center),
                                    size=list(oq.volume.size), frequencies=self.frequencies)
            obj_mons.append(oq.monitor)
        tr = self.method == "time_reversed"
        design = [] if tr else self._design_monitors()
        spec = self._spec_with(self.base.sources, obj_mons + design, rho=self.rho)
        self._close_forward()
        s = self._run(spec, checkpoints=self.checkpoint_every if tr else 0)
        n_obj = len(obj_mons)
        self.results_list = []
        for i, oq in enumerate(self.objective_arguments):
            oq.md = spec.dfts[i]
            oq.eval_data = s.engine.read_dft(i)
            self.results_list.append(oq.eval_data)
        self.fwd_design = [s.engine.read_dft(n_obj + j) for j in range(len(design))]
        self._t_end = s.round_time()
        self._fwd_spec = spec
        if tr:
            self._fwd_sim = s   # its checkpoints serve the adjoint sweep
        else:
            s.engine.close() if hasattr(s.engine, "close") else None
        f0 = [_eval_objective(f, self.results_list) for f in self.objective_functions]
        self.f0 = f0[0] if len(f0) == 1 else f0
        self.f_bank.append(self.f0)
        self.current_state = "FWD"
        return self.f0

    def injected_spectrum(self, component, freq):
        """F(w) = sum_n dt exp(i w t_n) f(t_n) of the waveform the time step
        injects for a source on ``component`` (E sources at t + dt/2, H
        sources at t; only while t <= cutoff), the same time stamps the DFT
        monitors of that component use."""
        R = self.base.real
        tp = self.base.sources[0].time_profile
        dt_r = self.sim.dt_r
        half = float(dt_r / R(2)) if component.electric else 0.0
        cut = self.base.cutoff
        out = 0j
        n = 0
        w = 2 * math.pi * freq
        while True:
            t = float(R(n) * dt_r)
            if t > cut or n > 10_000_000:
                break
            ts = t + half
            out += float(dt_r) * complex(math.cos(w * ts), math.sin(w * ts)) * complex(tp.eval(ts, R))
            n += 1
        return out

    def _adjoint_sources(self, fi):
        """Adjoint sources for frequency index ``fi``
        (src/Adjoint/OptimizationProblem.jl:178-195,
        src/Adjoint/ObjectiveQuantities.jl:292-345).

        Amplitude per monitor cell: conj(dJ) * 4 / F(w), with dJ the
        conjugate-Wirtinger derivative of the objective and F the spectrum of
        the injected waveform.  With sources entering as E = eps^-1 (D + S),
        the frequency-domain response to S = Re(f(t) A) is G (-w^2 A F / 2);
        that amplitude makes Re[deps_inv/drho D_adj . D_fwd] the derivative of
        the objective.  ``reference_scale=True`` uses the reference's
        dJ * adj_src_scale instead (src/Adjoint/AdjointSourceScale.jl:61-114),
        which differs from that by a complex conjugation and 1/(2 pi) in this
        discretisation (see the FD check)."""
        srcs = []
        fwd_tp = self.base.sources[0].time_profile
        R = self.base.real
        for func in self.objective_functions:
            for mi, oq in enumerate(self.objective_arguments):
                if not oq.component.electric:
                    raise NotImplementedError("adjoint sources for H-field objectives are not provided")
                dJ = wirtinger_derivative(func, self.results_list, mi)
                if not np.any(np.abs(dJ) > 0):
                    continue
                dJr = dJ.reshape(oq.eval_data.shape, order="C")[..., fi]
                if self.reference_scale:
                    scale = adj_src_scale(R, self.sim.dt_n, self._t_end, [self.frequencies[fi]], fwd_tp,
                                          include_resolution=False)
                    amp = (dJr * scale[0]).astype(np.complex128)
                else:
                    amp = (np.conj(dJr) * 4.0 / self.injected_spectrum(oq.component, self.frequencies[fi]))
                    amp = amp.astype(np.complex128)
                srcs.append(SourceData(component=oq.component, gv=oq.md.gv, amplitude=amp, time_profile=fwd_tp))
        return srcs

    def _close_forward(self):
        if self._fwd_sim is not None:
            eng = self._fwd_sim.engine
            eng.close() if hasattr(eng, "close") else None
            self._fwd_sim = None

    def adjoint_run(self):
        """src/Adjoint/OptimizationProblem.jl:169-330: one adjoint run per
        frequency (each reads its own frequency's D fields).  The
        time-reversed method has no separate run: its sweep happens in
        ``calculate_gradient``."""
        if self.method == "time_reversed":
            self.current_state = "ADJ"
            return
        self.adj_design = []
        for fi in range(len(self.frequencies)):
            srcs = self._adjoint_sources(fi)
            if not srcs:
                self.adj_design.append(None)
                continue
            spec = self._spec_with(srcs, self._design_monitors(), rho=self.rho, negate_bloch=True)
            s = self._run(spec)
            self.adj_design.append([s.engine.read_dft(j)[..., fi:fi + 1] for j in range(len(spec.dfts))])
            s.engine.close() if hasattr(s.engine, "close") else None
        self.current_state = "ADJ"

    def calculate_gradient(self, weights=None):
        """src/Adjoint/GradientKernel.jl:35-109 (first design region):
        (n_design, n_freq), one column per frequency.  Time-reversed method:
        (n_design, 1), the gradient of sum_i weights[i] * objective_i (all
        weights 1 by default) over all frequencies at once."""
        if self.method == "time_reversed":
            self.gradient = self._time_reversed_gradient(weights)[:, None]
            return self.gradient
        dr = self.design_regions[0]
        nd = len(dr.design_parameters)
        nf = len(self.frequencies)
        grad = np.zeros((nd, nf))
        for fi in range(nf):
            if self.adj_design[fi] is not None:
                grad[:, fi] = self._gradient_column(dr, fi)
        self.gradient = grad
        return grad

    def design_boxes(self, dr):
        """Per E component, the grid-index box ((lo), (n)) that update_design
        writes (clamped to cells 1..N and the array), or None."""
        N = self.base.N
        boxes = []
        for C in (G.Ex, G.Ey, G.Ez):
            gv = dr.gv[C.name]
            cnt = G.component_voxel_count(self.sim.Nx, self.sim.Ny, self.sim.Nz, C())
            st = gv.start_idx
            en = [st[0] + gv.Nx - 1, st[1] + gv.Ny - 1, st[2] + gv.Nz - 1]
            lo = [max(st[k], 1) for k in range(3)]
            hi = [min(en[k], N[k], cnt[k]) for k in range(3)]
            boxes.append(None if any(hi[k] < lo[k] for k in range(3))
                         else (lo, [hi[k] - lo[k] + 1 for k in range(3)]))
        return boxes

    def _time_reversed_gradient(self, weights=None):
        """dJ/drho from the library's dJ/d(eps^-1) per design voxel: summed
        over the extruded z range, times d(eps^-1)/dr = -(eps_max - eps_min) /
        eps^2, through the transpose of the bilinear interpolation."""
        if self._fwd_sim is None:
            raise RuntimeError("no forward run to differentiate (call forward_run first)")
        dr = self.design_regions[0]
        funcs = self.objective_functions
        w = np.ones(len(funcs)) if weights is None else np.atleast_1d(np.asarray(weights, dtype=np.float64))
        if len(w) != len(funcs):
            raise ValueError("one weight per objective function")
        cots = []
        for mi, oq in enumerate(self.objective_arguments):
            acc = np.zeros(oq.eval_data.shape, dtype=np.complex128)
            for fw, func in zip(w, funcs):
                dJ = wirtinger_derivative(func, self.results_list, mi)
                acc += fw * 2.0 * dJ.reshape(oq.eval_data.shape, order="C")
            cots.append(acc)
        boxes = self.design_boxes(dr)
        grads = self._fwd_sim.engine.adjoint_gradient(list(range(len(cots))), cots, boxes)
        de = dr.ε_max - dr.ε_min
        out = np.zeros(len(dr.design_parameters))
        for a, C in enumerate((G.Ex, G.Ey, G.Ez)):
            if grads[a] is None:
                continue
            gv = dr.gv[C.name]
            lo, n = boxes[a]
            st = gv.start_idx
            r = interpolate_to_yee(dr.design_parameters, dr.weights[C.name], gv.Nx * gv.Ny)
            deps_inv = -de / (dr.ε_min + r * de) ** 2
            gxy = grads[a].sum(axis=2)
            ix = np.arange(n[0]) + (lo[0] - st[0])
            iy = np.arange(n[1]) + (lo[1] - st[1])
            yi = iy[None, :] * gv.Nx + ix[:, None]
            gy = np.zeros(gv.Nx * gv.Ny)
            np.add.at(gy, yi.ravel(), gxy.ravel())
            gy *= deps_inv
            yee, des, wts = dr.weights[C.name]
            np.add.at(out, des, gy[yee] * wts)
        return out

    def _gradient_column(self, dr, fi):
        nd = len(dr.design_parameters)
        grad = np.zeros((nd, 1))
        de = dr.ε_max - dr.ε_min
        for ci, C in enumerate((G.Ex, G.Ey, G.Ez)):
            fwd, adj = self.fwd_design[ci][..., fi:fi + 1], self.adj_design[fi][ci]
            gv = dr.gv[C.name]
            w = dr.weights[C.name]
            rho_yee = interpolate_to_yee(dr.design_parameters, w, gv.Nx * gv.Ny)
            nx, ny, nz = fwd.shape[0], fwd.shape[1], fwd.shape[2]
            overlap = np.sum(adj * fwd, axis=2)  # (nx, ny, nf), summed over z
            if dr.volume.size[2] == 0.0:
                overlap = overlap / max(1, nz)
            yee, des, wts = w
            ix = yee % nx
            iy = yee // nx
            ok = (ix < nx) & (iy < ny)
            rl = np.clip(rho_yee[yee[ok]], 0.0, 1.0)
            eps = dr.ε_min + rl * de
            deps = -de / eps ** 2
            contrib = np.real((deps * wts[ok])[:, None] * overlap[ix[ok], iy[ok], :])
            np.add.at(grad, des[ok], contrib)
        return grad[:, 0]

    def __call__(self, rho, need_gradient: bool = True):
        """(f0, gradient) at design ``rho`` (src/Adjoint/OptimizationProblem.jl:63-82)."""
        self.update_design(rho)
        if self.current_state == "INIT":
            self.forward_run()
        if need_gradient and self.current_state == "FWD":
            self.adjoint_run()
            self.calculate_gradient()
        return self.f0, (self.gradient if need_gradient else None)

    def calculate_fd_gradient(self, rho, db: float = 1e-4, n_samples: int = 10, seed: int = 0):
        """Central differences at sampled parameters (src/Adjoint/GradientKernel.jl:117-152)."""
        rho = np.asarray(rho, dtype=np.float64)
        nd = len(rho)
        n_samples = min(n_samples, nd)
        idx = np.sort(np.random.default_rng(seed).permutation(nd)[:n_samples])
        out = np.zeros((n_samples, len(self.frequencies)))
        for si, k in enumerate(idx):
            left = rho.copy()
            left[k] = max(rho[k] - db, 0.0)
            right = rho.copy()
            right[k] = min(rho[k] + db, 1.0)
            fl, _ = self(left, need_gradient=False)
            fr, _ = self(right, need_gradient=False)
            step = right[k] - left[k]
            if step > 0:
                out[si, :] = np.real((np.asarray(fr) - np.asarray(fl)) / step)
        return out, idx


def simulate(opt: OptimizationProblem, designs):
    """Forward value of the problem (src/Adjoint/ChainRulesIntegration.jl:20-28)."""
    opt.update_design(designs)
    return opt.forward_run()


def simulate_torch(opt: OptimizationProblem, designs):
    """``simulate`` as a torch autograd node: backward runs the adjoint
    simulation and contracts the gradient with the incoming cotangent (the
    reverse rule of src/Adjoint/ChainRulesIntegration.jl:38-60, which returns
    the (n_design, n_freq) gradient itself)."""
    import torch

    class _Simulate(torch.autograd.Function):
        @staticmethod
        def forward(ctx, rho):
            f0 = simulate(opt, rho.detach().cpu().numpy())
            return torch.as_tensor(np.atleast_1d(np.asarray(f0, dtype=np.float64)))

        @staticmethod
        def backward(ctx, g):
            opt.adjoint_run()
            gw = g.detach().cpu().numpy()
            if opt.method == "time_reversed":
                # one sweep for the cotangent-weighted sum of the objectives
                return torch.as_tensor(opt.calculate_gradient(weights=gw.ravel())[:, 0])
            grad = opt.calculate_gradient()  # (n_design, n_freq)
            contracted = grad.sum(axis=1) * gw[0] if gw.size == 1 else grad @ gw
            return torch.as_tensor(contracted)

    return _Simulate.apply(designs)
