"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no field-value goldens (its tests assert no-NaN only), so
the oracle's time step is pinned by properties a plausible mistake breaks:

1. Discrete cavity spectrum.  With zero ghost planes, every axis of the
   reference grid has a Neumann end (ghost below index 1) and a Dirichlet end
   (index N+1), so the 1-D modes are cos(k (i - 1/2)) with
   k = (m + 1/2) pi / (N + 1/2).  The curl-curl operator then has eigenvalues
   sum_a 4 sin^2(k_a / 2) / delta_a^2 (twice each, times eps^-1 mu^-1), and the
   leapfrog map has cos(theta) = 1 - dt^2 lambda / 2.  Every eigenvalue of the
   oracle's one-step map must land on that set: a wrong sign, index, spacing
   or dt factor moves it.
2. PML cascade = textbook uniaxial PML.  With sigma constant per axis, the
   reference's three-stage U/T/W cascade (src/Kernels/Helpers.jl:113-139,
   :323-340) has the same z-domain transfer function,
   A = m l(s_own) K / (l(s_next) l(s_prev)) with l(s) = (1+s) - (1-s) z^-1,
   as Gedney's two-stage form ((d/dt + s_next) P = K,
   (d/dt + s_prev) A = m (d/dt + s_own) P) discretised the same way.  The
   oracle is compared with an independent numpy implementation of the
   two-stage form; a misplaced own-axis sigma or a wrong stage breaks it.
3. DFT accumulation = DTFT of the recorded field samples, with H sampled at
   t = n dt and E at t + dt/2, every D-th step (src/Kernels/Kernels.jl:465-497).
4. Planted mistake: an oracle built with one curl sign flipped fails (1).
"""
import math
import os
import subprocess

import numpy as np
import pytest

import khronos_hip as kh
from khronos_hip.simulation import SimSpec
from oracle_engine import OracleEngine, load as load_oracle, ORACLE_DIR


def _spec(N, delta, dt, real=np.float64, sigma=None, eps_inv=1.0, mu_inv=1.0):
    return SimSpec(N=tuple(N), real=real, delta=tuple(delta), dt=dt, sigma=sigma, eps_inv=[None] * 3,
                   mu_inv=[None] * 3, eps_inv_scalar=eps_inv, mu_inv_scalar=mu_inv, sources=[], dfts=[],
                   cutoff=-math.inf, mode="exact", dft_accum_f64=True)


def _interior(shape):
    return (slice(1, shape[0] - 1), slice(1, shape[1] - 1), slice(1, shape[2] - 1))


def _one_step_map(eng, N, eps_inv, mu_inv):
    """Dense matrix of one oracle step on the interior state (E, H); D and B
    are kept consistent (D = E / eps^-1, B = H / mu^-1)."""
    V = N[0] * N[1] * N[2]
    dim = 6 * V
    M = np.zeros((dim, dim))
    m = [eps_inv] * 3 + [mu_inv] * 3
    for j in range(dim):
        for c in range(6):
            shp = eng.field_shape(c)
            F = np.zeros(shp, order="F")
            blk = np.zeros(V)
            if j // V == c:
                blk[j % V] = 1.0
            F[1:N[0] + 1, 1:N[1] + 1, 1:N[2] + 1] = blk.reshape(N, order="F")
            eng.write(0, c, F)
            eng.write(1, c, F / m[c])
        eng.step(1)
        col = []
        for c in range(6):
            F = eng.read_field(c)
            col.append(F[1:N[0] + 1, 1:N[1] + 1, 1:N[2] + 1].ravel(order="F"))
        M[:, j] = np.concatenate(col)
    return M


def _analytic_lambdas(N, delta):
    ks = []
    for a in range(3):
        m = np.arange(N[a])
        k = (m + 0.5) * math.pi / (N[a] + 0.5)
        ks.append(4 * np.sin(k / 2) ** 2 / delta[a] ** 2)
    lam = (ks[0][:, None, None] + ks[1][None, :, None] + ks[2][None, None, :]).ravel()
    return np.sort(np.repeat(lam, 2))


def _spectrum_error(lib_real=np.float64, eng_factory=None):
    N = (3, 4, 2)
    delta = (0.1, 0.13, 0.07)
    dt = 0.5 * min(delta)
    eps_inv, mu_inv = 0.5, 0.8
    spec = _spec(N, delta, dt, eps_inv=eps_inv, mu_inv=mu_inv)
    eng = eng_factory(spec) if eng_factory else OracleEngine(spec)
    M = _one_step_map(eng, N, eps_inv, mu_inv)
    ev = np.linalg.eigvals(M)
    osc = ev[ev.imag > 1e-9]
    lam = np.sort((2 - 2 * osc.real) / dt ** 2) / (eps_inv * mu_inv)
    ref = _analytic_lambdas(N, delta)
    if len(lam) != len(ref):
        return math.inf
    return float(np.max(np.abs(lam - ref) / ref))


def test_oracle_cavity_spectrum():
    assert _spectrum_error() < 1e-9


def test_planted_curl_sign_fails_spectrum():
    """An oracle with one curl term's sign flipped must fail the spectrum pin."""
    import oracle_engine as oe
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "build/liboracle_plant_f64.so"], check=True)
    saved = oe._LIBS.pop("f64", None)
    try:
        oe._LIBS["f64"] = oe.load_path(np.float64, os.path.join(ORACLE_DIR, "build", "liboracle_plant_f64.so"))
        err = _spectrum_error()
    finally:
        oe._LIBS.pop("f64", None)
        if saved is not None:
            oe._LIBS["f64"] = saved
    assert err > 1e-3


def _curl_np(F, idx, dt, inv, N):
    """K = dt * curl on the interior, for the three components, from the other
    field's three arrays (reference layout), forward (idx=+1) or backward (-1)."""
    sl = lambda dx, dy, dz: (slice(1 + dx, N[0] + 1 + dx), slice(1 + dy, N[1] + 1 + dy), slice(1 + dz, N[2] + 1 + dz))
    c = sl(0, 0, 0)
    X, Y, Z = F
    d = lambda A, ax: (A[sl(*[idx if a == ax else 0 for a in range(3)])] - A[c]) * inv[ax]
    return [dt * (d(Y, 2) - d(Z, 1)), dt * (d(Z, 0) - d(X, 2)), dt * (d(X, 1) - d(Y, 0))]


def test_oracle_pml_cascade_equals_two_stage_upml():
    N = (6, 5, 7)
    delta = (0.1, 0.12, 0.09)
    dt = 0.5 * min(delta)
    s = (0.031, 0.017, 0.045)  # sigma * dt / 2 per axis, constant
    eps_inv, mu_inv = 0.5, 0.8
    sig = {"B": [np.full(2 * N[a] + 1, s[a]) for a in range(3)],
           "D": [np.full(2 * N[a] + 1, s[a]) for a in range(3)]}
    spec = _spec(N, delta, dt, sigma=sig, eps_inv=eps_inv, mu_inv=mu_inv)
    eng = OracleEngine(spec)
    rng = np.random.default_rng(7)
    inner = (slice(1, N[0] + 1), slice(1, N[1] + 1), slice(1, N[2] + 1))
    A = [None] * 6
    P = [None] * 6
    m = [eps_inv] * 3 + [mu_inv] * 3
    for c in range(6):
        a = c % 3
        own, nxt, prv = a, (a + 1) % 3, (a + 2) % 3
        shp = eng.field_shape(c)
        F = np.zeros(shp, order="F")
        F[inner] = rng.standard_normal(N)
        # consistent auxiliary start for the cascade: T = W = 0, U = s_prev A / (m s_own)
        U = np.zeros(shp, order="F")
        U[inner] = s[prv] * F[inner] / (m[c] * s[own])
        eng.write(0, c, F)
        eng.write(2, c, U)
        A[c] = F[inner].copy()
        P[c] = U[inner].copy()
    inv = [1.0 / d for d in delta]
    nsteps = 40
    for _ in range(nsteps):
        for electric in (False, True):
            base = 0 if electric else 3
            other = 3 if electric else 0
            full = []
            for q in range(3):
                G = np.zeros(eng.field_shape(other + q), order="F")
                G[inner] = A[other + q]
                full.append(G)
            K = _curl_np(full, -1 if electric else 1, dt, inv, N)
            for a in range(3):
                c = base + a
                sn, sp, so = s[(a + 1) % 3], s[(a + 2) % 3], s[a]
                Pn = ((1 - sn) * P[c] + K[a]) / (1 + sn)
                A[c] = ((1 - sp) * A[c] + m[c] * ((1 + so) * Pn - (1 - so) * P[c])) / (1 + sp)
                P[c] = Pn
    eng.step(nsteps)
    for c in range(6):
        got = eng.read_field(c)[inner]
        err = np.linalg.norm(got - A[c]) / np.linalg.norm(A[c])
        assert err < 1e-12, (c, err)


def test_oracle_dft_equals_dtft_of_samples():
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.5), component=kh.Ez(),
                            center=[0.0, 0.0, 0.0], size=[0.0, 0.0, 0.0])]
    freqs = [0.8, 1.0, 1.3]
    mons = [kh.DFTMonitor(component=kh.Ez(), center=[0.1, 0.0, 0.05], size=[0.2, 0.1, 0.0], frequencies=freqs, decimation=3),
            kh.DFTMonitor(component=kh.Hx(), center=[0.0, 0.1, 0.0], size=[0.1, 0.0, 0.2], frequencies=freqs, decimation=1)]
    sim = kh.Simulation(cell_size=[1.0, 1.0, 1.0], resolution=10, sources=src, monitors=mons,
                        boundaries=[[0.2, 0.2]] * 3, engine=OracleEngine)
    sim.prepare()
    nsteps = 120
    samples = [[] for _ in mons]
    for n in range(nsteps):
        sim.step(1)
        for i, md in enumerate(sim.spec.dfts):
            F = sim.engine.read_field(kh.grid.component_code(md.component))
            s0 = md.gv.start_idx
            blk = F[s0[0]:s0[0] + md.gv.Nx, s0[1]:s0[1] + md.gv.Ny, s0[2]:s0[2] + md.gv.Nz]
            samples[i].append(blk.copy())
    dt = sim.dt_n
    for i, md in enumerate(sim.spec.dfts):
        acc = np.zeros(tuple(md.gv.dims) + (len(freqs),), complex)
        for n in range(nsteps):
            if md.decimation > 1 and n % md.decimation:
                continue
            t = n * dt + (dt / 2 if md.component.electric else 0.0)
            for k, f in enumerate(freqs):
                acc[..., k] += dt * np.exp(1j * 2 * math.pi * f * t) * samples[i][n]
        got = sim.read_dft(md)
        assert np.linalg.norm(got - acc) / np.linalg.norm(acc) < 1e-12
