"""Pins of the oracle's periodic / Bloch wrap and complex fields.

The reference's single-chunk wrap copies plane N to ghost 0 times exp(-ikL)
and plane 1 to ghost N+1 times exp(+ikL), for the field group just updated
(src/Chunking.jl:1725-1768, :2138-2172).  Pinned here by:

1. Bloch plane waves.  With every axis Bloch-periodic, a field
   F_c(g) = A_c exp(i q.g delta) with q_a = (k_a L_a + 2 pi m_a) / (N_a delta_a)
   is mapped by the Yee update to the same spatial form (the wrapped ghosts
   equal the plane wave's own values there).  After several steps
   F_c(g) exp(-i q.g delta) must be the same at every cell; a wrap with the
   phase direction reversed breaks that at the boundary planes.
2. Spectrum.  With x, y Bloch and z closed, every eigenvalue of the complex
   one-step map lies on the analytic set: q from the Bloch condition on x, y,
   k = (m + 1/2) pi / (N + 1/2) on z, lambda = sum 4 sin^2(q delta / 2) / delta^2
   (two polarisations), cos(theta) = 1 - dt^2 lambda eps^-1 mu^-1 / 2, plus
   the static (gradient) modes at 1.
3. Periodic real fields are the real part of Bloch(k = 0) complex fields,
   value for value, with the imaginary part identically zero.
"""
import math

import numpy as np

import khronos_hip as kh
from khronos_hip.simulation import SimSpec
from oracle_engine import OracleEngine


def _spec(N, delta, dt, ks, periodic, eps_inv=1.0, mu_inv=1.0, cplx=True, swap=False):
    fwd, rev = [], []
    for a in range(3):
        L = N[a] * delta[a]
        x = ks[a] * L
        f, r = complex(math.cos(x), math.sin(x)), complex(math.cos(-x), math.sin(-x))
        if swap:
            f, r = r, f
        fwd.append(f if periodic[a] else 1 + 0j)
        rev.append(r if periodic[a] else 1 + 0j)
    return SimSpec(N=tuple(N), real=np.float64, delta=tuple(delta), dt=dt, sigma=None, eps_inv=[None] * 3,
                   mu_inv=[None] * 3, eps_inv_scalar=eps_inv, mu_inv_scalar=mu_inv, sources=[], dfts=[],
                   cutoff=-math.inf, mode="exact", dft_accum_f64=True, complex_fields=cplx,
                   periodic=tuple(periodic), bloch_fwd=tuple(fwd), bloch_rev=tuple(rev))


def _plane_wave_spread(swap=False):
    N = (6, 5, 4)
    delta = (0.1, 0.12, 0.09)
    dt = 0.5 * min(delta)
    ks = (1.3, -2.1, 0.7)
    m = (1, 2, 0)
    spec = _spec(N, delta, dt, ks, (True, True, True), eps_inv=0.7, mu_inv=0.9, swap=swap)
    eng = OracleEngine(spec)
    q = [(ks[a] * N[a] * delta[a] + 2 * math.pi * m[a]) / (N[a] * delta[a]) for a in range(3)]
    # the plane wave on grid indices 0..N+1: the ghost planes the first H
    # update reads hold the wave too (later steps refill them by the wrap)
    g = np.meshgrid(*[np.arange(0, N[a] + 2) for a in range(3)], indexing="ij")
    full = np.exp(1j * sum(q[a] * g[a] * delta[a] for a in range(3)))
    inner = (slice(1, N[0] + 1), slice(1, N[1] + 1), slice(1, N[2] + 1))
    wave = full[inner]
    rng = np.random.default_rng(5)
    for c in range(6):
        shp = eng.field_shape(c)
        F = np.zeros(shp, dtype=complex, order="F")
        if c < 3:
            F[:N[0] + 2, :N[1] + 2, :N[2] + 2] = (rng.standard_normal() + 1j * rng.standard_normal()) * full
        eng.write_state(0, c, F)
        eng.write_state(1, c, F / (0.7 if c < 3 else 0.9))
    eng.step(7)
    worst = 0.0
    for c in range(6):
        R = eng.read_field(c)[inner] / wave
        worst = max(worst, float(np.max(np.abs(R - R.mean())) / np.max(np.abs(R))))
    return worst


def test_bloch_plane_wave_keeps_its_form():
    assert _plane_wave_spread() < 1e-12


def test_reversed_bloch_phase_breaks_plane_wave():
    assert _plane_wave_spread(swap=True) > 1e-3


def test_bloch_spectrum_mixed_boundaries():
    N = (3, 4, 2)
    delta = (0.1, 0.13, 0.07)
    dt = 0.5 * min(delta)
    ks = (2.3, -1.1, 0.0)
    einv, minv = 0.5, 0.8
    spec = _spec(N, delta, dt, ks, (True, True, False), eps_inv=einv, mu_inv=minv)
    eng = OracleEngine(spec)
    V = N[0] * N[1] * N[2]
    inner = (slice(1, N[0] + 1), slice(1, N[1] + 1), slice(1, N[2] + 1))
    m = [einv] * 3 + [minv] * 3
    M = np.zeros((6 * V, 6 * V), dtype=complex)
    for j in range(6 * V):
        for c in range(6):
            blk = np.zeros(V)
            if j // V == c:
                blk[j % V] = 1.0
            F = np.zeros(eng.field_shape(c), dtype=complex, order="F")
            F[inner] = blk.reshape(N, order="F")
            eng.write_state(1, c, F / m[c])
            # the state includes the ghosts the wrap keeps consistent with it
            for a in (0, 1):
                src_hi = [slice(1, N[0] + 1), slice(1, N[1] + 1), slice(1, N[2] + 1)]
                dst_lo = list(src_hi)
                src_lo = list(src_hi)
                dst_hi = list(src_hi)
                src_hi[a], dst_lo[a] = N[a], 0
                src_lo[a], dst_hi[a] = 1, N[a] + 1
                F[tuple(dst_lo)] = F[tuple(src_hi)] * spec.bloch_rev[a]
                F[tuple(dst_hi)] = F[tuple(src_lo)] * spec.bloch_fwd[a]
            eng.write_state(0, c, F)
        eng.step(1)
        M[:, j] = np.concatenate([eng.read_field(c)[inner].ravel(order="F") for c in range(6)])
    ev = np.linalg.eigvals(M)
    assert np.allclose(np.abs(ev), 1.0, atol=1e-9)
    th = np.angle(ev)
    osc = np.sort(th[np.abs(th) > 1e-7])
    lam = []
    for mx in range(N[0]):
        qx = (ks[0] * N[0] * delta[0] + 2 * math.pi * mx) / (N[0] * delta[0])
        for my in range(N[1]):
            qy = (ks[1] * N[1] * delta[1] + 2 * math.pi * my) / (N[1] * delta[1])
            for mz in range(N[2]):
                kz = (mz + 0.5) * math.pi / (N[2] + 0.5)
                lam.append(4 * math.sin(qx * delta[0] / 2) ** 2 / delta[0] ** 2
                           + 4 * math.sin(qy * delta[1] / 2) ** 2 / delta[1] ** 2
                           + 4 * math.sin(kz / 2) ** 2 / delta[2] ** 2)
    ana = []
    for l in lam:
        t = math.acos(1 - dt * dt * l * einv * minv / 2)
        ana += [t, t, -t, -t]
    ana = np.sort(ana)
    assert len(osc) == len(ana) == 4 * V
    assert np.max(np.abs(osc - ana)) < 1e-9
    assert np.sum(np.abs(th) <= 1e-7) == 2 * V


def _slab_sim(bcs, precision="f64", engine=OracleEngine):
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.4), component=kh.Ex(),
                            center=[0.0, 0.0, -0.3], size=[0.4, 0.3, 0.0])]
    geo = [kh.Object(kh.Box([0.0, 0.0, 0.2], [0.2, 0.5, 0.2]), kh.Material(ε=3.0))]
    mon = [kh.DFTMonitor(component=kh.Ex(), center=[0.0, 0.0, 0.4], size=[0.4, 0.5, 0.0], frequencies=[0.9, 1.1])]
    return kh.Simulation(cell_size=[0.6, 0.5, 1.6], resolution=10, sources=src, geometry=geo, monitors=mon,
                         boundaries=[[0.0, 0.0], [0.0, 0.0], [0.3, 0.3]], boundary_conditions=bcs,
                         precision=precision, engine=engine)


def test_periodic_real_equals_bloch_k0_real_part():
    per = [[kh.Periodic(), kh.Periodic()], [kh.Periodic(), kh.Periodic()], [kh.PML(), kh.PML()]]
    blo = [[kh.Bloch(k=0.0), kh.Bloch(k=0.0)], [kh.Bloch(k=0.0), kh.Bloch(k=0.0)], [kh.PML(), kh.PML()]]
    a, b = _slab_sim(per), _slab_sim(blo)
    a.prepare()
    b.prepare()
    assert not a.spec.complex_fields and b.spec.complex_fields
    assert a.spec.periodic == (True, True, False) and b.spec.periodic == (True, True, False)
    a.step(60)
    b.step(60)
    for C in (kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz):
        fa, fb = a.read_field(C()), b.read_field(C())
        assert np.array_equal(fa, fb.real)
        assert np.all(fb.imag == 0)
    assert np.any(a.read_field(kh.Ex()) != 0)
    assert np.array_equal(a.engine.read_dft(0), b.engine.read_dft(0))


def test_host_bloch_phases_and_pml_override():
    bcs = [[kh.Bloch(k=1.5), kh.Bloch(k=1.5)], [kh.Periodic(), kh.Periodic()], [kh.PML(), kh.PML()]]
    sim = _slab_sim(bcs)
    spec = sim.build_spec()
    L = float(np.float64(0.6))
    assert spec.complex_fields and spec.periodic == (True, True, False)
    assert spec.bloch_fwd[0] == complex(math.cos(1.5 * L), math.sin(1.5 * L))
    assert spec.bloch_rev[0] == complex(math.cos(-1.5 * L), math.sin(-1.5 * L))
    assert spec.bloch_fwd[1] == 1 and spec.bloch_rev[2] == 1
    # periodic axes get no PML profile; z keeps its PML (src/Boundaries.jl:100-109)
    assert not np.any(spec.sigma["B"][0]) and not np.any(spec.sigma["D"][1])
    assert np.any(spec.sigma["B"][2]) and np.any(spec.sigma["D"][2])
