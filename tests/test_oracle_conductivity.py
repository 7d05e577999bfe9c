"""Pins of the oracle's conductive (C) stage and of the host conductivity setup.

The reference's C stage (src/Kernels/Helpers.jl:39-70) runs in front of the
PML cascade for every component with a conductivity array, with
sigma~ = 0.5 (dt sigma) (:278-279).  Two independent formulations pin it:

1. Uniform conductor, zero PML profiles: the C -> U -> T chain must equal the
   textbook lossy Yee update D' = ((1 - s) D + K) / (1 + s), E = eps^-1 D
   (one stage, no auxiliaries).  An oracle built with the 1/2 of sigma~
   dropped fails it.
2. Conductor inside a constant-sigma PML: per cell the cascade is a chain of
   first-order filters l(s) = (1 + s) - (1 - s) z^-1, whose transfer function
   A = m l'(s_own) (1 - z^-1) K / (l(s_D) l(s_next) l(s_prev)) does not
   depend on the order of the stages.  A numpy realisation with the
   conductive stage placed after the U stage (the reference puts it first)
   must give the same fields.

Host setup: absorber profiles against their closed form
(src/Geometry.jl:708-789), painting priority of the conductivity raster
(:233-249, :117-118), and the error for conductivity without PML arrays.
"""
import math
import os
import subprocess

import numpy as np
import pytest

import khronos_hip as kh
import oracle_engine as oe
from khronos_hip.grid import component_voxel_count
from khronos_hip.simulation import SimSpec
from oracle_engine import OracleEngine, ORACLE_DIR, load_path


def _curl_np(F, idx, dt, inv, N):
    """K = dt * curl on the interior (reference operand order), forward
    (idx = +1, H half step) or backward (idx = -1, E half step) differences."""
    sl = lambda dx, dy, dz: (slice(1 + dx, N[0] + 1 + dx), slice(1 + dy, N[1] + 1 + dy), slice(1 + dz, N[2] + 1 + dz))
    c = sl(0, 0, 0)
    X, Y, Z = F
    d = lambda A, ax: (A[sl(*[idx if a == ax else 0 for a in range(3)])] - A[c]) * inv[ax]
    return [dt * (d(Y, 2) - d(Z, 1)), dt * (d(Z, 0) - d(X, 2)), dt * (d(X, 1) - d(Y, 0))]


def _spec(N, delta, dt, sig, cond, eps_inv=1.0, mu_inv=1.0):
    return SimSpec(N=tuple(N), real=np.float64, delta=tuple(delta), dt=dt, sigma=sig, eps_inv=[None] * 3,
                   mu_inv=[None] * 3, eps_inv_scalar=eps_inv, mu_inv_scalar=mu_inv, sources=[], dfts=[],
                   cutoff=-math.inf, mode="exact", dft_accum_f64=True, cond=cond)


def _cond_arrays(N, sigma_e):
    """Uniform sigma_D on the three E components (geometry layout), none on H."""
    comps = [kh.Ex(), kh.Ey(), kh.Ez()]
    return [np.full(component_voxel_count(*N, c), sigma_e) for c in comps] + [None] * 3


def _lossy_yee_error(N=(6, 5, 4), delta=(0.1, 0.12, 0.09), sigma_e=3.7, eps_inv=0.5, steps=30):
    dt = 0.5 * min(delta)
    sig = {"B": [np.zeros(2 * N[a] + 1) for a in range(3)], "D": [np.zeros(2 * N[a] + 1) for a in range(3)]}
    eng = OracleEngine(_spec(N, delta, dt, sig, _cond_arrays(N, sigma_e), eps_inv=eps_inv))
    rng = np.random.default_rng(11)
    inner = (slice(1, N[0] + 1), slice(1, N[1] + 1), slice(1, N[2] + 1))
    E = [np.zeros(eng.field_shape(a), order="F") for a in range(3)]
    D = [np.zeros(eng.field_shape(a), order="F") for a in range(3)]
    H = []
    for a in range(3):
        h = np.zeros(eng.field_shape(3 + a), order="F")
        h[inner] = rng.standard_normal(N)
        eng.write(0, 3 + a, h)
        eng.write(1, 3 + a, h)      # B = H (mu^-1 = 1); E side starts at rest
        H.append(h)
    s = 0.5 * (dt * sigma_e)
    inv = [1.0 / d for d in delta]
    for _ in range(steps):
        K = _curl_np(E, 1, dt, inv, N)
        for a in range(3):
            H[a][inner] = H[a][inner] + K[a]
        K = _curl_np(H, -1, dt, inv, N)
        for a in range(3):
            D[a][inner] = ((1 - s) * D[a][inner] + K[a]) / (1 + s)
            E[a][inner] = eps_inv * D[a][inner]
    eng.step(steps)
    num = sum(float(np.sum((eng.read_field(c) - F) ** 2)) for c, F in enumerate(E + H))
    den = sum(float(np.sum(F ** 2)) for F in E + H)
    assert sum(float(np.sum(F ** 2)) for F in E) > 1e-3 * den
    return math.sqrt(num / den)


def test_oracle_conductor_equals_textbook_lossy_yee():
    assert _lossy_yee_error() < 1e-12


def test_planted_cond_scale_fails_lossy_yee():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "build/liboracle_plant_cond_f64.so"], check=True)
    saved = oe._LIBS.pop("f64", None)
    try:
        oe._LIBS["f64"] = load_path(np.float64, os.path.join(ORACLE_DIR, "build", "liboracle_plant_cond_f64.so"))
        err = _lossy_yee_error()
    finally:
        oe._LIBS.pop("f64", None)
        if saved is not None:
            oe._LIBS["f64"] = saved
    assert err > 1e-3


def test_oracle_conductor_in_pml_equals_reordered_stages():
    N = (6, 5, 7)
    delta = (0.1, 0.12, 0.09)
    dt = 0.5 * min(delta)
    s = (0.031, 0.017, 0.045)          # PML sigma * dt / 2 per axis, constant
    sigma_e = 2.3
    sd = 0.5 * (dt * sigma_e)
    eps_inv, mu_inv = 0.5, 0.8
    sig = {"B": [np.full(2 * N[a] + 1, s[a]) for a in range(3)],
           "D": [np.full(2 * N[a] + 1, s[a]) for a in range(3)]}
    eng = OracleEngine(_spec(N, delta, dt, sig, _cond_arrays(N, sigma_e), eps_inv=eps_inv, mu_inv=mu_inv))
    rng = np.random.default_rng(5)
    inner = (slice(1, N[0] + 1), slice(1, N[1] + 1), slice(1, N[2] + 1))
    m = [eps_inv] * 3 + [mu_inv] * 3
    A = [np.zeros(N) for _ in range(6)]
    P = [np.zeros(N) for _ in range(6)]
    Q = [np.zeros(N) for _ in range(3)]
    for a in range(3):
        c = 3 + a
        own, prv = a, (a + 2) % 3
        shp = eng.field_shape(c)
        F = np.zeros(shp, order="F")
        F[inner] = rng.standard_normal(N)
        # H at rest in its cascade: T = W = 0, U = s_prev A / (m s_own); E side all zero
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
            base, other = (0, 3) if electric else (3, 0)
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
                if electric:
                    # conductive stage after the U stage
                    Qn = ((1 - sd) * Q[a] + (Pn - P[c])) / (1 + sd)
                    A[c] = ((1 - sp) * A[c] + m[c] * ((1 + so) * Qn - (1 - so) * Q[a])) / (1 + sp)
                    Q[a] = Qn
                else:
                    A[c] = ((1 - sp) * A[c] + m[c] * ((1 + so) * Pn - (1 - so) * P[c])) / (1 + sp)
                P[c] = Pn
    eng.step(nsteps)
    for c in range(6):
        got = eng.read_field(c)[inner]
        err = np.linalg.norm(got - A[c]) / np.linalg.norm(A[c])
        assert err < 1e-12, (c, err)


# ------------------------------------------------------------- host setup
def _absorber_sim(absorbers, geometry=None, boundaries=((0.0, 0.0),) * 3, precision="f64"):
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.5), component=kh.Ez(),
                            center=[0.0, 0.0, 0.0], size=[0.0, 0.0, 0.0])]
    return kh.Simulation(cell_size=[2.0, 1.5, 1.2], resolution=10, sources=src, geometry=geometry or [],
                         boundaries=None if boundaries is None else [list(b) for b in boundaries],
                         absorbers=absorbers, precision=precision, engine=OracleEngine)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_absorber_profile_closed_form(precision):
    n, p = 6, 3
    sim = _absorber_sim([[None, kh.Absorber(num_layers=n, sigma_order=p)], None, None], precision=precision)
    spec = sim.build_spec()
    R = sim.real
    L = R(R(n) * sim.delta_r[0])
    smax = float(R(-(p + 1) * math.log(1e-6) / (2.0 * float(L))))
    assert abs(smax - 4 * math.log(1e6) / (2 * n * 0.1)) < 1e-5 * smax
    assert all(c is not None for c in spec.cond)      # absorbers allocate all six arrays
    for arr in spec.cond:
        Nx = arr.shape[0]
        prof = arr[:, 0, 0]
        assert np.all(arr == prof[:, None, None])
        want = np.zeros(Nx, dtype=R)
        for layer in range(1, n + 1):
            want[Nx - layer] = R(smax * ((n - layer + 1) / n) ** p)
        assert np.array_equal(prof, want)
        assert prof[-1] == R(smax) and prof[Nx - n] == R(smax * (1 / n) ** p) and prof[Nx - n - 1] == 0


def test_absorber_adds_to_material_conductivity():
    geo = [kh.Object(kh.Box([0.0, 0.0, 0.0], [10.0, 10.0, 10.0]), kh.Material(σD=0.25))]
    sim = _absorber_sim([None, [kh.Absorber(num_layers=3, sigma_order=2, sigma_max=2.0), None], None], geometry=geo)
    spec = sim.build_spec()
    for c in range(3):
        prof = spec.cond[c][0, :, 0]
        assert prof[0] == 0.25 + 2.0 and prof[1] == 0.25 + 2.0 * (2 / 3) ** 2 and prof[2] == 0.25 + 2.0 * (1 / 3) ** 2
        assert np.all(prof[3:] == 0.25)
    for c in range(3, 6):   # no material sigma_B: only the absorber
        prof = spec.cond[c][0, :, 0]
        assert prof[0] == 2.0 and np.all(prof[3:] == 0.0)


def test_conductivity_raster_priority_and_components():
    geo = [kh.Object(kh.Sphere([0.0, 0.0, 0.0], 0.3), kh.Material(ε=2.0)),            # paints 0 over the box
           kh.Object(kh.Box([0.0, 0.0, 0.0], [1.0, 1.0, 1.0]), kh.Material(σD=1.5, σDz=0.5))]
    sim = _absorber_sim(None, geometry=geo)
    spec = sim.build_spec()
    assert spec.cond[3:] == [None] * 3            # no sigma_B anywhere
    sx, sz = spec.cond[0], spec.cond[2]
    assert set(np.unique(sx)) == {0.0, 1.5} and set(np.unique(sz)) == {0.0, 0.5}
    c = [n // 2 for n in sx.shape]
    assert sx[tuple(c)] == 0.0                     # sphere (first object) wins at the centre


def test_conductivity_without_pml_arrays_is_rejected():
    geo = [kh.Object(kh.Box([0.0, 0.0, 0.0], [1.0, 1.0, 1.0]), kh.Material(σD=1.0))]
    sim = _absorber_sim(None, geometry=geo, boundaries=None)
    with pytest.raises(ValueError):
        sim.build_spec()
