"""GPU parity: libkhronos_hip (through its C ABI) against the CPU oracle.

Tolerances:
  EXACT mode   -- fields equal to the oracle value for value (same operation
                  order, no FMA contraction), DFTs within 1e-13.
  FAST mode    -- fields within rel-L2 1e-12 (float64) / 1e-5 (float32) of
                  the oracle's same-precision run; DFTs within 1e-10 / 1e-4.
Full-size configurations are checked through size-independent properties
(FAST vs EXACT on the device, linearity in the source amplitude, no NaN).
"""
import math
import sys

import numpy as np
import pytest

import khronos_hip as kh
from khronos_hip.grid import component_code
from oracle_engine import OracleEngine

pytestmark = pytest.mark.gpu

COMPS = [kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz]


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def field_errors(sa, sb):
    """rel-L2 of the whole field vector (all six components), and each
    component's error normalised by that same global norm.  A component that is
    zero by symmetry (Hz of a z-directed dipole) holds round-off on both sides,
    so its own norm is no yardstick."""
    fa = [sa.read_field(C()).astype(np.float64) for C in COMPS]
    fb = [sb.read_field(C()).astype(np.float64) for C in COMPS]
    nb = math.sqrt(sum(float(np.sum(b * b)) for b in fb))
    tot = math.sqrt(sum(float(np.sum((a - b) ** 2)) for a, b in zip(fa, fb))) / nb
    per = {C.__name__: float(np.linalg.norm(a - b)) / nb for C, a, b in zip(COMPS, fa, fb)}
    return tot, per


def _box_sim(precision, mode, engine=None, n=(24, 20, 18), pml=0.3, res=10, steps_monitor=True,
             geometry=True, source_comp=kh.Ez, amp=1.0):
    cell = [n[0] / res, n[1] / res, n[2] / res]
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.4), component=source_comp(),
                            center=[0.05, -0.1, 0.0], size=[0.0, 0.0, 0.0], amplitude=amp),
           kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=0.9, fwidth=0.3), component=kh.Hy(),
                            center=[-0.2, 0.1, 0.1], size=[0.0, 0.3, 0.0], amplitude=0.5 * amp)]
    geo = []
    if geometry:
        geo = [kh.Object(kh.Box([0.0, 0.0, 0.0], [0.6, 0.4, 0.3]), kh.Material(ε=4.0)),
               kh.Object(kh.Sphere([0.3, 0.2, -0.1], 0.25), kh.Material(ε=2.25))]
    mons = []
    if steps_monitor:
        mons = [kh.FluxMonitor(center=[0.35, 0.0, 0.0], size=[0.0, 0.8, 0.6], frequencies=[0.8, 1.0, 1.2]),
                kh.DFTMonitor(component=kh.Ez(), center=[0.0, 0.0, 0.1], size=[0.6, 0.6, 0.0],
                              frequencies=[0.9, 1.1], decimation=2)]
    return kh.Simulation(cell_size=cell, resolution=res, sources=src, geometry=geo, monitors=mons,
                         boundaries=[[pml, pml]] * 3, precision=precision, mode=mode, engine=engine)


def _run_pair(precision, mode, steps, **kw):
    g = _box_sim(precision, mode, **kw)
    o = _box_sim(precision, "exact", engine=OracleEngine, **kw)
    _note(f"[parity] {precision} {mode} grid {g.Nx}x{g.Ny}x{g.Nz} steps {steps}")
    g.step(steps)
    o.step(steps)
    g.engine.sync()
    return g, o


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_exact_mode_matches_oracle_value_for_value(precision):
    g, o = _run_pair(precision, "exact", 60)
    for C in COMPS:
        a = g.read_field(C())
        b = o.read_field(C())
        assert a.shape == b.shape
        assert np.array_equal(a, b), (C.__name__, float(np.max(np.abs(a - b))))
    for i, md in enumerate(g.spec.dfts):
        assert rel_l2(g.engine.read_dft(i), o.engine.read_dft(i)) < 1e-13


@pytest.mark.parametrize("precision,tol,dtol", [("f64", 1e-12, 1e-10), ("f32", 1e-5, 1e-4)])
def test_fast_mode_matches_oracle(precision, tol, dtol):
    g, o = _run_pair(precision, "fast", 150)
    tot, per = field_errors(g, o)
    assert tot < tol and max(per.values()) < tol, (tot, per)
    for i, md in enumerate(g.spec.dfts):
        assert rel_l2(g.engine.read_dft(i), o.engine.read_dft(i)) < dtol


def test_device_flux_and_norm_match_host_reduction():
    g, o = _run_pair("f64", "fast", 120)
    from khronos_hip.monitors import flux_from_dft
    fm = g.monitors[0]
    md = fm.monitor_data
    mons = md.tangential_E_monitors + md.tangential_H_monitors
    arrs = [g.read_dft(m.monitor_data) for m in mons]
    host = flux_from_dft(md, *arrs, [m.monitor_data.gv for m in mons])
    dev = kh.get_flux(fm)
    np.testing.assert_allclose(dev, host, rtol=1e-12, atol=1e-14 * np.max(np.abs(host)))
    norms = g.engine.dft_norms()
    for i, md in enumerate(g.spec.dfts):
        ref = math.sqrt(float(np.sum(np.abs(g.engine.read_dft(i)) ** 2)))
        assert abs(norms[i] - ref) <= 1e-12 * ref


def test_dipole_80_f64_500_steps():
    """The 80^3 dipole: vacuum, 10-cell PML, Ez Gaussian dipole (fcen 1,
    fwidth 0.2), Courant 0.5, float64, 500 steps, one DFT plane at 5 freqs."""
    def mk(engine=None, mode="fast"):
        src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.2), component=kh.Ez(),
                                center=[0.0, 0.0, 0.0], size=[0.0, 0.0, 0.0])]
        mon = [kh.DFTMonitor(component=kh.Ez(), center=[0.0, 0.0, 0.0], size=[4.0, 4.0, 0.0],
                             frequencies=list(np.linspace(0.9, 1.1, 5)))]
        return kh.Simulation(cell_size=[4.0, 4.0, 4.0], resolution=20, sources=src, monitors=mon,
                             boundaries=[[0.5, 0.5]] * 3, precision="f64", mode=mode, engine=engine)
    g, o = mk(), mk(OracleEngine, "exact")
    _note("[parity] dipole 80^3 f64 500 steps")
    g.step(500)
    o.step(500)
    tot, per = field_errors(g, o)
    assert tot < 1e-12 and max(per.values()) < 1e-12, (tot, per)
    assert rel_l2(g.engine.read_dft(0), o.engine.read_dft(0)) < 1e-10


def test_reset_reproduces_run_bit_for_bit():
    g = _box_sim("f32", "fast")
    g.step(40)
    first = [g.read_field(C()) for C in COMPS]
    d0 = g.engine.read_dft(0)
    kh.reset_fields(g)
    assert g.engine.timestep() == 0 and g.sources_active()
    for C in COMPS:
        assert not np.any(g.read_field(C()))
    g.step(40)
    for C, a in zip(COMPS, first):
        assert np.array_equal(g.read_field(C()), a)
    assert np.array_equal(g.engine.read_dft(0), d0)


def test_write_read_field_roundtrip():
    g = _box_sim("f64", "fast", steps_monitor=False)
    g.prepare()
    rng = np.random.default_rng(3)
    for C in COMPS:
        code = component_code(C())
        shp = g.engine.field_shape(code)
        F = np.zeros(shp, order="F")
        F[1:-1, 1:-1, 1:-1] = rng.standard_normal((shp[0] - 2, shp[1] - 2, shp[2] - 2))
        g.engine.write_field(code, F)
        back = g.engine.read_field(code)
        # updated cells are grid 1..N; a staggered axis's extra plane N+1 reads back as zero
        n = (g.Nx, g.Ny, g.Nz)
        inner = tuple(slice(1, n[a] + 1) for a in range(3))
        assert np.array_equal(back[inner], F[inner])


def test_sources_switch_off_after_cutoff():
    g = _box_sim("f64", "fast", steps_monitor=False)
    cut = g.last_source_time()
    n_cut = int(math.floor(cut / g.dt_n)) + 2
    g.step(n_cut)
    assert not g.sources_active()


def test_abi_errors_are_reported_not_fatal():
    import ctypes as C
    from khronos_hip import _lib
    lib = _lib.load()
    d = _lib.KhSimDesc()
    d.nx, d.ny, d.nz = 4, 4, 0  # invalid
    d.precision = 0
    d.dt = d.dx = d.dy = d.dz = 0.1
    ctx = C.c_void_p()
    rc = lib.kh_create(C.byref(d), C.byref(ctx))
    assert rc == -1 and b"extents" in lib.kh_last_error()
    g = _box_sim("f64", "fast", steps_monitor=False)
    g.prepare()
    st = (C.c_int32 * 3)(1, 1, 1)
    dm = (C.c_int32 * 3)(1, 1, 1)
    sid = C.c_int32()
    amp = np.zeros(2)
    g.step(1)
    rc = lib.kh_add_source(g.engine.ctx, 2, st, dm, amp.ctypes.data_as(C.c_void_p), C.byref(sid))
    assert rc == -5  # sources must be registered before stepping


def _strip_sim(precision, mode, amp=1.0, steps_flux=True):
    """The strip waveguide at full size (384 x 192 x 160, 12-cell PML,
    anisotropically smoothed strip)."""
    res = 20
    fcen = 1 / 1.55
    src = [kh.GaussianProfileSource(time_profile=kh.GaussianPulseSource(fcen=fcen, fwidth=0.2 * fcen),
                                    component=kh.Ez(), center=[-7.0, 0.0, 0.0], size=[0.0, 8.4, 6.8],
                                    waist=1.0, amplitude=amp)]
    geo = [kh.Object(kh.Box([0.0, 0.0, 0.0], [30.0, 0.5, 0.22]), kh.Material(ε=12.1)),
           kh.Object(kh.Box([0.0, 0.0, 0.0], [30.0, 30.0, 30.0]), kh.Material(ε=2.085))]
    freqs = list(np.linspace(0.8 * fcen, 1.2 * fcen, 50))
    mons = []
    if steps_flux:
        mons = [kh.FluxMonitor(center=[-6.0, 0.0, 0.0], size=[0.0, 8.4, 6.8], frequencies=freqs),
                kh.FluxMonitor(center=[6.0, 0.0, 0.0], size=[0.0, 8.4, 6.8], frequencies=freqs)]
    return kh.Simulation(cell_size=[19.2, 9.6, 8.0], resolution=res, sources=src, geometry=geo,
                         monitors=mons, boundaries=[[0.6, 0.6]] * 3, precision=precision, mode=mode,
                         subpixel_smoothing=kh.AnisotropicSmoothing())


def test_strip_waveguide_full_size_fast_vs_exact_and_linearity():
    steps = 300
    _note("[parity] strip waveguide 384x192x160 f32 fast/exact/2x-amplitude, 300 steps each")
    a = _strip_sim("f32", "fast")
    b = _strip_sim("f32", "exact")
    c = _strip_sim("f32", "fast", amp=2.0)
    for s in (a, b, c):
        s.step(steps)
    tot, per = field_errors(a, b)
    assert tot < 1e-5 and max(per.values()) < 1e-5, (tot, per)
    fas = [a.read_field(C()) for C in COMPS]
    fcs = [c.read_field(C()) for C in COMPS]
    norm = np.sqrt(sum(float(np.sum(2.0 * f.astype(np.float64) ** 2 * 2.0)) for f in fas))
    for C, fa, fc in zip(COMPS, fas, fcs):
        assert np.all(np.isfinite(fa))
        # scaling every input by 2 scales every rounding by 2, except where
        # values fall into the subnormal range (wavefront edge, deep PML); a
        # numerically zero cross-polarized component (|Hz| ~ 1e-19 here) is
        # measured against the whole field
        d = np.linalg.norm(fc.astype(np.float64) - 2.0 * fa.astype(np.float64))
        assert d / norm < 1e-6, C.__name__
    del a, b, c


@pytest.mark.parametrize("mode,tol", [("exact", 0.0), ("fast", 1e-12)])
def test_twelve_point_sources_and_line_sources(mode, tol):
    """Sixteen sources (src/Sources/Sources.jl:288-357 injects every source of
    its list): twelve Ez dipoles with different pulses, two of them on the
    same cell, more than one launch holds, and four Hx/Hy line sources, FP64,
    against the oracle (EXACT value for value, FAST within rel-L2 1e-12)."""
    def mk(engine=None, m="exact"):
        rng = np.random.default_rng(12)
        src = []
        for k in range(12):
            c = [0.0, 0.0, 0.0] if k in (3, 7) else [float(v) for v in rng.uniform(-0.6, 0.6, 3)]
            src.append(kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=0.8 + 0.05 * k, fwidth=0.4),
                                        component=kh.Ez(), center=c, size=[0.0, 0.0, 0.0],
                                        amplitude=1.0 + 0.1 * k))
        for k, C in enumerate((kh.Hx, kh.Hy, kh.Hx, kh.Hy)):
            src.append(kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.5), component=C(),
                                        center=[-0.3 + 0.2 * k, 0.1, -0.2], size=[0.0, 0.4, 0.0]))
        mon = [kh.DFTMonitor(component=kh.Ez(), center=[0.0, 0.0, 0.1], size=[1.4, 1.4, 0.0],
                             frequencies=[0.9, 1.1])]
        return kh.Simulation(cell_size=[2.0, 2.0, 2.0], resolution=10, sources=src, monitors=mon,
                             boundaries=[[0.3, 0.3]] * 3, precision="f64", mode=m, engine=engine)
    g, o = mk(m=mode), mk(OracleEngine)
    _note(f"[parity] 16 sources, {mode}")
    g.step(70)
    o.step(70)
    tot, per = field_errors(g, o)
    assert tot <= tol, (tot, per)
    assert rel_l2(g.engine.read_dft(0), o.engine.read_dft(0)) <= max(tol, 1e-13)


@pytest.mark.parametrize("mode,tol", [("exact", 0.0), ("fast", 1e-12)])
def test_tfsf_and_gaussian_beam_sources(mode, tol):
    """A TFSF box (six equivalent-source faces, src/Sources/SpatialSources.jl:
    297-425) and a Gaussian beam launched from a plane (:650-857) through the
    HIP library, against the oracle, FP64 (EXACT value for value, FAST within
    rel-L2 1e-12)."""
    def mk(engine=None, m="exact"):
        tp = kh.GaussianPulseSource(fcen=1.0, fwidth=0.4)
        src = kh.TFSFSource(time_profile=tp, size=[1.0, 0.8, 1.2], center=[0.1, 0.0, 0.0], injection_axis=3,
                            direction=-1, polarization_angle=0.4, n_grid=12)
        src.append(kh.GaussianBeamSource(time_profile=kh.GaussianPulseSource(fcen=0.9, fwidth=0.3),
                                         beam_center=[0.0, 0.1, -0.4], beam_waist=0.5, k_vector=[0.2, 0.0, 1.0],
                                         polarization=[1.0, 0.3, 0.0], center=[0.0, 0.0, -0.7],
                                         size=[1.6, 1.6, 0.0]))
        mon = [kh.FluxMonitor(center=[0.0, 0.0, 0.8], size=[1.6, 1.6, 0.0], frequencies=[0.9, 1.1])]
        return kh.Simulation(cell_size=[2.2, 2.2, 2.4], resolution=10, sources=src, monitors=mon,
                             boundaries=[[0.3, 0.3]] * 3, precision="f64", mode=m, engine=engine)
    g, o = mk(m=mode), mk(OracleEngine)
    _note(f"[parity] TFSF + Gaussian beam, {mode}")
    g.step(90)
    o.step(90)
    tot, per = field_errors(g, o)
    assert tot <= tol, (tot, per)
    for i in range(len(o.spec.dfts)):
        assert rel_l2(g.engine.read_dft(i), o.engine.read_dft(i)) <= max(tol, 1e-13)
