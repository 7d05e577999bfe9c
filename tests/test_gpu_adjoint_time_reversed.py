"""Time-reversed adjoint on the GPU (kh_set_checkpoints + kh_adjoint_gradient).

The gradient is the exact derivative of the discrete objective, so it is
pinned by central finite differences of device forward runs to far below the
reference's own acceptance (median < 10 %, test/test_adjoint_gradient.jl:255):
float64 directional derivatives agree to 1e-6 relative, in 2D (PML on x, y)
and in 3D with PML on every axis (every sigma combination of the cascade,
all three E components in the design box), in FAST and EXACT modes.  The
checkpoint interval changes nothing (bit-identical gradients), the forward
state and monitors are back on return, the torch reverse rule returns the
same gradient, the frequency-domain adjoint agrees to its own accuracy, and
float32 agrees with float64.
"""
import sys

import numpy as np
import pytest
import torch

import khronos_hip as kh
from khronos_hip import adjoint as A
from khronos_hip.grid import Volume
from khronos_hip._lib import KhronosError
from test_adjoint import slab2d_problem, slab_problem

pytestmark = pytest.mark.gpu


def _note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _time_reversed(opt, every):
    opt.method = "time_reversed"
    opt.checkpoint_every = every
    return opt


def box3d_problem(precision="f64", mode="fast", until=30.0, freqs=(0.45, 0.55)):
    """3D box with PML on all six faces: Ez sheet source, 5x4 design block,
    Ez/Ey objective plane behind it."""
    pml = 0.5
    cell = [3.0, 2.6, 1.6]
    geo = [kh.Object(kh.Box([0, 0, 0], [10.0, 10.0, 10.0]), kh.Material(ε=1.0))]
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=0.5, fwidth=0.15), component=kh.Ez(),
                            center=[-0.9, 0.0, 0.0], size=[0.0, 1.2, 0.5])]
    sim = kh.Simulation(cell_size=cell, resolution=10, sources=src, geometry=geo,
                        boundaries=[[pml, pml], [pml, pml], [pml, pml]], precision=precision, mode=mode)
    dr = A.DesignRegion(volume=Volume([0, 0, 0], [1.0, 1.2, 0.5]), design_parameters=np.full(20, 0.5),
                        grid_size=(5, 4), ε_min=1.0, ε_max=4.0)
    objs = [A.FourierFieldsObjective(volume=Volume([0.8, 0, 0], [0, 1.2, 0.5]), component=kh.Ez()),
            A.FourierFieldsObjective(volume=Volume([0.8, 0, 0], [0, 1.2, 0.5]), component=kh.Ey())]
    f = lambda E1, E2: torch.sum(torch.abs(E1) ** 2) + 0.5 * torch.sum(torch.abs(E2) ** 2)
    return A.OptimizationProblem(sim, [f], objs, [dr], list(freqs), until=until)


def _directional(opt, rho, n_dir=3, h=1e-4, seed=7):
    f0, g = opt(rho)
    g = g[:, 0]
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_dir):
        d = rng.standard_normal(len(rho))
        d /= np.linalg.norm(d)
        fp, _ = opt(rho + h * d, need_gradient=False)
        fm, _ = opt(rho - h * d, need_gradient=False)
        fd = (fp - fm) / (2 * h)
        ad = float(g @ d)
        out.append((ad, fd, abs(ad - fd) / max(abs(ad), abs(fd), 1e-300)))
    return out, g


@pytest.mark.parametrize("mode", ["fast", "exact"])
def test_2d_gradient_equals_finite_differences(mode):
    opt = _time_reversed(slab2d_problem(engine=None, precision="f64", mode=mode, freqs=(0.45, 0.55)), 37)
    rho = 0.3 + 0.4 * np.random.default_rng(42).random(24)
    _note(f"[adjoint-tr] 2D {mode}")
    res, g = _directional(opt, rho)
    _note(f"[adjoint-tr] 2D {mode} (adjoint, fd, rel): {res}")
    assert np.any(g != 0)
    assert max(r[2] for r in res) < 1e-6, res


@pytest.mark.parametrize("mode", ["fast", "exact"])
def test_3d_pml_every_axis_gradient_equals_finite_differences(mode):
    opt = _time_reversed(box3d_problem(mode=mode), 50)
    rho = 0.3 + 0.4 * np.random.default_rng(5).random(20)
    _note(f"[adjoint-tr] 3D {mode}")
    res, g = _directional(opt, rho)
    _note(f"[adjoint-tr] 3D {mode} (adjoint, fd, rel): {res}")
    assert max(r[2] for r in res) < 1e-6, res


def test_3d_per_pixel_finite_differences():
    """The reference's check: central differences at sampled pixels
    (src/Adjoint/GradientKernel.jl:117-152)."""
    opt = _time_reversed(box3d_problem(), 64)
    rho = 0.3 + 0.4 * np.random.default_rng(11).random(20)
    _, g = opt(rho)
    fd, idx = opt.calculate_fd_gradient(rho, db=1e-4, n_samples=6, seed=3)
    rel = np.abs(fd[:, 0] - g[idx, 0]) / np.max(np.abs(g))
    _note(f"[adjoint-tr] per-pixel rel errors {rel}")
    assert np.max(rel) < 1e-6, (fd[:, 0], g[idx, 0])


def test_checkpoint_interval_does_not_change_the_gradient():
    rho = 0.3 + 0.4 * np.random.default_rng(2).random(24)
    grads = []
    for every in (1000000, 100, 7):
        opt = _time_reversed(slab2d_problem(engine=None, precision="f64", mode="fast"), every)
        grads.append(opt(rho)[1])
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[0], grads[2])


@pytest.mark.parametrize("mode", ["fast", "exact"])
def test_captured_sweep_equals_eager_sweep(mode):
    """Each full checkpoint interval's re-run, design-box snapshots and walk
    back replay one captured graph (the seeds read the device clock); with
    graphs off every launch is eager.  The same gradient bit for bit, over
    intervals with the sources on, switching off and off, and a partial last
    interval."""
    rho = 0.3 + 0.4 * np.random.default_rng(6).random(24)
    for every in (7, 50):
        opt = _time_reversed(slab2d_problem(engine=None, precision="f64", mode=mode), every)
        opt.update_design(rho)
        opt.forward_run()
        eng = opt._fwd_sim.engine
        _note(f"[adjoint-tr] captured vs eager sweep, {mode}, {eng.timestep()} steps, intervals of {every}")
        g_graph = opt._time_reversed_gradient()
        eng.set_graph(0)
        g_eager = opt._time_reversed_gradient()
        assert np.array_equal(g_graph, g_eager)
        assert np.any(g_graph != 0)


def plane3d_problem(mode="fast", until=12.0):
    """3D, PML on every face: an Ez sheet source of 38 x 28 cells (too large
    for the end-of-lane form, so its corrections are computed per block of
    steps), a 5x4 design block and an Ez objective plane."""
    pml = 0.3
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=0.5, fwidth=0.15), component=kh.Ez(),
                            center=[-1.2, 0.0, 0.0], size=[0.0, 2.8, 2.8])]
    sim = kh.Simulation(cell_size=[3.2, 3.0, 3.0], resolution=10, sources=src,
                        boundaries=[[pml, pml]] * 3, precision="f64", mode=mode)
    dr = A.DesignRegion(volume=Volume([0, 0, 0], [1.0, 1.2, 0.5]), design_parameters=np.full(20, 0.5),
                        grid_size=(5, 4), ε_min=1.0, ε_max=4.0)
    objs = [A.FourierFieldsObjective(volume=Volume([0.9, 0, 0], [0, 1.2, 0.5]), component=kh.Ez())]
    f = lambda E1: torch.sum(torch.abs(E1) ** 2)
    return A.OptimizationProblem(sim, [f], objs, [dr], [0.45, 0.55], until=until)


def test_block_precomputed_sources_in_the_sweep(monkeypatch):
    """A plane source whose FAST corrections are computed per block of steps
    and added by the PML update: the forward run and the sweep (captured
    intervals, and the eager ones of a partial last interval) give the same
    objective and gradient bit for bit as with every source in a separate
    launch (KH_SRC_LAUNCH=1, read when a context plans its first step)."""
    rho = 0.3 + 0.4 * np.random.default_rng(8).random(20)
    out = []
    for env in (None, "1"):
        if env:
            monkeypatch.setenv("KH_SRC_LAUNCH", env)
        opt = _time_reversed(plane3d_problem(), 70)
        f, g = opt(rho)
        n = opt._fwd_sim.engine.timestep()
        out.append((f, g))
        opt._close_forward()
        monkeypatch.delenv("KH_SRC_LAUNCH", raising=False)
    _note(f"[adjoint-tr] plane source, block-precomputed vs separate launches, {n} steps, intervals of 70")
    assert n % 70 != 0
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1]) and np.any(out[0][1] != 0)


def test_forward_state_and_monitors_restored():
    opt = _time_reversed(slab2d_problem(engine=None, precision="f64", mode="fast"), 50)
    rho = 0.3 + 0.4 * np.random.default_rng(4).random(24)
    opt.update_design(rho)
    opt.forward_run()
    eng = opt._fwd_sim.engine
    ts = eng.timestep()
    before = [eng.read_dft(0)] + [eng.read_field(c) for c in range(6)]
    opt.adjoint_run()
    opt.calculate_gradient()
    after = [eng.read_dft(0)] + [eng.read_field(c) for c in range(6)]
    assert eng.timestep() == ts
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    # the restored context keeps stepping as the original run would
    eng.step(10)
    ref = _time_reversed(slab2d_problem(engine=None, precision="f64", mode="fast"), 50)
    ref.update_design(rho)
    ref.forward_run()
    ref._fwd_sim.engine.step(10)
    assert np.array_equal(eng.read_dft(0), ref._fwd_sim.engine.read_dft(0))


def test_torch_reverse_rule_time_reversed():
    opt = _time_reversed(slab2d_problem(engine=None, precision="f64", mode="fast", freqs=(0.45, 0.55)), 40)
    rho = torch.tensor(0.3 + 0.4 * np.random.default_rng(1).random(24), requires_grad=True)
    f = A.simulate_torch(opt, rho)
    (3.0 * f.sum()).backward()
    assert opt.gradient.shape == (24, 1)
    # the sweep ran with the cotangent as objective weight; an unweighted
    # sweep of the same forward run gives a third of it
    g1 = opt.calculate_gradient()[:, 0]
    assert np.allclose(rho.grad.numpy(), 3.0 * g1, rtol=1e-12, atol=0)


def test_agrees_with_frequency_domain_adjoint():
    """Both gradients of the same objective: the frequency-domain one is
    approximate for a finite run (its FD acceptance here is a 3 % median),
    the time-reversed one exact."""
    rho = 0.3 + 0.4 * np.random.default_rng(8).random(24)
    fr = slab2d_problem(engine=None, precision="f64", mode="fast", freqs=(0.45, 0.55))
    g_f = fr(rho)[1].sum(axis=1)
    tr = _time_reversed(slab2d_problem(engine=None, precision="f64", mode="fast", freqs=(0.45, 0.55)), 100)
    g_t = tr(rho)[1][:, 0]
    cos = float(g_f @ g_t / (np.linalg.norm(g_f) * np.linalg.norm(g_t)))
    rel = np.linalg.norm(g_f - g_t) / np.linalg.norm(g_t)
    _note(f"[adjoint-tr] frequency vs time-reversed: cos {cos:.6f} rel {rel:.3e}")
    assert cos > 0.99 and rel < 0.1


def test_float32_agrees_with_float64():
    """FP32 fields on the FP64 run's setup (the precision form: the same
    discrete problem) against the FP64 gradient; and the reference-style FP32
    run (FP32 setup, a slightly different discrete problem) against its own
    frequency-domain gradient."""
    rho = 0.3 + 0.4 * np.random.default_rng(9).random(24)
    g64 = _time_reversed(slab2d_problem(engine=None, precision="f64", mode="fast"), 100)(rho)[1][:, 0]
    g32 = _time_reversed(slab2d_problem(engine=None, precision="f32", mode="fast", setup_precision="f64"),
                         100)(rho)[1][:, 0]
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    _note(f"[adjoint-tr] f32 fields (f64 setup) vs f64 rel {rel:.3e}")
    gt = _time_reversed(slab2d_problem(engine=None, precision="f32", mode="fast"), 100)(rho)[1][:, 0]
    gf = slab2d_problem(engine=None, precision="f32", mode="fast")(rho)[1].sum(axis=1)
    rel_f = np.linalg.norm(gt - gf) / np.linalg.norm(gt)
    _note(f"[adjoint-tr] f32 time-reversed vs f32 frequency-domain rel {rel_f:.3e}")
    assert rel < 2e-3
    assert rel_f < 0.1


def test_rejects_what_it_does_not_cover():
    # periodic z (the quasi-1D slab): refused before any sweep
    opt = _time_reversed(slab_problem(engine=None, precision="f64", mode="fast"), 50)
    with pytest.raises(KhronosError):
        opt(0.5 * np.ones(24))
    # no checkpoints: the plain forward engine refuses the sweep
    opt2 = slab2d_problem(engine=None, precision="f64", mode="fast")
    opt2.update_design(0.5 * np.ones(24))
    spec = opt2._spec_with(opt2.base.sources, [], rho=opt2.rho)
    s = opt2._run(spec)
    with pytest.raises(KhronosError):
        s.engine.adjoint_gradient([], [], [None, None, None])


def test_bend_400x400_full_size():
    """The 2D bend at full size (400 x 400 cells, 120 x 120 design
    parameters rho ~ U[0, 1], 2000 steps, 3 frequencies, checkpoints every
    100 steps): FP32 fields (FP64 setup) and FP64 gradients against FP64
    central differences at 20 seeded pixels."""
    import time
    from khronos_hip import workloads as W
    opt64, rho = W.adjoint_bend(precision="f64")
    opt64.update_design(rho)
    t0 = time.perf_counter()
    f64 = opt64.forward_run()
    t1 = time.perf_counter()
    opt64.adjoint_run()
    g64 = opt64.calculate_gradient()[:, 0]
    t2 = time.perf_counter()
    _note(f"[adjoint-tr] bend f64: J {f64:.6e}, forward {t1 - t0:.3f} s, gradient {t2 - t1:.3f} s, "
          f"{opt64._fwd_sim.engine.timestep()} steps")
    opt32, _ = W.adjoint_bend(precision="f32", setup_precision="f64")
    opt32.update_design(rho)
    t0 = time.perf_counter()
    f32 = opt32.forward_run()
    t1 = time.perf_counter()
    opt32.adjoint_run()
    g32 = opt32.calculate_gradient()[:, 0]
    t2 = time.perf_counter()
    _note(f"[adjoint-tr] bend f32: J {f32:.6e}, forward {t1 - t0:.3f} s, gradient {t2 - t1:.3f} s")
    fd, idx = opt64.calculate_fd_gradient(rho, db=1e-3, n_samples=20, seed=0)
    scale = np.max(np.abs(g64))
    e64 = np.abs(fd[:, 0] - g64[idx]) / scale
    e32 = np.abs(fd[:, 0] - g32[idx]) / scale
    _note(f"[adjoint-tr] bend FD at 20 pixels: max rel err f64 {e64.max():.3e}, f32 {e32.max():.3e}; "
          f"f32 vs f64 gradient rel {np.linalg.norm(g32 - g64) / np.linalg.norm(g64):.3e}")
    assert abs(f32 - f64) <= 1e-4 * abs(f64)
    assert e64.max() < 1e-4, (fd[:, 0], g64[idx])
    assert e32.max() < 1e-3, (fd[:, 0], g32[idx])


def test_h_field_objective_and_decimated_monitor(monkeypatch):
    """An objective on Hy (seeded at the H time stamps) plus Ez, with the
    objective monitors sampling every 3rd step: still equal to central
    differences (the frequency-domain path has no H-field adjoint sources)."""
    import functools
    monkeypatch.setattr(A, "DFTMonitor", functools.partial(A.DFTMonitor, decimation=3))
    opt = slab2d_problem(engine=None, precision="f64", mode="fast", freqs=(0.45, 0.55))
    opt.objective_arguments = [
        A.FourierFieldsObjective(volume=Volume([0.75, 0, 0], [0, 0.6, 0.0]), component=kh.Hy()),
        A.FourierFieldsObjective(volume=Volume([0.75, 0, 0], [0, 0.6, 0.0]), component=kh.Ez())]
    opt.objective_functions = [lambda h, e: torch.sum(torch.abs(h) ** 2) * 100.0 + 0.1 * torch.sum(torch.abs(e) ** 2)]
    _time_reversed(opt, 45)
    rho = 0.3 + 0.4 * np.random.default_rng(12).random(24)
    res, _ = _directional(opt, rho)
    decs = [md.decimation for md in opt._fwd_spec.dfts]
    _note(f"[adjoint-tr] Hy + Ez objective, decimations {decs}: {res}")
    assert max(decs) > 1
    assert max(r[2] for r in res) < 1e-6, res


def test_overlapping_same_component_objectives():
    """Two Ez objectives whose planes overlap (and one listed twice): their
    seeds write common cells of one adjoint array, so they go to successive
    launches.  Central differences pin the sum, and two sweeps of the same
    forward run are bit-identical (no lost updates that vary run to run)."""
    opt = box3d_problem()
    plane = Volume([0.8, 0, 0], [0, 1.2, 0.5])
    part = Volume([0.8, 0.1, 0], [0, 0.6, 0.3])
    opt.objective_arguments = [A.FourierFieldsObjective(volume=plane, component=kh.Ez()),
                               A.FourierFieldsObjective(volume=part, component=kh.Ez()),
                               A.FourierFieldsObjective(volume=plane, component=kh.Ez())]
    opt.objective_functions = [lambda a, b, c: torch.sum(torch.abs(a) ** 2) + 3.0 * torch.sum(torch.abs(b) ** 2)
                               + 0.5 * torch.sum(torch.abs(c) ** 2)]
    _time_reversed(opt, 50)
    rho = 0.3 + 0.4 * np.random.default_rng(21).random(20)
    res, g = _directional(opt, rho)
    _note(f"[adjoint-tr] overlapping Ez objectives (adjoint, fd, rel): {res}")
    assert max(r[2] for r in res) < 1e-6, res
    g2 = opt.calculate_gradient()[:, 0]
    opt.adjoint_run()
    g3 = opt.calculate_gradient()[:, 0]
    assert np.array_equal(g2, g3)


def pml_design_2d(mode):
    """2D TE (Ey line source; Ex, Ey, Hz) with the design region reaching
    into the +x and both y PMLs, so Ex has a non-zero own-axis sigma in part
    of the design (the W stage carries the update there) and Ey in another."""
    pml, cx, cy = 1.0, 4.0, 2.6
    geo = [kh.Object(kh.Box([0, 0, 0], [cx + 1, cy + 1, 10.0]), kh.Material(ε=1.0))]
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=0.5, fwidth=0.15), component=kh.Ey(),
                            center=[-0.6, 0.0, 0.0], size=[0.0, 0.6, 0.0])]
    sim = kh.Simulation(cell_size=[cx, cy, 0.0], resolution=10, sources=src, geometry=geo,
                        boundaries=[[pml, pml], [pml, pml]], precision="f64", mode=mode)
    dr = A.DesignRegion(volume=Volume([1.0, 0, 0], [1.0, 0.8, 0.0]), design_parameters=np.full(24, 0.5),
                        grid_size=(6, 4), ε_min=1.0, ε_max=4.0)
    objs = [A.FourierFieldsObjective(volume=Volume([0.2, 0, 0], [0, 0.6, 0.0]), component=kh.Ey()),
            A.FourierFieldsObjective(volume=Volume([0.2, 0, 0], [0, 0.6, 0.0]), component=kh.Ex())]
    f = lambda ey, ex: torch.sum(torch.abs(ey) ** 2) + 10.0 * torch.sum(torch.abs(ex) ** 2)
    return A.OptimizationProblem(sim, [f], objs, [dr], [0.45, 0.55], until=30.0)


@pytest.mark.parametrize("mode", ["fast", "exact"])
def test_design_region_inside_the_pml(mode):
    opt = _time_reversed(pml_design_2d(mode), 40)
    rho = 0.3 + 0.4 * np.random.default_rng(17).random(24)
    res, g = _directional(opt, rho)
    _note(f"[adjoint-tr] design in the PML, {mode} (adjoint, fd, rel): {res}")
    assert np.any(g != 0)
    assert max(r[2] for r in res) < 1e-6, res


def slab2d_pml_design_problem(precision="f64", mode="fast", until=50.0, freqs=(0.45, 0.55)):
    """2D, Ex/Ey/Hz fields (Ey line source): a 6 x 4 design region from x = 0
    to x = 1.6 that reaches 0.6 into the x PML (PML from x = 1), so the Ex
    samples there sit where their own-axis sigma is not zero (the forward
    run carries them on W; the transposed half step takes the own-axis
    branch)."""
    pml, cx, cy = 1.0, 4.0, 2.6
    geo = [kh.Object(kh.Box([0, 0, 0], [cx + 1, cy + 1, 10.0]), kh.Material(ε=1.0))]
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=0.5, fwidth=0.15), component=kh.Ey(),
                            center=[-0.6, 0.0, 0.0], size=[0.0, cy - 2 * pml, 0.0])]
    sim = kh.Simulation(cell_size=[cx, cy, 0.0], resolution=10, sources=src, geometry=geo,
                        boundaries=[[pml, pml], [pml, pml]], precision=precision, mode=mode)
    dr = A.DesignRegion(volume=Volume([0.8, 0, 0], [1.6, cy - 2 * pml, 0.0]), design_parameters=np.full(24, 0.5),
                        grid_size=(6, 4), ε_min=1.0, ε_max=4.0)
    objs = [A.FourierFieldsObjective(volume=Volume([-0.8, 0, 0], [0, cy - 2 * pml, 0.0]), component=kh.Ey()),
            A.FourierFieldsObjective(volume=Volume([-0.8, 0, 0], [0, cy - 2 * pml, 0.0]), component=kh.Ex())]
    f = lambda E1, E2: torch.sum(torch.abs(E1) ** 2) + torch.sum(torch.abs(E2) ** 2)
    return A.OptimizationProblem(sim, [f], objs, [dr], list(freqs), until=until)


@pytest.mark.parametrize("mode", ["fast", "exact"])
def test_design_region_inside_the_x_pml(mode):
    opt = _time_reversed(slab2d_pml_design_problem(mode=mode), 41)
    rho = 0.3 + 0.4 * np.random.default_rng(19).random(24)
    _note(f"[adjoint-tr] 2D design region in the x PML, {mode}")
    res, g = _directional(opt, rho)
    _note(f"[adjoint-tr] 2D PML design {mode} (adjoint, fd, rel): {res}")
    assert np.any(g != 0)
    assert max(r[2] for r in res) < 1e-6, res


def _box3d_source_in_design(precision="f32"):
    """box3d_problem with its Ez sheet source moved inside the design block:
    the source writes design-box cells of Ez after the E update's stores, so
    the re-run keeps the separate snapshot copy."""
    opt = box3d_problem(precision=precision, until=20.0)
    opt.sim.sources[0] = kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=0.5, fwidth=0.15),
                                          component=kh.Ez(), center=[0.1, 0.0, 0.0], size=[0.0, 1.0, 0.4])
    return opt


@pytest.mark.parametrize("case", ["2d_x_pml_f64", "3d_f32", "3d_plane_source_f64", "3d_source_in_design_f32"])
def test_snapshots_written_by_the_update_equal_separate_copies(case, monkeypatch):
    """FAST mode: the re-run's E update writes the design-box snapshot from
    its registers (W where the own-axis sigma is non-zero, else E), instead
    of a box_copy_kernel launch after each step (KH_SNAP_LAUNCH=1, read per
    gradient).  The same gradient bit for bit, captured and eager, with the
    design region in the x PML, in 3D with PML on every axis, with
    block-precomputed source corrections, and with a source inside the
    design region (which keeps the separate copy)."""
    make, n, every = {
        "2d_x_pml_f64": (lambda: slab2d_pml_design_problem(mode="fast"), 24, 41),
        "3d_f32": (lambda: box3d_problem(precision="f32", until=20.0), 20, 30),
        "3d_plane_source_f64": (lambda: plane3d_problem(), 20, 70),
        "3d_source_in_design_f32": (_box3d_source_in_design, 20, 30),
    }[case]
    opt = _time_reversed(make(), every)
    rho = 0.3 + 0.4 * np.random.default_rng(23).random(n)
    opt.update_design(rho)
    opt.forward_run()
    eng = opt._fwd_sim.engine
    _note(f"[adjoint-tr] snapshots by the E update vs box copies, {case}, {eng.timestep()} steps, "
          f"intervals of {every}")
    g_fused = opt._time_reversed_gradient()
    monkeypatch.setenv("KH_SNAP_LAUNCH", "1")
    g_copy = opt._time_reversed_gradient()
    eng.set_graph(0)
    g_copy_eager = opt._time_reversed_gradient()
    monkeypatch.delenv("KH_SNAP_LAUNCH")
    g_fused_eager = opt._time_reversed_gradient()
    assert np.any(g_fused != 0)
    for g in (g_copy, g_copy_eager, g_fused_eager):
        assert np.array_equal(g_fused, g)


def test_seed_weight_table_equals_clock_weights(monkeypatch):
    """The objective's seeds take their weights from a table formed by one
    launch per interval (KH_SEED_TABLE=0: each lane forms them from the
    device clock): Hy and Ez objectives sampled every 3rd step, captured and
    eager intervals, the same gradient bit for bit."""
    import functools
    monkeypatch.setattr(A, "DFTMonitor", functools.partial(A.DFTMonitor, decimation=3))
    opt = slab2d_problem(engine=None, precision="f64", mode="fast", freqs=(0.45, 0.5, 0.55))
    opt.objective_arguments = [
        A.FourierFieldsObjective(volume=Volume([0.75, 0, 0], [0, 0.6, 0.0]), component=kh.Hy()),
        A.FourierFieldsObjective(volume=Volume([0.75, 0, 0], [0, 0.6, 0.0]), component=kh.Ez())]
    opt.objective_functions = [lambda h, e: torch.sum(torch.abs(h) ** 2) * 100.0 + 0.1 * torch.sum(torch.abs(e) ** 2)]
    _time_reversed(opt, 45)
    opt.update_design(0.3 + 0.4 * np.random.default_rng(12).random(24))
    opt.forward_run()
    eng = opt._fwd_sim.engine
    _note(f"[adjoint-tr] seed weight table vs clock weights, {eng.timestep()} steps, intervals of 45")
    g_tab = opt._time_reversed_gradient()
    monkeypatch.setenv("KH_SEED_TABLE", "0")
    g_clk = opt._time_reversed_gradient()
    eng.set_graph(0)
    g_clk_eager = opt._time_reversed_gradient()
    monkeypatch.delenv("KH_SEED_TABLE")
    g_tab_eager = opt._time_reversed_gradient()
    assert np.any(g_tab != 0)
    for g in (g_clk, g_clk_eager, g_tab_eager):
        assert np.array_equal(g_tab, g)


def test_more_than_sixteen_objective_monitors():
    """Eighteen Ez objective lines: past sixteen monitors of a group the
    seeds go to their own launches before the transposed half step.  Equal
    to central differences."""
    opt = slab2d_problem(engine=None, precision="f64", mode="fast", freqs=(0.45, 0.55))
    ys = np.linspace(-0.3, 0.3, 18)
    opt.objective_arguments = [A.FourierFieldsObjective(volume=Volume([0.75, float(y), 0], [0, 0.0, 0.0]),
                                                        component=kh.Ez()) for y in ys]
    opt.objective_functions = [lambda *e: sum((1.0 + 0.1 * i) * torch.sum(torch.abs(x) ** 2) for i, x in enumerate(e))]
    _time_reversed(opt, 45)
    rho = 0.3 + 0.4 * np.random.default_rng(13).random(24)
    res, g = _directional(opt, rho)
    _note(f"[adjoint-tr] 18 Ez objective monitors (adjoint, fd, rel): {res}")
    assert len(opt._fwd_spec.dfts) == 18 and np.any(g != 0)
    assert max(r[2] for r in res) < 1e-6, res


@pytest.mark.parametrize("case", ["2d_x_pml_f64", "3d_plane_source_f64"])
def test_two_stream_sweep_equals_one_stream(case, monkeypatch):
    """The walk back of each interval runs on a second stream while the next
    earlier interval is re-run (snapshots and seed weights in two buffers by
    interval parity; KH_SWEEP_STREAMS=0, read per gradient: one stream).
    The same gradient bit for bit, captured and eager, with a partial last
    interval."""
    make, n, every = {
        "2d_x_pml_f64": (lambda: slab2d_pml_design_problem(mode="fast"), 24, 41),
        "3d_plane_source_f64": (lambda: plane3d_problem(), 20, 70),
    }[case]
    opt = _time_reversed(make(), every)
    rho = 0.3 + 0.4 * np.random.default_rng(29).random(n)
    opt.update_design(rho)
    opt.forward_run()
    eng = opt._fwd_sim.engine
    _note(f"[adjoint-tr] two-stream vs one-stream sweep, {case}, {eng.timestep()} steps, intervals of {every}")
    assert eng.timestep() % every != 0 and eng.timestep() // every >= 3
    g_two = opt._time_reversed_gradient()
    g_two_again = opt._time_reversed_gradient()
    monkeypatch.setenv("KH_SWEEP_STREAMS", "0")
    g_one = opt._time_reversed_gradient()
    eng.set_graph(0)
    g_one_eager = opt._time_reversed_gradient()
    monkeypatch.delenv("KH_SWEEP_STREAMS")
    g_two_eager = opt._time_reversed_gradient()
    assert np.any(g_two != 0)
    for g in (g_two_again, g_one, g_one_eager, g_two_eager):
        assert np.array_equal(g_two, g)
