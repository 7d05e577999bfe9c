"""The reference's default 3D + PML chunk plan on the oracle (quirk Q1;
src/Chunking.jl:622-738, :907-923, :1090-1122, :1355-1378;
src/Kernels/Helpers.jl:200-223).

In the 27-chunk plan a chunk holds the U array of a component only when it
has PML along the component's next axis, and its profiles off its PML axes
are zero-filled; without U (and without a conductivity array) the curl stage
resolves to T' = T + K, dropping the previous axis's damping.  Pins:
* one step from a random state changes exactly the B values of the cells
  where that rule applies (a PML chunk, no PML on the next axis, a non-zero
  previous-axis profile) and leaves every other B value of the single-chunk
  step unchanged, bit for bit (the H half step comes first, from the same E;
  the E half step then differs wherever the new H does);
* a dipole run equals the single-chunk run bit for bit until the pulse
  reaches the PML, and then grows without bound: the x, y and z PML slabs
  damp one field component of each polarization and not the other, and the
  mixed system is not passive.
"""
import numpy as np
import pytest

import khronos_hip as kh
from oracle_engine import OracleEngine, chunk27_intervals

COMPS = [kh.Ex, kh.Ey, kh.Ez, kh.Hx, kh.Hy, kh.Hz]


def _box(n=16, pml=0.5, res=10):
    src = [kh.UniformSource(time_profile=kh.GaussianPulseSource(fcen=1.0, fwidth=0.5), component=kh.Ez(),
                            center=[0.0, 0.0, 0.0], size=[0.0, 0.0, 0.0])]
    s = kh.Simulation(cell_size=[n / res] * 3, resolution=res, sources=src, boundaries=[[pml, pml]] * 3,
                      precision="f64", mode="exact", engine=OracleEngine)
    s.prepare()
    return s


def test_one_step_changes_exactly_the_cells_without_u():
    a, b = _box(), _box()
    e, s = chunk27_intervals(b)
    assert e == [5, 5, 5] and s == [12, 12, 12]
    b.engine.set_chunk27(e, s)
    rng = np.random.default_rng(7)
    from khronos_hip.grid import component_code
    for C in COMPS:
        code = component_code(C())
        for which in (0, 1):  # field, D/B
            shape = a.engine.read(which, code).shape
            v = rng.standard_normal(shape)
            a.engine.write(which, code, v)
            b.engine.write(which, code, v)
    a.step(1)
    b.step(1)
    N = [a.Nx, a.Ny, a.Nz]
    sig = a.spec.sigma
    for C in (kh.Hx, kh.Hy, kh.Hz):
        c = C()
        code = component_code(c)
        electric = c.electric
        own = c.axis
        nxt, prv = (own + 1) % 3, (own + 2) % 3
        prof = sig["D" if electric else "B"]
        da, db = a.engine.read(1, code), b.engine.read(1, code)
        expect = np.zeros(da.shape, dtype=bool)
        for i in range(1, N[0] + 1):
            for j in range(1, N[1] + 1):
                for k in range(1, N[2] + 1):
                    g = (i, j, k)
                    hp = [g[x] <= e[x] or g[x] >= s[x] for x in range(3)]
                    sp = float(np.asarray(prof[prv])[2 * g[prv] - 2])
                    if any(hp) and not hp[nxt] and hp[prv] and sp != 0.0:
                        expect[i, j, k] = True
        differ = da != db
        assert np.array_equal(differ, expect), C.__name__
        assert expect.sum() > 0
        # the D/B values there are T + K, undamped: |D_27 - D_1| = |sp T + ...| is not tiny
        assert np.all(np.abs(da - db)[expect] > 0)


def test_dipole_equal_until_the_pml_then_unstable():
    a, b = _box(n=30, pml=0.8), _box(n=30, pml=0.8)
    e, s = chunk27_intervals(b)
    b.engine.set_chunk27(e, s)
    for n in (4, 300):
        a.step(n - a.timestep)
        b.step(n - b.timestep)
        same = all(np.array_equal(a.read_field(C()), b.read_field(C())) for C in COMPS)
        if n == 4:
            assert same  # the pulse has not left the PML-free chunk
        else:
            assert not same
    ea = sum(float(np.sum(a.read_field(C()) ** 2)) for C in COMPS)
    eb = sum(float(np.sum(b.read_field(C()) ** 2)) for C in COMPS)
    assert np.isfinite(ea) and ea < 1.0
    assert not np.isfinite(eb) or eb > 1e6 * ea
