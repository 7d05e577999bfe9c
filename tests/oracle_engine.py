"""Test-side engine that steps a SimSpec with the CPU oracle (oracle/khronos_oracle.c).

Same interface as khronos_hip.engine.HipEngine so that parity tests can run
one simulation description through both.  The oracle evaluates the source
waveforms itself (Gaussian / CW closed forms of src/Sources/TimeSources.jl),
independently of the host code that feeds the HIP library.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
_LIBS = {}


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load(real):
    key = "f32" if real is np.float32 else "f64"
    if key in _LIBS:
        return _LIBS[key]
    path = os.path.join(ORACLE_DIR, "build", f"liboracle_{key}.so")
    if not os.path.exists(path):
        build_oracle()
    _LIBS[key] = load_path(real, path)
    return _LIBS[key]


def load_path(real, path):
    lib = C.CDLL(path)
    P = C.c_void_p
    lib.orc_create.restype = P
    lib.orc_create.argtypes = [C.c_int] * 3 + [C.c_double] * 4 + [P] * 18 + [C.c_double] * 3
    lib.orc_destroy.argtypes = [P]
    lib.orc_add_source.restype = C.c_int
    lib.orc_add_source.argtypes = [P, C.c_int, P, P, P, C.c_int, P]
    lib.orc_set_table.argtypes = [P, C.c_int, C.c_int64, P]
    lib.orc_add_dft.restype = C.c_int
    lib.orc_add_dft.argtypes = [P, C.c_int, P, P, C.c_int, P, C.c_int, C.c_int]
    lib.orc_step.restype = C.c_int
    lib.orc_step.argtypes = [P, C.c_int64]
    lib.orc_timestep.restype = C.c_int64
    lib.orc_timestep.argtypes = [P]
    lib.orc_read.argtypes = [P, C.c_int, C.c_int, P]
    lib.orc_write.argtypes = [P, C.c_int, C.c_int, P]
    lib.orc_dims.argtypes = [P, C.c_int, P]
    lib.orc_read_dft.argtypes = [P, C.c_int, P]
    lib.orc_set_periodic.restype = C.c_int
    lib.orc_set_periodic.argtypes = [P, C.c_int, P, P, P]
    lib.orc_set_chi3.argtypes = [P, P]
    lib.orc_set_chunk27.restype = C.c_int
    lib.orc_set_chunk27.argtypes = [P, P, P]
    lib.orc_set_minv.restype = C.c_int
    lib.orc_set_minv.argtypes = [P, C.c_int, P]
    lib.orc_add_pole.restype = C.c_int
    lib.orc_add_pole.argtypes = [P, P, C.c_int, P]
    lib.orc_read_pole.argtypes = [P, C.c_int, C.c_int, C.c_int, P]
    lib.orc_write_pole.argtypes = [P, C.c_int, C.c_int, C.c_int, P]
    lib.orc_real_size.restype = C.c_int
    assert lib.orc_real_size() == np.dtype(real).itemsize
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleEngine:
    """Steps a SimSpec on the CPU oracle."""

    def __init__(self, spec, material_sigma=None):
        from khronos_hip.sources import GaussianPulseData, ContinuousWaveData
        self.spec = spec
        R = spec.real
        self.R = R
        self.lib = load(R)
        keep = []

        def arr(a, order="C"):
            if a is None:
                return None
            h = np.asarray(a, dtype=R)
            h = np.asfortranarray(h) if order == "F" else np.ascontiguousarray(h)
            keep.append(h)
            return _p(h)

        sig_b = [None] * 3
        sig_d = [None] * 3
        if spec.sigma is not None:
            sig_b = [arr(spec.sigma["B"][a]) for a in range(3)]
            sig_d = [arr(spec.sigma["D"][a]) for a in range(3)]
        minv = [arr(spec.eps_inv[a], "F") for a in range(3)] + [arr(spec.mu_inv[a], "F") for a in range(3)]
        msig = [None] * 6
        if material_sigma is None and any(c is not None for c in (getattr(spec, "cond", None) or [])):
            material_sigma = spec.cond
        if material_sigma is not None:
            msig = [arr(m, "F") for m in material_sigma]
        self.sim = self.lib.orc_create(*spec.N, *spec.delta, spec.dt, *sig_b, *sig_d, *minv, *msig,
                                       spec.eps_inv_scalar, spec.mu_inv_scalar, spec.cutoff)
        from khronos_hip.grid import component_code
        self.cplx = bool(getattr(spec, "complex_fields", False))
        per = tuple(getattr(spec, "periodic", (False, False, False)))
        if self.cplx or any(per):
            pa = np.array([1 if p else 0 for p in per], dtype=np.int32)
            fw = np.array([v for z in spec.bloch_fwd for v in (z.real, z.imag)], dtype=np.float64)
            rv = np.array([v for z in spec.bloch_rev for v in (z.real, z.imag)], dtype=np.float64)
            rc = self.lib.orc_set_periodic(self.sim, 1 if self.cplx else 0, _p(pa), _p(fw), _p(rv))
            assert rc == 0
        self.src_ids = []
        for sd in spec.sources:
            amp = np.asfortranarray(sd.amplitude)
            ri = np.empty(amp.size * 2, dtype=R)
            flat = amp.ravel(order="F")
            ri[0::2] = flat.real
            ri[1::2] = flat.imag
            st = np.array(sd.gv.start_idx, dtype=np.int32)
            dm = np.array(amp.shape, dtype=np.int32)
            tp = sd.time_profile
            if isinstance(tp, GaussianPulseData):
                kind, p = 0, np.array([tp.fcen, tp.width, tp.peak_time, tp.cutoff_t], dtype=np.float64)
            elif isinstance(tp, ContinuousWaveData):
                kind, p = 1, np.array([tp.fcen, 0, 0, 0], dtype=np.float64)
            else:
                kind, p = 2, np.zeros(4)
            sid = self.lib.orc_add_source(self.sim, component_code(sd.component), _p(st), _p(dm), _p(ri), kind, _p(p))
            self.src_ids.append((sid, kind, sd))
        self.dft_ids = []
        for md in spec.dfts:
            st = np.array(md.gv.start_idx, dtype=np.int32)
            dm = np.array(md.gv.dims, dtype=np.int32)
            fr = np.ascontiguousarray(md.frequencies, dtype=np.float64)
            did = self.lib.orc_add_dft(self.sim, component_code(md.component), _p(st), _p(dm), len(fr), _p(fr),
                                       int(md.decimation), 1 if spec.dft_accum_f64 else 0)
            self.dft_ids.append(did)
        if getattr(spec, "chi3", None) is not None:
            self.lib.orc_set_chi3(self.sim, arr(spec.chi3, "F"))
        self.pole_ids = []
        for pd in getattr(spec, "poles", []) or []:
            c = np.ascontiguousarray(pd.coeffs.as_array(), dtype=np.float64)
            self.pole_ids.append(self.lib.orc_add_pole(self.sim, _p(c), 1 if pd.coeffs.is_drude else 0,
                                                       arr(pd.sigma, "F")))
        self.ts = 0

    def __del__(self):
        if getattr(self, "sim", None):
            self.lib.orc_destroy(self.sim)
            self.sim = None

    def _tables(self, n_end):
        R = self.R
        for sid, kind, sd in self.src_ids:
            if kind != 2:
                continue
            dt = R(self.spec.dt)
            half = float(dt / R(2))
            vals = np.empty(2 * n_end)
            for n in range(n_end):
                t = float(R(n) * dt) + (half if sd.component.electric else 0.0)
                v = sd.time_profile.eval(t, R)
                vals[2 * n], vals[2 * n + 1] = v.real, v.imag
            self.lib.orc_set_table(self.sim, sid, n_end, _p(vals))

    def step(self, n):
        if n <= 0:
            return
        self._tables(self.ts + n)
        rc = self.lib.orc_step(self.sim, int(n))
        if rc != 0:
            raise RuntimeError("oracle: no reference method for this auxiliary-array combination")
        self.ts += n

    def sync(self):
        pass

    def timestep(self):
        return int(self.lib.orc_timestep(self.sim))

    def field_shape(self, code):
        d = np.zeros(3, dtype=np.int32)
        self.lib.orc_dims(self.sim, code, _p(d))
        return tuple(int(v) for v in d)

    def read(self, which, code):
        out = np.zeros(self.field_shape(code), dtype=self.R, order="F")
        self.lib.orc_read(self.sim, which, code, _p(out))
        return out

    def read_field(self, code):
        if self.cplx:
            return self.read(0, code) + 1j * self.read(6, code)
        return self.read(0, code)

    def write(self, which, code, arr):
        a = np.asfortranarray(arr, dtype=self.R)
        return self.lib.orc_write(self.sim, which, code, _p(a))

    def write_field(self, code, arr):
        if self.cplx:
            arr = np.asarray(arr)
            self.write(0, code, arr.real)
            self.write(6, code, arr.imag)
        else:
            self.write(0, code, arr)

    def write_state(self, which, code, arr):
        """Write field (which 0) or D/B (which 1), complex-aware."""
        if self.cplx:
            arr = np.asarray(arr)
            self.write(which, code, np.real(arr))
            self.write(which + 6, code, np.imag(arr))
        else:
            self.write(which, code, arr)

    def read_state(self, which, code):
        if self.cplx:
            return self.read(which, code) + 1j * self.read(which + 6, code)
        return self.read(which, code)

    def set_chunk27(self, pml_end, pml_start):
        """Step with the reference's default 27-chunk plan arithmetic
        (src/Chunking.jl:907-923, :1090-1122; src/Kernels/Helpers.jl:200-223);
        the PML cell intervals come from chunk27_intervals."""
        e = np.asarray(pml_end, dtype=np.int32)
        s = np.asarray(pml_start, dtype=np.int32)
        assert self.lib.orc_set_chunk27(self.sim, _p(e), _p(s)) == 0, "chunk plan needs 3D PML and no conductivity"

    def set_eps_inv(self, a, arr):
        h = np.asfortranarray(arr, dtype=self.R)
        assert self.lib.orc_set_minv(self.sim, a, _p(h)) == 0

    def close(self):
        if getattr(self, "sim", None):
            self.lib.orc_destroy(self.sim)
            self.sim = None

    def set_graph(self, steps):
        pass

    def pole_shape(self):
        return tuple(n + 2 for n in self.spec.N)

    def read_pole(self, pole, which, a):
        """P (which 0) or P_prev (which 1) of a pole, or the total fPD (pole -1);
        (Nx+2, Ny+2, Nz+2)."""
        out = np.zeros(self.pole_shape(), dtype=self.R, order="F")
        self.lib.orc_read_pole(self.sim, pole, which, a, _p(out))
        if self.cplx and which < 2:
            im = np.zeros(self.pole_shape(), dtype=self.R, order="F")
            self.lib.orc_read_pole(self.sim, pole, which + 2, a, _p(im))
            return out + 1j * im
        return out

    def write_pole(self, pole, which, a, arr):
        if self.cplx and which < 2:
            arr = np.asarray(arr)
            self.lib.orc_write_pole(self.sim, pole, which + 2, a, _p(np.asfortranarray(np.imag(arr), dtype=self.R)))
            arr = np.real(arr)
        h = np.asfortranarray(arr, dtype=self.R)
        return self.lib.orc_write_pole(self.sim, pole, which, a, _p(h))

    def read_dft(self, i):
        md = self.spec.dfts[i]
        shp = tuple(md.gv.dims) + (len(md.frequencies),)
        buf = np.zeros(int(np.prod(shp)) * 2)
        self.lib.orc_read_dft(self.sim, self.dft_ids[i], _p(buf))
        return (buf[0::2] + 1j * buf[1::2]).reshape(shp, order="F")

    def dft_norms(self):
        return [float(np.sqrt(np.sum(np.abs(self.read_dft(i)) ** 2))) for i in range(len(self.dft_ids))]


def chunk27_intervals(sim):
    """PML cell intervals of the reference's chunk plan (src/Chunking.jl:
    633-646): per axis the last cell of the lower PML (0: none) and the first
    cell of the upper one (N + 1: none), from the user's thicknesses and the
    backend-precision spacing."""
    import math
    N = [sim.Nx, sim.Ny, max(sim.Nz, 1)]
    end, start = [], []
    for a in range(3):
        lo, hi = sim.boundaries[a] if sim.boundaries and len(sim.boundaries) > a else (0.0, 0.0)
        d = float(sim.delta_n[a])
        end.append(math.ceil(lo / d) if lo > 0.0 else 0)
        start.append(N[a] - math.ceil(hi / d) + 1 if hi > 0.0 else N[a] + 1)
    return end, start
