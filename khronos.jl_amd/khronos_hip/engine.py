"""Time-step engine backed by libkhronos_hip (the only engine of the package).

Takes a :class:`~khronos_hip.simulation.SimSpec` (host setup in the
reference's layouts), creates the device context and exposes stepping and
readout.  Waveforms of the sources are evaluated on the host, in the backend
precision, exactly where the reference evaluates them (t for H sources,
t + dt/2 for E sources; src/Kernels/Kernels.jl:64-74), and uploaded in blocks
ahead of the steps that use them.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from .grid import component_voxel_count, component_from_code


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class HipEngine:
    WAVE_BLOCK = 4096

    def __init__(self, spec):
        self.lib = _lib.load()
        self.spec = spec
        # field precision; the host setup (and the waveform values) may be
        # in a wider one (the precision form: FP64 setup, FP32 fields)
        R = getattr(spec, "field_real", None) or spec.real
        self.R = R
        self._keep = []  # host arrays referenced during kh_create
        d = _lib.KhSimDesc()
        d.nx, d.ny, d.nz = spec.N
        d.precision = _lib.KH_F32 if R is np.float32 else _lib.KH_F64
        d.mode = _lib.KH_MODE_EXACT if spec.mode == "exact" else _lib.KH_MODE_FAST
        d.dft_accum_f64 = 1 if spec.dft_accum_f64 else 0
        d.dx, d.dy, d.dz = spec.delta
        d.dt = spec.dt
        if spec.sigma is not None:
            for a in range(3):
                if spec.sigma["B"][a] is None:   # 2D: no z profiles
                    continue
                sb = np.ascontiguousarray(spec.sigma["B"][a], dtype=R)
                sd = np.ascontiguousarray(spec.sigma["D"][a], dtype=R)
                self._keep += [sb, sd]
                d.sigma_b[a] = _ptr(sb)
                d.sigma_d[a] = _ptr(sd)
        for a in range(3):
            for arrs, slot in ((spec.eps_inv, d.eps_inv), (spec.mu_inv, d.mu_inv)):
                if arrs[a] is not None:
                    h = np.asfortranarray(arrs[a], dtype=R)
                    self._keep.append(h)
                    slot[a] = _ptr(h)
        for c, arr in enumerate(getattr(spec, "cond", None) or [None] * 6):
            if arr is not None:
                h = np.asfortranarray(arr, dtype=R)
                self._keep.append(h)
                d.sigma_mat[c] = _ptr(h)
        d.eps_inv_scalar = spec.eps_inv_scalar
        d.mu_inv_scalar = spec.mu_inv_scalar
        d.source_cutoff_time = spec.cutoff
        self.cplx = bool(getattr(spec, "complex_fields", False))
        d.complex_fields = 1 if self.cplx else 0
        per = tuple(getattr(spec, "periodic", (False, False, False)))
        for a in range(3):
            d.periodic[a] = 1 if per[a] else 0
            if per[a]:
                f, r = spec.bloch_fwd[a], spec.bloch_rev[a]
                d.bloch_fwd[a][0], d.bloch_fwd[a][1] = f.real, f.imag
                d.bloch_rev[a][0], d.bloch_rev[a][1] = r.real, r.imag
        self.desc = d
        ctx = C.c_void_p()
        _lib.check(self.lib.kh_create(C.byref(d), C.byref(ctx)))
        self.ctx = ctx
        self._keep = []
        # sources
        self.src_ids = []
        self.src_wave_len = []
        for sd in spec.sources:
            amp = np.asfortranarray(sd.amplitude.astype(np.complex64 if R is np.float32 else np.complex128))
            st = (C.c_int32 * 3)(*sd.gv.start_idx)
            dm = (C.c_int32 * 3)(*amp.shape)
            sid = C.c_int32()
            _lib.check(self.lib.kh_add_source(self.ctx, _code(sd.component), st, dm, _ptr(amp), C.byref(sid)))
            self.src_ids.append(sid.value)
            self.src_wave_len.append(0)
        # DFT monitors
        self.dft_ids = []
        for md in spec.dfts:
            st = (C.c_int32 * 3)(*md.gv.start_idx)
            dm = (C.c_int32 * 3)(*md.gv.dims)
            fr = np.ascontiguousarray(md.frequencies, dtype=np.float64)
            did = C.c_int32()
            _lib.check(self.lib.kh_add_dft(self.ctx, _code(md.component), st, dm, len(fr), _ptr(fr),
                                           int(md.decimation), C.byref(did)))
            self.dft_ids.append(did.value)
        # Kerr coefficient and dispersive poles
        if getattr(spec, "chi3", None) is not None:
            c3 = np.asfortranarray(spec.chi3, dtype=R)
            _lib.check(self.lib.kh_set_chi3(self.ctx, _ptr(c3)))
        self.pole_ids = []
        for pd in getattr(spec, "poles", []) or []:
            cf = np.ascontiguousarray(pd.coeffs.as_array(), dtype=np.float64)
            sg = np.asfortranarray(pd.sigma, dtype=R)
            pid = C.c_int32()
            _lib.check(self.lib.kh_add_pole(self.ctx, _ptr(cf), 1 if pd.coeffs.is_drude else 0, _ptr(sg),
                                            C.byref(pid)))
            self.pole_ids.append(pid.value)
        self.ts = 0
        self._norms, self._norms_ts = None, 0  # dft_norms of timestep _norms_ts

    def __del__(self):
        try:
            if getattr(self, "ctx", None):
                self.lib.kh_destroy(self.ctx)
                self.ctx = None
        except Exception:
            pass

    # -------------------------------------------------------------- waveforms
    def _wave_times(self, electric: bool, n0: int, n1: int):
        R = self.spec.real
        dt = R(self.spec.dt)
        half = float(dt / R(2))
        ts = [float(R(n) * dt) for n in range(n0, n1)]
        return [t + half for t in ts] if electric else ts

    def _ensure_waveforms(self, n_end: int):
        spec = self.spec
        # sources stop once t > cutoff: no values are needed past that step
        cap = int(math.floor(spec.cutoff / spec.dt)) + 3 if math.isfinite(spec.cutoff) else None
        if not hasattr(self, "_wave"):
            self._wave = [np.zeros(0) for _ in spec.sources]
        for q, sd in enumerate(spec.sources):
            have = self.src_wave_len[q]
            need = n_end if cap is None else min(n_end, cap)
            if have >= need:
                continue
            # with a cutoff the whole table goes up at once (one device table,
            # one captured graph); without, in blocks
            target = cap if cap is not None else max(need, have + self.WAVE_BLOCK)
            times = self._wave_times(sd.component.electric, have, target)
            new = np.empty(2 * len(times))
            for i, t in enumerate(times):
                v = sd.time_profile.eval(t, self.spec.real)
                new[2 * i] = v.real
                new[2 * i + 1] = v.imag
            self._wave[q] = np.concatenate([self._wave[q], new])
            _lib.check(self.lib.kh_set_waveform(self.ctx, self.src_ids[q], target, _ptr(self._wave[q])))
            self.src_wave_len[q] = target

    # -------------------------------------------------------------- stepping
    def step(self, n: int):
        if n <= 0:
            return
        self._ensure_waveforms(self.ts + n)
        _lib.check(self.lib.kh_step(self.ctx, int(n)))
        self.ts += n

    def set_graph(self, steps: int):
        """Steps per captured hipGraph (0: no graphs)."""
        _lib.check(self.lib.kh_set_graph(self.ctx, int(steps)))

    def prepare_graphs(self, n: int):
        """Capture the graphs the next ``step(n)`` replays, without stepping."""
        self._ensure_waveforms(self.ts + n)
        _lib.check(self.lib.kh_prepare_graphs(self.ctx, int(n)))

    def launch_counts(self) -> dict:
        """What the steps since the last call ran: graph replays, steps inside
        them, eager steps and DFT launches (then cleared)."""
        out = np.zeros(4, dtype=np.int64)
        _lib.check(self.lib.kh_get_launch_counts(self.ctx, _ptr(out)))
        return {"graph_replays": int(out[0]), "graph_steps": int(out[1]), "eager_steps": int(out[2]),
                "dft_launches": int(out[3])}

    def idle(self) -> bool:
        """All queued work finished (non-blocking)."""
        v = C.c_int32()
        _lib.check(self.lib.kh_query(self.ctx, C.byref(v)))
        return bool(v.value)

    def close(self):
        """Free the device context now (instead of at garbage collection)."""
        if getattr(self, "ctx", None):
            _lib.check(self.lib.kh_destroy(self.ctx))
            self.ctx = None

    def sync(self):
        _lib.check(self.lib.kh_sync(self.ctx))

    def sources_active(self) -> bool:
        v = C.c_int32()
        _lib.check(self.lib.kh_sources_active(self.ctx, C.byref(v)))
        return bool(v.value)

    def timestep(self) -> int:
        v = C.c_int64()
        _lib.check(self.lib.kh_get_timestep(self.ctx, C.byref(v)))
        return v.value

    # -------------------------------------------------------------- readout
    def field_shape(self, code: int):
        cnt = component_voxel_count(*self.spec.N, component_from_code(code))
        return tuple(c + 2 for c in cnt)

    def _cdtype(self):
        return np.complex64 if self.R is np.float32 else np.complex128

    def read_field(self, code: int) -> np.ndarray:
        shp = self.field_shape(code)
        out = np.zeros(shp, dtype=self._cdtype() if self.cplx else self.R, order="F")
        _lib.check(self.lib.kh_read_field(self.ctx, code, _ptr(out), out.nbytes))
        return out

    def write_field(self, code: int, arr: np.ndarray):
        a = np.asfortranarray(arr, dtype=self._cdtype() if self.cplx else self.R)
        assert a.shape == self.field_shape(code)
        _lib.check(self.lib.kh_write_field(self.ctx, code, _ptr(a), a.nbytes))

    def set_eps_inv(self, a: int, arr: np.ndarray):
        """Replace the per-voxel inverse permittivity of component a (reference
        geometry layout); between runs in FAST mode."""
        h = np.asfortranarray(arr, dtype=self.R)
        _lib.check(self.lib.kh_set_eps_inv(self.ctx, int(a), _ptr(h)))

    def pole_shape(self):
        return tuple(n + 2 for n in self.spec.N)

    def read_pole(self, pole: int, which: int, a: int) -> np.ndarray:
        """P (which 0) or P_prev (which 1) of a pole, (Nx+2, Ny+2, Nz+2)."""
        out = np.zeros(self.pole_shape(), dtype=self._cdtype() if self.cplx else self.R, order="F")
        _lib.check(self.lib.kh_read_pole(self.ctx, self.pole_ids[pole], which, a, _ptr(out), out.nbytes))
        return out

    def write_pole(self, pole: int, which: int, a: int, arr: np.ndarray):
        h = np.asfortranarray(arr, dtype=self._cdtype() if self.cplx else self.R)
        _lib.check(self.lib.kh_write_pole(self.ctx, self.pole_ids[pole], which, a, _ptr(h), h.nbytes))

    def read_dft(self, i: int) -> np.ndarray:
        md = self.spec.dfts[i]
        shp = tuple(md.gv.dims) + (len(md.frequencies),)
        buf = np.zeros(int(np.prod(shp)) * 2, dtype=np.float64)
        _lib.check(self.lib.kh_read_dft(self.ctx, self.dft_ids[i], _ptr(buf), buf.nbytes))
        z = buf[0::2] + 1j * buf[1::2]
        return z.reshape(shp, order="F")

    def flux(self, ids, normal_axis: int, dA: float) -> np.ndarray:
        nf = len(self.spec.dfts[ids[0]].frequencies)
        out = np.zeros(nf, dtype=np.float64)
        e1, e2, h1, h2 = (self.dft_ids[i] for i in ids)
        _lib.check(self.lib.kh_flux(self.ctx, e1, e2, h1, h2, int(normal_axis), float(dA), _ptr(out)))
        return out

    def mode_overlap(self, ids, normal_axis: int, mode: np.ndarray, dA: float) -> np.ndarray:
        """Overlap sums (n_freq, 2) complex [sum (s + t) dA, sum (s - t) dA] of
        kh_mode_overlap; ``mode`` is complex (4, n_freq, n2, n1)."""
        q, nf, n2, n1 = mode.shape
        buf = np.ascontiguousarray(mode, dtype=np.complex128)
        out = np.zeros(4 * nf, dtype=np.float64)
        e1, e2, h1, h2 = (self.dft_ids[i] for i in ids)
        _lib.check(self.lib.kh_mode_overlap(self.ctx, e1, e2, h1, h2, int(normal_axis), int(n1), int(n2),
                                            _ptr(buf), float(dA), _ptr(out)))
        o = out.reshape(nf, 4)
        return np.stack([o[:, 0] + 1j * o[:, 1], o[:, 2] + 1j * o[:, 3]], axis=1)

    def near2far(self, ids, normal_axis: int, normal_sign: float, base, d1: float, d2: float, dA: float,
                 n1: int, n2: int, freqs, eps: float, mu: float, obs) -> np.ndarray:
        """Far field (n_obs, 6, n_freq) complex of kh_near2far; ``base`` is the
        (4, 3) position of sample (0, 0) of e1, e2, h1, h2, ``obs`` (n_obs, 3)."""
        b = np.ascontiguousarray(base, dtype=np.float64).reshape(4, 3)
        f = np.ascontiguousarray(freqs, dtype=np.float64).ravel()
        o = np.ascontiguousarray(obs, dtype=np.float64).reshape(-1, 3)
        nf, n_obs = f.size, o.shape[0]
        out = np.zeros((nf, 6, n_obs), dtype=np.complex128)
        e1, e2, h1, h2 = (self.dft_ids[i] for i in ids)
        _lib.check(self.lib.kh_near2far(self.ctx, e1, e2, h1, h2, int(normal_axis), float(normal_sign), _ptr(b),
                                        float(d1), float(d2), float(dA), int(n1), int(n2), _ptr(f), nf,
                                        float(eps), float(mu), _ptr(o), n_obs, _ptr(out)))
        return np.transpose(out, (2, 1, 0))

    def near2far_spectrum(self, ids, normal_axis: int, n1: int, n2: int, k: int, pad1: int, pad2: int,
                          bins) -> np.ndarray:
        """(n_bins, 4) complex spectral samples of kh_near2far_spectrum; bins
        (n_bins, 2) integer frequencies."""
        b = np.ascontiguousarray(bins, dtype=np.int32).reshape(-1, 2)
        out = np.zeros(b.shape[0] * 8, dtype=np.float64)
        e1, e2, h1, h2 = (self.dft_ids[i] for i in ids)
        _lib.check(self.lib.kh_near2far_spectrum(self.ctx, e1, e2, h1, h2, int(normal_axis), int(n1), int(n2),
                                                 int(k), int(pad1), int(pad2), _ptr(b), b.shape[0], _ptr(out)))
        return (out[0::2] + 1j * out[1::2]).reshape(-1, 4)

    def dft_norms(self):
        """Per-monitor sqrt(sum |F|^2) (src/Simulation.jl:534-539).  A monitor's
        accumulator changes only on the steps it samples (timestep % decimation
        == 0, src/Kernels/Kernels.jl:470), so when none sampled since the last
        query its result is returned again without a device round trip; else
        all norms come from one kh_dft_norms call."""
        n = len(self.dft_ids)
        if self._norms is not None and not self._sampled_between(self._norms_ts, self.ts):
            return list(self._norms)
        out = np.zeros(n, dtype=np.float64)
        if n:
            assert self.dft_ids == list(range(n))
            _lib.check(self.lib.kh_dft_norms(self.ctx, n, _ptr(out)))
        self._norms, self._norms_ts = out, self.ts
        return list(out)

    def _sampled_between(self, t0: int, t1: int) -> bool:
        """Some monitor sampled a step n with t0 <= n < t1."""
        for md in self.spec.dfts:
            d = max(1, int(md.decimation))
            if -(-t0 // d) * d < t1:
                return True
        return False

    def reset(self):
        _lib.check(self.lib.kh_reset(self.ctx))
        self.ts = 0
        self._norms = None

    # ------------------------------------------------------------- adjoint
    def set_checkpoints(self, every: int):
        """Checkpoint the field state every ``every`` steps of the coming run
        (0: off), for ``adjoint_gradient``."""
        _lib.check(self.lib.kh_set_checkpoints(self.ctx, int(every)))

    def adjoint_gradient(self, monitors, cotangents, boxes):
        """dJ/d(eps^-1) by the time-reversed adjoint of the steps run so far.

        monitors: indices into ``spec.dfts``; cotangents: per monitor, a
        complex array shaped like ``read_dft(i)`` holding dJ/dRe F + i dJ/dIm F;
        boxes: per E component, None or ((lo_x, lo_y, lo_z), (n_x, n_y, n_z))
        in 1-based grid indices.  Returns per component None or the gradient
        as a float64 array of shape n (Fortran order)."""
        ids = np.array([self.dft_ids[i] for i in monitors] or [0], dtype=np.int32)
        parts = []
        for i, g in zip(monitors, cotangents):
            md = self.spec.dfts[i]
            shp = tuple(md.gv.dims) + (len(md.frequencies),)
            z = np.asarray(g, dtype=np.complex128).reshape(shp)
            flat = z.ravel(order="F")
            pair = np.empty(2 * flat.size, dtype=np.float64)
            pair[0::2] = flat.real
            pair[1::2] = flat.imag
            parts.append(pair)
        cot = np.ascontiguousarray(np.concatenate(parts) if parts else np.zeros(1))
        lo = np.ones(9, dtype=np.int32)
        n = np.zeros(9, dtype=np.int32)
        total = 0
        for a, b in enumerate(boxes):
            if b is None:
                continue
            lo[3 * a:3 * a + 3] = b[0]
            n[3 * a:3 * a + 3] = b[1]
            total += int(np.prod(b[1]))
        grad = np.zeros(max(total, 1), dtype=np.float64)
        _lib.check(self.lib.kh_adjoint_gradient(self.ctx, len(monitors), _ptr(ids), _ptr(cot), _ptr(lo), _ptr(n),
                                                _ptr(grad)))
        out, o = [], 0
        for b in boxes:
            if b is None or int(np.prod(b[1])) == 0:
                out.append(None)
                continue
            c = int(np.prod(b[1]))
            out.append(grad[o:o + c].reshape(tuple(b[1]), order="F"))
            o += c
        return out

    # -------------------------------------------------------------- timing
    def set_timing(self, on: bool):
        _lib.check(self.lib.kh_set_timing(self.ctx, 1 if on else 0))

    TIMING_SLOTS = ("h_inner", "h_full", "e_inner", "e_full", "dft", "material")

    def get_timing(self):
        """{slot: (total_ms, launches)} for the kernel classes of include/khronos_hip.h."""
        out = np.zeros(2 * len(self.TIMING_SLOTS), dtype=np.float64)
        _lib.check(self.lib.kh_get_timing(self.ctx, _ptr(out)))
        return {k: (float(out[2 * i]), int(out[2 * i + 1])) for i, k in enumerate(self.TIMING_SLOTS)}

    def plan_boxes(self):
        """Inner boxes (sigma zero on all axes) of the H and E half steps, as the
        library computes them: [(lo, hi) for H, (lo, hi) for E], grid indices."""
        out = []
        for key in ("B", "D"):
            lohi = []
            for a in range(3):
                n = self.spec.N[a]
                if self.spec.sigma is None or self.spec.sigma[key][a] is None:
                    lohi.append((1, n))
                    continue
                z = np.asarray(self.spec.sigma[key][a])[0:2 * n:2] == 0
                best, lo, hi, i = 0, 1, 0, 0
                while i < n:
                    if not z[i]:
                        i += 1
                        continue
                    j = i
                    while j + 1 < n and z[j + 1]:
                        j += 1
                    if j - i + 1 > best:
                        best, lo, hi = j - i + 1, i + 1, j + 1
                    i = j + 1
                lohi.append((lo, hi))
            out.append(([v[0] for v in lohi], [v[1] for v in lohi]))
        return out


def _code(component) -> int:
    from .grid import component_code
    return component_code(component)
