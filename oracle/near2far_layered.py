"""Layered-medium near-to-far projection, restated in numpy.

TEST INFRASTRUCTURE ONLY: the checker for khronos_hip's layered projection
(device spectral sums + host transfer matrices).  Only tests/ load it.

It restates, step by step:

  compute_layer_transmission ... src/Monitors/Near2Far.jl:565-669
  _bilinear_interp ............. src/Monitors/Near2Far.jl:677-703
  compute_far_field_layered .... src/Monitors/Near2Far.jl:723-985

with FFTW's forward FFT replaced by numpy's (the same unnormalised
sum_j f_j exp(-2 pi i j m / n)) and AbstractFFTs' fftfreq by its formula
((i - 1 - (i > n_nonneg) n) * fs / n).
"""
from __future__ import annotations

import cmath
import math

import numpy as np


def _kz(eps, mu, w, kt2):
    kz = cmath.sqrt(complex(eps * mu * w ** 2 - kt2))
    return -kz if kz.imag < 0 else kz


def layer_transmission(layers, z_monitor, freq, kx, ky):
    """(t_TE, t_TM) of tangential E from the monitor's layer to the top layer.
    layers: list of (z_min, z_max, eps, mu), bottom to top."""
    n = len(layers)
    w = 2 * math.pi * freq
    kt2 = kx ** 2 + ky ** 2
    mon = None
    for i, L in enumerate(layers):
        if L[0] <= z_monitor <= L[1]:
            mon = i
            break
    if mon is None:
        raise ValueError(f"Monitor z={z_monitor} is not inside any layer in the stack")
    if mon == n - 1:
        return 1.0 + 0j, 1.0 + 0j

    def p_te(L):
        return _kz(L[2], L[3], w, kt2) / (w * L[3])

    def p_tm(L):
        kz = _kz(L[2], L[3], w, kt2)
        if abs(kz) < 1e-30:
            return complex(math.inf, 0.0)
        return w * L[2] / kz

    res = [1.0 + 0j, 1.0 + 0j]
    for pol, get_p in enumerate((p_te, p_tm)):
        T = np.eye(2, dtype=complex)
        dead = False
        for i in range(mon, n - 1):
            Li = layers[i]
            kz_i = _kz(Li[2], Li[3], w, kt2)
            p_i = get_p(Li)
            d = (Li[1] - z_monitor) if i == mon else (Li[1] - Li[0])
            if d > 0:
                T = T @ np.array([[cmath.exp(1j * kz_i * d), 0], [0, cmath.exp(-1j * kz_i * d)]])
            p_n = get_p(layers[i + 1])
            if abs(p_n) < 1e-30 or math.isinf(abs(p_n)) or math.isinf(abs(p_i)):
                res[pol] = 0j
                dead = True
                break
            a = (p_i + p_n) / (2 * p_n)
            b = (p_i - p_n) / (2 * p_n)
            T = T @ np.array([[a, b], [b, a]])
        if dead:
            continue
        p_mon = get_p(layers[mon])
        p_top = get_p(layers[-1])
        if math.isinf(abs(p_top)) or abs(p_top) < 1e-30 or abs(T[1, 1]) < 1e-30:
            res[pol] = 0j
        else:
            res[pol] = (p_mon / p_top) / T[1, 1]
    return res[0], res[1]


def fftfreq(n, fs):
    nn = (n - 1) // 2 + 1
    m = fs / n
    return np.array([((i - 1) - (n if i > nn else 0)) * m for i in range(1, n + 1)])


def bilinear(data, kxg, kyg, kx, ky):
    nkx, nky = len(kxg), len(kyg)
    dkx = kxg[1] - kxg[0]
    dky = kyg[1] - kyg[0]
    fx = (kx - kxg[0]) / dkx + 1.0
    fy = (ky - kyg[0]) / dky + 1.0
    ix = math.floor(fx)
    iy = math.floor(fy)
    if ix < 1 or ix >= nkx or iy < 1 or iy >= nky:
        return 0j
    wx = fx - ix
    wy = fy - iy
    d = lambda a, b: data[a - 1, b - 1]
    return ((1 - wx) * (1 - wy) * d(ix, iy) + wx * (1 - wy) * d(ix + 1, iy)
            + (1 - wx) * wy * d(ix, iy + 1) + wx * wy * d(ix + 1, iy + 1))


def far_field_layered(planes, base, dx, dy, freqs, theta, phi, r, ns, layers):
    """planes: e1, e2, h1, h2 = Ex, Ey, Hx, Hy complex (Nx_q, Ny_q, Nz_q, nf)
    DFTs of a z-normal surface (averaged over two z planes when Nz_q >= 2);
    base: (4, 3) positions of their first samples.  Returns (n_obs, 6, nf)."""
    P = []
    for f in planes:
        f = np.asarray(f, dtype=np.complex128)
        if f.shape[2] >= 2:
            f = (f[:, :, 0:1, :] + f[:, :, 1:2, :]) / 2
        P.append(f)
    Nx = min(f.shape[0] for f in P)
    Ny = min(f.shape[1] for f in P)
    e1b, e2b, h1b, h2b = (np.asarray(b, dtype=np.float64) for b in base)
    z_mon = (e1b[2] + e2b[2] + h1b[2] + h2b[2]) / 4.0
    mon = None
    for i, L in enumerate(layers):
        if L[0] <= z_mon <= L[1]:
            mon = i
            break
    if mon is None:
        raise ValueError("monitor not inside any layer")
    eps_m, mu_m = layers[mon][2], layers[mon][3]
    eps_t, mu_t = layers[-1][2], layers[-1][3]
    Z_top = math.sqrt(mu_t / eps_t)
    z_E = (e1b[2] + e2b[2]) / 2.0
    z_H = (h1b[2] + h2b[2]) / 2.0
    dz = z_E - z_H
    n_obs = len(theta) * len(phi)
    out = np.zeros((n_obs, 6, len(freqs)), dtype=np.complex128)
    for k, freq in enumerate(freqs):
        w = 2 * math.pi * freq
        k_top = w * math.sqrt(eps_t * mu_t)
        Nxp, Nyp = 4 * Nx, 4 * Ny
        spec = []
        for f in P:
            pad = np.zeros((Nxp, Nyp), dtype=np.complex128)
            pad[:Nx, :Ny] = f[:Nx, :Ny, 0, k]
            spec.append(np.fft.fft2(pad) * (dx * dy))
        kxf = fftfreq(Nxp, 2 * math.pi / dx)
        kyf = fftfreq(Nyp, 2 * math.pi / dy)
        px = np.argsort(kxf, kind="stable")
        py = np.argsort(kyf, kind="stable")
        kxs, kys = kxf[px], kyf[py]
        S = [s[np.ix_(px, py)] for s in spec]
        pref = -1j * k_top * cmath.exp(1j * k_top * r) / (2 * math.pi * r)
        idx = 0
        for ph in phi:
            cp, sp = math.cos(ph), math.sin(ph)
            for th in theta:
                st, ct = math.sin(th), math.cos(th)
                kx = k_top * st * cp
                ky = k_top * st * sp
                kt2 = kx ** 2 + ky ** 2
                ex, ey, hx, hy = (bilinear(s, kxs, kys, kx, ky) for s in S)
                ex *= cmath.exp(-1j * kx * (e1b[0] - e2b[0]))
                ey *= cmath.exp(-1j * ky * (e2b[1] - e1b[1]))
                hx *= cmath.exp(-1j * ky * (h1b[1] - h2b[1]))
                hy *= cmath.exp(-1j * kx * (h2b[0] - h1b[0]))
                e_te = -sp * ex + cp * ey
                e_tm = cp * ex + sp * ey
                h_p = cp * hx + sp * hy
                h_s = -sp * hx + cp * hy
                kz = _kz(eps_m, mu_m, w, kt2)
                p_te = kz / (w * mu_m)
                p_tm = w * eps_m / kz if abs(kz) > 1e-30 else complex(math.inf, 0.0)
                phz = cmath.exp(-1j * kz * dz)
                ckz = cmath.cos(kz * dz)
                if abs(ckz) < 1e-30:
                    ckz = 1e-30 + 0j
                a_te = (e_te - ns * (h_p / p_te) * phz) / (2 * ckz) \
                    if (abs(p_te) > 1e-30 and not math.isinf(abs(p_te))) else 0j
                a_tm = (e_tm + ns * (h_s / p_tm) * phz) / (2 * ckz) \
                    if (abs(p_tm) > 1e-30 and not math.isinf(abs(p_tm))) else 0j
                t_te, t_tm = layer_transmission(layers, z_mon, freq, kx, ky)
                e_th = pref * (t_tm * a_tm)
                e_ph = pref * ct * (t_te * a_te)
                Ex = e_th * ct * cp + e_ph * (-sp)
                Ey = e_th * ct * sp + e_ph * cp
                Ez = e_th * (-st)
                rh = (st * cp, st * sp, ct)
                H = ((rh[1] * Ez - rh[2] * Ey) / Z_top, (rh[2] * Ex - rh[0] * Ez) / Z_top,
                     (rh[0] * Ey - rh[1] * Ex) / Z_top)
                out[idx, :, k] = (Ex, Ey, Ez, H[0], H[1], H[2])
                idx += 1
    return out
