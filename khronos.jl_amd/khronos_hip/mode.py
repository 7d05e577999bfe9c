"""Waveguide modes of a cross-section (host setup for mode sources and mode
monitors; the reference's get_mode_profiles / get_mode_profiles_sweep,
src/Mode.jl:41-390, which call the VectorModesolver package).

VectorModesolver is not vendored, so the eigenproblem is set up here from
Maxwell's equations.  With fields ~ exp(i (beta w - omega t)) in a right-handed
frame (u, v, w) whose w is the propagation axis, eliminating E_w and H_w gives
for the transverse electric field (k0 = omega, mu = 1)

    beta^2 E_u = k0^2 eps_u E_u - d_v Q + d_u (D / eps_w)
    beta^2 E_v = k0^2 eps_v E_v + d_u Q + d_v (D / eps_w)
    Q = d_u E_v - d_v E_u,   D = d_u (eps_u E_u) + d_v (eps_v E_v),

discretised on a two-dimensional Yee cell (E_u at (i+1/2, j), E_v at
(i, j+1/2), E_w and D at nodes, Q and H_w at (i+1/2, j+1/2)) inside a
perfectly conducting window.  The remaining components follow from
i beta eps_w E_w = -D, i k0 H_w = Q, i k0 H_u = d_v E_w - i beta E_v,
i k0 H_v = i beta E_u - d_u E_w (H_u sits with E_v, H_v with E_u).
Modes are the eigenvalues beta^2 closest above (k0 n_target)^2, found by
shift-invert; each is scaled to P = 1/2 Re int (E x H*) . w dA = 1/2 and its
phase set so the largest E component is real.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, List, Optional

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# (u, v, w) of a plane with the given normal axis, right-handed
FRAMES = {0: (1, 2, 0), 1: (2, 0, 1), 2: (0, 1, 2)}


@dataclass
class Mode:
    """One mode on its solver grid.  Coordinates in the cross-section frame:
    E_u, H_v on (uh, vn); E_v, H_u on (un, vh); E_w on (un, vn); H_w on (uh, vh)
    (n: nodes, h: half points).  ``axes`` = (u, v, w) simulation axes."""

    neff: float
    beta: float
    k0: float
    axes: tuple
    un: np.ndarray
    vn: np.ndarray
    uh: np.ndarray
    vh: np.ndarray
    Eu: np.ndarray
    Ev: np.ndarray
    Ew: np.ndarray
    Hu: np.ndarray
    Hv: np.ndarray
    Hw: np.ndarray

    def field(self, name: str, pu, pv):
        """Bilinear interpolation of component ``name`` ('Eu', 'Hw', ...) at
        the cross-section points (pu, pv); zero outside the window."""
        grids = {"Eu": (self.uh, self.vn), "Hv": (self.uh, self.vn), "Ev": (self.un, self.vh),
                 "Hu": (self.un, self.vh), "Ew": (self.un, self.vn), "Hw": (self.uh, self.vh)}
        gu, gv = grids[name]
        return bilinear(gu, gv, getattr(self, name), pu, pv)


def bilinear(gu, gv, F, pu, pv):
    pu = np.asarray(pu, float)
    pv = np.asarray(pv, float)
    out = np.zeros(np.broadcast(pu, pv).shape, dtype=np.complex128)
    PU, PV = np.broadcast_arrays(pu, pv)
    du = gu[1] - gu[0] if len(gu) > 1 else 1.0
    dv = gv[1] - gv[0] if len(gv) > 1 else 1.0
    fu = (PU - gu[0]) / du
    fv = (PV - gv[0]) / dv
    inside = (fu >= 0) & (fu <= len(gu) - 1) & (fv >= 0) & (fv <= len(gv) - 1)
    i0 = np.clip(np.floor(fu).astype(int), 0, max(len(gu) - 2, 0))
    j0 = np.clip(np.floor(fv).astype(int), 0, max(len(gv) - 2, 0))
    tu = fu - i0
    tv = fv - j0
    i1 = np.minimum(i0 + 1, len(gu) - 1)
    j1 = np.minimum(j0 + 1, len(gv) - 1)
    val = ((1 - tu) * (1 - tv) * F[i0, j0] + tu * (1 - tv) * F[i1, j0] + (1 - tu) * tv * F[i0, j1]
           + tu * tv * F[i1, j1])
    out[inside] = val[inside]
    return out


def _d(n_out, n_in, h, shift):
    """Difference operator: out[k] = (in[k + shift] - in[k + shift - 1]) / h."""
    rows, cols, vals = [], [], []
    for k in range(n_out):
        for c, s in ((k + shift, 1.0), (k + shift - 1, -1.0)):
            if 0 <= c < n_in:
                rows.append(k)
                cols.append(c)
                vals.append(s / h)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n_out, n_in))


def solve_modes(eps_fn: Callable, u0: float, u1: float, v0: float, v1: float, res: float, k0: float,
                num_modes: int = 1, target_neff: float = 0.0, axes=(0, 1, 2), tol: float = 1e-10) -> List[Mode]:
    """Modes of the cross-section [u0, u1] x [v0, v1] at resolution ``res``.

    ``eps_fn(comp, U, V)`` returns the relative permittivity of field
    component comp (0: u, 1: v, 2: w) at the points (U, V)."""
    nu = max(2, int(round((u1 - u0) * res)))
    nv = max(2, int(round((v1 - v0) * res)))
    du = (u1 - u0) / nu
    dv = (v1 - v0) / nv
    un = u0 + du * np.arange(nu + 1)
    vn = v0 + dv * np.arange(nv + 1)
    uh = u0 + du * (np.arange(nu) + 0.5)
    vh = v0 + dv * (np.arange(nv) + 0.5)
    # unknowns: E_u at (uh[i], vn[j]) for j = 1..nv-1; E_v at (un[i], vh[j]) for i = 1..nu-1
    UE, VE = np.meshgrid(uh, vn[1:-1], indexing="ij")
    UF, VF = np.meshgrid(un[1:-1], vh, indexing="ij")
    UN, VN = np.meshgrid(un[1:-1], vn[1:-1], indexing="ij")
    eu = np.asarray(eps_fn(0, UE, VE), float).ravel(order="F")
    ev = np.asarray(eps_fn(1, UF, VF), float).ravel(order="F")
    ew = np.asarray(eps_fn(2, UN, VN), float).ravel(order="F")
    nEu = nu * (nv - 1)
    nEv = (nu - 1) * nv
    I = lambda n: sp.identity(n, format="csr")
    # nodes (interior): (nu-1) x (nv-1); Hw points: nu x nv
    # d/du from Eu points (i+1/2) to interior nodes i = 1..nu-1: (Eu[i] - Eu[i-1]) / du
    Du_E2N = sp.kron(I(nv - 1), _d(nu - 1, nu, du, 1))
    # d/dv from Ev points (j+1/2) to interior nodes j = 1..nv-1
    Dv_F2N = sp.kron(_d(nv - 1, nv, dv, 1), I(nu - 1))
    # d/du from nodes to Eu points (i+1/2): (N[i+1] - N[i]) / du, nodes 0 and nu are zero
    Du_N2E = sp.kron(I(nv - 1), _d(nu, nu - 1, du, 0))
    # d/dv from nodes to Ev points (j+1/2)
    Dv_N2F = sp.kron(_d(nv, nv - 1, dv, 0), I(nu - 1))
    # Q at Hw points (i+1/2, j+1/2): d_u Ev - d_v Eu (boundary Ev, Eu are zero)
    Du_F2Q = sp.kron(I(nv), _d(nu, nu - 1, du, 0))
    Dv_E2Q = sp.kron(_d(nv, nv - 1, dv, 0), I(nu))
    # from Q back: d_v Q at Eu points (j interior), d_u Q at Ev points (i interior)
    Dv_Q2E = sp.kron(_d(nv - 1, nv, dv, 1), I(nu))
    Du_Q2F = sp.kron(I(nv), _d(nu - 1, nu, du, 1))
    Q = sp.hstack([-Dv_E2Q, Du_F2Q])
    D = sp.hstack([Du_E2N @ sp.diags(eu), Dv_F2N @ sp.diags(ev)])
    Dz = sp.diags(1.0 / ew) @ D
    A_u = k0 * k0 * sp.hstack([sp.diags(eu), sp.csr_matrix((nEu, nEv))]) - Dv_Q2E @ Q + Du_N2E @ Dz
    A_v = k0 * k0 * sp.hstack([sp.csr_matrix((nEv, nEu)), sp.diags(ev)]) + Du_Q2F @ Q + Dv_N2F @ Dz
    A = sp.vstack([A_u, A_v]).tocsc()
    nmax = math.sqrt(max(eu.max(), ev.max()))
    n_t = target_neff if target_neff > 0 else nmax
    sigma = (k0 * n_t) ** 2
    k = min(num_modes + 2, A.shape[0] - 2)
    vals, vecs = spla.eigs(A, k=k, sigma=sigma, which="LM", tol=tol)
    order = np.argsort(-vals.real)
    modes = []
    for idx in order:
        b2 = vals[idx].real
        if b2 <= 0:
            continue
        beta = math.sqrt(b2)
        x = vecs[:, idx]
        Eu = np.zeros((nu, nv + 1), complex)
        Ev = np.zeros((nu + 1, nv), complex)
        Eu[:, 1:-1] = x[:nEu].reshape((nu, nv - 1), order="F")
        Ev[1:-1, :] = x[nEu:].reshape((nu - 1, nv), order="F")
        Dn = (D @ x).reshape((nu - 1, nv - 1), order="F")
        Ew = np.zeros((nu + 1, nv + 1), complex)
        Ew[1:-1, 1:-1] = -Dn / (1j * beta * ew.reshape((nu - 1, nv - 1), order="F"))
        Qh = (Q @ x).reshape((nu, nv), order="F")
        Hw = Qh / (1j * k0)
        dvEw = (Ew[:, 1:] - Ew[:, :-1]) / dv          # at (un, vh)
        duEw = (Ew[1:, :] - Ew[:-1, :]) / du          # at (uh, vn)
        Hu = (dvEw - 1j * beta * Ev) / (1j * k0)
        Hv = (1j * beta * Eu - duEw) / (1j * k0)
        m = Mode(beta / k0, beta, k0, tuple(axes), un, vn, uh, vh, Eu, Ev, Ew, Hu, Hv, Hw)
        _normalize(m, du, dv)
        modes.append(m)
        if len(modes) == num_modes:
            break
    return modes


def _normalize(m: Mode, du: float, dv: float):
    # P = 1/2 Re int (Eu Hv* - Ev Hu*) dA with each product at its own points
    p = 0.5 * (np.sum(m.Eu * np.conj(m.Hv)) - np.sum(m.Ev * np.conj(m.Hu))).real * du * dv
    # phase: the largest transverse E sample real and positive
    flat = np.concatenate([m.Eu.ravel(), m.Ev.ravel()])
    ph = flat[np.argmax(np.abs(flat))]
    ph = ph / abs(ph) if abs(ph) > 0 else 1.0
    s = 1.0 / (math.sqrt(2.0 * abs(p)) * ph) if p != 0 else 1.0
    for name in ("Eu", "Ev", "Ew", "Hu", "Hv", "Hw"):
        setattr(m, name, getattr(m, name) * s)


def geometry_eps_fn(geometry, axes, w_pos: float):
    """eps of the simulation geometry on the plane w = w_pos (first object
    wins; 1 where no object sets one; static permittivity)."""
    def fn(comp, U, V):
        U = np.asarray(U, float)
        out = np.ones(U.shape)
        done = np.zeros(U.shape, dtype=bool)
        P = [None, None, None]
        P[axes[0]] = U
        P[axes[1]] = np.asarray(V, float)
        P[axes[2]] = np.full(U.shape, float(w_pos))
        a = axes[comp]
        for obj in geometry:
            inside = obj.shape.contains(P[0], P[1], P[2]) & ~done
            e = obj.material.perm(True, a)
            if e is not None:
                out[inside] = float(e)
            done |= inside
        return out
    return fn


@dataclass
class ModeSpec:
    """Mode solver settings of a mode monitor (src/DataStructures.jl:461-475)."""

    num_modes: int = 1
    target_neff: float = 0.0
    geometry: list = None
    mode_solver_resolution: int = 50
    solver_tolerance: float = 1e-6


def plane_modes(center, size, geometry, frequency: float, res: int, num_modes: int = 1,
                target_neff: float = 0.0) -> List[Mode]:
    """Modes of the planar region (center, size) of ``geometry``, propagating
    along +normal (the zero-size axis)."""
    normal = int([a for a in range(3) if float(size[a]) == 0.0][0])
    axes = FRAMES[normal]
    u, v = axes[0], axes[1]
    fn = geometry_eps_fn(geometry, axes, float(center[normal]))
    return solve_modes(fn, float(center[u]) - float(size[u]) / 2, float(center[u]) + float(size[u]) / 2,
                       float(center[v]) - float(size[v]) / 2, float(center[v]) + float(size[v]) / 2, res,
                       2 * math.pi * frequency, num_modes=num_modes, target_neff=target_neff, axes=axes)
