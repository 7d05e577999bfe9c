"""The algebra of the time-reversed adjoint, on the CPU.

One forward half step of the sigma cascade (src/Kernels/Helpers.jl:113-139,
:323-340, curls :286-298; sources dropped, they are additive) is a linear
map A of the state (U, T, W, A of the updated group; the other group's A
read through the curl).  adjoint_half_kernel applies its transpose in the
order written below.  The dot-product test <A x, y> = <x, A^T y> pins that
order (a wrong coefficient, sign, stage or curl direction breaks it), on a
small padded grid with random sigma profiles that are zero in places, so
every branch (o, n, p zero or not) is taken; the eps^-1 gradient term
(What * net') is checked against a central difference of <y, A(m) x> in m.
The device kernel itself is pinned by finite differences of whole runs
(tests/test_gpu_adjoint_time_reversed.py).
"""
import numpy as np
import pytest

N = (5, 4, 3)
SH = tuple(n + 2 for n in N)
INNER = tuple(slice(1, n + 1) for n in N)


def _sig(rng):
    # sampled sigma per axis: zero at the ghosts and at a few inner points
    s = []
    for n in N:
        v = np.zeros(n + 2)
        v[1:n + 1] = rng.uniform(0.05, 0.6, n) * (rng.random(n) < 0.6)
        s.append(v)
    return s


def _grid_sig(s, a):
    """sigma of axis a broadcast over the padded grid."""
    shape = [1, 1, 1]
    shape[a] = SH[a]
    return s[a].reshape(shape) * np.ones(SH)


def _shift(f, a, k):
    """g[i] = f[i + k] along axis a (zero past the padded grid)."""
    g = np.zeros_like(f)
    src = [slice(None)] * 3
    dst = [slice(None)] * 3
    if k > 0:
        src[a], dst[a] = slice(k, None), slice(None, -k)
    else:
        src[a], dst[a] = slice(None, k), slice(-k, None)
    g[tuple(dst)] = f[tuple(src)]
    return g


def _curl(F, sign, dt, inv):
    """H update (sign +1): forward differences; E update (-1): backward,
    d(A) = inv (A[i-1] - A[i])."""
    d = lambda f, a: inv[a] * (_shift(f, a, sign) - f)
    return [dt * (d(F[1], 2) - d(F[2], 1)), dt * (d(F[2], 0) - d(F[0], 2)), dt * (d(F[0], 1) - d(F[1], 0))]


def _mask():
    m = np.zeros(SH)
    m[INNER] = 1.0
    return m


def forward(state, F, sig, m, sign, dt, inv):
    """state: dict of lists U, T, W, A (3 components); returns the new one."""
    K = _curl(F, sign, dt, inv)
    M = _mask()
    out = {k: [v.copy() for v in state[k]] for k in state}
    for a in range(3):
        so, sn, sp = (_grid_sig(sig, a), _grid_sig(sig, (a + 1) % 3), _grid_sig(sig, (a + 2) % 3))
        U, T, W, A = state["U"][a], state["T"][a], state["W"][a], state["A"][a]
        Un = np.where(sn != 0, ((1 - sn) * U + K[a]) / (1 + sn), U)
        du = np.where(sn != 0, Un - U, K[a])
        Tn = ((1 - sp) * T + du) / (1 + sp)
        Wn = m[a] * Tn
        An = np.where(so != 0, A + (1 + so) * Wn - (1 - so) * W, Wn)
        Wk = np.where(so != 0, Wn, W)
        for key, v in (("U", Un), ("T", Tn), ("W", Wk), ("A", An)):
            out[key][a] = np.where(M > 0, v, state[key][a])
    return out


def transpose(lam, lamF, Kin_other, sig, m, sign, dt, inv, fwd_net=None):
    """adjoint_half_kernel's order: lam (U, T, W, A of this group after the
    step) -> before it, plus Kbar; the other group's adjoint gains minus the
    transposed curl of Kbar.  Returns (lam_before, lamF_after, grad)."""
    M = _mask()
    out = {k: [v.copy() for v in lam[k]] for k in lam}
    Kbar = [np.zeros(SH) for _ in range(3)]
    grad = [np.zeros(SH) for _ in range(3)]
    for a in range(3):
        so, sn, sp = (_grid_sig(sig, a), _grid_sig(sig, (a + 1) % 3), _grid_sig(sig, (a + 2) % 3))
        ap = lam["A"][a]
        wh = np.where(so == 0, ap, lam["W"][a] + (1 + so) * ap)
        newA = np.where(so == 0, 0.0, ap)
        newW = np.where(so == 0, lam["W"][a], -(1 - so) * ap)
        if fwd_net is not None:
            grad[a] = M * wh * fwd_net[a]
        th = lam["T"][a] + m[a] * wh
        newT = (1 - sp) / (1 + sp) * th
        dl = th / (1 + sp)
        uh = lam["U"][a] + dl
        newU = np.where(sn == 0, lam["U"][a], (1 - sn) / (1 + sn) * uh - dl)
        kb = np.where(sn == 0, dl, uh / (1 + sn))
        for key, v in (("U", newU), ("T", newT), ("W", newW), ("A", newA)):
            out[key][a] = np.where(M > 0, v, lam[key][a])
        Kbar[a] = M * kb
    # the other group's adjoint: minus the curl of the opposite direction
    C = _curl(Kbar, -sign, dt, inv)
    lamF2 = [lamF[a] - C[a] for a in range(3)]
    return out, lamF2, grad


def _rand_state(rng):
    return {k: [rng.standard_normal(SH) * _mask() for _ in range(3)] for k in ("U", "T", "W", "A")}


def _dot(s1, s2):
    return sum(float(np.sum(s1[k][a] * s2[k][a])) for k in s1 for a in range(3))


@pytest.mark.parametrize("sign", [1, -1])
def test_transposed_half_step_dot_product(sign):
    rng = np.random.default_rng(3 if sign > 0 else 4)
    sig = _sig(rng)
    m = [rng.uniform(0.2, 1.0, SH) for _ in range(3)]
    dt, inv = 0.37, (1.3, 0.9, 1.7)
    x = _rand_state(rng)
    xF = [rng.standard_normal(SH) * _mask() for _ in range(3)]
    y = _rand_state(rng)
    yF = [rng.standard_normal(SH) * _mask() for _ in range(3)]
    Ax = forward(x, xF, sig, m, sign, dt, inv)
    lhs = _dot(Ax, y) + sum(float(np.sum(xF[a] * yF[a])) for a in range(3))   # F passes through
    lam, lamF, _ = transpose(y, yF, None, sig, m, sign, dt, inv)
    rhs = _dot(x, lam) + sum(float(np.sum(xF[a] * lamF[a])) for a in range(3))
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # a planted mistake (the U stage's coefficient flipped) breaks it
    bad = [s.copy() for s in sig]
    lam_b, lamF_b, _ = transpose(y, yF, None, [2 * s for s in bad], m, sign, dt, inv)
    assert abs(_dot(x, lam_b) + sum(float(np.sum(xF[a] * lamF_b[a])) for a in range(3)) - lhs) > 1e-6 * abs(lhs)


def test_gradient_term_equals_difference_in_m():
    rng = np.random.default_rng(7)
    sig = _sig(rng)
    m = [rng.uniform(0.2, 1.0, SH) for _ in range(3)]
    dt, inv = 0.29, (1.1, 0.8, 1.4)
    x = _rand_state(rng)
    xF = [rng.standard_normal(SH) * _mask() for _ in range(3)]
    y = _rand_state(rng)
    # net' of the forward step: W' / m where the own-axis sigma is non-zero, A' / m elsewhere
    Ax = forward(x, xF, sig, m, -1, dt, inv)
    net = [np.where(_grid_sig(sig, a) != 0, Ax["W"][a], Ax["A"][a]) / m[a] for a in range(3)]
    _, _, grad = transpose(y, [np.zeros(SH)] * 3, None, sig, m, -1, dt, inv, fwd_net=net)
    d = [rng.standard_normal(SH) * _mask() for _ in range(3)]
    h = 1e-6
    fp = _dot(forward(x, xF, sig, [m[a] + h * d[a] for a in range(3)], -1, dt, inv), y)
    fm = _dot(forward(x, xF, sig, [m[a] - h * d[a] for a in range(3)], -1, dt, inv), y)
    fd = (fp - fm) / (2 * h)
    ad = sum(float(np.sum(grad[a] * d[a])) for a in range(3))
    assert ad == pytest.approx(fd, rel=1e-8)
