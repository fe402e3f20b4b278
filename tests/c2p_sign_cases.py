"""Inputs for the sign-proof tests of the con2prim residual.

Parameter sets (D, tau, S^2, B^2, S.B) spread over the ranges the solver
meets, each sampled at the Newton bracket end and its doublings (the bracket
loop's points), on a log grid over the bracket, and at relative offsets
2^-1 ... 2^-60 and one ulp around the root of the residual (the Brent loop's
last points), where a sign is hardest to prove.
"""
import ctypes

import numpy as np

GAMMA = 5.0 / 3.0
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _P(a):
    return a.ctypes.data_as(_dp)


def residual(L, pts, on_device):
    """(sign, f) of bh_c2p_sign_batch at pts[n][6]."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    n = len(pts)
    sign = np.zeros(n, dtype=np.int32)
    f = np.zeros(n)
    rc = L.bh_c2p_sign_batch(GAMMA, n, _P(pts), sign.ctypes.data_as(_ip), _P(f), on_device)
    assert rc == 0
    return sign, f


def parameter_sets(rng, n):
    D = 10.0 ** rng.uniform(-9, 0, n)
    tau = D * 10.0 ** rng.uniform(-6, 2, n)
    Ssq = (D + tau) ** 2 * rng.uniform(0, 1.2, n)
    Bsq = 10.0 ** rng.uniform(-12, 0, n)
    SB = np.sqrt(Ssq * Bsq) * rng.uniform(-1, 1, n)
    return np.stack([D, tau, Ssq, Bsq, SB], axis=1)


def points(L, rng, n_sets=1500):
    """Sample points [m][6] = (xi, Ssq, Bsq, SB, D, tau) and a mask of the
    near-root ones."""
    ps = parameter_sets(rng, n_sets)
    D, tau, Ssq, Bsq, SB = ps.T
    xi_h = tau + D + Bsq / 2.0 + 1e3
    rows, near = [], []

    def add(xi, k, is_near):
        rows.append(np.stack([xi, Ssq[k], Bsq[k], SB[k], D[k], tau[k]], axis=1))
        near.append(np.full(len(xi), is_near))

    idx = np.arange(n_sets)
    for j in range(21):  # bracket end and its doublings
        add(xi_h * 2.0 ** j, idx, False)
    for u in np.linspace(0, 1, 24):  # log grid over [D, xi_h]
        add(D * (xi_h / D) ** u, idx, False)
    # root of the residual by bisection where the bracket changes sign
    lo, hi = D.copy(), xi_h.copy()
    base = lambda x: np.stack([x, Ssq, Bsq, SB, D, tau], axis=1)
    _, flo = residual(L, base(lo), 0)
    _, fhi = residual(L, base(hi), 0)
    ok = np.isfinite(flo) & np.isfinite(fhi) & (flo * fhi < 0)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        _, fm = residual(L, base(mid), 0)
        left = fm * flo < 0
        hi = np.where(left, mid, hi)
        lo = np.where(left, lo, mid)
        flo = np.where(left, flo, fm)
    root = 0.5 * (lo + hi)
    k = idx[ok]
    for e in range(1, 61):
        for sgn in (-1.0, 1.0):
            add(root[k] * (1.0 + sgn * 2.0 ** -e), k, e > 20)
    add(np.nextafter(root[k], np.inf), k, True)
    add(np.nextafter(root[k], 0.0), k, True)
    add(root[k], k, True)
    return np.concatenate(rows), np.concatenate(near)


def check(sign, f):
    """Every proved sign agrees with the residual (which is then non-zero)."""
    bad = ((sign > 0) & ~(f > 0)) | ((sign < 0) & ~(f < 0))
    return int(bad.sum())
