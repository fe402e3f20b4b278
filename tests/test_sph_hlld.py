"""CPU pins of the spherical engine's HLLD Riemann solver (csrc/sph_physics.hpp:
hlld, hlld_frame and the face-frame maps), the relativistic five-wave solver
of Mignone, Ugliano & Bodo (2009) in the orthonormal frame of the normal
observer at each face, with the reference's wave structure and HLL fallback
(src/fluid/solver/hlld.cpp:230-312).

Each check is independent of the solver's own formulas:
  * consistency: equal states give the physical flux, computed by pt_flux
    (itself pinned to the stress-energy tensor in test_sph_oracle.py);
  * the frame maps: the solver's HLL branch, run in the face frame and mapped
    back, equals the coordinate-frame HLL (hll), at Kerr-Schild points, in
    all three directions -- a wrong triad vector or a missing shift term
    fails at O(1);
  * exactness: an isolated contact (only rho jumps) and an isolated
    tangential discontinuity (B^n = 0; v_x and the total pressure
    continuous) give the upwind physical flux, which HLL does not;
  * the Riemann fan: the Rankine-Hugoniot conditions
    lambda (U_b - U_a) = F(U_b) - F(U_a) hold across each of the five waves,
    with the intermediate states' fluxes formed from their own velocity,
    field and total pressure (the fan's states are those of the jump
    conditions, not of the equation of state), and the face flux is the
    fan's flux at x/t = 0.
"""
import math

import numpy as np

import sph_oracle as so
from sph_oracle import P
from test_sph_oracle import GAM, NMET, metric_rec, minkowski_rec, random_prim


def hlld(rec, wl, wr, n, hll_only=0, fan=False):
    out = np.zeros(7)
    fo = np.zeros(36) if fan else None
    rc = so.lib().os_hlld_point(P(rec[:NMET].copy()), GAM, P(wl.copy()), P(wr.copy()), n, hll_only, P(out),
                                P(fo) if fan else None)
    return (rc, out, fo) if fan else (rc, out)


def phys_flux(rec, w, n):
    """Densitised coordinate flux words (D, S_1..3, tau, B^t1, B^t2) and the
    undensitised conserved words and speeds of a primitive state."""
    f, c, lam = np.zeros(8), np.zeros(5), np.zeros(2)
    so.lib().os_flux_point(P(rec[:NMET].copy()), GAM, P(w.copy()), n, P(f), P(c), P(lam))
    t1, t2 = (n + 1) % 3, (n + 2) % 3
    return rec[16] * np.array([f[0], f[1], f[2], f[3], f[4], f[5 + t1], f[5 + t2]]), c, lam


def hll_coord(rec, wl, wr, n):
    out = np.zeros(7)
    so.lib().os_hll_point(P(rec[:NMET].copy()), GAM, P(wl.copy()), P(wr.copy()), n, P(out))
    return out


def ks_point(rng, a=0.9375):
    r = math.exp(rng.uniform(math.log(1.3), math.log(40)))
    th = rng.uniform(0.1, math.pi - 0.1)
    return metric_rec(a, 0.0, r, th)


def close(x, y, rtol):
    scale = max(np.abs(y).max(), 1e-300)
    return np.abs(x - y).max() <= rtol * scale


def test_hlld_consistency():
    rng = np.random.default_rng(21)
    hll_count = 0
    for it in range(300):
        rec = ks_point(rng, a=(0.9375, 0.5, 0.0)[it % 3])
        w = random_prim(rng, rec, magnetised=bool(it % 4))
        n = it % 3
        rc, out = hlld(rec, w, w, n)
        hll_count += rc
        ref, _, _ = phys_flux(rec, w, n)
        assert close(out, ref, 1e-11), (it, n, out, ref)
    assert hll_count <= 15, hll_count


def test_hlld_frame_hll_branch_equals_coordinate_hll():
    rng = np.random.default_rng(22)
    for it in range(300):
        rec = ks_point(rng, a=(0.9375, 0.99, 0.3)[it % 3])
        n = (it // 3) % 3
        wl = random_prim(rng, rec)
        wr = random_prim(rng, rec)
        wr[5 + n] = wl[5 + n]
        rc, out = hlld(rec, wl, wr, n, hll_only=1)
        assert rc == 1
        ref = hll_coord(rec, wl, wr, n)
        assert close(out, ref, 1e-12), (it, n, out, ref)


def test_hlld_resolves_isolated_contact():
    """Only rho jumps (B^n != 0): the contact moves with the flow and the
    flux through the face is the upwind state's physical flux."""
    rng = np.random.default_rng(23)
    worst_hlld, worst_hll = 0.0, 0.0
    for it in range(120):
        rec = minkowski_rec() if it % 2 == 0 else ks_point(rng)
        n = it % 3
        wl = random_prim(rng, rec)
        if abs(wl[5 + n]) < 1e-3 * math.sqrt(wl[0]):
            wl[5 + n] = 0.3 * math.sqrt(wl[0])
        wr = wl.copy()
        wr[0] = wl[0] * rng.uniform(0.2, 5.0)
        wr[1] = wl[1] * wl[0] / wr[0]          # same gas pressure
        rc, out = hlld(rec, wl, wr, n)
        assert rc == 0, it
        fl, _, laml = phys_flux(rec, wl, n)
        fr, _, _ = phys_flux(rec, wr, n)
        # contact speed in coordinates: alpha v^n - beta^n
        Wl = math.sqrt(1 + wl[2:5] @ np.array([[rec[4], rec[5], rec[6]], [rec[5], rec[7], rec[8]],
                                                [rec[6], rec[8], rec[9]]]) @ wl[2:5])
        vc = rec[0] * wl[2 + n] / Wl - rec[1 + n]
        ref = fl if vc > 0 else fr
        scale = np.abs(ref).max()
        worst_hlld = max(worst_hlld, np.abs(out - ref).max() / scale)
        worst_hll = max(worst_hll, np.abs(hll_coord(rec, wl, wr, n) - ref).max() / scale)
    assert worst_hlld < 1e-9, worst_hlld
    assert worst_hll > 1e-3, worst_hll    # the two-wave solver smears it


def test_hlld_resolves_tangential_discontinuity():
    """B^n = 0: v^n and the total pressure continuous, rho, v_t and B_t
    jump.  The face is static in flat space."""
    rng = np.random.default_rng(24)
    rec = minkowski_rec()
    worst = 0.0
    for it in range(100):
        n = it % 3
        t1, t2 = (n + 1) % 3, (n + 2) % 3
        vn = rng.uniform(-0.6, 0.6)
        def state(vn):
            rho = rng.uniform(0.1, 2.0)
            vt = rng.uniform(-0.5, 0.5, 2) * math.sqrt(max(1 - vn * vn, 0.0)) * 0.9
            Bt = rng.uniform(-1, 1, 2)
            v = np.zeros(3)
            v[n], v[t1], v[t2] = vn, vt[0], vt[1]
            B = np.zeros(3)
            B[t1], B[t2] = Bt
            return rho, v, B
        rl, vl, Bl = state(vn)
        rr, vr, Br = state(vn)
        def b2(v, B):
            W2 = 1 / (1 - v @ v)
            return B @ B / W2 + (v @ B) ** 2
        pl = rng.uniform(0.5, 2.0)
        pr = pl + 0.5 * (b2(vl, Bl) - b2(vr, Br))
        if pr <= 0.05:
            continue
        def prim(rho, p, v, B):
            w = np.zeros(8)
            W = 1 / math.sqrt(1 - v @ v)
            w[0], w[1] = rho, p / ((GAM - 1) * rho)
            w[2:5] = W * v
            w[5:8] = B
            return w
        wl, wr = prim(rl, pl, vl, Bl), prim(rr, pr, vr, Br)
        rc, out, fan = hlld(rec, wl, wr, n, fan=True)
        assert rc == 0 and fan[6] == 1.0, (it, rc, fan[:8])
        ref = phys_flux(rec, wl if vn > 0 else wr, n)[0]
        worst = max(worst, np.abs(out - ref).max() / np.abs(ref).max())
    assert worst < 1e-9, worst


def _cons_flat(w, n):
    """Frame conserved words (D, m_x, m_y, m_z, E, B_y, B_z) and fluxes of a
    primitive state in flat space (the frame is the coordinate frame with x
    along n)."""
    rec = minkowski_rec()
    f, c, _ = np.zeros(8), np.zeros(5), np.zeros(2)
    so.lib().os_flux_point(P(rec), GAM, P(w.copy()), n, P(f), P(c), P(_))
    t1, t2 = (n + 1) % 3, (n + 2) % 3
    perm = [n, t1, t2]
    U = np.array([c[0], c[1 + perm[0]], c[1 + perm[1]], c[1 + perm[2]], c[4] + c[0], w[5 + t1], w[5 + t2]])
    F = np.array([f[0], f[1 + perm[0]], f[1 + perm[1]], f[1 + perm[2]], f[4] + f[0], f[5 + t1], f[5 + t2]])
    return U, F


def fan_flux(U, p, bx):
    """The flux of a state of the fan from its conserved words U (D, m, E,
    B_y, B_z), its total pressure p and the normal field, without the
    equation of state (an HLL-type fan's states need not obey it): from
    m = (E + p) v - (v.B) B, v.B = B.m / (E + p - B^2) and
    v = (m + (v.B) B) / (E + p); then F(D) = D v_x, F(m) = m v_x -
    B_x (B / W^2 + (v.B) v) + p e_x, F(E) = m_x, F(B_t) = B_t v_x - B_x v_t."""
    B = np.array([bx, U[5], U[6]])
    m = U[1:4]
    E = U[4]
    vB = B @ m / (E + p - B @ B)
    v = (m + vB * B) / (E + p)
    assert v @ v < 1.0
    iW2 = 1.0 - v @ v
    F = np.zeros(7)
    F[0] = U[0] * v[0]
    F[1:4] = m * v[0] - bx * (B * iW2 + vB * v)
    F[1] += p
    F[4] = m[0]
    F[5] = B[1] * v[0] - bx * v[1]
    F[6] = B[2] * v[0] - bx * v[2]
    return F


def test_fan_flux_is_the_physical_flux():
    """fan_flux on physical states equals pt_flux (so the jump-condition
    checks below compare against the physics)."""
    rng = np.random.default_rng(27)
    rec = minkowski_rec()
    for it in range(100):
        n = it % 3
        w = random_prim(rng, rec)
        U, F = _cons_flat(w, n)
        v = w[2:5] / math.sqrt(1 + w[2:5] @ w[2:5])
        b2 = w[5:8] @ w[5:8] * (1 - v @ v) + (v @ w[5:8]) ** 2
        pt = (GAM - 1) * w[0] * w[1] + 0.5 * b2
        assert close(fan_flux(U, pt, w[5 + n]), F, 1e-10)


def _jumps(states, speeds, fluxes):
    worst = 0.0
    for k in range(len(speeds)):
        lhs = speeds[k] * (states[k + 1] - states[k])
        rhs = fluxes[k + 1] - fluxes[k]
        scale = max(np.abs(fluxes[k]).max(), np.abs(fluxes[k + 1]).max(), np.abs(states[k]).max())
        worst = max(worst, np.abs(lhs - rhs).max() / scale)
    return worst


def test_hlld_fan_satisfies_jump_conditions():
    """Random Riemann problems in flat space: the jump conditions
    lambda (U_b - U_a) = F(U_b) - F(U_a) hold across all five waves, with
    the fluxes of the intermediate states from their own velocity, field and
    total pressure (fan_flux), and every intermediate velocity sub-luminal."""
    rng = np.random.default_rng(25)
    rec = minkowski_rec()
    done, fell_back, worst = 0, 0, 0.0
    for it in range(400):
        n = it % 3
        wl = random_prim(rng, rec)
        wr = random_prim(rng, rec)
        wl[0] = rng.uniform(0.05, 2.0)
        wr[0] = rng.uniform(0.05, 2.0)
        wl[1], wr[1] = rng.uniform(0.05, 3.0), rng.uniform(0.05, 3.0)
        wl[2:5] *= 0.3
        wr[2:5] *= 0.3
        # a normal field clear of the degenerate range (B_x^2 < 1e-4 p)
        bx = rng.choice([-1.0, 1.0]) * rng.uniform(0.05, 1.0)
        wl[5 + n] = wr[5 + n] = bx
        rc, out, fan = hlld(rec, wl, wr, n, fan=True)
        if rc:
            fell_back += 1
            continue
        lL, laL, lc, laR, lR, p, deg = fan[:7]
        if not (lL < 0 < lR):
            continue
        assert deg == 0.0
        UL, FL = _cons_flat(wl, n)
        UR, FR = _cons_flat(wr, n)
        states = [UL, fan[8:15], fan[15:22], fan[22:29], fan[29:36], UR]
        fluxes = [FL] + [fan_flux(X, p, bx) for X in states[1:5]] + [FR]
        worst = max(worst, _jumps(states, [lL, laL, lc, laR, lR], fluxes))
        # the face flux is the fan's flux at x/t = 0
        k = int(np.searchsorted([lL, laL, lc, laR, lR], 0.0))
        t1, t2 = (n + 1) % 3, (n + 2) % 3
        Fk = fluxes[k]
        coord = np.zeros(7)
        coord[0] = Fk[0]
        coord[1 + n], coord[1 + t1], coord[1 + t2] = Fk[1], Fk[2], Fk[3]
        coord[4] = Fk[4] - Fk[0]
        coord[5], coord[6] = Fk[5], Fk[6]
        assert close(out, coord, 1e-9), (it, out, coord)
        done += 1
    assert done >= 250, (done, fell_back)
    assert worst < 1e-9, worst


def test_hlld_degenerate_fan_satisfies_jump_conditions():
    """B^n = 0: the Alfven waves merge with the contact (three waves)."""
    rng = np.random.default_rng(26)
    rec = minkowski_rec()
    done, worst = 0, 0.0
    for it in range(200):
        n = it % 3
        wl = random_prim(rng, rec)
        wr = random_prim(rng, rec)
        wl[0], wr[0] = rng.uniform(0.05, 2.0), rng.uniform(0.05, 2.0)
        wl[1], wr[1] = rng.uniform(0.05, 3.0), rng.uniform(0.05, 3.0)
        wl[2:5] *= 0.3
        wr[2:5] *= 0.3
        wl[5 + n] = wr[5 + n] = 0.0
        rc, out, fan = hlld(rec, wl, wr, n, fan=True)
        if rc:
            continue
        lL, laL, lc, laR, lR, p, deg = fan[:7]
        if not (lL < 0 < lR):
            continue
        assert deg == 1.0
        UL, FL = _cons_flat(wl, n)
        UR, FR = _cons_flat(wr, n)
        states = [UL, fan[8:15], fan[29:36], UR]
        fluxes = [FL, fan_flux(states[1], p, 0.0), fan_flux(states[2], p, 0.0), FR]
        worst = max(worst, _jumps(states, [lL, lc, lR], fluxes))
        done += 1
    assert done >= 100, done
    assert worst < 1e-8, worst
