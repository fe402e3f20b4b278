"""CPU pins of the spherical Kerr-Schild engine's shared physics
(csrc/sph_physics.hpp, sph_geom.hpp, sph_torus.hpp) and of its oracle.

Nothing in the reference covers this preset, so each function is checked
against what the mathematics fixes, by a route of its own:
  * the metric against the Cartesian Kerr-Schild form g = eta + f l l pulled
    back through the coordinate map, its derivatives and Gamma^0 against
    finite differences;
  * the conserved variables, fluxes and sources against the stress-energy
    tensor built from the four-velocity and the field four-vector obtained by
    projection (b = (B + (u.B) u) / W);
  * con2prim as the inverse of prim->cons, its fallback as momentum-keeping;
  * the torus against its own definition (u^t u_phi = l everywhere, the
    Keplerian l at r_max, h = 1 at r_in) and the stationary Euler equation
    d_i ln h = u^m u^n d_i g_mn / 2;
  * the PCG64 stream against numpy;
  * the oracle's orchestration against invariants (div B, axisymmetry,
    equilibrium of an unmagnetised torus).
"""
import ctypes
import math

import numpy as np
import pytest

import sph_oracle as so
from sph_oracle import P

NGC, NMET = so.NGC, so.NMET
GAM = 13.0 / 9.0


def metric_rec(a, Q, r, th):
    out = np.zeros(NGC)
    so.lib().os_metric_point(a, Q, r, th, P(out))
    return out


def four_metric(rec):
    """g_mn and g^mn from a metric record (alpha, beta^i, gamma_ij, gamma^ij)."""
    al = rec[0]
    b = rec[1:4]
    s = rec[4:10]
    gam = np.array([[s[0], s[1], s[2]], [s[1], s[3], s[4]], [s[2], s[4], s[5]]])
    si = rec[10:16]
    igam = np.array([[si[0], si[1], si[2]], [si[1], si[3], si[4]], [si[2], si[4], si[5]]])
    bl = gam @ b
    g = np.zeros((4, 4))
    g[0, 0] = -al * al + bl @ b
    g[0, 1:] = g[1:, 0] = bl
    g[1:, 1:] = gam
    gu = np.zeros((4, 4))
    gu[0, 0] = -1 / al ** 2
    gu[0, 1:] = gu[1:, 0] = b / al ** 2
    gu[1:, 1:] = igam - np.outer(b, b) / al ** 2
    return g, gu


def ks_cartesian_pullback(a, Q, r, th, ph=0.3):
    """Spherical Kerr-Schild g_mn from the Cartesian Kerr-Schild form
    g = eta + f l l, f = (2 r - Q^2) r^2 / (r^4 + a^2 z^2),
    l = (1, (r x + a y)/(r^2 + a^2), (r y - a x)/(r^2 + a^2), z / r), pulled
    back through x + i y = (r + i a) e^{i phi} sin(theta), z = r cos(theta)."""
    x = (r * math.cos(ph) - a * math.sin(ph)) * math.sin(th)
    y = (r * math.sin(ph) + a * math.cos(ph)) * math.sin(th)
    z = r * math.cos(th)
    f = (2 * r - Q * Q) * r * r / (r ** 4 + a * a * z * z)
    l = np.array([1.0, (r * x + a * y) / (r * r + a * a), (r * y - a * x) / (r * r + a * a), z / r])
    gc = np.diag([-1.0, 1.0, 1.0, 1.0]) + f * np.outer(l, l)
    # Jacobian d(t, x, y, z) / d(t, r, theta, phi)
    J = np.zeros((4, 4))
    J[0, 0] = 1.0
    J[1, 1] = math.cos(ph) * math.sin(th)
    J[2, 1] = math.sin(ph) * math.sin(th)
    J[3, 1] = math.cos(th)
    J[1, 2] = (r * math.cos(ph) - a * math.sin(ph)) * math.cos(th)
    J[2, 2] = (r * math.sin(ph) + a * math.cos(ph)) * math.cos(th)
    J[3, 2] = -r * math.sin(th)
    J[1, 3] = (-r * math.sin(ph) - a * math.cos(ph)) * math.sin(th)
    J[2, 3] = (r * math.cos(ph) - a * math.sin(ph)) * math.sin(th)
    J[3, 3] = 0.0
    return J.T @ gc @ J


@pytest.mark.parametrize("a,Q", [(0.0, 0.0), (0.9375, 0.0), (0.5, 0.3)])
def test_metric_is_kerr_schild(a, Q):
    rng = np.random.default_rng(1)
    for _ in range(40):
        r = math.exp(rng.uniform(math.log(1.1), math.log(60)))
        th = rng.uniform(0.05, math.pi - 0.05)
        rec = metric_rec(a, Q, r, th)
        g, gu = four_metric(rec)
        gk = ks_cartesian_pullback(a, Q, r, th)
        assert np.allclose(g, gk, rtol=1e-12, atol=1e-12 * np.abs(gk).max()), (r, th, g, gk)
        assert np.allclose(g @ gu, np.eye(4), atol=1e-11)
        # sqrt(gamma) = sqrt(det gamma_ij), sqrt(-g) = alpha sqrt(gamma) = Sigma sin(theta)
        gam = g[1:, 1:]
        assert rec[16] == pytest.approx(math.sqrt(np.linalg.det(gam)), rel=1e-12)
        assert rec[0] * rec[16] == pytest.approx(math.sqrt(-np.linalg.det(g)), rel=1e-12)


@pytest.mark.parametrize("a,Q", [(0.9375, 0.0), (0.5, 0.3)])
def test_metric_derivatives_and_gamma0_by_finite_differences(a, Q):
    rng = np.random.default_rng(2)
    pm = [0, 0, 0, 0, 1, 1, 1, 2, 2, 3]
    pn = [0, 1, 2, 3, 1, 2, 3, 2, 3, 3]
    for _ in range(20):
        r = math.exp(rng.uniform(math.log(1.3), math.log(40)))
        th = rng.uniform(0.1, math.pi - 0.1)
        rec = metric_rec(a, Q, r, th)
        h = 1e-5

        def g_at(rr, tt):
            return four_metric(metric_rec(a, Q, rr, tt))[0]

        dgr = (g_at(r + h * r, th) - g_at(r - h * r, th)) / (2 * h * r)
        dgt = (g_at(r, th + h) - g_at(r, th - h)) / (2 * h)
        for q in range(10):
            m, n = pm[q], pn[q]
            assert rec[NMET + q] == pytest.approx(dgr[m, n], rel=1e-6, abs=1e-8)
            assert rec[NMET + 10 + q] == pytest.approx(dgt[m, n], rel=1e-6, abs=1e-8)
        # Gamma^0_mn = g^0l (d_m g_ln + d_n g_lm - d_l g_mn) / 2
        _, gu = four_metric(rec)
        dg = np.zeros((4, 4, 4))
        dg[1], dg[2] = dgr, dgt
        for q in range(10):
            m, n = pm[q], pn[q]
            G = 0.5 * sum(gu[0, l] * (dg[m, l, n] + dg[n, l, m] - dg[l, m, n]) for l in range(4))
            assert rec[NMET + 20 + q] == pytest.approx(G, rel=1e-6, abs=1e-8)
        al = lambda rr, tt: metric_rec(a, Q, rr, tt)[0]
        assert rec[NMET + 30] == pytest.approx((al(r + h * r, th) - al(r - h * r, th)) / (2 * h * r), rel=1e-6,
                                               abs=1e-10)
        assert rec[NMET + 31] == pytest.approx((al(r, th + h) - al(r, th - h)) / (2 * h), rel=1e-6, abs=1e-10)


def random_prim(rng, rec, magnetised=True):
    w = np.zeros(8)
    w[0] = math.exp(rng.uniform(math.log(1e-6), 0))
    w[1] = math.exp(rng.uniform(math.log(1e-5), math.log(2.0)))
    # u~ with Lorentz factors up to ~5 in the cell's metric
    gam = np.array([[rec[4], rec[5], rec[6]], [rec[5], rec[7], rec[8]], [rec[6], rec[8], rec[9]]])
    d = rng.normal(size=3)
    d /= math.sqrt(d @ gam @ d)
    w[2:5] = d * rng.uniform(0, 4.0)
    if magnetised:
        e = rng.normal(size=3)
        e /= math.sqrt(e @ gam @ e)
        w[5:8] = e * math.sqrt(w[0]) * rng.uniform(0, 3.0)
    return w


def tmunu(rec, w, gam_ad=GAM):
    """u^m, b^m and T^mn from (rho, eps, u~, B) by the projection route."""
    g, gu = four_metric(rec)
    al, beta = rec[0], rec[1:4]
    ut = w[2:5]
    gam = g[1:, 1:]
    W = math.sqrt(1 + ut @ gam @ ut)
    u = np.zeros(4)
    u[0] = W / al
    u[1:] = ut - W * beta / al
    Bv = np.concatenate([[0.0], w[5:8]])     # Eulerian field as a four-vector (n . B = 0)
    ul = g @ u
    b = (Bv + (ul @ Bv) * u) / W
    b2 = b @ g @ b
    rho = w[0]
    p = (gam_ad - 1) * rho * w[1]
    h = 1 + gam_ad / (gam_ad - 1) * p / rho
    T = (rho * h + b2) * np.outer(u, u) + (p + b2 / 2) * gu - np.outer(b, b)
    return u, b, T, g, W


def test_conserved_fluxes_against_stress_energy():
    rng = np.random.default_rng(3)
    for a in (0.9375, 0.3):
        for _ in range(60):
            r = math.exp(rng.uniform(math.log(1.2), math.log(40)))
            th = rng.uniform(0.1, math.pi - 0.1)
            rec = metric_rec(a, 0.0, r, th)
            w = random_prim(rng, rec)
            u, b, T, g, W = tmunu(rec, w)
            al, sg = rec[0], rec[16]
            sqg = al * sg
            U = np.zeros(8)
            so.lib().os_ptoc(P(rec[:NMET].copy()), GAM, P(w), P(U))
            Tl = T @ g                       # T^m_n
            D = sg * w[0] * W
            assert U[0] == pytest.approx(D, rel=1e-12)
            for j in range(3):
                assert U[1 + j] == pytest.approx(sg * al * Tl[0, 1 + j], rel=1e-9, abs=1e-12 * abs(U).max())
            assert U[4] == pytest.approx(sg * al * al * T[0, 0] - D, rel=1e-8, abs=1e-11 * abs(U).max())
            for n in range(3):
                f = np.zeros(8)
                c = np.zeros(5)
                lam = np.zeros(2)
                so.lib().os_flux_point(P(rec[:NMET].copy()), GAM, P(w), n, P(f), P(c), P(lam))
                F = sg * f                  # densitised
                sc = np.abs(F).max() + 1e-300
                assert F[0] == pytest.approx(sqg * w[0] * u[1 + n], rel=1e-10, abs=1e-12 * sc)
                for j in range(3):
                    assert F[1 + j] == pytest.approx(sqg * Tl[1 + n, 1 + j], rel=1e-8, abs=1e-11 * sc)
                assert F[4] == pytest.approx(sqg * (al * T[0, 1 + n] - w[0] * u[1 + n]), rel=1e-8, abs=1e-11 * sc)
                for k in range(3):
                    ref = sqg * (b[1 + k] * u[1 + n] - b[1 + n] * u[1 + k])
                    assert F[5 + k] == pytest.approx(ref, rel=1e-8, abs=1e-11 * sc)


def test_sources_against_finite_difference_geometry():
    rng = np.random.default_rng(4)
    a = 0.9375
    pm = [0, 0, 0, 0, 1, 1, 1, 2, 2, 3]
    pn = [0, 1, 2, 3, 1, 2, 3, 2, 3, 3]
    for _ in range(40):
        r = math.exp(rng.uniform(math.log(1.3), math.log(40)))
        th = rng.uniform(0.15, math.pi - 0.15)
        rec = metric_rec(a, 0.0, r, th)
        w = random_prim(rng, rec)
        u, b, T, g, W = tmunu(rec, w)
        out = np.zeros(5)
        so.lib().os_source_point(P(rec.copy()), GAM, P(w), P(out))
        h = 1e-5
        g_at = lambda rr, tt: four_metric(metric_rec(a, 0.0, rr, tt))[0]
        dg = np.zeros((4, 4, 4))
        dg[1] = (g_at(r + h * r, th) - g_at(r - h * r, th)) / (2 * h * r)
        dg[2] = (g_at(r, th + h) - g_at(r, th - h)) / (2 * h)
        sqg = rec[0] * rec[16]
        sc = np.abs(out).max() + 1e-300
        for j in (1, 2):
            assert out[j] == pytest.approx(sqg * 0.5 * np.sum(T * dg[j]), rel=1e-5, abs=1e-7 * sc)
        assert out[3] == 0.0 and out[0] == 0.0
        _, gu = four_metric(rec)
        G0 = np.array([[0.5 * sum(gu[0, l] * (dg[m, l, n] + dg[n, l, m] - dg[l, m, n]) for l in range(4))
                        for n in range(4)] for m in range(4)])
        al = lambda rr, tt: metric_rec(a, 0.0, rr, tt)[0]
        da = np.array([0.0, (al(r + h * r, th) - al(r - h * r, th)) / (2 * h * r), (al(r, th + h) - al(r, th - h)) / (2 * h), 0.0])
        ref = sqg * (T[:, 0] @ da - rec[0] * np.sum(T * G0))
        assert out[4] == pytest.approx(ref, rel=1e-5, abs=1e-7 * sc)


def test_con2prim_inverts_prim_to_cons():
    rng = np.random.default_rng(5)
    worst = 0.0
    for a in (0.9375, 0.0):
        for _ in range(400):
            r = math.exp(rng.uniform(math.log(1.2), math.log(40)))
            th = rng.uniform(0.05, math.pi - 0.05)
            rec = metric_rec(a, 0.0, r, th)
            w = random_prim(rng, rec, magnetised=bool(rng.integers(2)))
            U = np.zeros(8)
            so.lib().os_ptoc(P(rec[:NMET].copy()), GAM, P(w), P(U))
            w2 = np.zeros(8)
            rc = so.lib().os_c2p_point(P(rec[:NMET].copy()), GAM, P(U), P(w2))
            assert rc == 0
            rel = np.abs(w2 - w) / np.maximum(np.abs(w), 1e-12 * np.abs(w).max())
            worst = max(worst, rel.max())
    assert worst < 1e-8, worst


def test_con2prim_fallback_keeps_mass_and_momentum():
    """U whose energy is below that of a cold gas with its momentum has no
    root; the fallback keeps D and S (a cold gas) and reports it."""
    rng = np.random.default_rng(6)
    for _ in range(50):
        r = math.exp(rng.uniform(math.log(2), math.log(30)))
        th = rng.uniform(0.3, math.pi - 0.3)
        rec = metric_rec(0.9375, 0.0, r, th)
        w = random_prim(rng, rec, magnetised=False)
        U = np.zeros(8)
        so.lib().os_ptoc(P(rec[:NMET].copy()), GAM, P(w), P(U))
        U[4] = -0.5 * abs(U[0])           # energy too small
        w2 = np.zeros(8)
        assert so.lib().os_c2p_point(P(rec[:NMET].copy()), GAM, P(U), P(w2)) == 1
        U2 = np.zeros(8)
        so.lib().os_ptoc(P(rec[:NMET].copy()), GAM, P(w2), P(U2))
        assert np.allclose(U2[:4], U[:4], rtol=1e-12, atol=1e-14 * abs(U[:4]).max())


def minkowski_rec():
    rec = np.zeros(NMET)
    rec[0] = 1.0
    rec[4] = rec[7] = rec[9] = 1.0
    rec[10] = rec[13] = rec[15] = 1.0
    rec[16] = 1.0
    return rec


def test_signal_speeds_relativistic_sound():
    """B = 0 in flat space: lambda = (v +- c_s) / (1 +- v c_s) along n."""
    rng = np.random.default_rng(7)
    rec = minkowski_rec()
    for _ in range(100):
        w = np.zeros(8)
        w[0] = rng.uniform(0.1, 1)
        w[1] = rng.uniform(0.01, 3)
        n = int(rng.integers(3))
        v = rng.uniform(-0.95, 0.95)
        W = 1 / math.sqrt(1 - v * v)
        w[2 + n] = W * v
        f, c, lam = np.zeros(8), np.zeros(5), np.zeros(2)
        so.lib().os_flux_point(P(rec), GAM, P(w), n, P(f), P(c), P(lam))
        p = (GAM - 1) * w[0] * w[1]
        cs = math.sqrt(GAM * p / (w[0] + GAM / (GAM - 1) * p))
        assert lam[1] == pytest.approx((v + cs) / (1 + v * cs), rel=1e-12)
        assert lam[0] == pytest.approx((v - cs) / (1 - v * cs), rel=1e-12, abs=1e-14)


def test_hll_consistency():
    rng = np.random.default_rng(8)
    for _ in range(60):
        r = math.exp(rng.uniform(math.log(1.3), math.log(40)))
        th = rng.uniform(0.1, math.pi - 0.1)
        rec = metric_rec(0.9375, 0.0, r, th)
        w = random_prim(rng, rec)
        n = int(rng.integers(3))
        wl = w.copy()
        out = np.zeros(7)
        so.lib().os_hll_point(P(rec[:NMET].copy()), GAM, P(wl), P(w.copy()), n, P(out))
        f, c, lam = np.zeros(8), np.zeros(5), np.zeros(2)
        so.lib().os_flux_point(P(rec[:NMET].copy()), GAM, P(w), n, P(f), P(c), P(lam))
        t1, t2 = (n + 1) % 3, (n + 2) % 3
        ref = rec[16] * np.array([f[0], f[1], f[2], f[3], f[4], f[5 + t1], f[5 + t2]])
        assert np.allclose(out, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())


def test_pcg64_matches_numpy():
    w = so.pcg_words(42)
    ref = np.random.Generator(np.random.PCG64(42)).random(200000)
    for n in list(range(50)) + [999, 65535, 123456, 199999]:
        assert so.lib().os_pcg64(*[int(x) for x in w], n) == ref[n]
    w7 = so.pcg_words(7)
    ref7 = np.random.Generator(np.random.PCG64(7)).random(10)
    assert [so.lib().os_pcg64(*[int(x) for x in w7], n) for n in range(10)] == list(ref7)


def torus_point(a, r, th, r_in=6.0, r_max=12.0, kappa=1e-3):
    out = np.zeros(7)
    so.lib().os_torus_point(a, GAM, r_in, r_max, kappa, r, th, P(out))
    return out


@pytest.mark.parametrize("a", [0.9375, 0.0, 0.6])
def test_torus_definition(a):
    # l = u^t u_phi of the prograde circular orbit at r_max (Bardeen, Press &
    # Teukolsky 1972: E, L of circular equatorial geodesics; u^t = -g^tt E + g^tphi L)
    r = 12.0
    sr = math.sqrt(r)
    den = r ** 0.75 * math.sqrt(r ** 1.5 - 3 * sr + 2 * a)
    E = (r ** 1.5 - 2 * sr + a) / den
    L = (r * r - 2 * a * sr + a * a) / den
    # Boyer-Lindquist inverse metric on the equator
    Del = r * r - 2 * r + a * a
    A = (r * r + a * a) ** 2 - Del * a * a
    gtt_u = -A / (r * r * Del)
    gtp_u = -2 * a * r / (r * r * Del)
    ut = -gtt_u * E + gtp_u * L
    out = torus_point(a, 12.0, math.pi / 2)
    assert out[0] == pytest.approx(ut * L, rel=1e-12)
    # h = 1 at the inner edge on the equator
    assert torus_point(a, 6.0, math.pi / 2)[2] == pytest.approx(0.0, abs=1e-13)
    # u.u = -1 and u^t u_phi = l inside the torus
    rng = np.random.default_rng(9)
    n = 0
    for _ in range(200):
        rr = rng.uniform(6.5, 20)
        th = rng.uniform(1.0, math.pi - 1.0)
        o = torus_point(a, rr, th)
        if o[3] <= 0:
            continue
        n += 1
        rec = metric_rec(a, 0.0, rr, th)
        g, _ = four_metric(rec)
        ut_ = o[5] / rec[1]                  # u~^r = u^t beta^r
        u = np.array([ut_, 0.0, 0.0, o[4]])
        assert u @ g @ u == pytest.approx(-1.0, rel=1e-10)
        assert ut_ * (g[3, 0] * ut_ + g[3, 3] * o[4]) == pytest.approx(out[0], rel=1e-10)
    assert n > 10


def test_torus_is_in_hydrostatic_equilibrium():
    """Stationary Euler equation of the torus: d_i ln h = u^m u^n d_i g_mn / 2
    (the pressure gradient of the polytrope balances gravity and rotation)."""
    a = 0.9375
    for rr, th in ((9.0, 1.5), (12.0, 1.3), (15.0, 1.6), (8.0, 1.45), (11.0, 1.75)):
        o = torus_point(a, rr, th)
        assert o[3] > 0
        rec = metric_rec(a, 0.0, rr, th)
        ut = o[5] / rec[1]
        u = np.array([ut, 0.0, 0.0, o[4]])
        h = 1e-5
        lnh = lambda r_, t_: torus_point(a, r_, t_)[2]
        dr = (lnh(rr + h * rr, th) - lnh(rr - h * rr, th)) / (2 * h * rr)
        dt = (lnh(rr, th + h) - lnh(rr, th - h)) / (2 * h)
        g_at = lambda r_, t_: four_metric(metric_rec(a, 0.0, r_, t_))[0]
        dgr = (g_at(rr + h * rr, th) - g_at(rr - h * rr, th)) / (2 * h * rr)
        dgt = (g_at(rr, th + h) - g_at(rr, th - h)) / (2 * h)
        assert dr == pytest.approx(0.5 * u @ dgr @ u, rel=1e-5, abs=1e-9)
        assert dt == pytest.approx(0.5 * u @ dgt @ u, rel=1e-5, abs=1e-9)


# ------------------------------------------------------------- the oracle

def _cfg(grid, bz, a=0.9375):
    return dict(bz=bz, a=a, Q=0.0, gamma=GAM, cfl=0.4, rho_fl=1e-7, eps_fl=1e-5, sigma_max=100.0, w_max=50.0,
                nr=grid["nr"], nth=grid["nth"], nph=grid["nph"], r_in=1.2, r_out=40.0, nlev=grid.get("nlevels", 1))


def _blocks(grid, bz):
    nbp = 1 if bz == 1 else grid["nph"] // bz
    return [(0, i, j, k) for k in range(nbp) for j in range(grid["nth"] // 16) for i in range(grid["nr"] // 16)]


TOR = dict(r_in=6.0, r_max=12.0, kappa=1e-3, beta=100.0, acut=0.2, pert=0.04)


def test_oracle_div_b_stays_at_round_off():
    g = dict(nr=64, nth=64, nph=1)
    o = so.OracleRun(_cfg(g, 1), _blocks(g, 1))
    o.init_torus(TOR, 42)
    d0 = o.diag()
    assert d0["divb_max"] <= 1e-14 * d0["divb_scale"]
    for _ in range(20):
        o.step(False)
    d = o.diag()
    assert d["divb_max"] <= 1e-13 * d["divb_scale"], d


def test_oracle_axisymmetric_3d_stays_axisymmetric():
    """Without the perturbation the torus is axisymmetric; the 3D step must
    keep every phi slice identical (bit for bit), fluxes, ghosts and EMFs
    included."""
    g = dict(nr=32, nth=32, nph=32)
    o = so.OracleRun(_cfg(g, 16), _blocks(g, 16))
    o.init_torus(dict(TOR, pert=0.0), 42)
    for _ in range(3):
        o.step(False)
    U, W, Bf = o.state()
    Ua = U.reshape(-1, 8, 20, 20, 20)[:, :, 2:18, 2:18, 2:18]
    nb = Ua.shape[0]
    # blocks ordered phi-slowest: compare every phi plane with the first
    per_k = nb // 2
    ref = Ua[:per_k, :, 0]
    for bk in range(2):
        for k in range(16):
            assert np.array_equal(Ua[bk * per_k:(bk + 1) * per_k, :, k], ref), (bk, k)


def test_oracle_unmagnetised_torus_is_stationary():
    """With no field and no perturbation the torus is an equilibrium: 40 steps
    change its density core by the truncation error only."""
    g = dict(nr=128, nth=128, nph=1)
    blocks = _blocks(g, 1)
    o = so.OracleRun(_cfg(g, 1), blocks)
    o.init_torus(dict(TOR, pert=0.0, acut=2.0), 42)
    U, W0, B = o.state()
    for _ in range(40):
        o.step(False)
    o.c2p()
    U, W1, B = o.state()
    r0 = W0.reshape(len(blocks), 8, 5, 20, 20)[:, 0, 2, 2:18, 2:18]
    r1 = W1.reshape(len(blocks), 8, 5, 20, 20)[:, 0, 2, 2:18, 2:18]
    core = r0 > 0.3
    assert core.sum() > 100
    rel = np.abs(r1[core] - r0[core]) / r0[core]
    assert rel.max() < 2e-2, rel.max()
    assert np.isfinite(W1).all()


def test_analysis_terms_against_stress_energy():
    """analysis_terms (the analyser quantities' per-zone terms) against u^m,
    b^m and T^mn built independently: b^2/2, rho h u^r u_phi, -b^r b_phi, the
    inflows -rho u^r and -T^r_phi, and the electromagnetic part of -T^r_t,
    with the rho_min cut and the horizon switch."""
    rng = np.random.default_rng(21)
    for a in (0.9375, 0.5):
        for _ in range(60):
            r = math.exp(rng.uniform(math.log(1.3), math.log(40)))
            th = rng.uniform(0.1, math.pi - 0.1)
            rec = metric_rec(a, 0.0, r, th)
            w = random_prim(rng, rec)
            u, b, T, g, W = tmunu(rec, w)
            out = np.zeros(9)
            so.lib().os_analysis_point(P(rec[:NMET].copy()), GAM, P(w), 0.0, 1, P(out))
            b2 = b @ g @ b
            rho, p = w[0], (GAM - 1) * w[0] * w[1]
            h = 1 + GAM / (GAM - 1) * p / rho
            ul, bl = g @ u, g @ b
            Tmix = T @ g                                   # T^m_n
            ref = [rho * w[1], b2 / 2, p, p + b2 / 2, rho * h * u[1] * ul[3], -b[1] * bl[3],
                   -rho * u[1], -Tmix[1, 3], -(b2 * u[1] * ul[0] - b[1] * bl[0])]
            for q in range(9):
                sc = abs(ref[q]) + 1e-12 * (abs(rho * h * u[1] * ul[3]) + b2 + p + abs(Tmix[1, 3]) + 1e-300)
                assert abs(out[q] - ref[q]) <= 1e-9 * sc, (q, out[q], ref[q])
            # the density cut and the horizon switch
            so.lib().os_analysis_point(P(rec[:NMET].copy()), GAM, P(w), 2 * rho, 0, P(out))
            assert out[0] == pytest.approx(ref[0]) and out[1] == pytest.approx(ref[1])
            assert (out[3:] == 0.0).all()
