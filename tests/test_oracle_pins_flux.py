"""Pins of the oracle's inviscid fluxes against published formulas and exact
properties, independent of the oracle's code:

* AUSM+ at subsonic and mixed faces against Liou's split Mach and pressure
  polynomials (Liou 1996, J. Comput. Phys. 129:364-382) with alpha = 3/16,
  beta = 1/8, evaluated here from the published definitions;
* HLLC: exact resolution of a stationary contact, a moving contact and a
  shear wave (air and the 9-species mixture), and the star region against
  Toro's HLLC star states (Riemann Solvers and Numerical Methods for Fluid
  Dynamics, 3rd ed., eqs. 10.33-10.39) with Batten's contact speed and the
  Roe-averaged wave-speed bounds;
* mixture Roe: consistency F(q, q) = physical flux with the species rows,
  and exact upwinding of an isobaric, isothermal composition contact (which
  is only exact when the species rows of the eigenvectors are right);
* the AWENO-5 + CDS flux difference of smooth data converges at fifth order
  over the cells where the central correction applies (AWENO.cu:215-264); the
  same data with any one CDS coefficient perturbed converges at second or
  fourth order, so the coefficient set is pinned.
"""
import numpy as np
import pytest

import pyoracle as po
from ostc import cases

R_U = 8314.462618
GAMMA = 1.4
MW_AIR = 28.9647


def _met(k):
    return np.array(list(k) + list(k), dtype=np.float64)


def _air_phys(q, k, J):
    """Metric-form Euler flux J (rho U, rho U u + k p, ..., U (rho E + p)), U = k . u."""
    rho, u, v, w, p = q[:5]
    E = 0.5 * rho * (u * u + v * v + w * w) + p / (GAMMA - 1)
    U = k[0] * u + k[1] * v + k[2] * w
    return J * np.array([rho * U, rho * U * u + k[0] * p, rho * U * v + k[1] * p, rho * U * w + k[2] * p,
                         U * (E + p)])


def _mix_rhoE(mech, q):
    rho, u, v, w, p = q[:5]
    Y = np.asarray(q[5:])
    W = np.array(mech.mw[:9])
    T = p / (rho * R_U * (Y / W).sum())
    h = po.enthalpy(mech, T)
    return rho * (Y * h).sum() - p + 0.5 * rho * (u * u + v * v + w * w), T


def _mix_phys(mech, q, k, J):
    rho, u, v, w, p = q[:5]
    rhoE, _ = _mix_rhoE(mech, q)
    U = k[0] * u + k[1] * v + k[2] * w
    f = [rho * U, rho * U * u + k[0] * p, rho * U * v + k[1] * p, rho * U * w + k[2] * p, U * (rhoE + p)]
    f += [rho * U * y for y in q[5:]]
    return J * np.array(f)


# ------------------------------------------------------------------ AUSM+

def _ausm_plus_liou(ql, qr, k, J):
    """Liou (1996) AUSM+: M+-(M) = +-1/4 (M +- 1)^2 +- beta (M^2 - 1)^2,
    P+-(M) = 1/4 (M +- 1)^2 (2 -+ M) +- alpha M (M^2 - 1)^2 for |M| < 1,
    (M +- |M|)/2 and (1 +- sign M)/2 otherwise; interface sound speed the
    mean of the two sides; mass flux c (rho_L m+ + rho_R m-), convected
    (1, u, H) of the upwind side; metric normal the face mean of J grad xi."""
    alpha, beta = 3.0 / 16.0, 1.0 / 8.0

    def mp(M):
        return 0.25 * (M + 1) ** 2 + beta * (M * M - 1) ** 2 if abs(M) < 1 else 0.5 * (M + abs(M))

    def mm(M):
        return -0.25 * (M - 1) ** 2 - beta * (M * M - 1) ** 2 if abs(M) < 1 else 0.5 * (M - abs(M))

    def pp(M):
        return 0.25 * (M + 1) ** 2 * (2 - M) + alpha * M * (M * M - 1) ** 2 if abs(M) < 1 else (1.0 if M > 0 else 0.0)

    def pm(M):
        return 0.25 * (M - 1) ** 2 * (2 + M) - alpha * M * (M * M - 1) ** 2 if abs(M) < 1 else (1.0 if M < 0 else 0.0)

    kk = J * np.asarray(k)
    gk = np.linalg.norm(kk)
    n = kk / gk
    rl, ul, pl_ = ql[0], np.asarray(ql[1:4]), ql[4]
    rr, ur, pr_ = qr[0], np.asarray(qr[1:4]), qr[4]
    c = 0.5 * (np.sqrt(GAMMA * pl_ / rl) + np.sqrt(GAMMA * pr_ / rr))
    Ml, Mr = ul @ n / c, ur @ n / c
    m = mp(Ml) + mm(Mr)
    mass = c * (rl * 0.5 * (m + abs(m)) + rr * 0.5 * (m - abs(m)))
    pf = pp(Ml) * pl_ + pm(Mr) * pr_
    up = (rl, ul, pl_) if mass >= 0 else (rr, ur, pr_)
    H = (0.5 * up[0] * up[1] @ up[1] + up[2] / (GAMMA - 1) + up[2]) / up[0]
    f = np.empty(5)
    f[0] = mass * gk
    f[1:4] = mass * gk * up[1] + pf * kk
    f[4] = mass * gk * H
    return f


@pytest.mark.parametrize("Ml,Mr", [(0.3, -0.5), (0.8, 0.4), (-0.2, -0.9), (0.05, 0.02), (1.4, 0.6), (0.7, -1.3)])
def test_ausm_plus_matches_liou_split_functions(Ml, Mr):
    k = np.array([0.6, -0.3, 0.2])
    J = 1.7e-3
    n = k / np.linalg.norm(k)
    rl, pl_, rr, pr_ = 1.2, 1.0e5, 0.9, 0.8e5
    cl, cr = np.sqrt(GAMMA * pl_ / rl), np.sqrt(GAMMA * pr_ / rr)
    t = np.cross(n, [0.0, 0.0, 1.0])
    ql = np.array([rl, *(Ml * cl * n + 20.0 * t), pl_])
    qr = np.array([rr, *(Mr * cr * n - 35.0 * t), pr_])
    got = po.face_flux(3, ql, qr, _met(k), np.array([J, J]))
    ref = _ausm_plus_liou(ql, qr, k, J)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_ausm_plus_alpha_beta_are_visible():
    """The published alpha and beta change a subsonic flux by per cent: the
    pin above would fail with alpha = 0 or beta = 0."""
    k, J = np.array([1.0, 0.0, 0.0]), 1.0
    ql = np.array([1.2, 150.0, 0.0, 0.0, 1.0e5])
    qr = np.array([0.9, -100.0, 0.0, 0.0, 0.8e5])
    got = po.face_flux(3, ql, qr, _met(k), np.array([J, J]))
    ref = _ausm_plus_liou(ql, qr, k, J)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    # the alpha/beta terms at this face are O(5 %) of the mass and pressure flux
    M = 150.0 / np.sqrt(GAMMA * 1e5 / 1.2)
    assert 0.125 * (M * M - 1) ** 2 > 0.05 and 3 / 16 * M * (M * M - 1) ** 2 > 0.01


# ------------------------------------------------------------------ HLLC

def _hllc_toro(ql, qr, k, J):
    """HLLC for air in metric form: wave speeds S_L = min(u_L - c_L, u~ - c~),
    S_R = max(u_R + c_R, u~ + c~) (Roe averages), contact speed (Batten et al.
    1997) S* = [p_R - p_L + rho_L u_L (S_L - u_L) - rho_R u_R (S_R - u_R)] /
    [rho_L (S_L - u_L) - rho_R (S_R - u_R)], star states (Toro eq. 10.39)
    U*_K = rho_K (S_K - u_K)/(S_K - S*) (1, S* n + tangential velocity,
    E_K/rho_K + (S* - u_K)(S* + p_K/(rho_K (S_K - u_K)))), flux F_K + S_K
    (U*_K - U_K), scaled by |k| and the face-mean J."""
    k = np.asarray(k)
    gk = np.linalg.norm(k)
    n = k / gk

    def parts(q):
        rho, vel, p = q[0], np.asarray(q[1:4]), q[4]
        E = 0.5 * rho * vel @ vel + p / (GAMMA - 1)
        un = vel @ n
        U = np.array([rho, *(rho * vel), E])
        F = np.array([rho * un, *(rho * un * vel + p * n), un * (E + p)])
        return rho, vel, p, E, un, U, F, np.sqrt(GAMMA * p / rho)

    rl, vl, pl_, El, ul, Ul, Fl, cl = parts(ql)
    rr, vr, pr_, Er, ur, Ur, Fr, cr = parts(qr)
    a, b = np.sqrt(rl) / (np.sqrt(rl) + np.sqrt(rr)), np.sqrt(rr) / (np.sqrt(rl) + np.sqrt(rr))
    vt = a * vl + b * vr
    Ht = a * (El + pl_) / rl + b * (Er + pr_) / rr
    ct = np.sqrt((GAMMA - 1) * (Ht - 0.5 * vt @ vt))
    SL = min(ul - cl, vt @ n - ct)
    SR = max(ur + cr, vt @ n + ct)
    if SL >= 0:
        F = Fl
    elif SR < 0:
        F = Fr
    else:
        Ss = (pr_ - pl_ + rl * ul * (SL - ul) - rr * ur * (SR - ur)) / (rl * (SL - ul) - rr * (SR - ur))

        def star(rho, vel, p, E, un, U, Fk, S):
            fac = rho * (S - un) / (S - Ss)
            Us = fac * np.array([1.0, *(vel + (Ss - un) * n), E / rho + (Ss - un) * (Ss + p / (rho * (S - un)))])
            return Fk + S * (Us - U)

        F = star(rl, vl, pl_, El, ul, Ul, Fl, SL) if Ss >= 0 else star(rr, vr, pr_, Er, ur, Ur, Fr, SR)
    return F * gk * J


@pytest.mark.parametrize("mix", [False, True])
def test_hllc_resolves_contacts_and_shear_exactly(mech, mix):
    k = np.array([0.7, 0.4, -0.2])
    J = 2.0e-3
    n = k / np.linalg.norm(k)
    t1 = np.cross(n, [0.0, 0.0, 1.0])
    t1 /= np.linalg.norm(t1)
    T, p = 600.0, 1.0e5
    if mix:
        Ya = np.zeros(9)
        Ya[2], Ya[8] = 0.233, 0.767
        Yb = np.zeros(9)
        Yb[0], Yb[8] = 0.3, 0.7
        W = np.array(mech.mw[:9])
        ra, rb = p / (R_U * T * (Ya / W).sum()), p / (R_U * T * (Yb / W).sum())
        extra_a, extra_b, m, phys = list(Ya), list(Yb), mech, lambda q: _mix_phys(mech, q, k, J)
    else:
        ra, rb = p * MW_AIR / (R_U * T), 0.45 * p * MW_AIR / (R_U * T)
        extra_a = extra_b = []
        m, phys = None, lambda q: _air_phys(q, k, J)
    jj = np.array([J, J])
    # stationary contact: only the pressure acts
    ql = np.array([ra, 0, 0, 0, p] + extra_a)
    qr = np.array([rb, 0, 0, 0, p] + extra_b)
    f = po.face_flux(4, ql, qr, _met(k), jj, m=m)
    ref = np.zeros_like(f)
    ref[1:4] = J * k * p
    np.testing.assert_allclose(f, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
    # contact moving with u_n = +60 m/s and a common tangential velocity: upwind flux
    vel = 60.0 * n + 25.0 * t1
    ql = np.array([ra, *vel, p] + extra_a)
    qr = np.array([rb, *vel, p] + extra_b)
    f = po.face_flux(4, ql, qr, _met(k), jj, m=m)
    np.testing.assert_allclose(f, phys(ql), rtol=1e-12, atol=1e-12 * np.abs(phys(ql)).max())
    # ... and moving with u_n = -60 m/s
    ql[1:4] = qr[1:4] = -60.0 * n + 25.0 * t1
    f = po.face_flux(4, ql, qr, _met(k), jj, m=m)
    np.testing.assert_allclose(f, phys(qr), rtol=1e-12, atol=1e-12 * np.abs(phys(qr)).max())
    # shear wave: the tangential velocity jumps, u_n, p, rho equal
    ql = np.array([ra, *(40.0 * n + 80.0 * t1), p] + extra_a)
    qr = np.array([ra, *(40.0 * n - 50.0 * t1), p] + extra_a)
    f = po.face_flux(4, ql, qr, _met(k), jj, m=m)
    np.testing.assert_allclose(f, phys(ql), rtol=1e-12, atol=1e-12 * np.abs(phys(ql)).max())


@pytest.mark.parametrize("seed", range(6))
def test_hllc_star_region_matches_toro(seed):
    rng = np.random.default_rng(seed)
    k = rng.standard_normal(3)
    J = 1e-3 * (1 + rng.random())
    n = k / np.linalg.norm(k)
    c0 = np.sqrt(GAMMA * 1e5 / 1.2)
    ql = np.array([1.2 * (1 + 0.3 * rng.random()), *(0.4 * c0 * rng.standard_normal(3)),
                   1e5 * (1 + 0.5 * rng.random())])
    qr = np.array([0.8 * (1 + 0.3 * rng.random()), *(0.4 * c0 * rng.standard_normal(3)),
                   0.7e5 * (1 + 0.5 * rng.random())])
    got = po.face_flux(4, ql, qr, _met(k), np.array([J, J]))
    ref = _hllc_toro(ql, qr, k, J)
    assert abs(ql[1:4] @ n) < 2 * c0 and abs(qr[1:4] @ n) < 2 * c0
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-11 * np.abs(ref).max())


# ------------------------------------------------------------- mixture Roe

def test_mixture_roe_consistency_and_composition_contact(mech):
    k = np.array([0.5, -0.6, 0.3])
    J = 1.5e-3
    jj = np.array([J, J])
    n = k / np.linalg.norm(k)
    t1 = np.cross(n, [1.0, 0.0, 0.0])
    t1 /= np.linalg.norm(t1)
    W = np.array(mech.mw[:9])
    rng = np.random.default_rng(7)
    Y = rng.random(9)
    Y /= Y.sum()
    T, p = 1400.0, 2.0e5
    rho = p / (R_U * T * (Y / W).sum())
    q = np.array([rho, *(300.0 * n + 50.0 * t1), p, *Y])
    ref = _mix_phys(mech, q, k, J)
    np.testing.assert_allclose(po.face_flux(2, q, q, _met(k), jj, m=mech), ref, rtol=1e-12,
                               atol=1e-12 * np.abs(ref).max())
    # isobaric, isothermal composition contact moving at +-80 m/s with a
    # common tangential velocity: the whole jump lies in the waves moving at
    # u_n, so the Roe flux is the upwind flux exactly
    Yb = np.zeros(9)
    Yb[0], Yb[8] = 0.3, 0.7
    rb = p / (R_U * T * (Yb / W).sum())
    for un in (80.0, -80.0):
        vel = un * n + 30.0 * t1
        ql = np.array([rho, *vel, p, *Y])
        qr = np.array([rb, *vel, p, *Yb])
        up = ql if un > 0 else qr
        ref = _mix_phys(mech, up, k, J)
        np.testing.assert_allclose(po.face_flux(2, ql, qr, _met(k), jj, m=mech), ref, rtol=1e-11,
                                   atol=1e-11 * np.abs(ref).max())


def test_mixture_hllc_and_ausm_consistency(mech):
    k = np.array([0.2, 0.9, -0.4])
    J = 3e-3
    W = np.array(mech.mw[:9])
    Y = np.linspace(1, 3, 9)
    Y /= Y.sum()
    T, p = 900.0, 1.5e5
    rho = p / (R_U * T * (Y / W).sum())
    q = np.array([rho, 120.0, -40.0, 30.0, p, *Y])
    ref = _mix_phys(mech, q, k, J)
    for scheme in (3, 4):
        np.testing.assert_allclose(po.face_flux(scheme, q, q, _met(k), np.array([J, J]), m=mech), ref, rtol=1e-12,
                                   atol=1e-12 * np.abs(ref).max())


# ---------------------------------------------------------------- AWENO-5 + CDS

CDS = (19.0 / 3840, -13.0 / 320, 17.0 / 256)


def _aweno_line(n, perturb=None):
    """AWENO-5 + CDS (AUSM+) right-hand side of smooth 1-D data on an n x 8
    Cartesian grid of spacing 1/n, every ghost filled from the same smooth
    functions.  Returns (dq / J, exact -dF/dx, node fluxes f with their
    point offsets) on one grid line."""
    import helpers
    h = 1.0 / n
    case = cases.c1_case(n, 8)
    case["perturb"] = None
    case["ngg"] = 3
    case["scheme"].update(reconstruction=4, inviscid_scheme=2)
    x = np.arange(n) * h
    case["x"] = np.broadcast_to(x, (1, 8, n)).copy()
    case["y"] = np.broadcast_to((np.arange(8) * h)[:, None], (1, 8, n)).copy()
    for b in case["bcs"]:
        b["type"] = 6
    z = helpers.oracle_zone(case)
    g = 3
    shp = cases.ghost_shape(n, 8, 1, g)
    xs = (np.arange(shp[2]) - g) * h
    tp = 2 * np.pi
    rho = 1.0 + 0.2 * np.sin(tp * xs)
    u = 300.0 + 50.0 * np.sin(tp * xs + 1.0)
    p = 1.0e5 * (1.0 + 0.1 * np.cos(tp * xs))
    E = 0.5 * rho * u * u + p / (GAMMA - 1)
    bv = z.field("bv").reshape(6, *shp)
    cv = z.field("cv").reshape(5, *shp)
    for arr, val in ((bv[0], rho), (bv[1], u), (bv[2], 0 * u), (bv[3], 0 * u), (bv[4], p),
                     (bv[5], p * MW_AIR / (rho * R_U))):
        arr[:] = val
    for arr, val in ((cv[0], rho), (cv[1], rho * u), (cv[2], 0 * u), (cv[3], 0 * u), (cv[4], E)):
        arr[:] = val
    z.rhs()
    dq = z.field("dq").reshape(5, 1, 8, n)[:, 0, 4, :]
    jac = z.field("jac").reshape(shp)[g, g + 4, g:g + n]
    # exact derivative of F = (rho u, rho u^2 + p, 0, 0, u (E + p))
    xi = x
    r_, u_, p_ = 1.0 + 0.2 * np.sin(tp * xi), 300.0 + 50.0 * np.sin(tp * xi + 1.0), 1.0e5 * (1 + 0.1 * np.cos(tp * xi))
    rx, ux, px = 0.2 * tp * np.cos(tp * xi), 50.0 * tp * np.cos(tp * xi + 1.0), -1.0e4 * tp * np.sin(tp * xi)
    E_ = 0.5 * r_ * u_ ** 2 + p_ / (GAMMA - 1)
    Ex = 0.5 * rx * u_ ** 2 + r_ * u_ * ux + px / (GAMMA - 1)
    dF = np.array([rx * u_ + r_ * ux, rx * u_ ** 2 + 2 * r_ * u_ * ux + px, 0 * xi, 0 * xi,
                   ux * (E_ + p_) + u_ * (Ex + px)])
    # node fluxes of the correction, J (rho U, rho U u + xi_x p, ., ., U (rho E + p)), U = u / h
    Jp = h * h * 1.0
    U = u / h
    fn = Jp * np.array([rho * U, rho * U * u + p / h, 0 * u, 0 * u, U * (E + p)])
    return dq / jac, -dF, fn, jac


def _cds(fn, n, c):
    """-(c1 (f+3 - f-3) + c2 (f+2 - f-2) + c3 (f+1 - f-1)) on the cells [0, n - 6)."""
    g = 3
    out = np.zeros((5, n - 2 * g))
    for s in range(n - 2 * g):
        q = s + g
        out[:, s] = -(c[0] * (fn[:, q + 3] - fn[:, q - 3]) + c[1] * (fn[:, q + 2] - fn[:, q - 2]) +
                      c[2] * (fn[:, q + 1] - fn[:, q - 1]))
    return out


def test_aweno_cds_flux_difference_is_fifth_order():
    errs, errs_bad = [], {0: [], 1: [], 2: []}
    for n in (48, 96, 192):
        num, ex, fn, jac = _aweno_line(n)
        cells = slice(0, n - 6)            # the reference applies CDS on [0, n - 2 ngg)
        scale = np.abs(ex[[0, 1, 4]]).max(axis=1)
        e = np.abs(num[[0, 1, 4], cells] - ex[[0, 1, 4], cells]).max(axis=1) / scale
        errs.append(e.max())
        J = jac[cells]
        for c in range(3):
            bad = list(CDS)
            bad[c] *= 1.1
            alt = num[:, cells] + (_cds(fn, n, bad) - _cds(fn, n, CDS)) / J
            eb = np.abs(alt[[0, 1, 4]] - ex[[0, 1, 4], cells]).max(axis=1) / scale
            errs_bad[c].append(eb.max())
    order = np.log2(errs[-2] / errs[-1])
    assert order > 4.5, (errs, order)
    for c in range(3):
        ob = np.log2(errs_bad[c][-2] / errs_bad[c][-1])
        assert ob < 4.5, (c, errs_bad[c], ob)


# ------------------------------------------------------------ reconstruction

def _recon(rec, q4):
    pv = np.ascontiguousarray(np.repeat(np.asarray(q4, dtype=np.float64)[:, None], 5, axis=1))
    pv[:, 4] = 1e5 + pv[:, 0]          # keep the pressure positive
    pl, pr = np.zeros(8), np.zeros(8)
    po.lib().orc_reconstruct_air(rec, po._p(pv.ravel().copy()), po._p(pl), po._p(pr))
    return pl[0], pr[0]


@pytest.mark.parametrize("rec", [2, 3])
def test_limited_reconstruction_tvd_properties(rec):
    """MUSCL (kappa = 1/3 on minmod-limited slopes) and NND2 of the
    reference both reduce to q_i + minmod(D-, D+)/2 (Harten's TVD slope
    limiting, minmod(a, b) = (sgn a + sgn b)/2 min(|a|, |b|)): exact face
    values for linear data, first order at an extremum, the smaller slope
    of a monotone kink."""
    # linear data: both one-sided reconstructions hit the face value
    pl, pr = _recon(rec, [1.0, 3.0, 5.0, 7.0])
    assert pl == pytest.approx(4.0, rel=1e-15) and pr == pytest.approx(4.0, rel=1e-15)
    # extremum at i: the left state is q_i
    pl, _ = _recon(rec, [2.0, 5.0, 3.0, 1.0])
    assert pl == 5.0
    # monotone kink: half the smaller slope
    pl, pr = _recon(rec, [0.0, 1.0, 5.0, 6.0])
    assert pl == pytest.approx(1.0 + 0.5 * 1.0, rel=1e-15)
    assert pr == pytest.approx(5.0 - 0.5 * 1.0, rel=1e-15)


def test_first_order_reconstruction_is_the_cell_value():
    pl, pr = _recon(1, [0.0, 1.0, 5.0, 6.0])
    assert (pl, pr) == (1.0, 5.0)


@pytest.mark.parametrize("seed", range(4))
def test_lax_friedrichs_is_rusanov(seed):
    """Rusanov's local Lax-Friedrichs flux: (F_L + F_R)/2 - s/2 (Q_R - Q_L)
    with s = max(|U_L| + c_L |k|, |U_R| + c_R |k|), metric-form, face-mean J."""
    rng = np.random.default_rng(seed)
    k = rng.standard_normal(3)
    J = 1e-3 * (1 + rng.random())
    ql = np.array([1.0 + rng.random(), *(200 * rng.standard_normal(3)), 1e5 * (1 + rng.random())])
    qr = np.array([1.0 + rng.random(), *(200 * rng.standard_normal(3)), 1e5 * (1 + rng.random())])

    def cons(q):
        return np.array([q[0], *(q[0] * q[1:4]), 0.5 * q[0] * q[1:4] @ q[1:4] + q[4] / (GAMMA - 1)])

    s = max(abs(k @ q[1:4]) + np.sqrt(GAMMA * q[4] / q[0]) * np.linalg.norm(k) for q in (ql, qr))
    ref = 0.5 * (_air_phys(ql, k, J) + _air_phys(qr, k, J)) - 0.5 * J * s * (cons(qr) - cons(ql))
    got = po.face_flux(1, ql, qr, _met(k), np.array([J, J]))
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
