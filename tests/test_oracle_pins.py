"""Pins of the CPU oracle against what the mathematics and the reference's
own data fix, independent of the oracle's code:

* NASA-7 thermo: textbook polynomials from the golden tables (h, cp, and
  g/RT = h/RT - s/R), cp = dh/dT, the below-T_low extrapolation rule;
* kinetics: Arrhenius closed form with Ea in cal/mol -> K, the published
  Troe fall-off form and its high/low-pressure limits, Kc from NASA-7
  free energies (detailed balance), mass and element conservation of omega;
* the EPI Jacobian against finite differences of omega at frozen rate
  constants; the LU solve against numpy; the Newton temperature inversion
  against the NASA-7 energy it inverts; the batch integrator against an
  independent dense restatement of one point-implicit step;
* transport: pure-species and binary-mixture limits of the Wilke / mixing
  rules, Sutherland at its reference point;
* fluxes: consistency F(U,U) = physical flux, supersonic upwinding, WENO5
  exact on quadratics and 5th-order convergent, L R = I of the AWENO
  eigenvectors, SSP-RK3 stability polynomial, uniform flow preserved;
* viscous flux: plane Couette shear, whose interior right-hand side is fixed
  in closed form (viscous dissipation mu a^2 heats the gas, nothing else
  changes), for air and the 9-species mixture.
"""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
from ostc import cases

R_U = 8314.462618
R_C = 1.98720425864083


def _coef(t, i, T):
    lo = t["low_coeff"].reshape(-1, 7)[i]
    hi = t["high_coeff"].reshape(-1, 7)[i]
    return lo if T < t["t_mid"][i] else hi


def nasa_h_cp_g(t, T):
    """Textbook NASA-7: h/RT, cp/R, s/R."""
    ns = t["n_spec"]
    h, cp, g = np.zeros(ns), np.zeros(ns), np.zeros(ns)
    for i in range(ns):
        a = _coef(t, i, T)
        cp_r = a[0] + a[1] * T + a[2] * T**2 + a[3] * T**3 + a[4] * T**4
        h_rt = a[0] + a[1] * T / 2 + a[2] * T**2 / 3 + a[3] * T**3 / 4 + a[4] * T**4 / 5 + a[5] / T
        s_r = a[0] * np.log(T) + a[1] * T + a[2] * T**2 / 2 + a[3] * T**3 / 3 + a[4] * T**4 / 4 + a[6]
        W = t["mw"][i]
        h[i] = h_rt * R_U * T / W
        cp[i] = cp_r * R_U / W
        g[i] = h_rt - s_r
    return h, cp, g


@pytest.mark.parametrize("T", [450.0, 999.0, 1001.0, 1800.0, 2700.0])
def test_nasa7_closed_form(mech, tables, T):
    h, cp = po.enthalpy_cp(mech, T)
    g = po.gibbs_rt(mech, T)
    hr, cpr, gr = nasa_h_cp_g(tables, T)
    np.testing.assert_allclose(h, hr, rtol=1e-13, atol=1e-6)
    np.testing.assert_allclose(cp, cpr, rtol=1e-13)
    np.testing.assert_allclose(g, gr, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(po.cp(mech, T), cpr, rtol=1e-13)
    # cp = dh/dT away from the range switch
    if abs(T - 1000.0) > 5:
        dT = 1e-3 * T
        dh = (po.enthalpy(mech, T + dT) - po.enthalpy(mech, T - dT)) / (2 * dT)
        np.testing.assert_allclose(dh, cp, rtol=1e-7)


def test_below_t_low_extrapolation(mech, tables):
    """Thermo.cu:8-16: below T_low, h(T) = h(T_low) + cp(T_low)(T - T_low) and
    cp, g/RT stay at their T_low values."""
    T = 150.0
    h = po.enthalpy(mech, T)
    cp = po.cp(mech, T)
    g = po.gibbs_rt(mech, T)
    for i in range(tables["n_spec"]):
        tl = tables["t_low"][i]
        h_l, cp_l, g_l = nasa_h_cp_g(tables, tl)
        assert h[i] == pytest.approx(h_l[i] + cp_l[i] * (T - tl), rel=1e-12, abs=1e-6)
        assert cp[i] == pytest.approx(cp_l[i], rel=1e-13)
        assert g[i] == pytest.approx(g_l[i], rel=1e-12, abs=1e-12)


def _state(tables, T=1500.0, p=2e5, seed=5):
    rng = np.random.default_rng(seed)
    Y = rng.uniform(0.01, 1.0, tables["n_spec"])
    Y /= Y.sum()
    mwinv = (Y / tables["mw"]).sum()
    rho = p / (R_U * T * mwinv)
    return rho, T, Y


def test_arrhenius_and_units(mech, tables):
    """Reaction 0 (H+O2=O+OH, plain Arrhenius): kf = A T^b exp(-Ea/(R_c T)) with
    Ea from the mechanism in cal/mol."""
    T = 1234.0
    c = np.full(9, 1e-6)
    kf, kb = po.rates(mech, T, c)
    A, b, Ea_cal = 3.55e15, -0.40, 1.66e4
    assert kf[0] == pytest.approx(A * T**b * np.exp(-Ea_cal / R_C / T), rel=1e-13)


def test_detailed_balance(mech, tables):
    """kf/kb = Kc = (p_atm/(R_u T) 1e-3)^dnu exp(-dG/RT) from textbook NASA-7
    free energies, for every reversible reaction without REV parameters."""
    T = 1700.0
    c = np.full(9, 2e-6)
    kf, kb = po.rates(mech, T, c)
    _, _, g = nasa_h_cp_g(tables, T)
    ns = tables["n_spec"]
    sf = tables["stoi_f"].reshape(-1, ns)
    sb = tables["stoi_b"].reshape(-1, ns)
    for r in range(tables["n_reac"]):
        if tables["label"][r] in (0, 2):
            continue
        dnu = (sb[r] - sf[r]).sum()
        kc = (101325.0 / R_U * 1e-3 / T) ** dnu * np.exp(-np.dot(sb[r] - sf[r], g))
        assert kf[r] / kb[r] == pytest.approx(kc, rel=1e-10), r


def test_troe_form_and_limits(mech, tables):
    """Published Troe fall-off (Gilbert, Luther, Troe 1983) for reaction 8
    (H+O2(+M)=HO2(+M)) and the high/low-pressure limits of the blend."""
    r = 8
    assert tables["label"][r] == 6
    T = 1400.0
    A, b, Ea = 1.48e12, 0.60, 0.0
    A0, b0, Ea0 = tables["A2"][r], tables["b2"][r], tables["Ea2"][r]
    alpha, T3, T1 = tables["troe_alpha"][r], tables["troe_t3"][r], tables["troe_t1"][r]
    eff = tables["third_body_coeff"].reshape(-1, 9)[r]

    def troe(c):
        kinf = A * T**b * np.exp(-Ea / T)
        k0 = A0 * T**b0 * np.exp(-Ea0 / T)
        Pr = k0 * np.dot(eff, c) / kinf
        Fc = (1 - alpha) * np.exp(-T / T3) + alpha * np.exp(-T / T1)
        cc = -0.4 - 0.67 * np.log10(Fc)
        nn = 0.75 - 1.27 * np.log10(Fc)
        f1 = (np.log10(Pr) + cc) / (nn - 0.14 * (np.log10(Pr) + cc))
        F = 10 ** (np.log10(Fc) / (1 + f1 * f1))
        return kinf * Pr / (1 + Pr) * F, kinf, k0

    for scale in (1e-9, 1e-6, 1e-3):
        c = np.full(9, scale)
        kf, _ = po.rates(mech, T, c)
        ref, kinf, k0 = troe(c)
        assert kf[r] == pytest.approx(ref, rel=1e-12)
    # limits: the broadening factor tends to 1 at both ends, so kf -> k_inf
    # at high pressure and kf -> k0 [M] at low pressure
    c = np.full(9, 1e3)
    kf, _ = po.rates(mech, T, c)
    assert kf[r] == pytest.approx(troe(c)[1], rel=2e-2)
    c = np.full(9, 1e-14)
    kf, _ = po.rates(mech, T, c)
    ref, kinf, k0 = troe(c)
    assert kf[r] / (k0 * np.dot(eff, c)) == pytest.approx(1.0, rel=2e-2)
    # at log10 Pr = -c the correction term vanishes and F equals F_cent exactly
    Fc = (1 - alpha) * np.exp(-T / T3) + alpha * np.exp(-T / T1)
    Pr = 10 ** (0.4 + 0.67 * np.log10(Fc))
    M = Pr * kinf / k0
    c = M / eff.sum() * np.ones(9)
    kf, _ = po.rates(mech, T, c)
    assert kf[r] == pytest.approx(kinf * Pr / (1 + Pr) * Fc, rel=1e-10)


def test_source_conserves_mass_and_elements(mech, tables):
    ns = tables["n_spec"]
    E = tables["elem_comp"].reshape(ns, -1)
    for seed, T in ((1, 900.0), (2, 1600.0), (3, 2600.0)):
        rho, T, Y = _state(tables, T=T, seed=seed)
        om, _, _ = po.chem_source(mech, rho, T, Y)
        scale = np.abs(om).sum()
        assert abs(om.sum()) < 1e-13 * scale
        el = (om / tables["mw"]) @ E
        for e in range(E.shape[1]):
            assert abs(el[e]) < 1e-13 * np.abs(om / tables["mw"]).sum()


def test_epi_jacobian_matches_finite_difference(mech, tables):
    """d omega_m / d(rho Y_n) at frozen kf, kb (the reference's EPI Jacobian
    neglects dk/dT and third-body derivatives) against central differences of
    omega_m = W_m 1e3 sum_r (nu''_rm - nu'_rm)(q1_r - q2_r)."""
    ns = tables["n_spec"]
    sf = tables["stoi_f"].reshape(-1, ns)
    sb = tables["stoi_b"].reshape(-1, ns)
    W = tables["mw"]
    rho, T, Y = _state(tables, T=1500.0, seed=11)
    c = rho * Y / W * 1e-3
    kf, kb = po.rates(mech, T, c)
    _, J, _ = po.chem_source(mech, rho, T, Y)

    def omega_frozen(cc):
        _, q1, q2 = po.progress(mech, kf, kb, cc)
        return W * 1e3 * ((sb - sf).T @ (q1 - q2))

    for n in range(ns):
        h = c[n] * 1e-3  # omega is at most quadratic in each c_n: central differences are exact
        cp, cm = c.copy(), c.copy()
        cp[n] += h
        cm[n] -= h
        dw_dc = (omega_frozen(cp) - omega_frozen(cm)) / (2 * h)
        fd = dw_dc * 1e-3 / W[n]
        np.testing.assert_allclose(J[:, n], fd, rtol=1e-6, atol=1e-8 * np.abs(fd).max(), err_msg=f"column {n}")


def test_lu_solve_matches_numpy():
    rng = np.random.default_rng(3)
    for n in (3, 9):
        for _ in range(5):
            A = rng.normal(size=(n, n)) + n * np.eye(n) * rng.uniform(0, 1)
            b = rng.normal(size=n)
            x = po.lu_solve(A.ravel(), b)
            np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    # a matrix that needs row exchanges
    A = np.array([[0.0, 2.0, 1.0], [1.0, 1.0, 0.0], [3.0, 0.0, 1.0]])
    b = np.array([1.0, 2.0, 3.0])
    np.testing.assert_allclose(po.lu_solve(A.ravel(), b), np.linalg.solve(A, b), rtol=1e-14)


def test_newton_temperature_inverts_energy(mech, tables):
    rho, T0, Y = _state(tables, T=1900.0, seed=4)
    h, _, _ = nasa_h_cp_g(tables, T0)
    mwinv = (Y / tables["mw"]).sum()
    vel = np.array([120.0, -30.0, 5.0])
    e = np.dot(Y, h) - R_U * mwinv * T0
    rhoE = rho * (e + 0.5 * np.dot(vel, vel))
    T, p, it = po.temperature_newton(mech, rho, rhoE, vel, Y, 0.8 * T0)
    assert T == pytest.approx(T0, rel=2e-6)
    assert p == pytest.approx(rho * R_U * mwinv * T, rel=1e-14)
    assert 2 <= it < 20


def test_chem_batch_equals_dense_point_implicit_step(mech, tables):
    """One cell of the batch integrator against numpy: omega and J from the
    oracle's pinned point routines, (I - dt J) x = omega by numpy, rho Y +=
    dt x, then T from the unchanged internal energy by bisection."""
    rho_a, T_a, p_a, Y_a = cases.chem_batch_inputs(16, seed=77)
    dt = 1e-7
    dY, dT = po.chem_batch(mech, rho_a, T_a, Y_a, dt)
    W = tables["mw"]
    for i in range(16):
        rho, T, Y = rho_a[i], T_a[i], Y_a[:, i]
        om, J, _ = po.chem_source(mech, rho, T, Y)
        x = np.linalg.solve(np.eye(9) - dt * J, om)
        Yn = (rho * Y + dt * x) / rho
        np.testing.assert_allclose(dY[:, i], Yn - Y, rtol=1e-9, atol=1e-14)
        h0, _, _ = nasa_h_cp_g(tables, T)
        e0 = np.dot(Y, h0) - R_U * (Y / W).sum() * T

        def e(Tt):
            hh, _, _ = nasa_h_cp_g(tables, Tt)
            return np.dot(Yn, hh) - R_U * (Yn / W).sum() * Tt - e0

        lo, hi = T - 500.0, T + 500.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if e(mid) > 0:
                hi = mid
            else:
                lo = mid
        assert T + dT[i] == pytest.approx(0.5 * (lo + hi), rel=1e-6)


def test_transport_limits(mech, tables):
    """Pure species: Wilke gives mu_k = vis_coeff sqrt(T)/Omega and the
    conductivity rule reduces to mu (cp + 1.25 R/W).  Binary mixture:
    rho D_1 = rho W_2 D_12 / W_mix up to the 1e-12 regularisation."""
    T, p = 1200.0, 1e5
    cp = po.cp(mech, T)
    for k in (0, 2, 8):
        Y = np.zeros(9)
        Y[k] = 1.0
        rho = p * tables["mw"][k] / (R_U * T)
        mu, lam, rhoD = po.transport(mech, T, p, rho, Y)
        tdl = T * tables["LJ_potent_inv"][k]
        om = 1.147 * tdl**-0.145 + (tdl + 0.5) ** -2
        mu_k = tables["vis_coeff"][k] * np.sqrt(T) / om
        assert mu == pytest.approx(mu_k, rel=1e-13)
        assert lam == pytest.approx(mu_k * (cp[k] + 1.25 * R_U / tables["mw"][k]), rel=1e-13)
    # binary H2/N2
    Y = np.zeros(9)
    Y[0], Y[8] = 0.3, 0.7
    W = tables["mw"]
    mwmix = 1 / (Y / W).sum()
    rho = p * mwmix / (R_U * T)
    mu, lam, rhoD = po.transport(mech, T, p, rho, Y)
    kb = tables["kb_over_eps_jk"].reshape(9, 9)[0, 8]
    tr = T * kb
    omd = 1 / tr**0.145 + 1 / (tr + 0.5) ** 2
    D12 = tables["binary_diffusivity_coeff"].reshape(9, 9)[0, 8] * T * np.sqrt(T) / p / omd
    assert rhoD[0] == pytest.approx(rho * W[8] * D12 / mwmix, rel=1e-9)
    assert po.sutherland(273.0) == pytest.approx(1.716e-5, rel=1e-15)


# ----------------------------------------------------------------- fluxes

def _physical_flux(q, k, J):
    rho, u, v, w, p = q
    E = 0.5 * rho * (u * u + v * v + w * w) + p / 0.4
    U = k[0] * u + k[1] * v + k[2] * w
    return J * np.array([rho * U, rho * U * u + k[0] * p, rho * U * v + k[1] * p, rho * U * w + k[2] * p,
                         U * (E + p)])


def _face(scheme, ql, qr, k, J):
    f = np.zeros(5)
    met = np.array(list(k) + list(k), dtype=np.float64)
    jac = np.array([J, J], dtype=np.float64)
    po.lib().orc_face_flux_air(scheme, po._p(np.asarray(ql, float)), po._p(np.asarray(qr, float)), po._p(met),
                               po._p(jac), po._p(f))
    return f


@pytest.mark.parametrize("scheme", [1, 3, 4])
def test_flux_consistency(scheme):
    q = np.array([1.2, 150.0, -40.0, 25.0, 9.0e4])
    k = np.array([0.8, 0.3, -0.2])
    J = 2.5e-3
    np.testing.assert_allclose(_face(scheme, q, q, k, J), _physical_flux(q, k, J), rtol=1e-12,
                               atol=1e-12 * np.abs(_physical_flux(q, k, J)).max())


@pytest.mark.parametrize("scheme", [3, 4])
def test_supersonic_upwinding(scheme):
    ql = np.array([1.0, 900.0, 10.0, 0.0, 1.0e5])     # Mach ~2.6 along k
    qr = np.array([1.3, 850.0, -5.0, 2.0, 1.2e5])
    k = np.array([1.0, 0.0, 0.0])
    J = 1.0
    np.testing.assert_allclose(_face(scheme, ql, qr, k, J), _physical_flux(ql, k, J), rtol=1e-12)
    np.testing.assert_allclose(_face(scheme, qr * np.array([1, -1, -1, -1, 1]), ql * np.array([1, -1, -1, -1, 1]),
                                     k, J), _physical_flux(ql * np.array([1, -1, -1, -1, 1]), k, J), rtol=1e-12)


def _roe(ql, qr, k, J, delta=(0.0, 0.0)):
    f = np.zeros(5)
    met = np.array(list(k) + list(k), dtype=np.float64)
    jac = np.array([J, J], dtype=np.float64)
    d = np.asarray(delta, dtype=np.float64)
    po.lib().orc_face_flux_roe_air(po._p(np.asarray(ql, float)), po._p(np.asarray(qr, float)), po._p(met),
                                   po._p(jac), po._p(d), po._p(f))
    return f


def _cons(q):
    rho, u, v, w, p = q
    return np.array([rho, rho * u, rho * v, rho * w, 0.5 * rho * (u * u + v * v + w * w) + p / 0.4])


def _prim(Q):
    rho = Q[0]
    u, v, w = Q[1:4] / rho
    return np.array([rho, u, v, w, 0.4 * (Q[4] - 0.5 * rho * (u * u + v * v + w * w))])


def test_roe_consistency():
    q = np.array([1.2, 150.0, -40.0, 25.0, 9.0e4])
    k = np.array([0.8, 0.3, -0.2])
    J = 2.5e-3
    ref = _physical_flux(q, k, J)
    np.testing.assert_allclose(_roe(q, q, k, J, (30.0, 40.0)), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_roe_upwinds_supersonic_and_contact():
    k = np.array([1.0, 0.0, 0.0])
    ql = np.array([1.0, 900.0, 10.0, 0.0, 1.0e5])
    qr = np.array([1.3, 850.0, -5.0, 2.0, 1.2e5])
    np.testing.assert_allclose(_roe(ql, qr, k, 1.0), _physical_flux(ql, k, 1.0), rtol=1e-11)
    # a contact (equal u and p) moving right: exact upwinding
    ql = np.array([1.0, 80.0, 0.0, 0.0, 1.0e5])
    qr = np.array([0.4, 80.0, 0.0, 0.0, 1.0e5])
    np.testing.assert_allclose(_roe(ql, qr, k, 1.0), _physical_flux(ql, k, 1.0), rtol=1e-11)


def test_roe_dissipation_matches_eigen_decomposition():
    """1/2 (F_L + F_R) - 1/2 R |Lambda| R^-1 (Q_R - Q_L) with the flux
    Jacobian of the Roe-averaged state taken by finite differences and
    decomposed by numpy (entropy fix inactive)."""
    rng = np.random.default_rng(4)
    for _ in range(6):
        ql = np.array([1.0, 50.0, 20.0, -10.0, 1.0e5]) * (1 + 0.2 * rng.standard_normal(5))
        qr = np.array([1.1, 30.0, -10.0, 5.0, 1.1e5]) * (1 + 0.2 * rng.standard_normal(5))
        k = rng.standard_normal(3)
        rl, rr = np.sqrt(ql[0]), np.sqrt(qr[0])
        Hl = (_cons(ql)[4] + ql[4]) / ql[0]
        Hr = (_cons(qr)[4] + qr[4]) / qr[0]
        uh = (rl * ql[1:4] + rr * qr[1:4]) / (rl + rr)
        Hh = (rl * Hl + rr * Hr) / (rl + rr)
        rh = rl * rr
        ph = rh * 0.4 / 1.4 * (Hh - 0.5 * uh @ uh)
        Qh = _cons(np.array([rh, *uh, ph]))
        F = lambda Q: _physical_flux(_prim(Q), k, 1.0)
        A = np.zeros((5, 5))
        for c in range(5):
            h = 1e-6 * max(abs(Qh[c]), 1.0)
            e = np.zeros(5)
            e[c] = h
            A[:, c] = (F(Qh + e) - F(Qh - e)) / (2 * h)
        lam, R = np.linalg.eig(A)
        absA = (R * np.abs(lam)) @ np.linalg.inv(R)
        ref = 0.5 * (F(_cons(ql)) + F(_cons(qr))) - 0.5 * np.real(absA) @ (_cons(qr) - _cons(ql))
        np.testing.assert_allclose(_roe(ql, qr, k, 1.0), ref, rtol=1e-6, atol=1e-6 * np.abs(ref).max())


def test_roe_entropy_fix():
    """Harten's fix: |lambda| < delta is replaced by (lambda^2/delta + delta)/2,
    so a stationary contact gets dissipation delta/2 instead of 0."""
    k = np.array([1.0, 0.0, 0.0])
    ql = np.array([1.0, 0.0, 0.0, 0.0, 1.0e5])
    qr = np.array([0.5, 0.0, 0.0, 0.0, 1.0e5])
    f0 = _roe(ql, qr, k, 1.0)
    f1 = _roe(ql, qr, k, 1.0, (10.0, 30.0))   # face delta 20
    # mass flux: only the entropy wave carries the density jump; -1/2 * (delta/2) * d rho
    assert f0[0] == pytest.approx(0.0, abs=1e-9)
    assert f1[0] == pytest.approx(-0.5 * (20.0 / 2) * (qr[0] - ql[0]), rel=1e-12)


def _weno(v):
    vm, vp = C.c_double(), C.c_double()
    v = np.ascontiguousarray(v, dtype=np.float64)
    po.lib().orc_weno5_scalar(po._p(v), C.byref(vm), C.byref(vp))
    return vm.value, vp.value


def test_weno5_exact_on_quadratics_and_fifth_order():
    s = np.arange(-2, 4, dtype=float)  # points i-2..i+3, face at +0.5
    for a, b, c in ((1.0, 0.0, 0.0), (2.0, 3.0, 0.0), (0.5, -1.0, 0.25)):
        v = a + b * s + c * s * s
        vm, vp = _weno(v)
        # point values; the face value of the interpolant: a + b/2 + c/4
        exact = a + b * 0.5 + c * 0.25
        assert vm == pytest.approx(exact, rel=1e-13, abs=1e-13)
        assert vp == pytest.approx(exact, rel=1e-13, abs=1e-13)
    errs = []
    for h in (0.1, 0.05, 0.025):
        x = (s + 0.3) * h
        vm, _ = _weno(np.sin(x) + 2.0)
        errs.append(abs(vm - (np.sin(0.8 * h) + 2.0)))
    order = np.log2(errs[1] / errs[2])
    assert order > 4.5


def test_aweno_eigenvectors_are_inverse():
    L, R = np.zeros(25), np.zeros(25)
    cl = np.array([1.1, 1.1 * 300.0, 1.1 * 40.0, -1.1 * 20.0, 0.0, 9e4, 300.0])
    cr = np.array([0.9, 0.9 * 280.0, 0.9 * 35.0, -0.9 * 10.0, 0.0, 8e4, 310.0])
    for c in (cl, cr):
        c[4] = 0.5 * (c[1] ** 2 + c[2] ** 2 + c[3] ** 2) / c[0] + c[5] / 0.4
    met = np.array([0.6, 0.5, -0.3, 0.62, 0.48, -0.31])
    po.lib().orc_aweno_matrices(po._p(cl), po._p(cr), po._p(met), po._p(L), po._p(R))
    np.testing.assert_allclose(L.reshape(5, 5) @ R.reshape(5, 5), np.eye(5), atol=1e-13)


def test_ssprk3_stability_polynomial():
    a, b, c = np.zeros(3), np.zeros(3), np.zeros(3)
    po.lib().orc_ssprk3_coeffs(po._p(a), po._p(b), po._p(c))
    for z in (-0.3, -1.1, 0.2 + 0.7j):
        qn = 1.0 + 0j
        u = qn
        for r in range(3):
            u = a[r] * qn + b[r] * u + c[r] * z * u
        assert abs(u - (1 + z + z * z / 2 + z**3 / 6)) < 1e-14


@pytest.mark.parametrize("variant", ["muscl_ausm", "muscl_hllc", "first_lf", "aweno"])
def test_uniform_flow_preserved(variant):
    """Free stream on a Cartesian grid: every scheme's RHS vanishes."""
    import helpers
    if variant == "aweno":
        case = cases.c2_case(20, 14, 12, perturb=False)
        case["x"], case["y"], case["z"] = np.meshgrid(np.arange(12) * 0.01, np.arange(14) * 0.01,
                                                      np.arange(20) * 0.01, indexing="ij")[::-1]
    else:
        case = cases.c1_case(24, 16)
        case["perturb"] = None
        case["scheme"]["inviscid_scheme"] = {"muscl_ausm": 3, "muscl_hllc": 4, "first_lf": 1}[variant]
        case["scheme"]["reconstruction"] = 1 if variant == "first_lf" else 2
    for b in case["bcs"]:
        if b["type"] in (2, 3):
            b["type"] = 6
    z = helpers.oracle_zone(case)
    z.rhs()
    dq = z.field("dq").reshape(5, -1)
    st = case["inflow"]
    scale = st["rho"] * st["vel"] ** 2 * 0.01 ** 2 + st["p"] * 0.01 ** 2
    assert np.abs(dq).max() < 1e-12 * scale * 100


def test_cartesian_metrics_closed_form():
    mx, my, mz, g = 12, 10, 8, 2
    dx, dy, dz = 0.01, 0.02, 0.005
    Z, Y, X = np.meshgrid(np.arange(mz) * dz, np.arange(my) * dy, np.arange(mx) * dx, indexing="ij")
    n = (mx + 2 * g) * (my + 2 * g) * (mz + 2 * g)
    nc = (mx + 2 * g + 2) * (my + 2 * g + 2) * (mz + 2 * g + 2)
    xg, yg, zg, jac, met = (np.zeros(k) for k in (nc, nc, nc, n, 9 * n))
    po.lib().orc_grid_metrics(mx, my, mz, g, po._p(X.ravel().copy()), po._p(Y.ravel().copy()),
                              po._p(Z.ravel().copy()), po._p(xg), po._p(yg), po._p(zg), po._p(jac), po._p(met))
    np.testing.assert_allclose(jac, dx * dy * dz, rtol=1e-12)
    m = met.reshape(-1, 9)
    np.testing.assert_allclose(m[:, 0], 1 / dx, rtol=1e-12)
    np.testing.assert_allclose(m[:, 4], 1 / dy, rtol=1e-12)
    np.testing.assert_allclose(m[:, 8], 1 / dz, rtol=1e-12)
    np.testing.assert_allclose(m[:, [1, 2, 3, 5, 6, 7]], 0, atol=1e-9)


# ----------------------------------------------------------- boundaries

def test_farfield_and_subsonic_inflow_keep_a_uniform_far_state():
    """With the interior at the far state, the Riemann-invariant farfield and
    the total-pressure subsonic inflow reproduce that state in their ghost
    layers (the invariants and totals are consistent), and the back-pressure
    ghosts carry the prescribed pressure."""
    import helpers
    from ostc import cases
    case = cases.subsonic_case(20, 12, back_ratio=0.97)
    case["perturb"] = None
    z = helpers.oracle_zone(case)
    g = case["ngg"]
    bv = z.field("bv").reshape(6, 1 + 2 * g, 12 + 2 * g, 20 + 2 * g)
    far = case["inflow"]
    ref = [far["rho"], far["u"], far["v"], far["w"], far["p"], far["t"]]
    k = g
    # farfield ghosts (j < 0 and j >= my) and subsonic-inflow ghosts (i < 0)
    for l in range(6):
        scale = max(abs(ref[l]), 1.0)
        assert np.max(np.abs(bv[l, k, :g, g:-g] - ref[l])) < 1e-12 * scale
        assert np.max(np.abs(bv[l, k, -g:, g:-g] - ref[l])) < 1e-12 * scale
        assert np.max(np.abs(bv[l, k, g:-g, :g] - ref[l])) < 1e-10 * scale
    # back-pressure ghosts (i >= mx): pressure prescribed, the rest extrapolated
    assert np.allclose(bv[4, k, g:-g, -g:], 0.97 * far["p"], rtol=1e-14)
    assert np.allclose(bv[0, k, g:-g, -g:], far["rho"], rtol=1e-14)


@pytest.mark.parametrize("variant", ["air_aweno", "mixture_muscl_hllc"])
def test_couette_shear_dissipation(variant, mech):
    """Plane Couette shear u = a (y - y0), v = w = 0 at uniform rho, p, T (and
    composition) on a Cartesian grid.  Exact solution of the flux balance in
    the interior: no mass, momentum or species change (the inviscid face
    fluxes are (0, 0, p, 0, 0) on every j-face, the viscous stress tau_xy =
    mu a is uniform, no heat or species diffusion), and the viscous
    dissipation alone heats the gas: d(rho E)/dt = mu a^2, so the energy row of
    the right-hand side equals mu a^2 times the cell volume.  Pins the signs,
    indices and the u . tau term of the viscous flux (air and mixture
    transport), independent of the oracle's formulas."""
    import helpers
    d = 0.01 if variant == "air_aweno" else 1e-4
    if variant == "air_aweno":
        case = cases.c2_case(20, 14, 12, perturb=False)
        case["x"], case["y"], case["z"] = np.meshgrid(np.arange(12) * d, np.arange(14) * d, np.arange(20) * d,
                                                      indexing="ij")[::-1]
        ns, m = 0, None
        Y = np.zeros(0)
        mw = cases.MW_AIR
        T, p = 300.0, 2.0e4
    else:
        case = cases.c3_case(24, 16, dx=d, reconstruction=2)
        case["scheme"]["reaction"] = 0
        ns, m = 9, mech
        Y = np.zeros(ns)
        Y[2], Y[8] = 0.233, 0.767   # O2, N2
        mw = 1.0 / sum(Y[l] / mech.mw[l] for l in range(ns))
        T, p = 300.0, 101325.0
    for b in case["bcs"]:
        b["type"] = 6
    mx, my, mz, g = case["mx"], case["my"], case["mz"], case["ngg"]
    shp = cases.ghost_shape(mx, my, mz, g)
    rho = p * mw / (R_U * T)
    a = 2.0e3 if variant == "air_aweno" else 2.0e5       # 40 m/s across the zone
    yj = (np.arange(shp[1]) - g) * d
    bv = np.empty((6,) + shp)
    bv[0], bv[2], bv[3], bv[4], bv[5] = rho, 0.0, 0.0, p, T
    bv[1] = a * (yj - 0.5 * (my - 1) * d)[None, :, None]
    sv = np.empty((ns,) + shp)
    for l in range(ns):
        sv[l] = Y[l]
    z = helpers.oracle_zone(case, mech=m, bv=bv, sv=sv)
    z.rhs()
    nv = 5 + ns
    dq = z.field("dq").reshape(nv, mz, my, mx)
    mu = z.field("mul").reshape(shp)
    jac = z.field("jac").reshape(shp)
    inner = (slice(4, mz - 4) if mz > 1 else slice(0, 1), slice(4, my - 4), slice(4, mx - 4))
    ginner = tuple(slice(s.start + g, s.stop + g) for s in inner)
    mu_c, vol = mu[ginner], jac[ginner]
    assert np.ptp(mu_c) <= 1e-15 * mu_c.max()
    heat = mu_c * a * a * vol
    assert np.allclose(dq[4][inner], heat, rtol=1e-9, atol=0), (dq[4][inner] / heat).ravel()[:4]
    tol = 1e-9 * heat.max()
    for l in (0, 1, 2, 3) + tuple(range(5, nv)):
        assert np.abs(dq[l][inner]).max() < tol, (l, np.abs(dq[l][inner]).max(), tol)
