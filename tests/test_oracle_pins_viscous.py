"""Pins of the oracle's transport properties, viscous flux and SST source
against published formulas and closed forms, independent of the oracle's
code:

* mixture transport from the raw molecular weights and the species
  viscosities: Wilke's rule (Wilke 1950, J. Chem. Phys. 18:517) with
  Phi_ij = [1 + (mu_i/mu_j)^1/2 (W_j/W_i)^1/4]^2 / sqrt(8 (1 + W_i/W_j)), the
  Mason-Saxena conductivity (Mason and Saxena 1958, Phys. Fluids 1:361)
  lambda = sum X_i lambda_i / (X_i + 1.065 sum_{j != i} X_j Phi_ij) with
  Eucken's lambda_i = mu_i (cp_i + 5/4 R/W_i), and the mixture-averaged
  diffusivity rho D_k = rho (1 - Y_k) / sum_{l != k} X_l / D_kl
  (Bird, Stewart and Lightfoot; Kee et al.);
* heat conduction: a linear temperature slab at rest and uniform pressure
  (air and a 9-species mixture): the energy right-hand side is the discrete
  closed form dy a (lambda(T_{i+1}) - lambda(T_{i-1}))/2 with lambda(T) from the
  published formulas above (Sutherland and Pr for air);
* species diffusion: for a binary mixture with a linear mass-fraction slab
  at rest the mixture-averaged flux with its correction velocity reduces to
  Fick's law J_1 = rho D_12 dY_1/dx; the right-hand side converges at second
  order to d/dx (rho D_12 dY_1/dx) and the energy row to (h_1 - h_2) times it;
  for a 9-species composition the diffusive mass fluxes sum to zero;
* SST source terms (Menter 1994, AIAA J. 32:1598) on linear fields with
  analytic gradients near a wall: F1, F2, the cross-diffusion term, mu_t =
  a1 rho k / max(a1 omega, Omega F2), the blended constants, and both
  branches of the production switch of SST.cu:230-236.
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import oracle_zone
from ostc import cases

R_U = 8314.462618
GAMMA, MW_AIR = 1.4, 28.9647


def _omega_mu(ts):
    return 1.147 * ts ** -0.145 + (ts + 0.5) ** -2


def mixture_transport(t, T, p, Y, cp_k):
    """Published mixing rules from the raw W, the species viscosities
    mu_k = vis_coeff_k sqrt(T) / Omega(T eps_k^-1) and binary diffusivities
    D_kl = coeff_kl T^1.5 / (p Omega_D(T k/eps_kl)).  Returns mu, lambda, rho D."""
    W = t["mw"]
    ns = W.size
    Wbar = 1.0 / (Y / W).sum()
    X = Y * Wbar / W
    mu_k = t["vis_coeff"] * np.sqrt(T) / _omega_mu(T * t["LJ_potent_inv"])
    phi = np.ones((ns, ns))
    for i in range(ns):
        for j in range(ns):
            if i != j:
                phi[i, j] = (1 + np.sqrt(mu_k[i] / mu_k[j]) * (W[j] / W[i]) ** 0.25) ** 2 / np.sqrt(
                    8 * (1 + W[i] / W[j]))
    mu = sum(X[i] * mu_k[i] / (X @ phi[i]) for i in range(ns))
    lam_k = mu_k * (cp_k + 1.25 * R_U / W)
    lam = sum(X[i] * lam_k[i] / (X[i] + 1.065 * (X @ phi[i] - X[i])) for i in range(ns))
    kb = t["kb_over_eps_jk"].reshape(ns, ns)
    bc = t["binary_diffusivity_coeff"].reshape(ns, ns)
    rho = p * Wbar / (R_U * T)
    rhoD = np.zeros(ns)
    for k in range(ns):
        s = 0.0
        for l in range(ns):
            if l != k:
                tr = T * kb[k, l]
                D = bc[k, l] * T ** 1.5 / p / (tr ** -0.145 + (tr + 0.5) ** -2)
                s += X[l] / D
        rhoD[k] = rho * (1 - Y[k]) / s
    return mu, lam, rhoD


@pytest.mark.parametrize("T", [300.0, 1200.0, 2400.0])
def test_mixture_transport_matches_published_rules(mech, tables, T):
    rng = np.random.default_rng(int(T))
    Y = rng.random(9) + 0.05
    Y /= Y.sum()
    p = 2.0e5
    W = tables["mw"]
    rho = p / (R_U * T * (Y / W).sum())
    cp = po.cp(mech, T)
    mu, lam, rhoD = po.transport(mech, T, p, rho, Y)
    mu_r, lam_r, rhoD_r = mixture_transport(tables, T, p, Y, cp)
    assert mu == pytest.approx(mu_r, rel=1e-12)
    assert lam == pytest.approx(lam_r, rel=1e-12)
    np.testing.assert_allclose(rhoD, rhoD_r, rtol=1e-9)


def test_air_conductivity_is_mu_cp_over_pr():
    case = cases.c1_case(12, 10)
    case["scheme"]["viscous_order"] = 2
    z = oracle_zone(case)
    T = z.field("bv").reshape(6, -1)[5]
    mu = 1.716e-5 * (T / 273) ** 1.5 * (273 + 111) / (T + 111)
    cp = GAMMA * R_U / ((GAMMA - 1) * MW_AIR)
    g = case["ngg"]
    shp = cases.ghost_shape(12, 10, 1, g)
    inner = (g, slice(g - 1, g + 11), slice(g - 1, g + 13))
    np.testing.assert_allclose(z.field("mul").reshape(shp)[inner], mu.reshape(shp)[inner], rtol=1e-14)
    np.testing.assert_allclose(z.field("cond").reshape(shp)[inner], (mu * cp / 0.72).reshape(shp)[inner],
                               rtol=1e-14)


def _slab_case(variant, mech, n=24, ny=12):
    """2-D Cartesian slab at rest, every face an outflow."""
    if variant == "air":
        case = cases.c1_case(n, ny)
        case["scheme"]["viscous_order"] = 2
        m = None
        d = 1e-3
    else:
        d = 1e-4
        case = cases.c3_case(n, ny, dx=d, reconstruction=2)
        case["scheme"]["reaction"] = 0
        m = mech
    case["perturb"] = None
    for b in case["bcs"]:
        b["type"] = 6
    return case, m, d


@pytest.mark.parametrize("variant", ["air", "mixture"])
def test_linear_temperature_slab_heat_conduction(variant, mech, tables):
    case, m, d = _slab_case(variant, mech)
    mx, my, g = case["mx"], case["my"], case["ngg"]
    shp = cases.ghost_shape(mx, my, 1, g)
    p = 1.0e5
    a = 2.0e5 if variant == "air" else 2.0e6        # K/m: 200 K across the slab
    xs = (np.arange(shp[2]) - g) * d
    T = 700.0 + a * xs
    ns = 0 if variant == "air" else 9
    Y = np.zeros(ns)
    if ns:
        Y[2], Y[8] = 0.233, 0.767
    W = MW_AIR if not ns else 1.0 / (Y / tables["mw"]).sum()
    bv = np.zeros((6,) + shp)
    bv[5] = T[None, None, :]
    bv[4] = p
    bv[0] = p * W / (R_U * bv[5])
    sv = np.zeros((ns,) + shp)
    for l in range(ns):
        sv[l] = Y[l]
    z = oracle_zone(case, mech=m, bv=bv, sv=sv)
    z.rhs()
    nv = 5 + ns
    dq = z.field("dq").reshape(nv, 1, my, mx)
    Ti = 700.0 + a * np.arange(-1, mx + 1) * d       # T at i-1 .. i+1 of every interior cell

    def lam(Tv):
        if not ns:
            mu = 1.716e-5 * (Tv / 273) ** 1.5 * (273 + 111) / (Tv + 111)
            return mu * GAMMA * R_U / ((GAMMA - 1) * MW_AIR) / 0.72
        return np.array([mixture_transport(tables, t, p, Y, po.cp(mech, t))[1] for t in Tv])

    L = lam(Ti)
    ref = d * a * 0.5 * (L[2:] - L[:-2])               # dy * (lambda_{i+1/2} - lambda_{i-1/2}) * dT/dx
    cells = slice(4, mx - 4)
    got = dq[4, 0, 4:my - 4, cells]
    np.testing.assert_allclose(got, np.broadcast_to(ref[cells], got.shape), rtol=1e-9)
    tol = 1e-9 * np.abs(ref).max()
    for l in [0, 1, 2, 3] + list(range(5, nv)):
        assert np.abs(dq[l, 0, 4:my - 4, cells]).max() < max(tol, 1e-9 * p * d), l


def _binary_rhs(mech, tables, n):
    """H2/N2 slab with Y_H2 linear in x at rest, uniform T and p."""
    case, m, d = _slab_case("mixture", mech, n=n, ny=10)
    L = 2.4e-3
    d = L / n
    mx, my, g = case["mx"], case["my"], case["ngg"]
    xi = np.arange(mx) * d
    case["x"] = np.broadcast_to(xi, (1, my, mx)).copy()
    case["y"] = np.broadcast_to((np.arange(my) * d)[:, None], (1, my, mx)).copy()
    shp = cases.ghost_shape(mx, my, 1, g)
    xs = (np.arange(shp[2]) - g) * d
    a = 0.4 / L
    Y1 = 0.2 + a * xs
    T, p = 500.0, 1.0e5
    W = tables["mw"]
    sv = np.zeros((9,) + shp)
    sv[0] = Y1[None, None, :]
    sv[8] = 1 - sv[0]
    Wbar = 1.0 / (sv[0] / W[0] + sv[8] / W[8])
    bv = np.zeros((6,) + shp)
    bv[5], bv[4] = T, p
    bv[0] = p * Wbar / (R_U * T)
    z = oracle_zone(case, mech=m, bv=bv, sv=sv)
    z.rhs()
    dq = z.field("dq").reshape(14, 1, my, mx)
    jac = z.field("jac").reshape(shp)[g, g:g + my, g:g + mx]
    # d/dx (rho D12 dY1/dx) with rho D12 = Wbar p D12 / (R T), dWbar/dx = -Wbar^2 (1/W1 - 1/W2) a
    kb = tables["kb_over_eps_jk"].reshape(9, 9)[0, 8]
    tr = T * kb
    D12 = tables["binary_diffusivity_coeff"].reshape(9, 9)[0, 8] * T ** 1.5 / p / (tr ** -0.145 + (tr + 0.5) ** -2)
    y1 = 0.2 + a * xi
    wb = 1.0 / (y1 / W[0] + (1 - y1) / W[8])
    div = p * D12 / (R_U * T) * a * (-wb * wb * (1 / W[0] - 1 / W[8]) * a)
    h = po.enthalpy(mech, T)
    return dq / jac[None, None], div, (h[0] - h[8]) * div


def test_binary_diffusion_converges_to_ficks_law(mech, tables):
    errs, errs_e = [], []
    for n in (24, 48):
        num, div, ediv = _binary_rhs(mech, tables, n)
        cells = slice(n // 4, 3 * n // 4)       # the same physical interval on both grids
        errs.append(np.abs(num[5, 0, 5, cells] - div[cells]).max() / np.abs(div).max())
        errs_e.append(np.abs(num[4, 0, 5, cells] - ediv[cells]).max() / np.abs(ediv).max())
        # N2 carries the opposite flux, the mass row is untouched
        assert np.abs(num[5, 0, 5, cells] + num[13, 0, 5, cells]).max() < 1e-8 * np.abs(div).max()
        assert np.abs(num[0, 0, 5, cells]).max() < 1e-9 * np.abs(div).max()
    assert errs[-1] < 1e-3 and errs_e[-1] < 1e-3, (errs, errs_e)
    assert 1.7 < np.log2(errs[0] / errs[1]) < 2.3, errs
    assert 1.7 < np.log2(errs_e[0] / errs_e[1]) < 2.3, errs_e


def test_multispecies_diffusion_conserves_mass(mech, tables):
    """A smooth 9-species composition at rest, uniform T and p: the
    correction velocity makes the diffusive fluxes sum to zero, so the mass
    row and the sum of the species rows vanish while each species row does
    not."""
    case, m, d = _slab_case("mixture", mech, n=20, ny=16)
    mx, my, g = case["mx"], case["my"], case["ngg"]
    shp = cases.ghost_shape(mx, my, 1, g)
    X = (np.arange(shp[2]) - g)[None, :] / mx
    Yy = (np.arange(shp[1]) - g)[:, None] / my
    W = tables["mw"]
    sv = np.zeros((9,) + shp)
    for l in range(9):
        sv[l, 0] = 0.05 + 0.04 * np.sin(2.0 * X * (l + 1) + 3.0 * Yy * (9 - l)) + 0.2 * (l == 8)
    sv[:, 1:] = sv[:, :1]
    sv /= sv.sum(axis=0)
    T, p = 1100.0, 1.0e5
    bv = np.zeros((6,) + shp)
    bv[5], bv[4] = T, p
    bv[0] = p / (R_U * T * np.tensordot(1 / W, sv, axes=1))
    z = oracle_zone(case, mech=m, bv=bv, sv=sv)
    z.rhs()
    dq = z.field("dq").reshape(14, 1, my, mx)[:, 0, 4:my - 4, 4:mx - 4]
    scale = np.abs(dq[5:]).max()
    assert scale > 0
    assert np.abs(dq[5:].sum(axis=0)).max() < 1e-12 * 9 * scale
    assert np.abs(dq[0]).max() < 1e-12 * scale


# ------------------------------------------------------------------- SST

SIG_W2, BETA_STAR, A1 = 0.856, 0.09, 0.31
BETA1, BETA2, KAPPA = 0.075, 0.0828, 0.41
SIG_W1 = 0.5


def _menter_source(rho, mu, k, w, d, grads):
    """Menter (1994) SST source terms of one point with the production switch
    of SST.cu:230-236 (P replaced by mu_t |Omega|^2 when it exceeds ten times
    that).  grads: du/dx_j (3x3, row = velocity component), grad k, grad w."""
    G, gk, gw = grads
    vort = np.sqrt((G[1, 0] - G[0, 1]) ** 2 + (G[2, 0] - G[0, 2]) ** 2 + (G[2, 1] - G[1, 2]) ** 2)
    nu = mu / rho
    arg2 = max(2 * np.sqrt(k) / (BETA_STAR * w * d), 500 * nu / (d * d * w))
    F2 = np.tanh(arg2 ** 2)
    cross = 2 * rho * SIG_W2 / w * (gk @ gw)
    CD = max(cross, 1e-20)
    arg1 = min(max(np.sqrt(k) / (BETA_STAR * w * d), 500 * nu / (d * d * w)), 4 * rho * SIG_W2 * k / (CD * d * d))
    F1 = np.tanh(arg1 ** 4)
    mut = A1 * rho * k / max(A1 * w, vort * F2)
    S = 0.5 * (G + G.T)
    div = np.trace(G)
    P = mut * (2 * np.sum(S * S) - 2.0 / 3 * div * div) - 2.0 / 3 * rho * k * div
    if P > 10 * mut * vort * vort:
        P = mut * vort * vort
    g1 = BETA1 / BETA_STAR - SIG_W1 * KAPPA ** 2 / np.sqrt(BETA_STAR)
    g2 = BETA2 / BETA_STAR - SIG_W2 * KAPPA ** 2 / np.sqrt(BETA_STAR)
    beta = F1 * BETA1 + (1 - F1) * BETA2
    gam = F1 * g1 + (1 - F1) * g2
    sk = P - BETA_STAR * rho * k * w
    sw = gam * rho / mut * P + (1 - F1) * cross - beta * rho * w * w
    return sk, sw, mut, F1, F2


@pytest.mark.parametrize("flow", ["shear", "strain"])
def test_sst_source_on_linear_fields(flow):
    """Linear u, v, k, omega on a Cartesian grid with a wall along j = 0:
    central differences are exact, the wall distance is y, so every term of
    the source has a closed form.  'shear' keeps the production below ten
    times mu_t Omega^2, 'strain' (almost irrotational) triggers the switch."""
    nx, ny, dxy = 16, 14, 1e-4
    case = cases.sst_plate_case(nx, ny, compressibility=0)
    X, Y = np.meshgrid(np.arange(nx) * dxy, np.arange(ny) * dxy, indexing="xy")
    case["x"], case["y"], case["z"] = X[None].copy(), Y[None].copy(), np.zeros((1, ny, nx))
    case["bcs"] = [b for b in case["bcs"] if b["type"] != 3]
    for b in case["bcs"]:
        if b["type"] == 2:
            b["raw"] = (0, nx - 1, 0, 0, 0, 0)      # the whole j = 0 side is wall: d = y
        else:
            b["type"] = 6
    g = case["ngg"]
    shp = cases.ghost_shape(nx, ny, 1, g)
    xs = (np.arange(shp[2]) - g) * dxy
    ys = (np.arange(shp[1]) - g) * dxy
    Xg, Yg = np.meshgrid(xs, ys, indexing="xy")
    T, p = 300.0, 1.0e5
    rho = p * MW_AIR / (R_U * T)
    if flow == "shear":
        ux, uy, vx, vy = 2.0e3, 4.0e4, -1.0e3, -1.5e3
    else:
        ux, uy, vx, vy = 3.0e4, 2.0e2, 1.9e2, -2.9e4
    u = 100.0 + ux * Xg + uy * Yg
    v = vx * Xg + vy * Yg
    kk = 20.0 + 5.0e3 * Xg + 2.0e4 * Yg
    ww = 3.0e4 + 4.0e7 * Yg + 1.0e7 * Xg
    bv = np.zeros((6,) + shp)
    bv[0], bv[1], bv[2], bv[4], bv[5] = rho, u, v, p, T
    sv = np.zeros((2,) + shp)
    sv[0], sv[1] = kk, ww
    for b in case["bcs"]:
        if b["type"] == 5:
            b["state"] = case["inflow"]
    z = oracle_zone(case, bv=bv, sv=sv)
    # the initialisation's boundary conditions rewrote ghosts and the wall
    # row: put the linear fields back on every point before the source
    zb = z.field("bv").reshape((6,) + shp)
    zs = z.field("sv").reshape((2,) + shp)
    zb[:] = bv
    zs[:] = sv
    z.source()
    dq = z.field("dq").reshape(7, 1, ny, nx)
    jac = z.field("jac").reshape(shp)[g, g:g + ny, g:g + nx]
    mut_f = z.field("mut").reshape(shp)[g, g:g + ny, g:g + nx]
    wd = z.field("wall_distance").reshape(shp)[g, g:g + ny, g:g + nx]
    mu = 1.716e-5 * (T / 273) ** 1.5 * (273 + 111) / (T + 111)
    G = np.array([[ux, uy, 0.0], [vx, vy, 0.0], [0.0, 0.0, 0.0]])
    f1s, switched = [], []
    for j in range(2, ny - 2):
        for i in range(2, nx - 2):
            x, y = i * dxy, j * dxy
            assert wd[j, i] == pytest.approx(y, rel=1e-12)
            kv, wv = 20.0 + 5.0e3 * x + 2.0e4 * y, 3.0e4 + 4.0e7 * y + 1.0e7 * x
            sk, sw, mut, F1, F2 = _menter_source(rho, mu, kv, wv, y, (G, np.array([5.0e3, 2.0e4, 0.0]),
                                                                      np.array([1.0e7, 4.0e7, 0.0])))
            f1s.append(F1)
            vort = abs(vx - uy)
            S = 0.5 * (G + G.T)
            P = mut * (2 * np.sum(S * S) - 2.0 / 3 * np.trace(G) ** 2) - 2.0 / 3 * rho * kv * np.trace(G)
            switched.append(P > 10 * mut * vort * vort)
            assert mut_f[j, i] == pytest.approx(mut, rel=1e-11)
            assert dq[5, 0, j, i] == pytest.approx(sk * jac[j, i], rel=1e-10, abs=1e-12 * abs(sk * jac[j, i]))
            assert dq[6, 0, j, i] == pytest.approx(sw * jac[j, i], rel=1e-10)
    f1s = np.array(f1s)
    # the blending is exercised away from its limits, and the switch on the intended branch
    assert ((f1s > 0.05) & (f1s < 0.95)).mean() > 0.2, f1s
    assert (np.mean(switched) > 0.5) if flow == "strain" else not any(switched), np.mean(switched)
