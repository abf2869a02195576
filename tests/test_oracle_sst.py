"""Pins of the oracle's k-omega SST model (oracle/orc_flow.c) against what
the mathematics fixes, independent of the oracle's code:

* the derived constants gamma1, gamma2 against Menter's published values;
* the compressibility corrections: zero below their threshold Mach numbers,
  continuous at it, Zeman's saturating at 0.75 beta*, Wilcox's quadratic;
* the wall distance on a rectilinear grid: |y| above the wall, the distance
  to the end node beyond its ends;
* decay of uniform free-stream turbulence: with no gradients and no wall,
  mu_t = rho k / omega, dq_k = -beta* rho k omega J, dq_omega = -beta2 rho
  omega^2 J (F1 -> 0 far from any wall), and the flow rows stay zero;
* the wall value omega_w = 60 mu_w / (rho_w beta1 d1^2) and k_w = 0;
* plane Couette shear: Menter's production mu_t |vorticity|^2, the k-epsilon
  constants far from walls and the (mu + mu_t) a^2 dissipation in closed form;
* the laminar limit: with k = 0 the flow rows of the turbulent right-hand
  side equal the laminar ones bit for bit (mu_t = 0, lambda_t = 0, no
  -2/3 rho k), and the k row vanishes.
"""
import numpy as np

import pyoracle as po
from ostc import cases
from helpers import oracle_zone

GAMMA, R_U, MW = 1.4, 8314.462618, 28.9647


def _zone(case, bv, sv):
    for b in case["bcs"]:
        if b["type"] == 5:
            b["state"] = case["inflow"]
    return oracle_zone(case, bv=bv, sv=sv)


def _shape(case):
    return cases.ghost_shape(case["mx"], case["my"], case["mz"], case["ngg"])


def test_sst_constants_match_menter():
    g1, g2 = (po.C.c_double(), po.C.c_double())
    po.lib().orc_sst_constants(po.C.byref(g1), po.C.byref(g2))
    # Menter (1994) quotes gamma1 = 0.5532, gamma2 = 0.4403 (4 decimals)
    assert abs(g1.value - 0.5532) < 1e-4
    assert abs(g2.value - 0.4403) < 1e-4


def test_compressibility_corrections():
    L = po.lib()
    gp1 = 2.4
    mt0 = 0.25 * np.sqrt(2.0 / gp1)
    assert L.orc_sst_zeman(0.0, gp1) == 0.0 and L.orc_sst_zeman(mt0, gp1) == 0.0
    assert L.orc_sst_zeman(mt0 + 1e-6, gp1) < 1e-12
    assert abs(L.orc_sst_zeman(50.0, gp1) - 0.75 * 0.09) < 1e-12
    zs = [L.orc_sst_zeman(m, gp1) for m in np.linspace(mt0, 3, 50)]
    assert np.all(np.diff(zs) >= 0)
    assert L.orc_sst_wilcox(0.25) == 0.0 and L.orc_sst_wilcox(0.1) == 0.0
    # quadratic above the threshold: second difference constant
    w = np.array([L.orc_sst_wilcox(m) for m in (0.5, 0.6, 0.7)])
    assert abs((w[2] - 2 * w[1] + w[0]) - 2 * 0.09 * 2 * 0.01) < 1e-15


def test_wall_distance_rectilinear():
    case = cases.sst_plate_case(nx=24, ny=16, le=6)
    bv, sv = cases.initial_fields(case)
    z = _zone(case, bv, sv)
    g = case["ngg"]
    d = z.field("wall_distance").reshape(_shape(case))[g]
    xs, ys = case["x"][0, 0], case["y"][0, :, 0]
    dx = xs[1] - xs[0]
    xg = np.concatenate([xs[0] - dx * np.arange(g, 0, -1), xs, xs[-1] + dx * np.arange(1, g + 1)])
    # reflected ghost nodes in y
    yg = np.concatenate([-ys[g:0:-1], ys, 2 * ys[-1] - ys[-2:-g - 2:-1]])
    x_le, x_end = xs[case["le"]], xs[-1]
    X, Y = np.meshgrid(xg, yg, indexing="xy")
    ref = np.where(X < x_le, np.hypot(X - x_le, Y), np.where(X > x_end, np.hypot(X - x_end, Y), np.abs(Y)))
    np.testing.assert_allclose(d, ref, rtol=1e-13, atol=1e-15)


def test_uniform_turbulence_decay():
    case = cases.sst_plate_case(nx=20, ny=14, compressibility=0)
    case["bcs"][-1] = dict(type=3, raw=case["bcs"][-1]["raw"], label=3)  # no wall anywhere
    bv, sv = cases.initial_fields(dict(case, perturb=None))
    z = _zone(case, bv, sv)
    z.rhs()
    st = case["inflow"]
    rho, k, w = st["rho"], st["sv"][0], st["sv"][1]
    mx, my = case["mx"], case["my"]
    g = case["ngg"]
    J = z.field("jac").reshape(_shape(case))[g, g:g + my, g:g + mx]
    dq = z.field("dq").reshape(7, my, mx)
    mut = z.field("mut").reshape(_shape(case))[g, g:g + my, g:g + mx]
    np.testing.assert_allclose(mut, rho * k / w, rtol=1e-14)
    np.testing.assert_allclose(dq[5], -0.09 * rho * k * w * J, rtol=1e-10)
    np.testing.assert_allclose(dq[6], -0.0828 * rho * w * w * J, rtol=1e-10)
    scale = np.max(np.abs(dq[5]))
    assert np.max(np.abs(dq[:5])) < 1e-9 * scale


def test_wall_values():
    case = cases.sst_plate_case(nx=24, ny=16, le=6)
    bv, sv = cases.sst_plate_fields(case)
    z = _zone(case, bv, sv)
    g, le = case["ngg"], case["le"]
    shp = _shape(case)
    B = z.field("bv").reshape((6,) + shp)[:, g]
    S = z.field("sv").reshape((2,) + shp)[:, g]
    tw = case["bcs"][-1]["t_wall"]
    d1 = case["y"][0, 1, le:]
    p = B[4, g + 1, g + le:g + case["mx"]]
    rho_w = p * MW / (R_U * tw)
    mu_w = cases.sutherland(tw)
    np.testing.assert_allclose(S[0, g, g + le:g + case["mx"]], 0.0, atol=0)
    np.testing.assert_allclose(S[1, g, g + le:g + case["mx"]], 60 * mu_w / (rho_w * 0.075 * d1 ** 2), rtol=1e-13)
    # wall ghosts: k = 0, omega copied from the wall point
    for gg in range(1, g + 1):
        np.testing.assert_array_equal(S[0, g - gg, g + le:g + case["mx"]], 0.0)
        np.testing.assert_array_equal(S[1, g - gg, g + le:g + case["mx"]], S[1, g, g + le:g + case["mx"]])


def test_laminar_limit():
    case = cases.sst_plate_case(nx=24, ny=16, le=6)
    bv, sv = cases.sst_plate_fields(case)
    sv[0] = 0.0  # k = 0 everywhere
    st = dict(case["inflow"], sv=[0.0, case["inflow"]["sv"][1]])
    case["inflow"] = st
    zt = _zone(case, bv, sv)
    lam = cases.sst_plate_case(nx=24, ny=16, le=6)
    lam["scheme"] = dict(lam["scheme"], turbulence=0)
    lam["inflow"] = dict(st, sv=[])
    zl = _zone(lam, bv, np.zeros((0,) + bv.shape[1:]))
    zt.rhs()
    zl.rhs()
    mx, my = case["mx"], case["my"]
    dqt = zt.field("dq").reshape(7, my, mx)
    dql = zl.field("dq").reshape(5, my, mx)
    np.testing.assert_array_equal(dqt[:5], dql)
    g = case["ngg"]
    # mu_t vanishes wherever it was computed ([-1, m] at initialisation, the interior by the source)
    assert not zt.field("mut").reshape(_shape(case))[g, g - 1:g + my + 1, g - 1:g + mx + 1].any()
    np.testing.assert_array_equal(dqt[5], 0.0)


def test_couette_shear_production():
    """Plane Couette shear u = a (y - y0) with uniform rho, p, T, k, omega on
    a uniform Cartesian grid and no wall (F1, F2 -> 0, so the k-epsilon set
    gamma2, beta2 applies and mu_t = rho k / omega).  |vorticity| = a, so
    Menter's production is P_k = mu_t a^2 (below its 10 beta* rho k omega
    limit here) and, with grad k = 0 (no diffusion, no cross-diffusion):
    dq_k = (mu_t a^2 - beta* rho k omega) J and dq_omega = (gamma2 rho a^2 -
    beta2 rho omega^2) J, gamma2 = beta2 / beta* - sigma_w2 kappa^2 /
    sqrt(beta*) (Menter 1994, set 2).  The total stress (mu + mu_t) a is
    uniform, so the momentum rows vanish, and the dissipation heats the gas:
    dq_E = (mu + mu_t) a^2 J.  Pins the production, the blended constants, the
    eddy-viscosity stress and the u . tau work of the SST path."""
    case = cases.sst_plate_case(nx=20, ny=16, compressibility=3)
    for b in case["bcs"]:
        b["type"] = 6
    mx, my, g, d = case["mx"], case["my"], case["ngg"], 0.01
    X, Y = np.meshgrid(np.arange(mx) * d, np.arange(my) * d, indexing="xy")
    case["x"], case["y"], case["z"] = X[None].copy(), Y[None].copy(), np.zeros((1, my, mx))
    shp = _shape(case)
    T, p, a, k, w = 300.0, 2.0e4, 2.0e3, 1.0, 1.0e4
    rho = p * MW / (R_U * T)
    bv = np.empty((6,) + shp)
    bv[0], bv[2], bv[3], bv[4], bv[5] = rho, 0.0, 0.0, p, T
    bv[1] = a * ((np.arange(shp[1]) - g) * d - 0.5 * (my - 1) * d)[None, :, None]
    sv = np.empty((2,) + shp)
    sv[0], sv[1] = k, w
    z = _zone(case, bv, sv)
    z.rhs()
    inner = (slice(4, my - 4), slice(4, mx - 4))
    gin = (g,) + tuple(slice(s.start + g, s.stop + g) for s in inner)
    J = z.field("jac").reshape(shp)[gin]
    mut = z.field("mut").reshape(shp)[gin]
    mu = cases.sutherland(T)
    dq = z.field("dq").reshape(7, my, mx)[(slice(None),) + inner]
    np.testing.assert_allclose(mut, rho * k / w, rtol=1e-12)
    beta_s, beta2, sigma_w2, kappa = 0.09, 0.0828, 0.856, 0.41
    gamma2 = beta2 / beta_s - sigma_w2 * kappa**2 / np.sqrt(beta_s)
    mut0 = rho * k / w
    assert mut0 * a * a < 10 * beta_s * rho * k * w        # production limiter inactive
    np.testing.assert_allclose(dq[5], (mut0 * a * a - beta_s * rho * k * w) * J, rtol=1e-9)
    np.testing.assert_allclose(dq[6], (gamma2 * rho * a * a - beta2 * rho * w * w) * J, rtol=1e-9)
    heat = (mu + mut0) * a * a * J
    np.testing.assert_allclose(dq[4], heat, rtol=1e-9)
    assert np.abs(dq[:4]).max() < 1e-9 * heat.max()
