"""Pins of the oracle's step-level routines against closed forms, independent
of the oracle's code:

* DPLUR's Jacobian products (A + lambda I) dq (DPLUR.cuh:89-148) against
  central finite differences of the metric-form Euler flux, for air and for
  the 9-species mixture (pressure from the NASA-7 energy inverted to
  round-off), species rows included;
* wall and symmetry ghosts on the curvilinear 3-D grid: about the boundary
  node, each ghost and its mirror interior point average to the wall values
  (u = 0 and T = T_w for the isothermal wall, equal pressure), and to zero
  normal velocity with equal tangential velocity, density, pressure and
  temperature at a symmetry plane whose normal is not a grid axis;
* the CFL time step on a uniform state: dt = CFL / (sum_i (|u_i| + a)/dx_i
  + max(gamma, 4/3)/rho (mu/Pr) sum_i 1/dx_i^2) in 2-D and 3-D, for air
  and for a mixture (gamma and a from its NASA-7 cp);
* the residual of a hand-built field: sqrt(sum (q - q_last)^2 / N) per
  variable (rho, |V|, p, T), scaled by the values of step 1.
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import oracle_zone
from ostc import cases

R_U = 8314.462618
GAMMA, MW_AIR = 1.4, 28.9647


# ------------------------------------------------------------------ DPLUR

def _air_flux(Q, m):
    rho, mom, rE = Q[0], Q[1:4], Q[4]
    u = mom / rho
    p = (GAMMA - 1) * (rE - 0.5 * mom @ mom / rho)
    U = m @ u
    return np.array([rho * U, *(mom * U + m * p), U * (rE + p)])


def _mix_T(mech, W, rY, e_int, T0):
    """Newton to round-off on sum rho Y h(T) - R T sum rho Y / W = e_int."""
    T = T0
    for _ in range(60):
        h, cp = po.enthalpy_cp(mech, T)
        f = rY @ h - R_U * T * (rY / W).sum() - e_int
        fp = rY @ cp - R_U * (rY / W).sum()
        dT = f / fp
        T -= dT
        if abs(dT) < 1e-13 * T:
            break
    return T


def _mix_flux(Q, m, mech, W, T0):
    rho, mom, rE, rY = Q[0], Q[1:4], Q[4], Q[5:]
    u = mom / rho
    T = _mix_T(mech, W, rY, rE - 0.5 * mom @ mom / rho, T0)
    p = R_U * T * (rY / W).sum()
    U = m @ u
    return np.array([rho * U, *(mom * U + m * p), U * (rE + p), *(rY * U)])


@pytest.mark.parametrize("lam", [0.0, 450.0, -450.0])
def test_dplur_air_jacobian_product_matches_finite_differences(lam):
    rng = np.random.default_rng(3)
    m = np.array([120.0, -35.0, 60.0])
    rho, vel, p = 0.8, np.array([420.0, -90.0, 35.0]), 7.0e4
    T = p * MW_AIR / (rho * R_U)
    Q = np.array([rho, *(rho * vel), 0.5 * rho * vel @ vel + p / (GAMMA - 1)])
    for _ in range(4):
        dq = Q * rng.standard_normal(5) * 0.3
        got = po.jac_times_dq([rho, *vel, p, T], [], m, lam, dq)
        eps = 1e-6
        fd = (_air_flux(Q + eps * dq, m) - _air_flux(Q - eps * dq, m)) / (2 * eps) + lam * dq
        np.testing.assert_allclose(got, fd, rtol=1e-7, atol=1e-8 * np.abs(fd).max())


def test_dplur_mixture_jacobian_product_matches_finite_differences(mech, tables):
    rng = np.random.default_rng(5)
    W = tables["mw"]
    m = np.array([-50.0, 140.0, 20.0])
    Y = rng.random(9) + 0.02
    Y /= Y.sum()
    T, p = 1300.0, 1.5e5
    rho = p / (R_U * T * (Y / W).sum())
    vel = np.array([250.0, 60.0, -30.0])
    h, cp_k = po.enthalpy_cp(mech, T)
    rE = rho * (Y @ h) - p + 0.5 * rho * vel @ vel
    Q = np.array([rho, *(rho * vel), rE, *(rho * Y)])
    cp = Y @ cp_k
    gam = cp / (cp - R_U * (Y / W).sum())
    for lam in (0.0, 700.0):
        dq = Q * (0.2 * rng.standard_normal(14))
        got = po.jac_times_dq([rho, *vel, p, T], Y, m, lam, dq, m=mech, gam=gam)
        eps = 1e-6
        fd = (_mix_flux(Q + eps * dq, m, mech, W, T) - _mix_flux(Q - eps * dq, m, mech, W, T)) / (2 * eps) + lam * dq
        np.testing.assert_allclose(got, fd, rtol=2e-7, atol=2e-8 * np.abs(fd).max())


# ---------------------------------------------------------- wall / symmetry

def test_wall_and_symmetry_ghosts_mirror_about_the_boundary_node():
    case = cases.c2_case(20, 14, 12)
    case["z"] = case["z"] + 0.3 * case["x"]      # tilt the k-planes: the symmetry normal leaves the z axis
    g = case["ngg"]
    mx, my, mz = case["mx"], case["my"], case["mz"]
    z = oracle_zone(case)          # seeded density noise, boundary conditions applied
    shp = cases.ghost_shape(mx, my, mz, g)
    bv = z.field("bv").reshape((6,) + shp)
    met = z.field("metric").reshape(shp + (9,))
    T_w = 300.0
    # isothermal wall on j = 0: pairs j = -l, +l about the wall node
    I, K = slice(g + 1, g + mx - 1), slice(g + 1, g + mz - 1)
    for l in range(1, g + 1):
        gh, it = bv[:, K, g - l, I], bv[:, K, g + l, I]
        assert np.abs(0.5 * (gh[1:4] + it[1:4])).max() < 1e-12 * np.abs(it[1:4]).max()
        clipped = 2 * T_w - it[5] <= 0.1 * T_w
        assert not clipped.any()
        np.testing.assert_allclose(0.5 * (gh[5] + it[5]), T_w, rtol=1e-14)
        np.testing.assert_allclose(gh[4], it[4], rtol=0)
    # the wall node itself carries the wall values
    assert np.abs(bv[1:4, K, g, I]).max() == 0.0
    np.testing.assert_allclose(bv[5, K, g, I], T_w, rtol=0)
    # symmetry on k = 0: the normal is the zeta-metric row of the boundary node
    J, I2 = slice(g + 1, g + my - 1), slice(g + 1, g + mx - 1)
    n = met[g, J, I2][..., 6:9]
    n = n / np.linalg.norm(n, axis=-1, keepdims=True)
    assert np.abs(n[..., 0]).max() > 1e-2          # not a grid axis
    for l in range(1, g + 1):
        gh, it = bv[:, g - l, J, I2], bv[:, g + l, J, I2]
        ug, ui = np.moveaxis(gh[1:4], 0, -1), np.moveaxis(it[1:4], 0, -1)
        un_g, un_i = (ug * n).sum(-1), (ui * n).sum(-1)
        assert np.abs(0.5 * (un_g + un_i)).max() < 1e-12 * np.abs(ui).max()
        np.testing.assert_allclose(ug - un_g[..., None] * n, ui - un_i[..., None] * n, rtol=0,
                                   atol=1e-12 * np.abs(ui).max())
        for c in (0, 4, 5):
            np.testing.assert_allclose(gh[c], it[c], rtol=0)


# ---------------------------------------------------------------- CFL dt

@pytest.mark.parametrize("dim", [2, 3])
def test_cfl_time_step_closed_form_air(dim):
    dx, dy, dz = 2e-3, 1e-3, 4e-3
    if dim == 3:
        nx, ny, nz = 10, 9, 8
        case = cases.c2_case(nx, ny, nz, perturb=False)
        Z, Y, X = np.meshgrid(np.arange(nz) * dz, np.arange(ny) * dy, np.arange(nx) * dx, indexing="ij")
        case["x"], case["y"], case["z"] = X, Y, Z
        steps = (dx, dy, dz)
    else:
        nx, ny = 12, 10
        case = cases.c1_case(nx, ny)
        case["perturb"] = None
        Y, X = np.meshgrid(np.arange(ny) * dy, np.arange(nx) * dx, indexing="ij")
        case["x"], case["y"] = X[None], Y[None]
        steps = (dx, dy)
    for b in case["bcs"]:
        b["type"] = 6
    z = oracle_zone(case)
    st = case["inflow"]
    rho, T = st["rho"], st["t"]
    vel = np.array([st["u"], st["v"], st["w"]])
    a = np.sqrt(GAMMA * R_U / MW_AIR * T)
    mu = 1.716e-5 * (T / 273) ** 1.5 * (273 + 111) / (T + 111)
    inv = sum((abs(vel[c]) + a) / steps[c] for c in range(dim))
    visc = max(GAMMA, 4.0 / 3.0) / rho * mu / case["scheme"].get("prandtl_number", 0.72) * sum(
        1 / s ** 2 for s in steps)
    ref = 0.5 / (inv + visc)
    assert z.global_dt() == pytest.approx(ref, rel=1e-13)
    np.testing.assert_allclose(z.field("dt_local"), ref, rtol=1e-13)


def test_cfl_time_step_closed_form_mixture(mech, tables):
    d = 1e-4
    case = cases.c3_case(12, 10, dx=d, reconstruction=2)
    case["scheme"]["reaction"] = 0
    case["perturb"] = None
    for b in case["bcs"]:
        b["type"] = 6
    shp = cases.ghost_shape(12, 10, 1, case["ngg"])
    Y = np.zeros(9)
    Y[0], Y[2], Y[8] = 0.05, 0.2, 0.75
    W = tables["mw"]
    T, p, u, v = 900.0, 1.2e5, 300.0, -80.0
    Wb = 1.0 / (Y / W).sum()
    rho = p * Wb / (R_U * T)
    bv = np.zeros((6,) + shp)
    bv[0], bv[1], bv[2], bv[4], bv[5] = rho, u, v, p, T
    sv = np.zeros((9,) + shp)
    for l in range(9):
        sv[l] = Y[l]
    z = oracle_zone(case, mech=mech, bv=bv, sv=sv)
    cp = Y @ po.cp(mech, T)
    gam = cp / (cp - R_U / Wb)
    a = np.sqrt(gam * R_U / Wb * T)
    mu = po.transport(mech, T, p, rho, Y)[0]
    dy = case["y"][0, 1, 0] - case["y"][0, 0, 0]
    dx = case["x"][0, 0, 1] - case["x"][0, 0, 0]
    inv = (abs(u) + a) / dx + (abs(v) + a) / dy
    visc = max(gam, 4.0 / 3.0) / rho * mu / 0.72 * (1 / dx ** 2 + 1 / dy ** 2)
    assert z.global_dt() == pytest.approx(0.5 / (inv + visc), rel=1e-12)


# --------------------------------------------------------------- residual

def test_residual_of_a_hand_built_field():
    case = cases.c1_case(16, 10)
    case["perturb"] = None
    z = oracle_zone(case)
    g = case["ngg"]
    shp = cases.ghost_shape(16, 10, 1, g)
    bv = z.field("bv").reshape((6,) + shp)
    vel = z.field("vel").reshape(shp)
    N = 16 * 10
    z.store_last()
    # step 1: 12 cells change by (d_rho, d_|V|, d_p, d_T)
    d1 = np.array([0.01, 3.0, 250.0, 2.0])
    blk = (g, slice(g + 2, g + 5), slice(g + 3, g + 7))
    bv[0][blk] += d1[0]
    vel[blk] += d1[1]
    bv[4][blk] += d1[2]
    bv[5][blk] += d1[3]
    r1 = z.residual(1)
    np.testing.assert_allclose(r1[:4], 1.0, rtol=1e-15)      # scaled by itself
    scale = np.abs(d1) * np.sqrt(12 / N)
    # step 2 against a new store: 5 cells change, one variable stays
    z.store_last()
    d2 = np.array([-0.02, 0.0, 100.0, 7.0])
    blk2 = (g, slice(g + 6, g + 7), slice(g + 1, g + 6))
    bv[0][blk2] += d2[0]
    bv[4][blk2] += d2[2]
    bv[5][blk2] += d2[3]
    r2 = z.residual(2)
    ref = np.abs(d2) * np.sqrt(5 / N) / scale
    np.testing.assert_allclose(r2[:4], ref, rtol=1e-12, atol=1e-15)
    assert r2[4] == pytest.approx(ref.max(), rel=1e-12)
