"""The oracle's AWENO-5 for mixtures and transported scalars (this build's
definition: the reference's mixture interpolation is an empty stub,
AWENO.cu:10-15, and its CDS correction prints "Not implemented" for
mixtures, AWENO.cu:209-212).  Pinned by:

* L R = I for the characteristic matrices of the 9-species system and of
  air with two transported scalars, at random Roe states and normals;
* free-stream preservation: a uniform 9-species flow on a Cartesian grid
  has a vanishing right-hand side;
* fifth-order convergence of the AWENO-5 + CDS flux difference (HLLC and
  AUSM+) on smooth 1-D mixture data (composition, temperature and velocity
  all varying) to the exact -dF/dx, over the cells where the correction
  applies;
* reduction to air: with transported scalars on an air flow, the flow rows
  of the right-hand side equal those of the air AWENO (the scalar waves do
  not enter the acoustic and entropy fields when there are no species).
"""
import numpy as np
import pytest

import pyoracle as po
from helpers import oracle_zone
from ostc import cases

R_U = 8314.462618
GAMMA, MW_AIR = 1.4, 28.9647


def _mats(mech, nsc, cl, cr, k0, k1):
    nv = 5 + nsc
    L, R = np.zeros(nv * nv), np.zeros(nv * nv)
    m = None if mech is None else po.C.cast(po.C.byref(mech), po.C.c_void_p)
    po.lib().orc_aweno_mix_matrices(m, nsc, po._p(np.ascontiguousarray(cl)), po._p(np.ascontiguousarray(cr)),
                                    po._p(np.ascontiguousarray(k0)), po._p(np.ascontiguousarray(k1)), po._p(L),
                                    po._p(R))
    return L.reshape(nv, nv), R.reshape(nv, nv)


def _cons(mech, W, rho, vel, T, Y, sc=()):
    p = rho * R_U * T * (Y / W).sum() if mech is not None else rho * R_U * T / MW_AIR
    if mech is not None:
        E = rho * (Y @ po.enthalpy(mech, T)) - p + 0.5 * rho * vel @ vel
        q = [rho, *(rho * vel), E, *(rho * Y)]
    else:
        q = [rho, *(rho * vel), p / (GAMMA - 1) + 0.5 * rho * vel @ vel]
    q += [rho * s for s in sc]
    return np.array(q + [p, T])


@pytest.mark.parametrize("seed", range(4))
def test_characteristic_matrices_are_inverse(mech, tables, seed):
    rng = np.random.default_rng(seed)
    W = tables["mw"]
    for sp in (True, False):
        Ya, Yb = rng.random(9) + 0.01, rng.random(9) + 0.01
        Ya /= Ya.sum()
        Yb /= Yb.sum()
        m = mech if sp else None
        sc = () if sp else (rng.random() * 50, 1e4 * (1 + rng.random()))
        cl = _cons(m, W, 0.5 + rng.random(), 300 * rng.standard_normal(3), 800 + 900 * rng.random(), Ya, sc)
        cr = _cons(m, W, 0.5 + rng.random(), 300 * rng.standard_normal(3), 800 + 900 * rng.random(), Yb, sc)
        k0, k1 = rng.standard_normal(3), rng.standard_normal(3)
        L, R = _mats(m, 9 if sp else 2, cl, cr, k0, k0 + 0.1 * k1)
        nv = L.shape[0]
        np.testing.assert_allclose(L @ R, np.eye(nv), atol=2e-12)
        np.testing.assert_allclose(R @ L, np.eye(nv), atol=2e-9 * np.abs(R).max() * np.abs(L).max())


def _mix_case(n, ny=8, h=None, scheme=4):
    h = h if h is not None else 1.0 / n
    case = cases.c3_case(n, ny, inviscid_scheme=scheme)
    case["scheme"].update(reconstruction=4, reaction=0, viscous_order=0)
    case["setup"]["int reconstruction"] = 4
    case["ngg"] = 3
    case["perturb"] = None
    case["x"] = np.broadcast_to(np.arange(n) * h, (1, ny, n)).copy()
    case["y"] = np.broadcast_to((np.arange(ny) * h)[:, None], (1, ny, n)).copy()
    for b in case["bcs"]:
        b["type"] = 6
    return case


def test_uniform_mixture_flow_is_preserved(mech, tables):
    case = _mix_case(24, 12, h=1e-4)
    shp = cases.ghost_shape(24, 12, 1, 3)
    W = tables["mw"]
    Y = np.linspace(1.0, 2.0, 9)
    Y /= Y.sum()
    T, p = 1200.0, 1.0e5
    rho = p / (R_U * T * (Y / W).sum())
    bv = np.zeros((6,) + shp)
    bv[0], bv[1], bv[2], bv[4], bv[5] = rho, 700.0, -150.0, p, T
    sv = np.zeros((9,) + shp)
    for l in range(9):
        sv[l] = Y[l]
    z = oracle_zone(case, mech=mech, bv=bv, sv=sv)
    z.rhs()
    dq = z.field("dq").reshape(14, -1)
    scale = rho * 700.0 ** 2 * 1e-4 + p * 1e-4
    assert np.abs(dq).max() < 1e-11 * scale


def _profiles(x):
    tp = 2 * np.pi
    rho = 0.6 + 0.1 * np.sin(tp * x)
    u = 250.0 + 40.0 * np.sin(tp * x + 1.0)
    T = 1300.0 + 150.0 * np.cos(tp * x + 0.3)         # stays above T_mid: one NASA-7 range
    yh2 = 0.04 + 0.02 * np.sin(tp * x + 2.0)
    yo2 = 0.20 + 0.05 * np.cos(tp * x)
    yh2o = 0.10 + 0.03 * np.sin(tp * x + 0.5)
    Y = np.zeros((9,) + np.shape(x))
    Y[0], Y[2], Y[7] = yh2, yo2, yh2o
    Y[8] = 1.0 - yh2 - yo2 - yh2o
    return rho, u, T, Y


def _flux(mech, W, x):
    rho, u, T, Y = _profiles(x)
    p = rho * R_U * T * (Y / W[:, None]).sum(axis=0)
    h = np.array([po.enthalpy(mech, t) for t in T]).T
    E = rho * (Y * h).sum(axis=0) - p + 0.5 * rho * u * u
    return np.array([rho * u, rho * u * u + p, 0 * u, 0 * u, u * (E + p), *(rho * Y * u)])


def _aweno_mix_line(mech, tables, n, scheme):
    W = tables["mw"]
    h = 1.0 / n
    case = _mix_case(n, scheme=scheme)
    g = 3
    shp = cases.ghost_shape(n, 8, 1, g)
    xs = (np.arange(shp[2]) - g) * h
    rho, u, T, Y = _profiles(xs)
    p = rho * R_U * T * (Y / W[:, None]).sum(axis=0)
    bv = np.zeros((6,) + shp)
    bv[0], bv[1], bv[4], bv[5] = rho, u, p, T
    sv = np.zeros((9,) + shp)
    sv[:] = Y[:, None, None, :]
    z = oracle_zone(case, mech=mech, bv=bv, sv=sv)
    # the boundary conditions of the initialisation rewrote the ghosts: put
    # the smooth fields back everywhere (bv, sv and cv)
    zb, zs, zc = (z.field(f).reshape((-1,) + shp) for f in ("bv", "sv", "cv"))
    zb[:] = bv
    zs[:] = sv
    hh = np.array([po.enthalpy(mech, t) for t in T]).T
    E = rho * (Y * hh).sum(axis=0) - p + 0.5 * rho * u * u
    zc[:] = 0
    zc[0], zc[1], zc[4] = rho, rho * u, E
    zc[5:] = (rho * Y)[:, None, None, :]
    z.rhs()
    dq = z.field("dq").reshape(14, 1, 8, n)[:, 0, 4, :]
    jac = z.field("jac").reshape(shp)[g, g + 4, g:g + n]
    x = np.arange(n) * h
    d = 2e-3
    c8 = [(1, 4 / 5), (2, -1 / 5), (3, 4 / 105), (4, -1 / 280)]
    dF = sum(w * (_flux(mech, W, x + k * d) - _flux(mech, W, x - k * d)) for k, w in c8) / d
    return dq / jac, -dF


@pytest.mark.parametrize("scheme", [4, 3])
def test_mixture_aweno_cds_is_fifth_order(mech, tables, scheme):
    errs = []
    for n in (48, 96, 192):
        num, ex = _aweno_mix_line(mech, tables, n, scheme)
        cells = slice(0, n - 6)
        rows = [0, 1, 4, 5, 7, 12, 13]
        scale = np.abs(ex[rows]).max(axis=1)
        errs.append((np.abs(num[rows, cells] - ex[rows, cells]).max(axis=1) / scale).max())
    orders = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    print("mixture AWENO errors", errs, "orders", orders)
    assert orders[-1] > 4.5, (errs, orders)


def test_scalar_awenos_flow_rows_reduce_to_air(tmp_path):
    nx, ny = 40, 16
    base = cases.sst_plate_case(nx, ny, compressibility=0)
    out = {}
    for nt in (0, 2):
        case = dict(base)
        case["scheme"] = dict(base["scheme"], reconstruction=4, viscous_order=0, turbulence=1 if nt else 0)
        case["ngg"] = 3
        case["bcs"] = [dict(b) for b in base["bcs"]]
        for b in case["bcs"]:
            if b["type"] == 5:
                b["state"] = dict(case["inflow"], sv=list(case["inflow"]["sv"]) if nt else [])
        shp = cases.ghost_shape(nx, ny, 1, 3)
        rng = np.random.default_rng(9)
        st = case["inflow"]
        bv = np.zeros((6,) + shp)
        bv[0] = st["rho"] * (1 + 0.05 * rng.standard_normal(shp[1:]))
        bv[1] = st["u"] * (1 + 0.02 * rng.standard_normal(shp[1:]))
        bv[2] = 5.0 * rng.standard_normal(shp[1:])
        bv[4] = st["p"]
        bv[5] = st["p"] * MW_AIR / (bv[0] * R_U)
        sv = np.zeros((2,) + shp) if nt else None
        if nt:
            sv[0] = 10.0 * (1 + rng.random(shp[1:]))
            sv[1] = 1e4 * (1 + rng.random(shp[1:]))
        z = oracle_zone(case, bv=bv, sv=sv)
        z.rhs()
        out[nt] = z.field("dq").reshape(5 + nt, -1)[:5]
    scale = np.abs(out[0]).max(axis=1, keepdims=True)
    assert np.abs(out[2] - out[0]).max() < 1e-12 * scale.max(), np.abs(out[2] - out[0]).max() / scale.max()


def _sides(mech, nsc, pts, k0, k1):
    c6 = np.ascontiguousarray(np.concatenate(pts))
    nv = 5 + nsc
    pl, pr = np.zeros(nv + 2), np.zeros(nv + 2)
    m = None if mech is None else po.C.cast(po.C.byref(mech), po.C.c_void_p)
    fell_back = po.lib().orc_aweno_mix_sides(m, nsc, po._p(c6), po._p(np.ascontiguousarray(k0, dtype=float)),
                                             po._p(np.ascontiguousarray(k1, dtype=float)), po._p(pl), po._p(pr))
    return pl, pr, bool(fell_back)


def test_admissibility_fallback_at_a_two_cell_jet(mech, tables):
    """A jet port two cells wide in a wall row (air at 300 K, pure H2 at 300 K
    and twice the pressure, sonic normal velocity): every candidate stencil of
    the face between the two jet points crosses a jump, the WENO overshoot
    drives the reconstructed density towards zero and the mass fractions far
    outside [0, 1], and both sides take the first-order values -- the jet
    state itself, with its own rho E and gamma."""
    W = tables["mw"]
    Ya = np.zeros(9)
    Ya[2], Ya[7], Ya[8] = 0.25, 0.12, 0.63
    Yj = np.zeros(9)
    Yj[0] = 1.0
    air = _cons(mech, W, 1.099, np.zeros(3), 300.0, Ya)
    jet = _cons(mech, W, 0.1638, np.array([0.0, 1318.0, 0.0]), 300.0, Yj)
    k = np.array([1.0, 0.0, 0.0])
    pl, pr, fb = _sides(mech, 9, [air, air, jet, jet, air, air], k, k)
    assert fb
    p_jet = jet[14]
    for side in (pl, pr):
        np.testing.assert_allclose(side[:5], [0.1638, 0.0, 1318.0, 0.0, p_jet], rtol=1e-14)
        np.testing.assert_allclose(side[5:14], Yj, atol=1e-15)
        assert side[14] == pytest.approx(jet[4], rel=1e-12)       # rho E of the jet
        assert side[15] == pytest.approx(1.405, abs=0.01)         # gamma of H2 at 300 K


def test_admissible_smooth_faces_keep_the_weno_states(mech, tables):
    """Smooth data: no fallback, and the sides agree with the midpoint value
    of the smooth profile to the reconstruction's accuracy."""
    W = tables["mw"]
    xs = np.arange(-2, 4) * 0.02
    pts = []
    for x in xs:
        Y = np.array([0.02 + 0.01 * x, 0.001, 0.22 - 0.02 * x, 0.002, 0.003, 0.0, 0.0, 0.1 + 0.05 * x, 0.0])
        Y[8] = 1.0 - Y[:8].sum()
        pts.append(_cons(mech, W, 0.4 + 0.1 * x, np.array([600.0 + 100 * x, 20.0, 0.0]), 1500.0 + 200 * x, Y))
    k = np.array([1.0, 0.0, 0.0])
    pl, pr, fb = _sides(mech, 9, pts, k, k)
    assert not fb
    mid = 0.5 * (pts[2] + pts[3])
    assert pl[0] == pytest.approx(mid[0], rel=1e-3) and pr[0] == pytest.approx(mid[0], rel=1e-3)
    assert pl[1] == pytest.approx(mid[1] / mid[0], rel=1e-3)
    assert abs(pl[5:14].sum() - 1) < 1e-10 and abs(pr[5:14].sum() - 1) < 1e-10
