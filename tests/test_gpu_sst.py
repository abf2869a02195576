"""GPU parity of the k-omega SST path against the FP64 oracle on identical
seeded inputs: the 2-D turbulent supersonic flat plate in air (wall
distance, mu_t initialisation, right-hand side, SSP-RK3 steps, the steady
path with DPLUR and with the explicit implicit-diagonal step, FP32 fluxes),
and a reduced 3-D reacting channel with wall injection (9 species + k,
omega, HLLC/MUSCL, DPLUR with exact point-implicit chemistry).

Bars: 1e-12 (normwise per component group) after one evaluation / step,
1e-9 after 20 RK3 steps or 10 pseudo-iterations."""
import numpy as np
import pytest

import ostc
from ostc import cases
from helpers import BV_GROUPS, CV_GROUPS, grouped, oracle_zone, relerr

pytestmark = pytest.mark.gpu

NS = 9


def _plate(tmp_path, method=None, nx=64, ny=48, **kw):
    case = cases.sst_plate_case(nx=nx, ny=ny, **kw)
    if method is not None:
        case["setup"]["int implicit_method"] = method
        case["scheme"]["implicit_method"] = method
    s = ostc.solver_for_case(case, str(tmp_path / "plate"), initial=False)
    o = s.inflow_state("freestream", 7 + 2 + 1)
    st = dict(rho=o[0], u=o[1], v=o[2], w=o[3], p=o[4], t=o[5], vel=o[6], sv=list(o[7:9]), mut=o[9])
    for key in ("rho", "u", "p", "t", "mut"):
        assert abs(st[key] - case["inflow"][key]) <= 1e-12 * abs(case["inflow"][key]), key
    case["inflow"] = st
    bv, sv = cases.sst_plate_fields(case)
    s.upload("bv", bv)
    s.upload("sv", sv)
    s.initialize()
    for b in case["bcs"]:
        if b["type"] == 5:
            b["state"] = st
    return case, s, oracle_zone(case, bv=bv, sv=sv)


def _compare(s, z, tol, label, ns=0, fields=("mul", "mut", "acoustic")):
    out = dict(grouped(s.download("bv"), z.field("bv"), 6, BV_GROUPS))
    sv_p, sv_o = s.download("sv").reshape(ns + 2, -1), z.field("sv").reshape(ns + 2, -1)
    if ns:
        out["Y"] = relerr(sv_p[:ns], sv_o[:ns])
    out["k"] = relerr(sv_p[ns], sv_o[ns])
    out["omega"] = relerr(sv_p[ns + 1], sv_o[ns + 1])
    for f in fields:
        out[f] = relerr(s.download(f), z.field(f))
    print(label, {k: "%.2e" % v for k, v in out.items()}, flush=True)
    bad = {k: v for k, v in out.items() if not v < tol}
    assert not bad, (label, bad)


def test_plate_init_and_rhs(tmp_path):
    case, s, z = _plate(tmp_path)
    assert relerr(s.download("wall_distance"), z.field("wall_distance")) < 1e-14
    _compare(s, z, 1e-13, "plate-init")
    s.rhs()
    z.rhs()
    dq_p, dq_o = s.download("dq").reshape(7, -1), z.field("dq").reshape(7, -1)
    errs = grouped(dq_p[:5], dq_o[:5], 5, CV_GROUPS)
    errs["k"] = relerr(dq_p[5], dq_o[5])
    errs["omega"] = relerr(dq_p[6], dq_o[6])
    errs["mut"] = relerr(s.download("mut"), z.field("mut"))
    errs["turb_cond"] = relerr(s.download("turb_cond"), z.field("turb_cond"))
    errs["turb_jac"] = relerr(s.download("turb_jac"), z.field("turb_jac"))
    print("plate rhs", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    assert max(errs.values()) < 1e-12, errs
    dt_p, dt_o = s.global_dt(), z.global_dt()
    assert abs(dt_p - dt_o) / dt_o < 1e-14


def test_plate_rk3(tmp_path):
    case, s, z = _plate(tmp_path)
    s.step(1)
    z.run(1)
    _compare(s, z, 1e-12, "plate-rk3-1")
    dt_p, _ = s.step(19)
    dt_o, _ = z.run(19)
    assert relerr(dt_p, dt_o) < 1e-9
    _compare(s, z, 1e-9, "plate-rk3-20")


@pytest.mark.parametrize("method", [1, 0])
def test_plate_steady(tmp_path, method):
    case, s, z = _plate(tmp_path, method=method)
    s.step_steady(1)
    z.run_steady(1)
    _compare(s, z, 1e-12, f"plate-steady-m{method}-1")
    s.step_steady(9)
    z.run_steady(9)
    _compare(s, z, 1e-9, f"plate-steady-m{method}-10")


def test_plate_fp32_flux(tmp_path):
    case, s, z = _plate(tmp_path)
    s.set_flux_fp32(True)
    s.step(5)
    z.run(5)
    out = dict(grouped(s.download("bv"), z.field("bv"), 6, BV_GROUPS))
    sv_p, sv_o = s.download("sv").reshape(2, -1), z.field("sv").reshape(2, -1)
    out["k"] = relerr(sv_p[0], sv_o[0])
    out["omega"] = relerr(sv_p[1], sv_o[1])
    print("plate fp32 flux, 5 steps", {k: "%.2e" % v for k, v in out.items()}, flush=True)
    # FP32 reconstruction / Riemann / viscous arithmetic: per-field bound
    assert max(out.values()) < 1e-4, out


def _c5(tmp_path, mech, nx=28, ny=20, nz=16, recon=2):
    case = cases.c5_case(nx=nx, ny=ny, nz=nz, lx=0.03, ly=0.01, lz=0.008, first=2e-5, reconstruction=recon)
    s = ostc.solver_for_case(case, str(tmp_path / "c5"), initial=False)
    states = {}
    for name in ("vitiated_air", "h2_jet"):
        o = s.inflow_state(name, 7 + NS + 2 + 1)
        states[name] = dict(rho=o[0], u=o[1], v=o[2], w=o[3], p=o[4], t=o[5], vel=o[6], sv=list(o[7:7 + NS + 2]),
                            mut=o[7 + NS + 2], mw=1.0 / sum(o[7 + l] / mech.mw[l] for l in range(NS)))
    bv, sv = cases.c5_fields(case, states["vitiated_air"])
    s.upload("bv", bv)
    s.upload("sv", sv)
    s.initialize()
    for b in case["bcs"]:
        if b["type"] == 5:
            b["state"] = states[b["stream"]]
    return case, s, oracle_zone(case, mech=mech, bv=bv, sv=sv)


def test_c5_reduced_steady(tmp_path, mech):
    case, s, z = _c5(tmp_path, mech)
    assert relerr(s.download("wall_distance"), z.field("wall_distance")) < 1e-14
    _compare(s, z, 1e-13, "c5-init", ns=NS)
    s.step_steady(1)
    z.run_steady(1)
    _compare(s, z, 1e-12, "c5-steady-1", ns=NS)
    s.step_steady(4)
    z.run_steady(4)
    _compare(s, z, 1e-9, "c5-steady-5", ns=NS)
