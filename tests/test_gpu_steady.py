"""GPU parity of the steady path (local time steps, implicit treatment,
update_bv) against the FP64 oracle: DPLUR for air in 2-D and for a 3-D
viscous case with an isothermal wall (wall dq zeroed between the inner
sweeps), and for the reacting 9-species mixing layer with the exact point
implicit (EPI) chemistry; the explicit steady step with EPI and with the
diagonal approximation (DA).  Bars: 1e-12 after one iteration, 1e-9 after
several."""
import numpy as np
import pytest

import ostc
from ostc import cases
from helpers import BV_GROUPS, grouped, oracle_zone, relerr

pytestmark = pytest.mark.gpu

NS = 9


def _steady(case, method, chem=None, inner=3, cfl=None):
    case["setup"]["int implicit_method"] = method
    case["setup"]["int DPLUR_inner_step"] = inner
    case["scheme"].update(implicit_method=method, DPLUR_inner_step=inner)
    if chem is not None:
        case["setup"]["int chemSrcMethod"] = chem
        case["scheme"]["chemSrcMethod"] = chem
    if cfl is not None:
        case["setup"]["real cfl"] = cfl
        case["scheme"]["cfl"] = cfl
    return case


def _air(case, tmp_path):
    s = ostc.solver_for_case(case, str(tmp_path / case["name"]))
    z = oracle_zone(case)
    return s, z


def _mixture(case, tmp_path, mech):
    s = ostc.solver_for_case(case, str(tmp_path / "c3"), initial=False)
    states = {}
    for name in ("air_stream", "fuel_stream"):
        o = s.inflow_state(name, 7 + NS)
        states[name] = dict(rho=o[0], u=o[1], v=o[2], w=o[3], p=o[4], t=o[5], vel=o[6], sv=o[7:7 + NS])
    bv, sv = cases.c3_fields(case, states, np.array([mech.mw[i] for i in range(NS)]))
    s.upload("bv", bv)
    s.upload("sv", sv)
    s.initialize()
    for b in case["bcs"]:
        if b["type"] == 5:
            b["state"] = states[b["stream"]]
    return s, oracle_zone(case, mech=mech, bv=bv, sv=sv)


def _compare(s, z, tol, label, species=False):
    out = dict(grouped(s.download("bv"), z.field("bv"), 6, BV_GROUPS))
    if species:
        out["Y"] = relerr(s.download("sv"), z.field("sv"))
    for f in ("mul", "acoustic"):
        out[f] = relerr(s.download(f), z.field(f))
    print(label, {k: "%.2e" % v for k, v in out.items()}, flush=True)
    bad = {k: v for k, v in out.items() if not v < tol}
    assert not bad, (label, bad)


def test_c1_dplur(tmp_path):
    case = _steady(cases.c1_case(60, 30), 1, cfl=2.0)
    s, z = _air(case, tmp_path)
    rp = s.step_steady(1)
    ro = z.run_steady(1)
    _compare(s, z, 1e-12, "c1-dplur-1")
    rp = s.step_steady(19)
    ro = z.run_steady(19)
    _compare(s, z, 1e-9, "c1-dplur-20")
    rel = np.abs(rp[:, :4] - ro[:, :4]) / np.abs(ro[:, :4])
    assert np.nanmax(rel[:, [0, 3]]) < 1e-9


def test_c2_viscous_wall_dplur(tmp_path):
    case = _steady(cases.c2_case(24, 16, 12, reconstruction=2, inviscid_scheme=3), 1, cfl=1.0)
    s, z = _air(case, tmp_path)
    s.step_steady(1)
    z.run_steady(1)
    _compare(s, z, 1e-12, "c2-dplur-1")
    s.step_steady(9)
    z.run_steady(9)
    _compare(s, z, 1e-9, "c2-dplur-10")


@pytest.mark.parametrize("method,chem", [(1, 1), (0, 1), (0, 2)])
def test_c3_reacting_steady(tmp_path, mech, method, chem):
    case = _steady(cases.c3_case(32, 16, reconstruction=2), method, chem=chem)
    s, z = _mixture(case, tmp_path, mech)
    s.step_steady(1)
    z.run_steady(1)
    _compare(s, z, 1e-12, f"c3-steady-m{method}-c{chem}-1", species=True)
    s.step_steady(7)
    z.run_steady(7)
    _compare(s, z, 1e-9, f"c3-steady-m{method}-c{chem}-8", species=True)


@pytest.mark.parametrize("method", [1, 0])
def test_c3_steady_inline_chemistry(tmp_path, mech, method):
    """The steady path with the chemistry operator-split inline (one FP64
    point-implicit step of 2e-8 s per pseudo-iteration after the flow
    update, no source or chemistry Jacobian in the flow part) against the
    oracle's same split: 1 iteration at 1e-12, 8 at 1e-9."""
    case = _steady(cases.c3_case(32, 16, reconstruction=2), method, chem=1)
    case["scheme"].update(split_chem=1, steady_chem_dt=2e-8)
    s, z = _mixture(case, tmp_path, mech)
    s.set_steady_chem(2e-8, 0)
    s.step_steady(1)
    z.run_steady(1)
    _compare(s, z, 1e-12, f"c3-inline-chem-m{method}-1", species=True)
    s.step_steady(7)
    z.run_steady(7)
    _compare(s, z, 1e-9, f"c3-inline-chem-m{method}-8", species=True)
    cells, fb = s.steady_chem_stats()
    assert cells == 8 * 32 * 16 and fb == 0
    s.set_steady_chem(0.0)
    assert s.steady_chem_stats() == (0, 0)
