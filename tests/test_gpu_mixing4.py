"""GPU parity of the 4-species path: the mixing layer with the reference's
own mixing mechanism (chemistry/H2_mech_mixing.inp: H2, H2O, O2, N2, no
reactions), HLLC with AWENO-5 (the multi-species characteristic
interpolation) and with MUSCL, mixture-averaged viscous and diffusive
fluxes, SSP-RK3, against the FP64 CPU oracle built from the reference
parser's tables of that mechanism (tests/golden/h2_mixing_4sp_tables.txt).
Bars: 1e-12 after one RHS evaluation, 1e-9 after 20 steps; a 3-D block of
the same mixture with k-omega SST on the steady DPLUR path for 5
pseudo-iterations at 1e-9."""
import os

import numpy as np
import pytest

import ostc
import pyoracle as po
from ostc import cases
from helpers import BV_GROUPS, grouped, oracle_zone, relerr

pytestmark = pytest.mark.gpu

NS = 4
NV = 5 + NS
CV = {"rho": [0], "mom": [1, 2, 3], "rhoE": [4], "rhoY": list(range(5, NV))}
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def mech4():
    return po.make_mech(po.load_tables(os.path.join(ROOT, "tests", "golden", "h2_mixing_4sp_tables.txt")))


def _setup(tmp_path, mech, recon):
    case = cases.mixing4_case(48, 24, reconstruction=recon)
    s = ostc.solver_for_case(case, str(tmp_path / "m4"), initial=False)
    states = {}
    for name in ("air_stream", "fuel_stream"):
        o = s.inflow_state(name, 7 + NS)
        states[name] = dict(rho=o[0], u=o[1], v=o[2], w=o[3], p=o[4], t=o[5], vel=o[6], sv=o[7:7 + NS])
    mw = np.array([mech.mw[i] for i in range(NS)])
    bv, sv = cases.c3_fields(case, states, mw)
    s.upload("bv", bv)
    s.upload("sv", sv)
    s.initialize()
    for b in case["bcs"]:
        if b["type"] == 5:
            b["state"] = states[b["stream"]]
    z = oracle_zone(case, mech=mech, bv=bv, sv=sv)
    return case, s, z, states


@pytest.mark.parametrize("recon", [4, 2])
def test_mixing4_rhs_and_twenty_steps(tmp_path, mech4, recon):
    print("4-species mixing layer, reconstruction", recon, flush=True)
    case, s, z, states = _setup(tmp_path, mech4, recon)
    # the streams: air (O2 0.233, N2 0.767) and fuel (H2 0.3, N2 0.7) in this
    # mechanism's species order H2, H2O, O2, N2
    assert np.allclose(states["air_stream"]["sv"], [0, 0, 0.233, 0.767])
    assert np.allclose(states["fuel_stream"]["sv"], [0.3, 0, 0, 0.7])
    s.rhs()
    z.rhs()
    errs = grouped(s.download("dq"), z.field("dq"), NV, CV)
    print("rhs err", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    assert max(errs.values()) < 1e-12
    dt_p, res_p = s.step(20)
    dt_o, res_o = z.run(20)
    out = grouped(s.download("bv"), z.field("bv"), 6, BV_GROUPS)
    out["Y"] = relerr(s.download("sv"), z.field("sv"))
    for f in ("mul", "cond", "gamma", "rhoD"):
        out[f] = relerr(s.download(f), z.field(f))
    print("20 steps", {k: "%.2e" % v for k, v in out.items()}, flush=True)
    assert max(out.values()) < 1e-9
    assert relerr(dt_p, dt_o) < 1e-9
    s.close()


def test_mixing4_sst_steady(tmp_path, mech4):
    """A 3-D channel block of the 4-species mixture with k-omega SST: HLLC +
    AWENO-5 (k and omega on their own characteristic waves), viscous,
    steady DPLUR with 3 inner sweeps, 5 pseudo-iterations."""
    case = cases.c5_case(nx=28, ny=20, nz=16, lx=0.03, ly=0.01, lz=0.008, first=2e-5, chem=0)
    case["setup"].update({"int reaction": 0, "string mechanism_file": "chemistry/H2_mech_mixing.inp"})
    case["scheme"] = dict(case["scheme"], reaction=0)
    case["name"] = "c5m4"
    print("4-species SST channel", flush=True)
    s = ostc.solver_for_case(case, str(tmp_path / "c5m4"), initial=False)
    states = {}
    for name in ("vitiated_air", "h2_jet"):
        o = s.inflow_state(name, 7 + NS + 2 + 1)
        states[name] = dict(rho=o[0], u=o[1], v=o[2], w=o[3], p=o[4], t=o[5], vel=o[6], sv=list(o[7:7 + NS + 2]),
                            mut=o[7 + NS + 2], mw=1.0 / sum(o[7 + l] / mech4.mw[l] for l in range(NS)))
    bv, sv = cases.c5_fields(case, states["vitiated_air"])
    s.upload("bv", bv)
    s.upload("sv", sv)
    s.initialize()
    for b in case["bcs"]:
        if b["type"] == 5:
            b["state"] = states[b["stream"]]
    z = oracle_zone(case, mech=mech4, bv=bv, sv=sv)
    res_p = s.step_steady(5)
    res_o = z.run_steady(5)
    out = grouped(s.download("bv"), z.field("bv"), 6, BV_GROUPS)
    out["sv"] = relerr(s.download("sv"), z.field("sv"))
    print("5 pseudo-iterations", {k: "%.2e" % v for k, v in out.items()}, flush=True)
    assert max(out.values()) < 1e-9
    assert np.max(np.abs(res_p[:, :4] - res_o[:, :4]) / np.abs(res_o[:, :4])) < 1e-8
    s.close()
