"""GPU parity of the boundary conditions beyond inflow / outflow /
symmetry / wall: the Riemann-invariant farfield (air, and the mixture with
its gamma(T) iteration), the total-pressure subsonic inflow and the back
pressure, applied in the reference order farfield -> outflow -> back
pressure -> inflow -> subsonic inflow -> symmetry -> wall."""
import numpy as np
import pytest

import ostc
import pyoracle as po
from ostc import cases
from helpers import BV_GROUPS, grouped, oracle_zone, relerr

pytestmark = pytest.mark.gpu


def test_subsonic_channel(tmp_path):
    case = cases.subsonic_case(60, 30, back_ratio=0.98)
    s = ostc.solver_for_case(case, str(tmp_path / "sub"))
    z = oracle_zone(case)
    e0 = grouped(s.download("bv"), z.field("bv"), 6, BV_GROUPS)
    print("subsonic init (ghosts included)", {k: "%.2e" % v for k, v in e0.items()}, flush=True)
    assert max(e0.values()) < 1e-13
    s.rhs()
    z.rhs()
    assert relerr(s.download("dq"), z.field("dq")) < 1e-12
    dt_p, _ = s.step(20)
    dt_o, _ = z.run(20)
    assert relerr(dt_p, dt_o) < 1e-10
    e = grouped(s.download("bv"), z.field("bv"), 6, BV_GROUPS)
    print("subsonic 20 steps", {k: "%.2e" % v for k, v in e.items()}, flush=True)
    assert max(e.values()) < 1e-9


def _mixture_farfield(tmp_path, mech, sides, name):
    NS = 9
    case = cases.c3_case(32, 16, reconstruction=2)
    Ya = np.zeros(NS)
    Ya[2], Ya[8] = 0.233, 0.767
    st = po.inflow_state(2.0, 101325.0, 1500.0, Y=Ya, m=mech)
    far = dict(st, a_inf=st["vel"] / 2.0, gamma_inf=st["gamma"], entropy=st["p"] / st["rho"] ** st["gamma"],
               mw_inf=st["mw"], sv=list(Ya))
    case["setup"]["array string boundary_conditions"] = ["air_stream", "fuel_stream", "outflow", "symmetry",
                                                         "farfield"]
    case["setup"]["struct farfield"] = [("string", "type", "farfield"), ("int", "label", 4), ("real", "mach", 2.0),
                                        ("real", "pressure", 101325.0), ("real", "temperature", 1500.0),
                                        ("real", "u", 1.0), ("real", "v", 0.0), ("real", "w", 0.0),
                                        ("real", "O2", 0.233), ("real", "N2", 0.767)]
    for b in case["bcs"]:
        if b["type"] == 3 and b["raw"][2] in sides:
            b["type"], b["label"], b["state"] = 4, 4, far
    s = ostc.solver_for_case(case, str(tmp_path / name), initial=False)
    states = {}
    for nm in ("air_stream", "fuel_stream"):
        o = s.inflow_state(nm, 7 + NS)
        states[nm] = dict(rho=o[0], u=o[1], v=o[2], w=o[3], p=o[4], t=o[5], vel=o[6], sv=o[7:7 + NS])
    bv, sv = cases.c3_fields(case, states, np.array([mech.mw[i] for i in range(NS)]))
    s.upload("bv", bv)
    s.upload("sv", sv)
    s.initialize()
    for b in case["bcs"]:
        if b["type"] == 5:
            b["state"] = states[b["stream"]]
    z = oracle_zone(case, mech=mech, bv=bv, sv=sv)
    e0 = grouped(s.download("bv"), z.field("bv"), 6, BV_GROUPS)
    e0["Y"] = relerr(s.download("sv"), z.field("sv"))
    print(f"mixture farfield {name} init", {k: "%.2e" % v for k, v in e0.items()}, flush=True)
    assert max(e0.values()) < 1e-13
    return s, z


def test_mixture_farfield(tmp_path, mech):
    """Reacting mixing layer with farfield boundaries at j=0 and j=max whose
    far state is the air stream: the mixture branch with its gamma(T)
    iteration on both the fuel and the air side (the fuel side leaves by the
    subsonic-outflow branch).  The boundary points' gamma comes from
    properties computed before the first boundary application."""
    s, z = _mixture_farfield(tmp_path, mech, (0, 15), "both")
    s.step(10)
    z.run(10)
    e = grouped(s.download("bv"), z.field("bv"), 6, BV_GROUPS)
    e["Y"] = relerr(s.download("sv"), z.field("sv"))
    print("mixture farfield 10 steps", {k: "%.2e" % v for k, v in e.items()}, flush=True)
    assert max(e.values()) < 1e-9
