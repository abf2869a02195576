"""GPU parity of the reacting 9-species path (reduced C3 mixing layer:
HLLC + MUSCL-minmod, mixture-averaged viscous/diffusive fluxes, explicit
finite-rate source, SSP-RK3) against the FP64 CPU oracle on identical seeded
inputs.  Bars: 1e-12 after one RHS evaluation / one step, 1e-9 after 20
steps; per named group of components."""
import numpy as np
import pytest

import ostc
import pyoracle as po
from ostc import cases
from helpers import BV_GROUPS, grouped, oracle_zone, relerr

pytestmark = pytest.mark.gpu

NS = 9
CV14 = {"rho": [0], "mom": [1, 2, 3], "rhoE": [4], "rhoY": list(range(5, 14))}
SPECIES = {"Y": list(range(NS))}


def _setup(tmp_path, mech, nx=48, ny=24, fp32=False, scheme=4, recon=2):
    case = cases.c3_case(nx, ny, inviscid_scheme=scheme, reconstruction=recon)
    s = ostc.solver_for_case(case, str(tmp_path / "c3"), initial=False)
    states = {}
    for name in ("air_stream", "fuel_stream"):
        o = s.inflow_state(name, 7 + NS)
        states[name] = dict(rho=o[0], u=o[1], v=o[2], w=o[3], p=o[4], t=o[5], vel=o[6], sv=o[7:7 + NS])
    mw = np.array([mech.mw[i] for i in range(NS)])
    bv, sv = cases.c3_fields(case, states, mw)
    s.upload("bv", bv)
    s.upload("sv", sv)
    s.initialize()
    if fp32:
        s.set_flux_fp32(True)
    for b in case["bcs"]:
        if b["type"] == 5:
            b["state"] = states[b["stream"]]
    z = oracle_zone(case, mech=mech, bv=bv, sv=sv)
    return case, s, z, states


def _compare(s, z, tol, label):
    out = {}
    out.update(grouped(s.download("bv"), z.field("bv"), 6, BV_GROUPS))
    out.update(grouped(s.download("sv"), z.field("sv"), NS, SPECIES))
    out.update({"cv_" + k: v for k, v in grouped(s.download("cv"), z.field("cv"), 14, CV14).items()})
    for f in ("mul", "cond", "acoustic", "gamma", "cp", "rhoD"):
        out[f] = relerr(s.download(f), z.field(f))
    print(label, {k: "%.2e" % v for k, v in out.items()}, flush=True)
    bad = {k: v for k, v in out.items() if not v < tol}
    assert not bad, (label, bad)


def test_inflow_states_match_oracle(tmp_path, mech):
    _, s, _, states = _setup(tmp_path, mech)
    Ya = np.zeros(NS)
    Ya[2], Ya[8] = 0.233, 0.767
    Yf = np.zeros(NS)
    Yf[0], Yf[8] = 0.3, 0.7
    ra = po.inflow_state(2.0, 101325.0, 1500.0, Y=Ya, m=mech)
    rf = po.inflow_state(1.2, 101325.0, 300.0, Y=Yf, m=mech)
    for k in ("rho", "u", "p", "t", "vel"):
        assert abs(states["air_stream"][k] / ra[k] - 1) < 1e-14
        assert abs(states["fuel_stream"][k] / rf[k] - 1) < 1e-14


def test_c3_init_and_rhs(tmp_path, mech):
    case, s, z, _ = _setup(tmp_path, mech)
    _compare(s, z, 1e-13, "c3-init")
    s.rhs()
    z.rhs()
    errs = grouped(s.download("dq"), z.field("dq"), 14, CV14)
    print("c3 rhs err", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    assert max(errs.values()) < 1e-12
    dt_p, dt_o = s.global_dt(), z.global_dt()
    assert abs(dt_p - dt_o) / dt_o < 1e-13


def test_c3_one_step(tmp_path, mech):
    case, s, z, _ = _setup(tmp_path, mech)
    dt_p, res_p = s.step(1)
    dt_o, res_o = z.run(1)
    assert relerr(dt_p, dt_o) < 1e-12
    _compare(s, z, 1e-12, "c3-1step")


def test_c3_twenty_steps(tmp_path, mech):
    case, s, z, _ = _setup(tmp_path, mech)
    dt_p, res_p = s.step(20)
    dt_o, res_o = z.run(20)
    print("c3 20-step dt err %.2e" % relerr(dt_p, dt_o), flush=True)
    assert relerr(dt_p, dt_o) < 1e-9
    _compare(s, z, 1e-9, "c3-20step")
    assert np.isfinite(res_p).all()


def test_c3_mixed_precision_errors(tmp_path, mech):
    """FP32 flux arithmetic (FP64 Newton T, chemistry and accumulation):
    per-field L2 / Linf errors against the FP64 oracle after 10 steps."""
    case, s, z, _ = _setup(tmp_path, mech, fp32=True)
    s.step(10)
    z.run(10)
    bp = s.download("bv").reshape(6, -1)
    bo = z.field("bv").reshape(6, -1)
    sp = s.download("sv").reshape(NS, -1)
    so = z.field("sv").reshape(NS, -1)
    rep = {}
    for name, a, b in (("rho", bp[0], bo[0]), ("u", bp[1], bo[1]), ("p", bp[4], bo[4]), ("T", bp[5], bo[5]),
                       ("Y", sp, so)):
        linf = np.max(np.abs(a - b)) / np.max(np.abs(b))
        l2 = np.linalg.norm(a - b) / np.linalg.norm(b)
        rep[name] = (l2, linf)
    print("c3 mixed precision L2/Linf", {k: "%.1e/%.1e" % v for k, v in rep.items()}, flush=True)
    assert max(v[1] for v in rep.values()) < 1e-4


def test_c3_roe_steps(tmp_path, mech):
    """Roe flux for the mixture (species rows in the characteristic
    decomposition, gamma and c from the Roe-averaged temperature)."""
    case, s, z, _ = _setup(tmp_path, mech, scheme=2)
    s.rhs()
    z.rhs()
    errs = grouped(s.download("dq"), z.field("dq"), 14, CV14)
    print("c3 roe rhs err", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    assert max(errs.values()) < 1e-12
    s.step(10)
    z.run(10)
    _compare(s, z, 1e-10, "c3-roe-10step")


def test_c3_operator_split_chemistry(tmp_path, mech):
    """Split reacting step: a non-reacting SSP-RK3 flow step, then one
    point-implicit chemistry step of every cell over the step's dt at its
    density and total energy, boundaries and properties re-applied."""
    case, s, z, _ = _setup(tmp_path, mech)
    case["scheme"]["split_chem"] = 1
    z = oracle_zone(case, mech=mech, bv=z.field("bv").copy(), sv=z.field("sv").copy())
    s.set_chem_split(True)
    for it in range(5):
        dt_p, _ = s.step(1)
        dt_o, _ = z.run(1)
        assert abs(dt_p[0, 2] - dt_o[0, 2]) <= 1e-12 * dt_o[0, 2]
        nf = s.chem_step(float(dt_o[0, 2]), mode=0)
        z.chem_step(float(dt_o[0, 2]))
        assert nf == 0
        if it == 0:
            _compare(s, z, 1e-12, "c3-split-1")
    _compare(s, z, 1e-9, "c3-split-5")
