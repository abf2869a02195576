"""Multi-block cases on the GPU against the oracle: blocks read from the
reference's grid / boundary / inner-face files, metrics computed on the
device, the inner-face exchange (DataCommunication.cuh:58-110) every stage,
the CFL dt and the residual over all blocks.

Bars: device metrics bit-identical to the host computation; 1e-12 after
the initialisation and one step, 1e-9 after 100 steps with the dt and
residual histories compared entry by entry; and, on the GPU alone, a cut
case reproducing its single-block run."""
import numpy as np
import pytest

import ostc
import pyoracle as po
from ostc import cases
from helpers import BV_GROUPS, CV_GROUPS, grouped, oracle_mesh, oracle_zone, relerr

pytestmark = pytest.mark.gpu

NS = 9
SPECIES = {"Y": list(range(NS))}


def _errs(s, M, nb, nv, ns=0):
    out = {}
    for b in range(nb):
        z = M.zones[b]
        e = grouped(s.download("bv", blk=b), z.field("bv"), 6, BV_GROUPS)
        e.update({"cv_" + k: v for k, v in grouped(s.download("cv", blk=b), z.field("cv"), nv,
                                                   {"all": list(range(nv))}).items()})
        if ns:
            e["sv"] = relerr(s.download("sv", blk=b), z.field("sv"))
        for f in ("vel", "acoustic", "mul"):
            e[f] = relerr(s.download(f, blk=b), z.field(f))
        for k, v in e.items():
            out[k] = max(out.get(k, 0.0), v)
    return out


def _check(s, M, nb, nv, tol, label, ns=0):
    e = _errs(s, M, nb, nv, ns)
    print(label, {k: "%.2e" % v for k, v in e.items()}, flush=True)
    assert max(e.values()) < tol, (label, e)


def _hist(dt_p, dt_o, res_p, res_o, label, tol=1e-9, cols=(0, 1, 2, 3, 4)):
    e_dt = (np.abs(dt_p - dt_o) / dt_o).max()
    cols = list(cols)
    rel = np.abs(res_p[:, cols] - res_o[:, cols]) / np.abs(res_o[:, cols])
    print(f"{label}: dt history {e_dt:.2e}, residual history {rel.max():.2e}", flush=True)
    assert np.isfinite(res_p).all()
    assert e_dt < tol and rel.max() < tol, (label, e_dt, rel.max(axis=0))


@pytest.mark.parametrize("shape", [(40, 24, 1), (20, 14, 12)])
def test_device_metrics_match_host(tmp_path, shape):
    mx, my, mz = shape
    case = cases.c1_case(mx, my) if mz == 1 else cases.c2_case(mx, my, mz)
    print("device metrics", shape, flush=True)
    s = ostc.solver_for_case(case, str(tmp_path / "g"), initial=False)
    jd, md = s.geometry()
    h = ostc.Solver(str(tmp_path / "g"), device=-1)
    h.add_block(mx, my, mz, case["x"], case["y"], case["z"])
    jh, mh = h.geometry()
    np.testing.assert_array_equal(jd, jh)
    np.testing.assert_array_equal(md, mh)
    s.close()
    h.close()


def _c1_split(tmp_path, axis, rotate, nx=40, ny=24):
    case = cases.c1_case(nx, ny)
    mc = cases.split_case(case, axis=axis, at=(nx if axis == 0 else ny) * 2 // 5, rotate=rotate)
    bv, sv = cases.initial_fields(case)
    fields = cases.split_fields(mc, bv, sv)
    print("c1 split axis", axis, "rotate", rotate, flush=True)
    s = ostc.solver_for_mesh_case(mc, str(tmp_path / "mb"), fields=fields)
    return case, mc, s, oracle_mesh(mc, fields), (bv, sv)


@pytest.mark.parametrize("axis,rotate", [(0, True), (1, False)])
def test_c1_split_hundred_steps(tmp_path, axis, rotate):
    case, mc, s, M, _ = _c1_split(tmp_path, axis, rotate)
    assert s.block_count() == 2
    for b in range(2):
        jo = M.zones[b].field("jac")
        np.testing.assert_array_equal(s.geometry(b)[0], jo)
    _check(s, M, 2, 5, 1e-14, "c1-split-init")
    dt_p, res_p = s.step(1)
    dt_o, res_o = M.run(1)
    _check(s, M, 2, 5, 1e-12, "c1-split-1")
    dt_p2, res_p2 = s.step(99)
    dt_o2, res_o2 = M.run(99)
    _check(s, M, 2, 5, 1e-9, "c1-split-100")
    # the density perturbation is advected at uniform velocity and pressure:
    # the |V| and p residuals are rounding noise, rho and T carry the signal
    _hist(np.vstack([dt_p, dt_p2]), np.vstack([dt_o, dt_o2]), np.vstack([res_p, res_p2]),
          np.vstack([res_o, res_o2]), "c1-split-100", cols=(0, 3))


def test_c1_split_reproduces_single_block_on_gpu(tmp_path):
    case, mc, s, _, (bv, sv) = _c1_split(tmp_path, 0, True)
    one = ostc.solver_for_case(case, str(tmp_path / "one"), initial=False)
    one.upload("bv", bv)
    one.initialize()
    dt_m, _ = s.step(20)
    dt_1, _ = one.step(20)
    assert relerr(dt_m, dt_1) < 1e-13
    g = mc["ngg"]
    for b in range(2):
        blk = mc["blocks"][b]
        mine = s.download("bv", blk=b).reshape((6,) + cases.ghost_shape(blk["mx"], blk["my"], blk["mz"], g))
        mine = mine[:, g:g + blk["mz"], g:g + blk["my"], g:g + blk["mx"]]
        ref = cases.block_interior_of_full(mc, b, one.download("bv"), 6)
        e = relerr(mine, ref)
        print(f"block {b} against the single block after 20 steps: {e:.2e}", flush=True)
        assert e < 1e-13


def test_exchange_stage_call(tmp_path):
    """ostc_exchange on its own against the oracle's exchange, with and without cv."""
    case, mc, s, M, _ = _c1_split(tmp_path, 0, True)
    s.step(3)
    M.run(3)
    for with_cv in (False, True):
        s.exchange(with_cv)
        M.exchange(with_cv)
        _check(s, M, 2, 5, 1e-12, f"exchange with_cv={with_cv}")


def test_c2_split_aweno_viscous(tmp_path):
    """3-D curvilinear viscous AWENO-5 + CDS with the isothermal wall, cut
    along i with the second block turned: 1 step at 1e-12, 20 at 1e-9."""
    case = cases.c2_case(28, 16, 12)
    mc = cases.split_case(case, axis=0, at=13, rotate=True)
    bv, sv = cases.initial_fields(case)
    fields = cases.split_fields(mc, bv, sv)
    print("c2 split", flush=True)
    s = ostc.solver_for_mesh_case(mc, str(tmp_path / "c2mb"), fields=fields)
    M = oracle_mesh(mc, fields)
    _check(s, M, 2, 5, 1e-13, "c2-split-init")
    s.step(1)
    M.run(1)
    _check(s, M, 2, 5, 1e-12, "c2-split-1")
    dt_p, res_p = s.step(19)
    dt_o, res_o = M.run(19)
    _check(s, M, 2, 5, 1e-9, "c2-split-20")
    _hist(dt_p, dt_o, res_p, res_o, "c2-split-20")


def test_c3_split_mixture(tmp_path, mech):
    """The 9-species mixing layer (HLLC + AWENO-5, viscous, finite-rate
    source) cut along j: 10 steps at 1e-9."""
    case = cases.c3_case(40, 24)
    h = ostc.Solver(str(cases.write_case_dir(case, str(tmp_path / "st"))), device=-1)
    states = {}
    for name in ("air_stream", "fuel_stream"):
        o = h.inflow_state(name, 7 + NS)
        states[name] = dict(rho=o[0], u=o[1], v=o[2], w=o[3], p=o[4], t=o[5], vel=o[6], sv=o[7:7 + NS])
    h.close()
    mw = np.array([mech.mw[i] for i in range(NS)])
    bv, sv = cases.c3_fields(case, states, mw)
    for b in case["bcs"]:
        if b["type"] == 5:
            b["state"] = states[b["stream"]]
    mc = cases.split_case(case, axis=1, at=11)
    fields = cases.split_fields(mc, bv, sv)
    print("c3 split", flush=True)
    s = ostc.solver_for_mesh_case(mc, str(tmp_path / "c3mb"), fields=fields)
    M = oracle_mesh(mc, fields, mech=mech)
    _check(s, M, 2, 14, 1e-12, "c3-split-init", ns=NS)
    dt_p, res_p = s.step(10)
    dt_o, res_o = M.run(10)
    _check(s, M, 2, 14, 1e-9, "c3-split-10", ns=NS)
    _hist(dt_p, dt_o, res_p, res_o, "c3-split-10")
