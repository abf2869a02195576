"""Conservative symmetric metrics on the GPU (the setup option
conservative_metric): device metrics equal to the host build bit for bit,
a uniform flow on a grid curved in all three directions kept to rounding by
the fused sweeps (AUSM+, HLLC, LF and Roe with MUSCL; AUSM+ and HLLC with
AWENO-5), and the perturbed flow on that grid stepping as the oracle does."""
import numpy as np
import pytest

import ostc
from ostc import cases
from helpers import BV_GROUPS, CV_GROUPS, grouped, oracle_zone, relerr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rec,nz", [(4, 12), (2, 12), (2, 1)])
def test_device_metrics_equal_host(tmp_path, rec, nz):
    c = cases.uniform_flow_case(20, 14, nz, rec, 4, conservative_metric=1)
    print("conservative metrics", rec, nz, flush=True)
    s = ostc.solver_for_case(c, str(tmp_path / "d"), initial=False)
    jd, md = s.geometry()
    h = ostc.Solver(str(tmp_path / "d"), device=-1)
    h.add_block(c["mx"], c["my"], c["mz"], c["x"], c["y"], c["z"])
    jh, mh = h.geometry()
    np.testing.assert_array_equal(jd, jh)
    np.testing.assert_array_equal(md, mh)
    s.close()
    h.close()


@pytest.mark.parametrize("rec,scheme", [(2, 3), (2, 4), (2, 1), (2, 2), (4, 3), (4, 4)])
def test_uniform_flow_kept_on_gpu(tmp_path, rec, scheme):
    c = cases.uniform_flow_case(40, 24, 20, rec, scheme, conservative_metric=1)
    print("uniform flow", rec, scheme, flush=True)
    s = ostc.solver_for_case(c, str(tmp_path / "u"))
    s.rhs()
    dq = s.download("dq").reshape(5, -1)
    st = c["inflow"]
    area = np.mean(s.download("jac")) ** (2 / 3)
    mom = st["rho"] * st["vel"]
    scale = np.array([st["rho"], mom, mom, mom, st["p"] / 0.4 + 0.5 * mom * st["vel"] + st["p"]]) * st["vel"] * area
    err = float(np.max(np.abs(dq) / scale[:, None]))
    print(f"uniform flow rhs {err:.2e}", flush=True)
    assert err < 1e-13
    s.close()


@pytest.mark.parametrize("rec,scheme", [(4, 4), (2, 1)])
def test_perturbed_flow_steps_match_oracle(tmp_path, rec, scheme):
    """The density-perturbed stream on the curved grid, viscous, HLLC with
    AWENO-5 and LF with MUSCL (the face vector of both changes with the
    option): RHS at 1e-12, 10 steps at 1e-9."""
    c = cases.uniform_flow_case(40, 24, 20, rec, scheme, conservative_metric=1)
    c["perturb"] = "density"
    print("perturbed flow", rec, scheme, flush=True)
    s = ostc.solver_for_case(c, str(tmp_path / "p"))
    z = oracle_zone(c)
    s.rhs()
    z.rhs()
    errs = grouped(s.download("dq"), z.field("dq"), 5, CV_GROUPS)
    print("rhs err", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    assert max(errs.values()) < 1e-12
    dt_p, _ = s.step(10)
    dt_o, _ = z.run(10)
    errs = grouped(s.download("bv"), z.field("bv"), 6, BV_GROUPS)
    print("10 steps bv err", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    assert max(errs.values()) < 1e-9
    assert relerr(dt_p, dt_o) < 1e-9
    s.close()
