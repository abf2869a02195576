"""The drop-in boundary's caller-memory and stage-level entry points on the GPU.

* ostc_bind_zone: a block whose state lives in caller-owned device arrays
  (DZone layout) evolves bit for bit like one in the library's own memory,
  and the kernels write the caller's arrays in place.
* The stage calls ostc_rhs / ostc_global_dt / ostc_rk3_stage /
  ostc_apply_bc / ostc_update_properties with ostc_store_last and
  ostc_residual reproduce the fused ostc_step_rk3 bit for bit (dt of every
  stage, residuals, state): the same kernels on the same data.
* ostc_local_dt writes the per-cell dt whose minimum ostc_global_dt returns.
* On 3-D air, where ostc_step_rk3 computes the CFL dt inside the k-sweep,
  the stage calls (stand-alone time-step pass) still match it bit for bit.
"""
import numpy as np
import pytest

import ostc
from ostc import cases

pytestmark = pytest.mark.gpu


def test_bind_zone_matches_library_memory(tmp_path):
    case = cases.c1_case(60, 30)
    a = ostc.solver_for_case(case, str(tmp_path / "a"))
    b = ostc.solver_for_case(case, str(tmp_path / "b"), initial=False)
    nbv, g, _ = b.field_info("bv")
    ncv, _, n_var = b.field_info("cv")
    ndq, _, _ = b.field_info("dq")
    mine = {k: ostc.DeviceArray(b, 8 * n) for k, n in (("bv", nbv), ("cv", ncv), ("qn", ncv), ("dq", ndq))}
    b.bind_zone(case["mx"], case["my"], case["mz"], g, n_var, 0, bv=mine["bv"], cv=mine["cv"], qn=mine["qn"],
                dq=mine["dq"])
    # a mismatched view is refused and leaves the binding alone
    with pytest.raises(ostc.OstcError):
        b.bind_zone(case["mx"] + 1, case["my"], case["mz"], g, n_var, 0, bv=mine["bv"], cv=mine["cv"])
    bv0, _ = cases.initial_fields(case)
    b.upload("bv", bv0)
    b.initialize()
    # the library works in the caller's arrays
    assert np.array_equal(mine["bv"].to_host((nbv,)), b.download("bv"))
    a.step(5)
    b.step(5)
    for f in ("bv", "cv", "dq"):
        assert np.array_equal(a.download(f), mine[f].to_host((b.field_info(f)[0],))), f
    assert np.array_equal(a.download("mul"), b.download("mul"))
    b.close()  # frees the library's arrays; the bound ones belong to the caller (here ostc_dev_alloc)
    a.close()


def test_stage_calls_match_fused_step(tmp_path):
    case = cases.c1_case(60, 30)
    a = ostc.solver_for_case(case, str(tmp_path / "a"))
    b = ostc.solver_for_case(case, str(tmp_path / "b"))
    dts, res = a.step(3)
    for step in (1, 2, 3):
        b.store_last()
        for rk in range(3):
            b.rhs()
            dt = b.global_dt()
            assert dt == dts[step - 1, rk], (step, rk)
            b.rk3_stage(dt, rk)
            b.apply_bc(True)
            b.update_properties()
        r, e = b.residual(step)
        print("step %d residual %s err_max %.3e" % (step, np.array2string(r, precision=3), e), flush=True)
        assert np.array_equal(r, res[step - 1, :4]) and e == res[step - 1, 4], step
    for f in ("bv", "cv", "mul", "acoustic"):
        assert np.array_equal(a.download(f), b.download(f)), f
    # ostc_local_dt alone writes the per-cell dt that ostc_global_dt minimises
    b.rhs()
    b.local_dt()
    loc = b.download("dt_local")
    g = b.global_dt()
    assert np.array_equal(loc, b.download("dt_local"))
    assert loc.min() == g
    a.close()
    b.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_stage_calls_match_fused_step_3d(tmp_path, fp32):
    """3-D air (AWENO-5 + CDS, viscous, curvilinear, ragged i extent): in
    ostc_step_rk3 the k-sweep computes each stage's CFL dt and stage 0 of
    the update stores qn; the stage calls run the stand-alone time-step pass
    (ostc_global_dt) instead.  Both give the same dt bits every stage, the
    same residuals and the same state."""
    case = cases.c2_case(70, 20, 12)
    print("3-D stage calls vs fused step, fp32 flux", fp32, flush=True)
    a = ostc.solver_for_case(case, str(tmp_path / "a"))
    b = ostc.solver_for_case(case, str(tmp_path / "b"))
    a.set_flux_fp32(fp32)
    b.set_flux_fp32(fp32)
    dts, res = a.step(3)
    for step in (1, 2, 3):
        b.store_last()
        for rk in range(3):
            b.rhs()
            dt = b.global_dt()
            assert dt == dts[step - 1, rk], (step, rk, dt, dts[step - 1, rk])
            b.rk3_stage(dt, rk)
            b.apply_bc(True)
            b.update_properties()
        r, e = b.residual(step)
        assert np.array_equal(r, res[step - 1, :4]) and e == res[step - 1, 4], step
    # dt_local of the last stage: written by the k-sweep in one solver, by
    # the stand-alone pass in the other, from the same state
    for f in ("bv", "cv", "mul", "acoustic", "dt_local"):
        assert np.array_equal(a.download(f), b.download(f)), f
    a.close()
    b.close()
