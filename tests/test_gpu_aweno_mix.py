"""GPU parity of the multi-species / transported-scalar AWENO-5 path
(characteristic interpolation on the frozen multi-species Euler eigensystem,
CDS correction on every conservative component) against the FP64 oracle on
identical seeded inputs: the reduced C3 mixing layer with HLLC + AWENO-5 as
configured (RHS, one step, 100 steps with the residual history), AUSM+ on the
same layer, the FP32-flux variant with its per-field errors, the reduced C5
channel (9 species + k, omega, steady DPLUR) with AWENO-5, and laminar air
carrying k and omega (the SST plate).

Bars: 1e-12 after one RHS evaluation or step, 1e-9 after 100 steps (C3) and
20 pseudo-iterations (C5)."""
import numpy as np
import pytest

from helpers import grouped, relerr
from test_gpu_mixture import CV14, _compare as _compare_mix, _setup as _setup_c3
from test_gpu_sst import _c5, _compare as _compare_sst, _plate

pytestmark = pytest.mark.gpu


def _residual_history(res_p, res_o, label, tol=1e-9):
    rel = np.abs(res_p[:, :5] - res_o[:, :5]) / np.abs(res_o[:, :5])
    print(f"{label} residual history: {res_o.shape[0]} entries, max rel {rel.max():.2e}", flush=True)
    assert np.isfinite(res_p).all()
    assert rel.max() < tol, (label, rel.max(axis=0))


def test_c3_aweno_rhs_and_one_step(tmp_path, mech):
    case, s, z, _ = _setup_c3(tmp_path, mech, recon=4)
    assert case["ngg"] == 3
    _compare_mix(s, z, 1e-13, "c3-aweno-init")
    s.rhs()
    z.rhs()
    errs = grouped(s.download("dq"), z.field("dq"), 14, CV14)
    print("c3 aweno rhs err", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    assert max(errs.values()) < 1e-12
    dt_p, _ = s.step(1)
    dt_o, _ = z.run(1)
    assert relerr(dt_p, dt_o) < 1e-12
    _compare_mix(s, z, 1e-12, "c3-aweno-1step")


def test_c3_aweno_hundred_steps(tmp_path, mech):
    case, s, z, _ = _setup_c3(tmp_path, mech, recon=4)
    dt_p, res_p = s.step(100)
    dt_o, res_o = z.run(100)
    print("c3 aweno 100-step dt history max rel %.2e" % (np.abs(dt_p - dt_o) / dt_o).max(), flush=True)
    assert (np.abs(dt_p - dt_o) / dt_o).max() < 1e-9
    _compare_mix(s, z, 1e-9, "c3-aweno-100")
    _residual_history(res_p, res_o, "c3-aweno-100")


def test_c3_aweno_ausm_plus(tmp_path, mech):
    case, s, z, _ = _setup_c3(tmp_path, mech, scheme=3, recon=4)
    s.rhs()
    z.rhs()
    errs = grouped(s.download("dq"), z.field("dq"), 14, CV14)
    print("c3 aweno ausm+ rhs err", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    assert max(errs.values()) < 1e-12
    s.step(10)
    z.run(10)
    _compare_mix(s, z, 1e-10, "c3-aweno-ausm-10")


def test_c3_aweno_mixed_precision_errors(tmp_path, mech):
    """FP32 characteristic interpolation, Riemann and viscous arithmetic (FP64
    side temperatures, chemistry and accumulation): per-field L2 / Linf
    errors against the FP64 oracle after 10 steps."""
    case, s, z, _ = _setup_c3(tmp_path, mech, fp32=True, recon=4)
    s.step(10)
    z.run(10)
    bp, bo = s.download("bv").reshape(6, -1), z.field("bv").reshape(6, -1)
    sp, so = s.download("sv").reshape(9, -1), z.field("sv").reshape(9, -1)
    rep = {}
    for name, a, b in (("rho", bp[0], bo[0]), ("u", bp[1], bo[1]), ("p", bp[4], bo[4]), ("T", bp[5], bo[5]),
                       ("Y", sp, so)):
        rep[name] = (np.linalg.norm(a - b) / np.linalg.norm(b), np.max(np.abs(a - b)) / np.max(np.abs(b)))
    print("c3 aweno mixed precision L2/Linf", {k: "%.1e/%.1e" % v for k, v in rep.items()}, flush=True)
    assert max(v[1] for v in rep.values()) < 1e-4


def test_c5_aweno_steady(tmp_path, mech):
    case, s, z = _c5(tmp_path, mech, recon=4)
    assert case["ngg"] == 3
    _compare_sst(s, z, 1e-13, "c5-aweno-init", ns=9)
    res_p = s.step_steady(1)
    res_o = z.run_steady(1)
    _compare_sst(s, z, 1e-12, "c5-aweno-1", ns=9)
    res_p = s.step_steady(19)
    res_o = z.run_steady(19)
    _compare_sst(s, z, 1e-9, "c5-aweno-20", ns=9)
    _residual_history(res_p, res_o, "c5-aweno-20")


def test_sst_plate_aweno(tmp_path):
    """Laminar air carrying k and omega through the AWENO path (one
    characteristic wave per scalar, gamma = 1.4): RHS, one step and 20 RK3
    steps."""
    case, s, z = _plate(tmp_path, nx=48, ny=32, reconstruction=4)
    assert case["ngg"] == 3
    s.rhs()
    z.rhs()
    dq_p, dq_o = s.download("dq").reshape(7, -1), z.field("dq").reshape(7, -1)
    errs = {"flow": relerr(dq_p[:5], dq_o[:5]), "k": relerr(dq_p[5], dq_o[5]), "omega": relerr(dq_p[6], dq_o[6])}
    print("sst plate aweno rhs err", {k: "%.2e" % v for k, v in errs.items()}, flush=True)
    assert max(errs.values()) < 1e-12
    s.step(1)
    z.run(1)
    _compare_sst(s, z, 1e-12, "plate-aweno-1")
    s.step(19)
    z.run(19)
    _compare_sst(s, z, 1e-9, "plate-aweno-20")
