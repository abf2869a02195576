"""Output and restart files on the GPU path: the Tecplot flow-field file,
step and residual-scale files written in the reference's layout
(FieldIO.h:58-600, Residual.cuh:123-139) checked with an independent
parser, and a restart from them (Initialize.cuh:64-313, Driver.cu:21-27)
continuing a run as the uninterrupted run does."""
import numpy as np
import pytest

import ostc
from ostc import cases
from helpers import relerr
import plt_format

pytestmark = pytest.mark.gpu

NS = 9


def _interior(a, mx, my, mz, g, ncomp):
    return np.asarray(a).reshape((ncomp,) + cases.ghost_shape(mx, my, mz, g))[:, g:g + mz, g:g + my, g:g + mx]


def _split_solver(tmp_path, name):
    case = cases.c1_case(40, 24)
    mc = cases.split_case(case, axis=0, at=16, rotate=True)
    bv, sv = cases.initial_fields(case)
    fields = cases.split_fields(mc, bv, sv)
    return mc, ostc.solver_for_mesh_case(mc, str(tmp_path / name), fields=fields), fields


def test_flowfield_file_layout(tmp_path):
    mc, s, _ = _split_solver(tmp_path, "w")
    s.step(7)
    out = tmp_path / "run"
    s.write_flowfield(out)
    f = plt_format.read(str(out / "output" / "field" / "flowfield.plt"))
    assert f["names"] == ["x", "y", "z", "density", "u", "v", "w", "pressure", "temperature", "mach"]
    t, step = s.time()
    g = mc["ngg"]
    for b, (blk, z) in enumerate(zip(mc["blocks"], f["zones"])):
        assert z["dims"] == (blk["mx"], blk["my"], blk["mz"]) and z["time"] == t
        np.testing.assert_array_equal(z["data"]["x"], np.asarray(blk["x"]).ravel())
        bv = _interior(s.download("bv", blk=b), blk["mx"], blk["my"], 1, g, 6).reshape(6, -1)
        for l, n in enumerate(("density", "u", "v", "w", "pressure", "temperature")):
            np.testing.assert_array_equal(z["data"][n], bv[l])
        mach = _interior(s.download("mach", blk=b), blk["mx"], blk["my"], 1, g, 1).ravel()
        np.testing.assert_array_equal(z["data"]["mach"], mach)
        assert z["minmax"][3].tolist() == [bv[0].min(), bv[0].max()]
    assert int(open(out / "output" / "message" / "step.txt").read()) == step == 7
    scale = np.loadtxt(out / "output" / "message" / "residual_scale.txt")
    assert scale.shape == (4,) and np.all(scale > 0)


def test_restart_continues_the_run(tmp_path):
    """10 steps, write, restart a fresh solver from the files, 10 more steps:
    the same state, step count, time and residual history as 20 steps of
    one solver (the restart rebuilds cv from bv, so agreement is to
    rounding, 1e-12)."""
    mc, s, fields = _split_solver(tmp_path, "a")
    s.step(10)
    out = tmp_path / "run"
    s.write_flowfield(out)
    dt_a, res_a = s.step(10)
    r = ostc.solver_for_mesh_case(mc, str(tmp_path / "b"), fields=None)
    for b in range(2):
        r.initialize_uniform(b)
    r.read_flowfield(out)
    r.initialize()
    assert r.time()[1] == 10 and r.time()[0] == pytest.approx(s.time()[0] - dt_a[:, 2].sum(), rel=1e-14)
    dt_b, res_b = r.step(10)
    g = mc["ngg"]
    for b, blk in enumerate(mc["blocks"]):
        e = relerr(_interior(r.download("bv", blk=b), blk["mx"], blk["my"], 1, g, 6),
                   _interior(s.download("bv", blk=b), blk["mx"], blk["my"], 1, g, 6))
        print(f"restart block {b}: bv {e:.2e}", flush=True)
        assert e < 1e-12
    assert r.time()[1] == s.time()[1] == 20
    assert relerr(dt_b, dt_a) < 1e-12
    # rho and T carry the signal of the advected density field (|V| and p
    # residuals are rounding noise at uniform velocity and pressure)
    assert np.max(np.abs(res_b[:, [0, 3]] - res_a[:, [0, 3]]) / np.abs(res_a[:, [0, 3]])) < 1e-9


def test_restart_maps_variables_by_name(tmp_path):
    """A file from another writer with renamed, reordered and extra
    variables: density as RHO, pressure as P, temperature as T, an unknown
    variable skipped, and no Mach number."""
    case = cases.c1_case(20, 12)
    s = ostc.solver_for_case(case, str(tmp_path / "c"), initial=False)
    s.initialize_uniform()
    rng = np.random.default_rng(5)
    n = 20 * 12
    vals = {"RHO": 1 + 0.1 * rng.random(n), "U": 600 + rng.random(n), "V": rng.random(n), "W": np.zeros(n),
            "P": 1e5 * (1 + 0.1 * rng.random(n)), "T": 300 + 10 * rng.random(n), "junk": rng.random(n)}
    names = ["x", "y", "z", "junk", "T", "P", "V", "U", "W", "RHO"]
    data = dict(vals, x=np.asarray(case["x"]).ravel(), y=np.asarray(case["y"]).ravel(),
                z=np.asarray(case["z"]).ravel())
    (tmp_path / "in" / "output" / "field").mkdir(parents=True)
    (tmp_path / "in" / "output" / "message").mkdir(parents=True)
    plt_format.write(str(tmp_path / "in" / "output" / "field" / "flowfield.plt"), names, [((20, 12, 1), data)],
                     time=2.5e-4)
    (tmp_path / "in" / "output" / "message" / "step.txt").write_text("42")
    s.read_flowfield(tmp_path / "in")
    bv = _interior(s.download("bv"), 20, 12, 1, case["ngg"], 6).reshape(6, -1)
    for l, key in enumerate(("RHO", "U", "V", "W", "P", "T")):
        np.testing.assert_array_equal(bv[l], vals[key])
    s.initialize()
    assert s.time() == (2.5e-4, 42)


def test_mixture_sst_file_names_and_fill(tmp_path, mech):
    """The 9-species + SST channel writes species by name, tke, omega and mut;
    a laminar file read into it leaves k and omega at the inflow state."""
    from test_gpu_sst import _c5
    case, s, _ = _c5(tmp_path, mech)
    s.step_steady(2)
    s.write_flowfield(tmp_path / "run")
    f = plt_format.read(str(tmp_path / "run" / "output" / "field" / "flowfield.plt"))
    names = f["names"]
    assert names[10:19] == ["H2", "H", "O2", "O", "OH", "HO2", "H2O2", "H2O", "N2"]
    assert names[19:] == ["tke", "omega", "mut"]
    g = case["ngg"]
    sv = _interior(s.download("sv"), case["mx"], case["my"], case["mz"], g, NS + 2).reshape(NS + 2, -1)
    np.testing.assert_array_equal(f["zones"][0]["data"]["omega"], sv[NS + 1])
    # drop tke / omega / mut: a laminar result read into the SST case
    z = f["zones"][0]
    keep = names[:19]
    (tmp_path / "lam" / "output" / "field").mkdir(parents=True)
    plt_format.write(str(tmp_path / "lam" / "output" / "field" / "flowfield.plt"), keep,
                     [(z["dims"], {n: z["data"][n] for n in keep})])
    s.read_flowfield(tmp_path / "lam")
    o = s.inflow_state("vitiated_air", 7 + NS + 2)
    sv2 = _interior(s.download("sv"), case["mx"], case["my"], case["mz"], g, NS + 2).reshape(NS + 2, -1)
    assert np.all(sv2[NS] == o[7 + NS]) and np.all(sv2[NS + 1] == o[8 + NS])
    np.testing.assert_array_equal(sv2[:NS], sv[:NS])
