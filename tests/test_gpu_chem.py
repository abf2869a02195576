"""GPU parity of the FP64 point-implicit chemistry batch against the oracle."""
import numpy as np
import pytest

import ostc
import pyoracle as po
from ostc import cases
from helpers import relerr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def chem_solver(tmp_path_factory):
    case = cases.c1_case(20, 12)
    root = str(tmp_path_factory.mktemp("chem"))
    case["setup"]["int species"] = 1
    case["setup"]["int reaction"] = 1
    case["setup"]["int chemSrcMethod"] = 1
    cases.write_case_dir(case, root)
    return ostc.Solver(root, 0)


def test_chem_batch_parity(chem_solver, mech):
    s = chem_solver
    n = 4096
    rho, T, p, Y = cases.chem_batch_inputs(n, seed=1004)
    dt = 1e-7
    d = {k: ostc.DeviceArray.from_host(s, v) for k, v in dict(rho=rho, T=T, Y=Y).items()}
    dY = ostc.DeviceArray(s, Y.nbytes)
    dT = ostc.DeviceArray(s, T.nbytes)
    nf = s.chem_batch(n, d["rho"], d["T"], d["Y"], dt, dY, dT, mode=0)
    assert nf == 0
    gY = dY.to_host(Y.shape)
    gT = dT.to_host(T.shape)
    oY, oT = po.chem_batch(mech, rho, T, Y, dt)
    eY = relerr(gY, oY)
    eT = relerr(gT, oT)
    # per-cell relative error of dT
    cell = np.abs(gT - oT) / np.maximum(np.abs(oT), 1e-300)
    print("chem batch dY err %.2e dT err %.2e worst cell %.2e" % (eY, eT, cell.max()), flush=True)
    assert eY < 1e-11
    assert eT < 1e-11
    assert np.all(np.isfinite(gY)) and np.all(np.isfinite(gT))


@pytest.mark.parametrize("n", [1, 129, 1000])
def test_chem_batch_ragged_sizes(chem_solver, mech, n):
    """Batch sizes that leave a partial workgroup (the kernel runs 128 cells
    per workgroup): every cell matches the oracle; n = 0 is a no-op."""
    s = chem_solver
    rho, T, p, Y = cases.chem_batch_inputs(n, seed=2000 + n)
    dt = 1e-7
    d = {k: ostc.DeviceArray.from_host(s, v) for k, v in dict(rho=rho, T=T, Y=Y).items()}
    dY = ostc.DeviceArray(s, Y.nbytes)
    dT = ostc.DeviceArray(s, T.nbytes)
    assert s.chem_batch(n, d["rho"], d["T"], d["Y"], dt, dY, dT, mode=0) == 0
    gY, gT = dY.to_host(Y.shape), dT.to_host(T.shape)
    oY, oT = po.chem_batch(mech, rho, T, Y, dt)
    print("chem batch n=%d dY err %.2e dT err %.2e" % (n, relerr(gY, oY), relerr(gT, oT)), flush=True)
    assert relerr(gY, oY) < 1e-11 and relerr(gT, oT) < 1e-11
    # an empty batch leaves the outputs alone
    assert s.chem_batch(0, d["rho"], d["T"], d["Y"], dt, dY, dT, mode=0) == 0
    assert np.array_equal(dT.to_host(T.shape), gT)
