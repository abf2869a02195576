"""The allocation-free chemistry batch (ostc_chem_batch_noalloc) against
ostc_chem_batch through the C ABI: the same kernels on a fixed per-device
scratch array, so the results must be bit-identical, on two calls in a row
(the second smaller, so stale scratch contents of the first must not leak),
in both modes, and sizes above the array's limit are refused."""
import numpy as np
import pytest

import ostc
from ostc import cases

pytestmark = pytest.mark.gpu

DT = 1e-7
LIMIT = 1 << 20


@pytest.fixture(scope="module")
def solver(tmp_path_factory):
    case = cases.c1_case(20, 12)
    root = str(tmp_path_factory.mktemp("noalloc"))
    case["setup"]["int species"] = 1
    case["setup"]["int reaction"] = 1
    case["setup"]["int chemSrcMethod"] = 1
    cases.write_case_dir(case, root)
    s = ostc.Solver(root, 0)
    n = 1 << 16
    rho, t, _, Y = cases.chem_batch_inputs(n, seed=77)
    d = {k: ostc.DeviceArray.from_host(s, v) for k, v in dict(rho=rho, T=t, Y=Y).items()}
    s.surrogate_init(5)
    s.surrogate_train(n, d["rho"], d["T"], d["Y"], DT, epochs=2, batch=8192, seed=77)
    for v in d.values():
        v.free()
    yield s
    s.close()


def _call(s, fn, n, seed, mode):
    rho, t, _, Y = cases.chem_batch_inputs(n, seed=seed)
    d = {k: ostc.DeviceArray.from_host(s, v) for k, v in dict(rho=rho, T=t, Y=Y).items()}
    dY = ostc.DeviceArray.from_host(s, np.full(Y.shape, 7.0))
    dT = ostc.DeviceArray.from_host(s, np.full(t.shape, 7.0))
    nf = fn(n, d["rho"], d["T"], d["Y"], DT, dY, dT, mode=mode)
    out = dY.to_host(Y.shape), dT.to_host(t.shape), nf
    for v in list(d.values()) + [dY, dT]:
        v.free()
    return out


@pytest.mark.parametrize("mode", [1, 0])
def test_noalloc_matches_allocating_call_twice(solver, mode):
    for n, seed in ((300_001, 11), (1_000, 12)):
        print(f"noalloc mode {mode}: n={n} seed={seed}", flush=True)
        y_ref, t_ref, nf_ref = _call(solver, solver.chem_batch, n, seed, mode)
        y_na, t_na, nf_na = _call(solver, solver.chem_batch_noalloc, n, seed, mode)
        assert nf_na == nf_ref
        assert np.array_equal(y_na, y_ref)
        assert np.array_equal(t_na, t_ref)
        assert np.isfinite(y_na).all() and np.isfinite(t_na).all()
        if mode == 1:
            assert 0 <= nf_na <= n


def test_noalloc_refuses_batches_above_its_scratch(solver):
    small = ostc.DeviceArray.from_host(solver, np.zeros(16))
    with pytest.raises(RuntimeError, match="at most"):
        solver.chem_batch_noalloc(LIMIT + 1, small, small, small, DT, small, small, mode=1)
    small.free()
