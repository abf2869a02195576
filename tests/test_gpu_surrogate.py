"""GPU tests of the MLP chemistry surrogate: the BF16-MFMA forward and
backward kernels against the numpy oracle, training with the project's own
kernels, accuracy against the FP64 point-implicit integrator, the gate and
its fallback, and the ablation against the integrator alone."""
import os
import time

import numpy as np
import pytest

import mlp_oracle as mo
import ostc
from ostc import cases

pytestmark = pytest.mark.gpu

DT = 1e-7
MAJOR = [0, 2, 7, 8]  # H2, O2, H2O, N2


def relerr(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def solver(tmp_path_factory):
    case = cases.c1_case(20, 12)
    root = str(tmp_path_factory.mktemp("sur"))
    case["setup"]["int species"] = 1
    case["setup"]["int reaction"] = 1
    case["setup"]["int chemSrcMethod"] = 1
    cases.write_case_dir(case, root)
    return ostc.Solver(root, 0)


def _batch(s, n, seed, T=None):
    rho, t, p, Y = cases.chem_batch_inputs(n, seed=seed)
    if T is not None:
        t = np.full(n, T)
        mwinv = (Y / np.array([2.016, 1.008, 31.998, 15.999, 17.007, 33.006, 34.014, 18.015, 28.014])[:, None]).sum(0)
        rho = p / (cases.R_U * t * mwinv)
    d = {k: ostc.DeviceArray.from_host(s, v) for k, v in dict(rho=rho, T=t, Y=Y).items()}
    return d, (rho, t, p, Y)


def _run(s, d, host, mode):
    rho, t, p, Y = host
    n = t.size
    dY = ostc.DeviceArray(s, Y.nbytes)
    dT = ostc.DeviceArray(s, t.nbytes)
    nf = s.chem_batch(n, d["rho"], d["T"], d["Y"], DT, dY, dT, mode=mode)
    out = dY.to_host(Y.shape), dT.to_host(t.shape)
    dY.free()
    dT.free()
    return out + (nf,)


@pytest.mark.parametrize("B", [64, 256])
def test_forward_matches_oracle(solver, B):
    p = mo.init_params(7).astype(np.float32)
    solver.surrogate_set_params(p)
    assert np.array_equal(solver.surrogate_params(), p)
    x = np.random.default_rng(B).standard_normal((B, 11)).astype(np.float32)
    y = solver.surrogate_forward(x)
    yb, _ = mo.forward(p.astype(float), x.astype(float), bf=True)
    yf, _ = mo.forward(p.astype(float), x.astype(float), bf=False)
    e_b, e_f = relerr(y, yb), relerr(y, yf)
    print("B=%d forward vs BF16-rounded oracle %.2e, vs FP64 oracle %.2e" % (B, e_b, e_f), flush=True)
    assert e_b < 1e-3
    assert e_f < 3e-2


@pytest.mark.parametrize("B", [64, 2048])  # 2048 runs the split-K weight-gradient path
def test_gradient_matches_oracle(solver, B):
    p = mo.init_params(8)
    layers = mo.unpack(p)
    rng = np.random.default_rng(100 + B)
    layers = [(W, 0.1 * rng.standard_normal(b.shape)) for W, b in layers]
    p = mo.pack(layers).astype(np.float32)
    solver.surrogate_set_params(p)
    x = rng.standard_normal((B, 11)).astype(np.float32)
    t = rng.standard_normal((B, 8)).astype(np.float32)
    g, loss = solver.surrogate_gradient(x, t)
    lb, gb = mo.loss_and_grad(p.astype(float), x.astype(float), t.astype(float), bf=True)
    lf, gf = mo.loss_and_grad(p.astype(float), x.astype(float), t.astype(float), bf=False)
    assert abs(loss - lb) <= 1e-5 * lb
    G, Gb, Gf = mo.unpack(g), mo.unpack(gb), mo.unpack(gf)
    for l in range(4):
        ew, eb = relerr(G[l][0], Gb[l][0]), relerr(G[l][1], Gb[l][1])
        print("B=%d layer %d: dW %.2e db %.2e (vs FP64 oracle %.2e %.2e)" % (
            B, l, ew, eb, relerr(G[l][0], Gf[l][0]), relerr(G[l][1], Gf[l][1])), flush=True)
        assert ew < 2e-3 and eb < 2e-3
        assert relerr(G[l][0], Gf[l][0]) < 5e-2
    # the padding of the packed layout receives no gradient
    live = mo.pack([(np.ones_like(W), np.ones_like(b)) for W, b in layers]) != 0
    assert not g[~live].any()


@pytest.fixture(scope="module")
def trained(solver):
    n = 1 << 22
    d, _ = _batch(solver, n, seed=2004)
    solver.surrogate_init(11)
    t0 = time.time()
    hist = solver.surrogate_train(n, d["rho"], d["T"], d["Y"], DT, epochs=12, batch=8192, seed=2004)
    print("trained on %d pairs x 12 epochs in %.2f s; loss %s" % (n, time.time() - t0, np.array2string(hist, precision=4)),
          flush=True)
    for v in d.values():
        v.free()
    return hist


def test_training_reduces_loss(trained):
    assert np.all(np.isfinite(trained))
    assert trained[-1] < 0.2 * trained[0]


def test_surrogate_accuracy_and_conservation(solver, trained, mech):
    n = 1 << 17
    d, host = _batch(solver, n, seed=1004)
    rho, T, p, Y = host
    Y0, T0, _ = _run(solver, d, host, 0)
    Y1, T1, nf = _run(solver, d, host, 1)
    err = np.abs(Y1 - Y0)
    eT = np.abs((T + T1) - (T + T0)) / (T + T0)
    print("surrogate vs FP64 integrator: fallback %.4f, major Linf %.3e, all Linf %.3e, T rel max %.3e, "
          "peak T %.2f vs %.2f" % (nf / n, err[MAJOR].max(), err.max(), eT.max(), (T + T1).max(), (T + T0).max()),
          flush=True)
    assert err[MAJOR].max() <= 1e-3
    assert eT.max() <= 1e-2
    assert abs((T + T1).max() - (T + T0).max()) <= 1e-2 * (T + T0).max()
    # about 0.1 after the benchmark's 2^24 pairs x 16 epochs; this shorter
    # training leaves more cells to the integrator
    assert nf < 0.3 * n
    # cells that fell back carry the integrator's result exactly
    same = np.all(Y1 == Y0, axis=0) & (T1 == T0)
    assert same.sum() >= nf
    # element and mass conservation of the accepted increments
    nm = solver.surrogate_norm()
    acc = ~same
    mw = np.array([2.016, 1.008, 31.998, 15.999, 17.007, 33.006, 34.014, 18.015, 28.014])
    comp = np.array([[2, 0, 0], [1, 0, 0], [0, 2, 0], [0, 1, 0], [1, 1, 0], [1, 2, 0], [2, 2, 0], [2, 1, 0], [0, 0, 2]])
    aw = np.array([1.008, 15.999, 14.007])
    rows = (comp * aw[None, :] / mw[:, None]).T
    assert np.abs(Y1[:, acc].sum(axis=0)).max() < 1e-13
    assert np.abs(rows @ Y1[:, acc]).max() < 1e-13
    assert np.allclose(nm["basis"].T @ nm["basis"], np.eye(6), atol=1e-14)
    for v in d.values():
        v.free()


def test_fused_network_matches_layer_path(solver, trained):
    """The fused inference kernels (128 and 64 cells per workgroup, hidden
    tiles computed transposed) against one GEMM launch per layer on the same
    trained weights: BF16 operands and FP32 accumulation in both, so the
    standardised outputs agree to BF16 rounding (ragged count: not a
    multiple of either workgroup size)."""
    n = 5000
    d, _ = _batch(solver, n, seed=606)
    outs = {}
    for kernel in (0, 1, 2, 3):
        solver.surrogate_set_kernel(kernel)
        outs[kernel] = solver.surrogate_network_outputs(n, d["rho"], d["T"], d["Y"])
    solver.surrogate_set_kernel(3)
    assert np.array_equal(outs[3], outs[1])  # the pipelined kernel reorders nothing
    for kernel in (1, 2, 3):
        e = relerr(outs[kernel], outs[0])
        print("fused kernel %d vs layer path: rel. error %.2e" % (kernel, e), flush=True)
        assert np.all(np.isfinite(outs[kernel]))
        assert e < 1e-2
    for v in d.values():
        v.free()


def test_gate_sends_out_of_range_and_wrong_dt_to_fallback(solver, trained):
    n = 4096
    d, host = _batch(solver, n, seed=77, T=3000.0)  # outside the 900-2400 K training range
    Y0, T0, _ = _run(solver, d, host, 0)
    Y1, T1, nf = _run(solver, d, host, 1)
    assert nf == n
    assert np.array_equal(Y0, Y1) and np.array_equal(T0, T1)
    for v in d.values():
        v.free()
    d, host = _batch(solver, n, seed=78)
    rho, T, p, Y = host
    dY = ostc.DeviceArray(solver, Y.nbytes)
    dT = ostc.DeviceArray(solver, T.nbytes)
    nf = solver.chem_batch(n, d["rho"], d["T"], d["Y"], 2 * DT, dY, dT, mode=1)
    assert nf == n
    for v in list(d.values()) + [dY, dT]:
        v.free()


def test_save_load_round_trip(solver, trained, tmp_path, tables):
    path = str(tmp_path / "sur.bin")
    solver.surrogate_save(path)
    n = 8192
    d, host = _batch(solver, n, seed=91)
    Ya, Ta, nfa = _run(solver, d, host, 1)
    other = ostc.Solver(solver.root, 0)
    other.surrogate_load(path)
    dd = {k: ostc.DeviceArray.from_host(other, v) for k, v in zip(("rho", "T", "Y"), (host[0], host[1], host[3]))}
    Yb, Tb, nfb = _run(other, dd, host, 1)
    assert nfa == nfb
    assert np.array_equal(Ya, Yb) and np.array_equal(Ta, Tb)
    other.close()
    for v in d.values():
        v.free()


def test_ablation_network_versus_integrator_alone(solver, trained):
    """What the trained network gains: the same batch through the FP64
    integrator alone and through network + gate + fallback; an untrained
    network with the same standardisation is far off (the gain is learned)."""
    n = 1 << 20
    d, host = _batch(solver, n, seed=1005)
    rho, T, p, Y = host
    dY = ostc.DeviceArray(solver, Y.nbytes)
    dT = ostc.DeviceArray(solver, T.nbytes)
    times = {}
    for mode in (0, 1, 0, 1):
        solver.sync()
        t0 = time.perf_counter()
        nf = solver.chem_batch(n, d["rho"], d["T"], d["Y"], DT, dY, dT, mode=mode)
        times[mode] = time.perf_counter() - t0
    Y1, T1 = dY.to_host(Y.shape), dT.to_host(T.shape)
    solver.chem_batch(n, d["rho"], d["T"], d["Y"], DT, dY, dT, mode=0)
    Y0, T0 = dY.to_host(Y.shape), dT.to_host(T.shape)
    e_trained = np.abs(Y1 - Y0)[MAJOR].max()
    print("ablation n=%d: integrator %.2f ms, surrogate %.2f ms (x%.2f), fallback %.3f, major Linf %.2e" % (
        n, times[0] * 1e3, times[1] * 1e3, times[0] / times[1], nf / n, e_trained), flush=True)
    assert times[1] < times[0]
    # untrained network: keep the trained standardisation, replace the weights
    w = solver.surrogate_params()
    solver.surrogate_set_params(mo.init_params(12).astype(np.float32))
    nf_u = solver.chem_batch(n, d["rho"], d["T"], d["Y"], DT, dY, dT, mode=1)
    Yu = dY.to_host(Y.shape)
    e_untrained = np.abs(Yu - Y0)[MAJOR].max()
    acc = np.any(Yu != Y0, axis=0)
    e_med = np.median(np.abs(Yu - Y0)[:, acc].max(axis=0)) if acc.any() else 0.0
    e_med_t = np.median(np.abs(Y1 - Y0)[:, np.any(Y1 != Y0, axis=0)].max(axis=0))
    print("untrained network: fallback %.3f, major Linf %.2e, median accepted-cell error %.2e (trained %.2e)" % (
        nf_u / n, e_untrained, e_med, e_med_t), flush=True)
    assert e_med > 10 * e_med_t or nf_u > 0.9 * n
    solver.surrogate_set_params(w)
    for v in list(d.values()) + [dY, dT]:
        v.free()
