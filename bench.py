"""Benchmark: Mcell-updates/s of the SSP-RK3 step on the 3-D laminar
Navier-Stokes air case (256x128x128 curvilinear, AWENO-5 + CDS, AUSM+,
second-order viscous, CFL dt recomputed every stage, residual every step) on
one MI355X, FP64.  Also reported: the same step with FP32 flux arithmetic,
the FP64 point-implicit chemistry batch (2^22 cells, dt = 1e-7 s) in
cells/s, the roofline of the dominant kernel and a serial CPU baseline
(the oracle) on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  Inputs are synthetic (seeded), resident in HBM before
the timed region.
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "openstc-dev_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# compulsory HBM bytes per cell of one j- (or k-) direction sweep launch:
# reads cv(5) + bv(6) + mu + lambda + J + metric(9) = 23 doubles, read and
# write dq(5) = 10 doubles; halo re-reads are not counted
SWEEP_JK_BYTES_PER_CELL = 8 * (23 + 10)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline_flow(case_fn, steps=2):
    """Serial FP64 oracle on a bounded sample of the same workload (reduced
    grid, same scheme and boundaries)."""
    import helpers
    case = case_fn()
    z = helpers.oracle_zone(case)
    t0 = time.perf_counter()
    z.run(steps)
    t1 = time.perf_counter()
    cells = case["mx"] * case["my"] * case["mz"]
    return cells * steps / (t1 - t0) / 1e6, f"{case['mx']}x{case['my']}x{case['mz']} grid, {steps} RK3 steps", t1 - t0


def cpu_baseline_chem(n=1 << 20):
    import pyoracle
    from ostc import cases
    m = pyoracle.make_mech()
    rho, T, p, Y = cases.chem_batch_inputs(n, seed=1004)
    t0 = time.perf_counter()
    pyoracle.chem_batch(m, rho, T, Y, 1e-7)
    t1 = time.perf_counter()
    return n / (t1 - t0), f"{n} cells of the seeded batch", t1 - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--chem-cells", type=int, default=1 << 22)
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; multi-GPU sharding is described in DESIGN.md "
              "and not launched here", file=sys.stderr)
        sys.exit(2)

    import ostc
    from ostc import cases

    cells = a.nx * a.ny * a.nz
    result = {}
    with tempfile.TemporaryDirectory() as d:
        case = cases.c2_case(a.nx, a.ny, a.nz)
        log(f"bench: building case {a.nx}x{a.ny}x{a.nz}")
        s = ostc.solver_for_case(case, d)
        log("bench: warmup")
        s.step(a.warmup)
        s.sync()
        log(f"bench: timing {a.steps} FP64 steps")
        s.profile(True)
        t0 = time.perf_counter()
        dt_h, res_h = s.step(a.steps)
        s.sync()
        t1 = time.perf_counter()
        prof = s.profile_report()
        s.profile(False)
        fp64_s = t1 - t0
        # FP32-flux variant on the same state path
        s.set_flux_fp32(True)
        s.step(a.warmup)
        s.sync()
        log(f"bench: timing {a.steps} mixed-precision steps")
        t2 = time.perf_counter()
        s.step(a.steps)
        s.sync()
        t3 = time.perf_counter()
        mixed_s = t3 - t2
        finite = bool(np.isfinite(res_h).all())
        s.close()

        # chemistry batch (C4 inputs), FP64 point-implicit
        chem = None
        if a.chem_cells > 0:
            cc = cases.c1_case(20, 12)
            cc["setup"]["int species"] = 1
            cc["setup"]["int reaction"] = 1
            cc["setup"]["int chemSrcMethod"] = 1
            root = os.path.join(d, "chem")
            cases.write_case_dir(cc, root)
            sc = ostc.Solver(root, 0)
            n = a.chem_cells
            rho, T, p, Y = cases.chem_batch_inputs(n, seed=1004)
            dr, dT_, dY_ = (ostc.DeviceArray.from_host(sc, v) for v in (rho, T, Y))
            oY, oT = ostc.DeviceArray(sc, Y.nbytes), ostc.DeviceArray(sc, T.nbytes)
            sc.chem_batch(n, dr, dT_, dY_, 1e-7, oY, oT, mode=0)
            sc.profile(True)
            reps = 3
            t4 = time.perf_counter()
            for _ in range(reps):
                sc.chem_batch(n, dr, dT_, dY_, 1e-7, oY, oT, mode=0)
            sc.sync()
            t5 = time.perf_counter()
            cp = sc.profile_report()
            sc.close()
            chem = dict(cells_per_s=n * reps / (t5 - t4), kernel_ms=cp["chem_batch_fp64"][1] / cp["chem_batch_fp64"][0],
                        cells=n)

    ms_per_step = fp64_s / a.steps * 1e3
    value = cells * a.steps / fp64_s / 1e6
    # dominant kernel: the largest total event time among the sweeps
    sweeps = {k: v for k, v in prof.items() if k.startswith("sweep_")}
    dom = max(sweeps, key=lambda k: sweeps[k][1])
    calls, total_ms = sweeps[dom]
    avg_ms = total_ms / calls
    algo_bytes = SWEEP_JK_BYTES_PER_CELL * cells
    achieved = algo_bytes / (avg_ms * 1e-3) / 1e9
    stage_ms = {k: v[1] / v[0] for k, v in prof.items()}

    cpu = None
    if not a.no_cpu:
        log("bench: CPU baseline (serial oracle)")
        v, sample, secs = cpu_baseline_flow(lambda: cases.c2_case(128, 64, 64), steps=3)
        cpu = {"value": v, "unit": "Mcell-updates/s", "cores": 1, "kind": "port",
               "sample": f"FP64 serial oracle, {sample} of the same scheme/boundaries ({secs:.1f} s)"}
        if chem is not None:
            cv_, csample, csecs = cpu_baseline_chem()
            cpu["chem_cells_per_s"] = cv_
            cpu["chem_sample"] = f"{csample} ({csecs:.1f} s)"

    out = {
        "metric": "Mcell-updates/s per RK3 step; chem cells/s surrogate vs point-implicit",
        "value": value,
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 1% density perturbation on the uniform Mach 2.5 inflow)",
        "config": {"workload": f"3-D laminar NS air, curvilinear {a.nx}x{a.ny}x{a.nz}, AWENO-5+CDS, AUSM+ "
                               f"(Roe+WENO5 request), 2nd-order viscous, SSP-RK3 CFL 0.5, isothermal wall",
                   "cells": cells, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": algo_bytes, "avg_launch_ms": avg_ms},
        "cpu_baseline": cpu,
        "mixed_fp32_flux": {"value": cells * a.steps / mixed_s / 1e6, "unit": "Mcell-updates/s",
                            "ms_per_step": mixed_s / a.steps * 1e3},
        "chem_point_implicit_fp64": chem,
        "kernel_avg_ms": stage_ms,
        "residual_finite": finite,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
