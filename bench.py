"""Benchmark: Mcell-updates/s of the SSP-RK3 step on the 3-D laminar
Navier-Stokes air case (256x128x128 curvilinear, AWENO-5 + CDS, AUSM+,
second-order viscous, CFL dt recomputed every stage, residual every step) on
one MI355X, FP64.  Also reported: the same step with FP32 flux arithmetic;
the 2-D reacting 9-species mixing layer (512x256, HLLC + AWENO-5, viscous,
finite-rate source) in FP64 and mixed precision; the chemistry batch (2^22
cells, dt = 1e-7 s) with the FP64 point-implicit integrator and with the
BF16-MFMA surrogate + FP64 fallback (trained first, untimed, on 2^24 pairs
from the integrator) and the surrogate's error against the integrator; the
rooflines of the dominant sweep (HBM), the plain chemistry (FP64) and the
surrogate (MFMA); a serial CPU baseline (the oracle) on a bounded sample.

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
# compulsory HBM bytes of one fully fused SSP-RK3 step of a 3-D air cell:
# 8 (5 n_var reads + 3 n_var writes + 3 n_geom), n_var = 5, n_geom = 10
STEP_COMPULSORY_BYTES = 8 * (5 * 5 + 3 * 5 + 3 * 10)
BF16_MFMA_PEAK_TFLOPS = 2500.0

# surrogate network FLOP per cell: 2 (11*256 + 256*256 + 256*256 + 256*8)
SURROGATE_FLOP_PER_CELL = 2 * (11 * 256 + 2 * 256 * 256 + 256 * 8)

# compulsory HBM bytes per cell of one j- (or k-) direction sweep launch:
# reads cv(5) + bv(6) + mu + lambda + J + metric(9) = 23 doubles, read and
# write dq(5) = 10 doubles; halo re-reads are not counted
SWEEP_JK_BYTES_PER_CELL = 8 * (23 + 10)


def chem_flop_per_cell(labels, ns=9):
    """Static operation count of one point-implicit cell step as the FP64
    kernel executes it (+ - * / = 1, FMA = 2, exp/log/pow = 1): state and
    energy 154, concentrations and g/RT 143, per reaction 325 (Arrhenius 5,
    kb 21, progress with pow over all species 38, creation/destruction 72,
    Jacobian row updates 189) plus the third-body (19), duplicate (6) and
    Troe fall-off (78) extras, source and Jacobian scaling 279, I - dt J 162,
    9x9 LU with substitution 650, new Y 27, two Newton iterations 542."""
    per = {0: 325, 1: 325, 2: 325, 3: 331, 4: 344, 5: 403, 6: 403, 7: 406}
    return 154 + 143 + sum(per.get(int(l), 325) for l in labels) + 279 + 162 + 650 + 27 + 542 + 10


# kernel behind each timer name of the air case (AWENO-5 + CDS, FP64)
# (the k-sweep of the one-block RK3 step also computes the CFL time step)
SWEEP_KERNELS = {"sweep_i": "sweep_i_aweno<double>", "sweep_j": "sweep_jk_aweno<1, 4, double, false>",
                 "sweep_k": "sweep_jk_aweno<2, 4, double, true>"}
# algorithmic bytes per cell of each sweep: the i-sweep reads cv 5, bv 6,
# mu, lambda, J, metric 9 and writes dq 5 (23 + 5 doubles); the j-sweep
# also reads dq (23 + 10); the k-sweep with the time step reads the other
# two metric rows and the sound speed and writes dt (23 + 10 + 8)
SWEEP_BYTES_PER_CELL = {"sweep_i": 8 * (23 + 5), "sweep_j": 8 * (23 + 10), "sweep_k": 8 * (23 + 10 + 8)}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "air_256x128x128_hbm_traffic.md")


def measured(key):
    """A figure recorded in the committed counter summary as a line
    "<key> = <value>", or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith(key + " = "):
                    return float(line.rsplit("=", 1)[1])
    except OSError:
        pass
    return None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), or None."""
    return measured("traffic_bytes_per_launch " + kernel)


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


def c3_solver(d, nx, ny, fixed_dt=None):
    """Reacting mixing-layer solver with its seeded initial field (CFL dt,
    or a fixed dt)."""
    import ostc
    from ostc import cases
    case = cases.c3_case(nx, ny)
    if fixed_dt is not None:
        case["setup"]["bool fixed_time_step"] = 1
        case["setup"]["real dt"] = fixed_dt
        case["name"] = "c3_fixed"
    s = ostc.solver_for_case(case, os.path.join(d, case["name"]), initial=False)
    states = {}
    for name in ("air_stream", "fuel_stream"):
        o = s.inflow_state(name, 16)
        states[name] = dict(rho=o[0], u=o[1], v=o[2], w=o[3], p=o[4], t=o[5], vel=o[6], sv=o[7:16])
    bv, sv = cases.c3_fields(case, states, s.table("mw", 9))
    s.upload("bv", bv)
    s.upload("sv", sv)
    s.initialize()
    return s


def c5_solver(d, nx, ny, nz):
    """Reacting RANS channel with wall injection and its seeded initial field."""
    import ostc
    from ostc import cases
    case = cases.c5_case(nx, ny, nz)
    s = ostc.solver_for_case(case, os.path.join(d, "c5"), initial=False)
    o = s.inflow_state("vitiated_air", 7 + 11 + 1)
    mw = s.table("mw", 9)
    air = dict(rho=o[0], u=o[1], v=o[2], w=o[3], p=o[4], t=o[5], vel=o[6], sv=list(o[7:18]), mut=o[18],
               mw=1.0 / sum(o[7 + l] / mw[l] for l in range(9)))
    bv, sv = cases.c5_fields(case, air)
    s.upload("bv", bv)
    s.upload("sv", sv)
    s.initialize()
    return case, s, air


def cpu_baseline_c5(nx=96, ny=48, nz=32, iters=2):
    """Serial FP64 oracle on a reduced grid of the channel, same scheme,
    boundaries and port layout."""
    import helpers
    import pyoracle
    from ostc import cases
    import ostc
    with tempfile.TemporaryDirectory() as d:
        case = cases.c5_case(nx, ny, nz)
        cases.write_case_dir(case, d)
        h = ostc.Solver(d, -1)
        m = pyoracle.make_mech()
        states = {}
        for name in ("vitiated_air", "h2_jet"):
            o = h.inflow_state(name, 19)
            states[name] = dict(rho=o[0], u=o[1], v=o[2], w=o[3], p=o[4], t=o[5], vel=o[6], sv=list(o[7:18]),
                                mut=o[18], mw=1.0 / sum(o[7 + l] / m.mw[l] for l in range(9)))
        h.close()
    bv, sv = cases.c5_fields(case, states["vitiated_air"])
    for b in case["bcs"]:
        if b["type"] == 5:
            b["state"] = states[b["stream"]]
    z = helpers.oracle_zone(case, mech=m, bv=bv, sv=sv)
    t0 = time.perf_counter()
    z.run_steady(iters)
    t1 = time.perf_counter()
    cells = nx * ny * nz
    return cells * iters / (t1 - t0) / 1e6, f"{nx}x{ny}x{nz} grid, {iters} pseudo-iterations", t1 - t0


def c5_inline(s5, cells, a):
    """The channel with inline (operator-split) chemistry, FP32 fluxes: one
    chemistry step of a.c5_chem_dt per pseudo-iteration by the FP64
    point-implicit integrator, then by the surrogate with its fallback
    (trained first, untimed, on the channel's own states at that dt), both
    from the same state; times, the fallback fraction and the surrogate
    run's errors against the integrator run."""
    import ostc
    dt = a.c5_chem_dt
    n = cells
    st = {k: ostc.DeviceArray(s5, 8 * n * c) for k, c in (("rho", 1), ("T", 1), ("Y", 9))}
    s5.chem_states(st["rho"], st["T"], st["Y"])
    log(f"bench: channel inline chemistry: training the surrogate on {n} channel states at dt = {dt:g} s")
    s5.surrogate_init(31)
    t_tr = time.perf_counter()
    hist = s5.surrogate_train(n, st["rho"], st["T"], st["Y"], dt, epochs=a.c5_epochs, batch=8192, seed=31)
    t_tr = time.perf_counter() - t_tr
    for v in st.values():
        v.free()
    snap = {f: s5.download(f) for f in ("bv", "sv", "cv")}
    out = {"chem_dt": dt, "train_states": n, "train_epochs": a.c5_epochs, "train_s": t_tr,
           "train_final_loss": float(hist[-1])}
    res = {}
    for mode, tag in ((0, "integrator"), (1, "surrogate")):
        for f, v in snap.items():
            s5.upload(f, v)
        s5.update_properties()
        s5.set_steady_chem(dt, mode)
        s5.step_steady(1)
        s5.sync()
        t0 = time.perf_counter()
        s5.step_steady(a.c5_iters)
        s5.sync()
        t1 = time.perf_counter()
        c, fb = s5.steady_chem_stats()
        res[mode] = (s5.download("bv"), s5.download("sv"))
        if mode == 1:
            out["gate_fallback_by_condition_last_iteration"] = s5.surrogate_gate_reasons()
        out[tag + "_ms_per_iter"] = (t1 - t0) / a.c5_iters * 1e3
        out[tag + "_mcell_iter_per_s"] = cells * a.c5_iters / (t1 - t0) / 1e6
        out[tag + "_fallback_fraction"] = fb / max(c, 1)
    s5.set_steady_chem(0.0)
    (bv0, sv0), (bv1, sv1) = res[0], res[1]
    T0, T1 = bv0.reshape(6, -1)[5], bv1.reshape(6, -1)[5]
    Y0, Y1 = sv0.reshape(11, -1), sv1.reshape(11, -1)
    out["iters_compared"] = a.c5_iters + 1
    out["peak_T_rel_err"] = float(abs(T1.max() - T0.max()) / T0.max())
    out["T_rel_linf"] = float(np.max(np.abs(T1 - T0) / T0))
    out["major_species_linf"] = float(np.abs(Y1[[0, 2, 7, 8]] - Y0[[0, 2, 7, 8]]).max())
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(a, d_chem=True, d_c5=True):
    """The oracle built for this host (-O3 -march=native, no FMA contraction)
    and timed pinned to one core on bounded samples of the workloads; the
    chemistry batch is also timed on all the cores this process may use
    (OpenMP build, in a child process)."""
    import subprocess
    import pyoracle
    from ostc import cases
    share = sorted(os.sched_getaffinity(0))
    threads = min(len(share), int(os.environ.get("OMP_NUM_THREADS", len(share)) or len(share)))
    info = {"cores": 1, "kind": "port", "cpu_model": cpu_model(), "cpus_available": len(share)}
    with tempfile.TemporaryDirectory() as bd:
        try:
            serial = pyoracle.build_native(bd)
            omp = pyoracle.build_native(bd, openmp=True)
            pyoracle.use_library(serial)
            info["build"] = "gcc -O3 -march=native -ffp-contract=off, FP64, serial"
        except (OSError, subprocess.CalledProcessError) as e:
            omp = None
            info["build"] = f"prebuilt oracle (-O3, no -march): native build failed ({e})"
        os.sched_setaffinity(0, {share[0]})
        try:
            log(f"bench: CPU baseline (oracle pinned to CPU {share[0]})")
            v, sample, secs = cpu_baseline_flow(lambda: cases.c2_case(128, 64, 64), steps=7)
            info.update(value=v, unit="Mcell-updates/s",
                        sample=f"FP64 serial oracle, {sample} of the same scheme/boundaries ({secs:.1f} s)")
            v1, s1, t1 = cpu_baseline_flow(lambda: cases.c1_case(200, 100), steps=100)
            info["c1_mcell_per_s"] = v1
            info["c1_sample"] = f"all 100 steps of the 200x100 case ({t1:.1f} s)"
            if d_chem:
                cv_, csample, csecs = cpu_baseline_chem()
                info["chem_cells_per_s"] = cv_
                info["chem_sample"] = f"{csample} ({csecs:.1f} s)"
            if d_c5:
                v5, s5s, t5s = cpu_baseline_c5()
                info["c5_mcell_iter_per_s"] = v5
                info["c5_sample"] = f"FP64 serial oracle, {s5s} of the channel ({t5s:.1f} s)"
        finally:
            os.sched_setaffinity(0, set(share))
        if d_chem and omp is not None and threads > 1:
            log(f"bench: CPU baseline, chemistry batch on {threads} threads")
            env = dict(os.environ, OSTC_ORACLE_LIB=omp, OMP_NUM_THREADS=str(threads))
            r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-chem-worker", str(1 << 22)],
                               env=env, capture_output=True, text=True)
            if r.returncode == 0:
                w = json.loads(r.stdout.strip().splitlines()[-1])
                info["chem_allcore"] = {"cells_per_s": w["cells_per_s"], "threads": threads,
                                        "sample": f"{w['sample']} ({w['seconds']:.1f} s), OpenMP over cells"}
            else:
                info["chem_allcore"] = {"error": r.stderr.strip()[-300:]}
    return info


def time_steps(s, steps, warmup):
    s.step(warmup)
    s.sync()
    t0 = time.perf_counter()
    s.step(steps)
    s.sync()
    return time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=256)
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--nz", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--cpu-chem-worker", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--no-split", action="store_true")
    ap.add_argument("--split-dt", type=float, default=1e-8)
    ap.add_argument("--split-pairs", type=int, default=1 << 22)
    ap.add_argument("--c5-iters", type=int, default=5)
    ap.add_argument("--no-c5-inline", dest="c5_inline", action="store_false")
    ap.add_argument("--c5-chem-dt", type=float, default=1e-7)
    ap.add_argument("--c5-epochs", type=int, default=4)
    ap.add_argument("--chem-cells", type=int, default=1 << 22)
    ap.add_argument("--train-pairs", type=int, default=1 << 24)
    ap.add_argument("--epochs", type=int, default=16)
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="bytes of one device-to-device copy run at the end (HBM counter calibration)")
    a = ap.parse_args()
    if a.cpu_chem_worker > 0:
        # child process of the CPU baseline: the oracle named by
        # OSTC_ORACLE_LIB on all the threads OMP_NUM_THREADS allows
        v, sample, secs = cpu_baseline_chem(a.cpu_chem_worker)
        print(json.dumps({"cells_per_s": v, "sample": sample, "seconds": secs}))
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; multi-GPU sharding is described in DESIGN.md "
              "and not launched here", file=sys.stderr)
        sys.exit(2)

    import ostc
    from ostc import cases

    cells = a.nx * a.ny * a.nz
    with tempfile.TemporaryDirectory() as d:
        case = cases.c2_case(a.nx, a.ny, a.nz)
        log(f"bench: building case {a.nx}x{a.ny}x{a.nz}")
        s = ostc.solver_for_case(case, d)
        log("bench: warmup")
        s.step(a.warmup)
        s.sync()
        log(f"bench: timing {a.steps} FP64 steps")
        t0 = time.perf_counter()
        dt_h, res_h = s.step(a.steps)
        s.sync()
        t1 = time.perf_counter()
        fp64_s = t1 - t0
        # per-kernel HIP-event times from a separate, profiled run (the
        # events would otherwise sit inside the timed region above)
        prof_steps = max(2, min(a.steps, 10))
        log(f"bench: profiled run of {prof_steps} FP64 steps (per-kernel times)")
        s.profile(True)
        s.step(prof_steps)
        prof = s.profile_report()
        s.profile(False)
        # FP32-flux variant on the same state path
        s.set_flux_fp32(True)
        log(f"bench: timing {a.steps} mixed-precision steps")
        mixed_s = time_steps(s, a.steps, a.warmup)
        finite = bool(np.isfinite(res_h).all())
        s.close()

        c1 = None
        if not a.no_c1:
            # the reference's own 2-D case, all 100 steps from the seeded
            # initial field (after an untimed warm-up run from the same field)
            log("bench: 2-D air 200x100, 100 RK3 steps")
            case1 = cases.c1_case(200, 100)
            s1 = ostc.solver_for_case(case1, os.path.join(d, "c1"))
            bv1, _ = cases.initial_fields(case1)
            s1.step(5)
            s1.upload("bv", bv1)
            s1.initialize()
            t0 = time.perf_counter()
            _, res1 = s1.step(100)
            s1.sync()
            t1 = time.perf_counter()
            s1.close()
            c1 = {"workload": "2-D inviscid air 200x100 Cartesian, AUSM+ with MUSCL-minmod, SSP-RK3 CFL 0.5, "
                              "Mach 2 inflow / outflow / symmetry, 100 steps from the seeded initial field",
                  "mcell_per_s": 200 * 100 * 100 / (t1 - t0) / 1e6, "ms_per_step": (t1 - t0) / 100 * 1e3,
                  "residual_finite": bool(np.isfinite(res1[:, 4]).all())}

        c3 = None
        if not a.no_c3:
            log("bench: reacting mixing layer 512x256")
            s3 = c3_solver(d, 512, 256)
            c3_cells = 512 * 256
            t64 = time_steps(s3, a.steps, a.warmup)
            s3.set_flux_fp32(True)
            t32 = time_steps(s3, a.steps, a.warmup)
            s3.close()
            c3 = {"workload": "2-D reacting H2/air mixing layer 512x256, 9 species / 19 reactions, HLLC + AWENO-5 "
                              "(multi-species characteristic interpolation), viscous, explicit finite-rate source, "
                              "SSP-RK3",
                  "fp64_mcell_per_s": c3_cells * a.steps / t64 / 1e6, "fp64_ms_per_step": t64 / a.steps * 1e3,
                  "mixed_mcell_per_s": c3_cells * a.steps / t32 / 1e6, "mixed_ms_per_step": t32 / a.steps * 1e3}

        split = None
        if not a.no_split and a.split_pairs > 0:
            # operator-split reacting step on the mixing layer at a fixed dt:
            # FP64 flow + FP64 point-implicit chemistry, against FP32-flux flow
            # + surrogate chemistry (network trained untimed on mixing-layer
            # states at the same dt); both runs start from the same state
            log("bench: split reacting step 512x256, fixed dt")
            s4 = c3_solver(d, 512, 256, fixed_dt=a.split_dt)
            n4 = 512 * 256
            s4.set_chem_split(True)
            bv0, sv0 = s4.download("bv").copy(), s4.download("sv").copy()
            mw9 = s4.table("mw", 9)

            def run(mode, nsteps):
                for _ in range(nsteps):
                    s4.step(1)
                    s4.chem_step(a.split_dt, mode=mode)
                s4.sync()

            run(0, a.warmup)
            t0 = time.perf_counter()
            run(0, a.steps)
            t1 = time.perf_counter()
            bv_ref, sv_ref = s4.download("bv").reshape(6, -1), s4.download("sv").reshape(9, -1)
            # training states: the flow's own chemistry inputs at every split
            # step of the run (replayed untimed from the initial state) plus
            # synthetic mixing-layer states for coverage
            s4.upload("bv", bv0)
            s4.upload("sv", sv0)
            s4.initialize()
            g_st = {k: ostc.DeviceArray(s4, 8 * n4 * c) for k, c in (("rho", 1), ("T", 1), ("Y", 9))}
            hv = {"rho": [], "T": [], "Y": []}
            for _ in range(a.warmup + a.steps):
                s4.step(1)
                s4.chem_states(g_st["rho"], g_st["T"], g_st["Y"])
                hv["rho"].append(g_st["rho"].to_host((n4,)))
                hv["T"].append(g_st["T"].to_host((n4,)))
                hv["Y"].append(g_st["Y"].to_host((9, n4)))
                s4.chem_step(a.split_dt, mode=0)
            for v in g_st.values():
                v.free()
            r2, t2, p2, y2 = cases.mixing_layer_states(a.split_pairs, mw=mw9)
            r2 = np.concatenate(hv["rho"] + [r2])
            t2 = np.concatenate(hv["T"] + [t2])
            y2 = np.ascontiguousarray(np.concatenate(hv["Y"] + [y2], axis=1))
            n_tr = r2.size
            log(f"bench: surrogate training for the split step on {n_tr} states ({n_tr - a.split_pairs} from the "
                f"flow) (untimed)")
            tr = [ostc.DeviceArray.from_host(s4, v) for v in (r2, t2, y2)]
            s4.surrogate_init(13)
            hist4 = s4.surrogate_train(n_tr, tr[0], tr[1], tr[2], a.split_dt, epochs=8, batch=8192, seed=2003)
            for v in tr:
                v.free()
            s4.upload("bv", bv0)
            s4.upload("sv", sv0)
            s4.initialize()
            s4.set_flux_fp32(True)
            nfb = []
            reasons = {}

            def run_sur(nsteps, count=False):
                for _ in range(nsteps):
                    s4.step(1)
                    nfb.append(s4.chem_step(a.split_dt, mode=1))
                    if count:
                        for k, v in s4.surrogate_gate_reasons().items():
                            reasons[k] = reasons.get(k, 0) + v
                s4.sync()

            run_sur(a.warmup)
            t2_ = time.perf_counter()
            run_sur(a.steps)
            t3_ = time.perf_counter()
            bv_s, sv_s = s4.download("bv").reshape(6, -1), s4.download("sv").reshape(9, -1)
            # gate statistics from one more, untimed step
            run_sur(1, count=True)
            s4.close()
            Tr, Ts = bv_ref[5], bv_s[5]
            split = {"workload": "2-D reacting H2/air mixing layer 512x256, operator-split step: SSP-RK3 flow step "
                                 "(HLLC + AWENO-5, viscous) + one point-implicit chemistry step, fixed dt",
                     "dt": a.split_dt,
                     "fp64_plain_ms_per_step": (t1 - t0) / a.steps * 1e3,
                     "mixed_surrogate_ms_per_step": (t3_ - t2_) / a.steps * 1e3,
                     "speedup": (t1 - t0) / (t3_ - t2_),
                     "fp64_plain_mcell_per_s": n4 * a.steps / (t1 - t0) / 1e6,
                     "mixed_surrogate_mcell_per_s": n4 * a.steps / (t3_ - t2_) / 1e6,
                     "fallback_fraction": float(np.mean(nfb)) / n4,
                     "steps_compared": a.warmup + a.steps,
                     "peak_T_rel_err": float(abs(Ts.max() - Tr.max()) / Tr.max()),
                     "T_rel_linf": float(np.max(np.abs(Ts - Tr) / Tr)),
                     "major_species_linf": float(np.abs(sv_s[[0, 2, 7, 8]] - sv_ref[[0, 2, 7, 8]]).max()),
                     "train_states": n_tr, "train_states_from_flow": n_tr - a.split_pairs,
                     "train_final_loss": float(hist4[-1]),
                     "gate_fallback_by_condition": reasons}

        c5 = None
        if not a.no_c5:
            log("bench: reacting RANS channel 384x192x128 (steady DPLUR, SST, EPI)")
            case5, s5, _ = c5_solver(d, 384, 192, 128)
            c5_cells = 384 * 192 * 128
            s5.step_steady(2)
            s5.sync()
            t0 = time.perf_counter()
            res5 = s5.step_steady(a.c5_iters)
            s5.sync()
            t1 = time.perf_counter()
            # per-kernel breakdown from a separate profiled pseudo-iteration
            s5.profile(True)
            s5.step_steady(1)
            p5 = s5.profile_report()
            s5.profile(False)
            s5.set_flux_fp32(True)
            s5.step_steady(1)
            s5.sync()
            t2 = time.perf_counter()
            s5.step_steady(a.c5_iters)
            s5.sync()
            t3 = time.perf_counter()
            c5 = {"workload": "3-D reacting RANS (k-omega SST) H2/air channel 384x192x128 with 4 sonic H2 wall "
                              "jets, 9 species + k, omega, HLLC + AWENO-5 (multi-species characteristic "
                              "interpolation carrying k and omega), viscous, steady DPLUR (3 inner sweeps), "
                              "exact point-implicit chemistry, implicit SST source",
                  "cells": c5_cells, "iters": a.c5_iters,
                  "fp64_mcell_iter_per_s": c5_cells * a.c5_iters / (t1 - t0) / 1e6,
                  "fp64_ms_per_iter": (t1 - t0) / a.c5_iters * 1e3,
                  "mixed_mcell_iter_per_s": c5_cells * a.c5_iters / (t3 - t2) / 1e6,
                  "mixed_ms_per_iter": (t3 - t2) / a.c5_iters * 1e3,
                  "residual_finite": bool(np.isfinite(res5[:, 4]).all()),
                  "kernel_avg_ms": {k: v[1] / v[0] for k, v in p5.items()}}
            if a.c5_inline:
                c5["inline_chemistry"] = c5_inline(s5, c5_cells, a)
            s5.close()

        # chemistry batch: FP64 point-implicit, then the surrogate
        chem = sur = None
        if a.chem_cells > 0:
            cc = cases.c1_case(20, 12)
            cc["setup"]["int species"] = 1
            cc["setup"]["int reaction"] = 1
            cc["setup"]["int chemSrcMethod"] = 1
            root = os.path.join(d, "chem")
            cases.write_case_dir(cc, root)
            sc = ostc.Solver(root, 0)
            labels = sc.table("label", 19)
            n = a.chem_cells
            rho, T, p, Y = cases.chem_batch_inputs(n, seed=1004)
            dr, dT_, dY_ = (ostc.DeviceArray.from_host(sc, v) for v in (rho, T, Y))
            oY, oT = ostc.DeviceArray(sc, Y.nbytes), ostc.DeviceArray(sc, T.nbytes)
            log("bench: chemistry batch FP64")
            sc.chem_batch(n, dr, dT_, dY_, 1e-7, oY, oT, mode=0)
            sc.profile(True)
            reps = 3
            t4 = time.perf_counter()
            for _ in range(reps):
                sc.chem_batch(n, dr, dT_, dY_, 1e-7, oY, oT, mode=0)
            sc.sync()
            t5 = time.perf_counter()
            cp = sc.profile_report()
            sc.profile(False)
            Y0, T0 = oY.to_host(Y.shape), oT.to_host(T.shape)
            k_ms = cp["chem_batch_fp64"][1] / cp["chem_batch_fp64"][0]
            fpc = chem_flop_per_cell(labels)
            chem = dict(cells=n, cells_per_s=n * reps / (t5 - t4), kernel_ms=k_ms, flop_per_cell_est=fpc,
                        roofline={"bound": "fp64", "achieved": fpc * n / (k_ms * 1e-3) / 1e12,
                                  "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                  "frac": fpc * n / (k_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS})
            if a.train_pairs > 0:
                log(f"bench: surrogate training on {a.train_pairs} pairs x {a.epochs} epochs (untimed)")
                r2, t2, p2, y2 = cases.chem_batch_inputs(a.train_pairs, seed=2004)
                tr = [ostc.DeviceArray.from_host(sc, v) for v in (r2, t2, y2)]
                del r2, t2, p2, y2
                sc.surrogate_init(11)
                t6 = time.perf_counter()
                hist = sc.surrogate_train(a.train_pairs, tr[0], tr[1], tr[2], 1e-7, epochs=a.epochs, batch=8192,
                                          seed=2004)
                t7 = time.perf_counter()
                for v in tr:
                    v.free()
                log("bench: surrogate inference")
                sc.chem_batch(n, dr, dT_, dY_, 1e-7, oY, oT, mode=1)
                net_ms, gate_ms, fb_ms, nfs = [], [], [], []
                t8 = time.perf_counter()
                for _ in range(reps):
                    nf = sc.chem_batch(n, dr, dT_, dY_, 1e-7, oY, oT, mode=1)
                    st = sc.surrogate_last_stats()
                    net_ms.append(st["net_ms"])
                    gate_ms.append(st["gate_ms"])
                    fb_ms.append(st["fallback_ms"])
                    nfs.append(nf)
                sc.sync()
                t9 = time.perf_counter()
                Y1, T1 = oY.to_host(Y.shape), oT.to_host(T.shape)
                err = np.abs(Y1 - Y0)
                Tn0, Tn1 = T + T0, T + T1
                net = float(np.mean(net_ms))
                achieved = SURROGATE_FLOP_PER_CELL * n / (net * 1e-3) / 1e12
                sur = dict(cells=n, cells_per_s=n * reps / (t9 - t8),
                           speedup_vs_fp64=(n * reps / (t9 - t8)) / chem["cells_per_s"],
                           fallback_fraction=nfs[-1] / n, network_ms=net, gate_ms=float(np.mean(gate_ms)),
                           fallback_ms=float(np.mean(fb_ms)),
                           major_species_linf=float(err[[0, 2, 7, 8]].max()), all_species_linf=float(err.max()),
                           T_rel_linf=float(np.max(np.abs(Tn1 - Tn0) / Tn0)),
                           peak_T_rel_err=float(abs(Tn1.max() - Tn0.max()) / Tn0.max()),
                           train_pairs=a.train_pairs, epochs=a.epochs, train_s=t7 - t6, final_loss=float(hist[-1]),
                           gate=sc.surrogate_gate(),
                           roofline={"bound": "mfma", "kernel": "k_sur_net_pipe<1>", "achieved": achieved,
                                     "peak": BF16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                                     "frac": achieved / BF16_MFMA_PEAK_TFLOPS,
                                     "flop_per_cell": SURROGATE_FLOP_PER_CELL})
            sc.close()

        if a.calibrate_copy > 0:
            # one device-to-device copy of known size, for scaling the
            # FETCH_SIZE / WRITE_SIZE counters of a profiled run
            cc = cases.c1_case(20, 12)
            root = os.path.join(d, "copy")
            cases.write_case_dir(cc, root)
            sk = ostc.Solver(root, 0)
            src, dst = ostc.DeviceArray(sk, a.calibrate_copy), ostc.DeviceArray(sk, a.calibrate_copy)
            ostc.lib().ostc_memcpy_d2d(sk.h, dst.ptr, src.ptr, a.calibrate_copy)
            sk.close()

    ms_per_step = fp64_s / a.steps * 1e3
    value = cells * a.steps / fp64_s / 1e6
    # dominant kernel: the largest total event time among the sweeps
    sweeps = {k: v for k, v in prof.items() if k.startswith("sweep_")}
    dom = max(sweeps, key=lambda k: sweeps[k][1])
    calls, total_ms = sweeps[dom]
    avg_ms = total_ms / calls
    algo_bytes = SWEEP_BYTES_PER_CELL.get(dom, SWEEP_JK_BYTES_PER_CELL) * cells
    achieved = algo_bytes / (avg_ms * 1e-3) / 1e9
    kname = SWEEP_KERNELS.get(dom, dom)
    full = (a.nx, a.ny, a.nz) == (256, 128, 128)
    traffic = measured_traffic(kname) if full else None
    # FP64 operations of one launch (SQ_INSTS_VALU_FLOPS_FP64 + _TRANS) over
    # this run's launch time: the sweep's FP64 fraction beside its HBM one
    flop = measured("fp64_flop_per_launch " + kname) if full else None
    fp64 = None
    if flop:
        fp64 = {"flop_per_launch": flop, "achieved": flop / (avg_ms * 1e-3) / 1e12, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": flop / (avg_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
    # copy-calibrated HBM bytes of every launch of one whole RK3 step per
    # cell, against the compulsory bytes of a fully fused step
    step_b = measured("step_traffic_bytes_per_cell") if full else None
    step_traffic = None
    if step_b:
        step_traffic = {"bytes_per_cell_step": step_b, "compulsory_bytes_per_cell_step": STEP_COMPULSORY_BYTES,
                        "ratio": step_b / STEP_COMPULSORY_BYTES,
                        "achieved_GBs": step_b * cells / (ms_per_step * 1e-3) / 1e9}
    stage_ms = {k: v[1] / v[0] for k, v in prof.items()}

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(a, d_chem=chem is not None, d_c5=c5 is not None)

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
        "roofline": {"bound": "hbm", "kernel": kname, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": algo_bytes, "avg_launch_ms": avg_ms, "fp64": fp64,
                     "step_traffic": step_traffic},
        "cpu_baseline": cpu,
        "c1_air_2d": c1,
        "mixed_fp32_flux": {"value": cells * a.steps / mixed_s / 1e6, "unit": "Mcell-updates/s",
                            "ms_per_step": mixed_s / a.steps * 1e3},
        "reacting_mixing_layer": c3,
        "reacting_split_step": split,
        "reacting_rans_channel": c5,
        "chem_point_implicit_fp64": chem,
        "chem_surrogate": sur,
        "kernel_avg_ms": stage_ms,
        "residual_finite": finite,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
