"""ctypes front-end of the C oracle (oracle/orc_*.c).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product never does.  It loads oracle/_build/liboracle.so
(built by oracle/build.sh), fills the mechanism struct from the golden dump
of the reference's own CHEMKIN readers (tests/golden/h2o2_9sp_tables.txt),
restates the reference's inflow-state construction (BoundCond.cpp:7-122) and
drives the flow oracle.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "_build", "liboracle.so")
TABLES = os.path.join(ROOT, "tests", "golden", "h2o2_9sp_tables.txt")

MAXS, MAXR, MAXBC = 20, 50, 32
R_U = 8314.462618
GAMMA_AIR = 1.4
MW_AIR = 28.9647


def build():
    subprocess.run(["bash", os.path.join(HERE, "build.sh")], check=True)


def build_native(out_dir, openmp=False):
    """Compiles the oracle for the host it runs on (-O3 -march=native, no FMA
    contraction; -fopenmp splits the independent cells of orc_chem_batch over
    threads) into out_dir and returns the library path.  Used by the CPU
    baseline, which times the oracle on the machine that also runs the GPU."""
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, "liboracle_native%s.so" % ("_omp" if openmp else ""))
    cmd = ["gcc", "-std=c11", "-O3", "-march=native", "-ffp-contract=off", "-fPIC", "-shared", "-w"]
    if openmp:
        cmd.append("-fopenmp")
    cmd += ["-o", path, os.path.join(HERE, "orc_chem.c"), os.path.join(HERE, "orc_flow.c"), "-lm"]
    subprocess.run(cmd, check=True)
    return path


def use_library(path):
    """Loads the oracle from `path` instead of oracle/_build (before first use)."""
    global LIB, _lib
    LIB = path
    _lib = None


class Mech(C.Structure):
    _fields_ = [("n_spec", C.c_int), ("n_reac", C.c_int),
                ("mw", C.c_double * MAXS), ("t_low", C.c_double * MAXS), ("t_mid", C.c_double * MAXS),
                ("t_high", C.c_double * MAXS), ("hi", C.c_double * (MAXS * 7)), ("lo", C.c_double * (MAXS * 7)),
                ("LJ_potent_inv", C.c_double * MAXS), ("vis_coeff", C.c_double * MAXS),
                ("WjDivWi_to_One4th", C.c_double * (MAXS * MAXS)),
                ("sqrt_WiDivWjPl1Mul8", C.c_double * (MAXS * MAXS)),
                ("binary_diffusivity_coeff", C.c_double * (MAXS * MAXS)),
                ("kb_over_eps_jk", C.c_double * (MAXS * MAXS)),
                ("label", C.c_int * MAXR), ("stoi_f", C.c_int * (MAXR * MAXS)), ("stoi_b", C.c_int * (MAXR * MAXS)),
                ("order", C.c_int * MAXR),
                ("A", C.c_double * MAXR), ("b", C.c_double * MAXR), ("Ea", C.c_double * MAXR),
                ("A2", C.c_double * MAXR), ("b2", C.c_double * MAXR), ("Ea2", C.c_double * MAXR),
                ("third_body_coeff", C.c_double * (MAXR * MAXS)),
                ("troe_alpha", C.c_double * MAXR), ("troe_t3", C.c_double * MAXR), ("troe_t1", C.c_double * MAXR),
                ("troe_t2", C.c_double * MAXR)]


class BC(C.Structure):
    _fields_ = [("type", C.c_int), ("raw", C.c_int * 6),
                ("rho", C.c_double), ("u", C.c_double), ("v", C.c_double), ("w", C.c_double), ("p", C.c_double),
                ("t", C.c_double), ("vel", C.c_double), ("sv", C.c_double * (MAXS + 4)),
                ("t_wall", C.c_double), ("isothermal", C.c_int),
                ("a_inf", C.c_double), ("gamma_inf", C.c_double), ("entropy", C.c_double), ("mw_inf", C.c_double),
                ("p_total", C.c_double), ("t_total", C.c_double), ("p_back", C.c_double), ("mut", C.c_double)]


class Cfg(C.Structure):
    _fields_ = [("mix", C.c_int), ("reaction", C.c_int), ("chemSrcMethod", C.c_int), ("inviscid_scheme", C.c_int),
                ("reconstruction", C.c_int), ("limiter", C.c_int), ("viscous_order", C.c_int),
                ("fixed_time_step", C.c_int), ("output_screen", C.c_int), ("gradPInDiffusionFlux", C.c_int),
                ("implicit_method", C.c_int), ("dplur_inner_step", C.c_int), ("turbulence", C.c_int),
                ("turb_implicit", C.c_int), ("compressibility_correction", C.c_int), ("split_chem", C.c_int),
                ("cfl", C.c_double), ("Pr", C.c_double), ("Sc", C.c_double), ("Prt", C.c_double),
                ("Sct", C.c_double), ("dt", C.c_double), ("entropy_fix_factor", C.c_double),
                ("n_bc", C.c_int), ("bc", BC * MAXBC), ("steady_chem_dt", C.c_double),
                ("conservative_metric", C.c_int)]


_lib = None
if os.environ.get("OSTC_ORACLE_LIB"):
    LIB = os.environ["OSTC_ORACLE_LIB"]


class InnerFace(C.Structure):
    _fields_ = [("rs", C.c_int * 3), ("re", C.c_int * 3), ("face", C.c_int), ("dir", C.c_int),
                ("ts", C.c_int * 3), ("te", C.c_int * 3), ("tblk", C.c_int), ("tface", C.c_int),
                ("tdir", C.c_int), ("src_tar", C.c_int * 3), ("loop", C.c_int * 3), ("tloop", C.c_int * 3),
                ("n", C.c_int * 3)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        L.orc_enthalpy.argtypes = [C.c_double, C.POINTER(Mech), dp]
        L.orc_enthalpy_and_cp.argtypes = [C.c_double, C.POINTER(Mech), dp, dp]
        L.orc_cp.argtypes = [C.c_double, C.POINTER(Mech), dp]
        L.orc_gibbs_rt.argtypes = [C.c_double, C.POINTER(Mech), dp]
        L.orc_sutherland.argtypes = [C.c_double]
        L.orc_sutherland.restype = C.c_double
        L.orc_transport.argtypes = [C.POINTER(Mech), C.c_double, C.c_double, C.c_double, dp, dp, dp, dp, dp]
        L.orc_forward_rate.argtypes = [C.POINTER(Mech), C.c_double, dp, dp]
        L.orc_backward_rate.argtypes = [C.POINTER(Mech), C.c_double, dp, dp]
        L.orc_rate_of_progress.argtypes = [C.POINTER(Mech), dp, dp, dp, dp, dp, dp]
        L.orc_chem_source_point.argtypes = [C.POINTER(Mech), C.c_double, C.c_double, dp, dp, dp, dp]
        L.orc_lu_solve.argtypes = [dp, dp, C.c_int]
        L.orc_temperature_newton.argtypes = [C.POINTER(Mech), C.c_double, C.c_double, dp, dp, C.c_double, dp,
                                             C.POINTER(C.c_int)]
        L.orc_temperature_newton.restype = C.c_double
        L.orc_chem_batch.argtypes = [C.POINTER(Mech), C.c_int64, dp, dp, dp, C.c_double, dp, dp]
        L.orc_zone_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Cfg), C.POINTER(Mech), dp, dp,
                                      dp]
        L.orc_zone_create.restype = C.c_void_p
        L.orc_zone_destroy.argtypes = [C.c_void_p]
        L.orc_zone_field.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64)]
        L.orc_zone_field.restype = dp
        L.orc_zone_initialize.argtypes = [C.c_void_p]
        L.orc_zone_rhs.argtypes = [C.c_void_p]
        L.orc_zone_global_dt.argtypes = [C.c_void_p]
        L.orc_zone_global_dt.restype = C.c_double
        L.orc_zone_run_rk3.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.orc_zone_run_steady.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_zone_implicit.argtypes = [C.c_void_p]
        L.orc_zone_chem_step.argtypes = [C.c_void_p, C.c_double]
        L.orc_zone_state.argtypes = [C.c_void_p, dp, C.POINTER(C.c_int)]
        L.orc_grid_metrics.argtypes = [C.c_int] * 4 + [dp] * 7
        L.orc_sst_wilcox.argtypes = [C.c_double]
        L.orc_sst_wilcox.restype = C.c_double
        L.orc_sst_zeman.argtypes = [C.c_double, C.c_double]
        L.orc_sst_zeman.restype = C.c_double
        L.orc_sst_constants.argtypes = [dp, dp]
        L.orc_face_flux.argtypes = [C.c_int, C.c_void_p, dp, dp, dp, dp, C.c_void_p, dp]
        L.orc_jac_times_dq.argtypes = [C.c_void_p, C.c_int, dp, dp, C.c_double, dp, C.c_double, dp, dp]
        L.orc_reconstruct_air.argtypes = [C.c_int, dp, dp, dp]
        L.orc_aweno_mix_matrices.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, dp, dp, dp]
        L.orc_aweno_mix_sides.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, dp, dp]
        L.orc_zone_store_last.argtypes = [C.c_void_p]
        L.orc_zone_source.argtypes = [C.c_void_p]
        L.orc_zone_residual.argtypes = [C.c_void_p, C.c_int, dp]
        ip = C.POINTER(C.c_int)
        L.orc_inner_face_register.argtypes = [ip, ip, C.c_int, C.c_int, C.POINTER(InnerFace)]
        L.orc_mesh_create.argtypes = [C.c_int, ip, C.c_int, C.POINTER(Cfg), C.POINTER(Mech),
                                      C.POINTER(C.POINTER(C.c_double)), ip, ip]
        L.orc_mesh_create.restype = C.c_void_p
        L.orc_mesh_destroy.argtypes = [C.c_void_p]
        L.orc_mesh_zone.argtypes = [C.c_void_p, C.c_int]
        L.orc_mesh_zone.restype = C.c_void_p
        L.orc_mesh_initialize.argtypes = [C.c_void_p]
        L.orc_mesh_run_rk3.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.orc_mesh_run_steady.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_mesh_exchange.argtypes = [C.c_void_p, C.c_int]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def face_flux(scheme, pl, pr, met6, jac2, m=None, delta2=None):
    """Face flux of one Riemann solver (1 LF, 2 Roe, 3 AUSM+, 4 HLLC) for the
    primitive states pl, pr = (rho, u, v, w, p[, Y...]) of air (m None) or a
    mixture; met6 the two points' metric rows, jac2 their Jacobians."""
    pl = np.ascontiguousarray(pl, dtype=np.float64)
    pr = np.ascontiguousarray(pr, dtype=np.float64)
    met6 = np.ascontiguousarray(met6, dtype=np.float64)
    jac2 = np.ascontiguousarray(jac2, dtype=np.float64)
    f = np.zeros(pl.size)
    d = None if delta2 is None else np.ascontiguousarray(delta2, dtype=np.float64)
    lib().orc_face_flux(scheme, None if m is None else C.cast(C.byref(m), C.c_void_p), _p(pl), _p(pr), _p(met6),
                        _p(jac2), None if d is None else d.ctypes.data_as(C.c_void_p), _p(f))
    return f


def jac_times_dq(b6, sv, m3, lam, dq, m=None, gam=0.0):
    """DPLUR's (A + lam I) dq at one state (b6 = rho, u, v, w, p, T)."""
    b6 = np.ascontiguousarray(b6, dtype=np.float64)
    sv = np.ascontiguousarray(sv if len(sv) else [0.0], dtype=np.float64)
    m3 = np.ascontiguousarray(m3, dtype=np.float64)
    dq = np.ascontiguousarray(dq, dtype=np.float64)
    out = np.zeros(dq.size)
    lib().orc_jac_times_dq(None if m is None else C.cast(C.byref(m), C.c_void_p), dq.size - 5, _p(b6), _p(sv), gam,
                           _p(m3), lam, _p(dq), _p(out))
    return out


# ----------------------------------------------------------------- tables

def load_tables(path=TABLES):
    """Parse the golden dump into a dict of numpy arrays / lists."""
    t = {}
    with open(path) as f:
        for line in f:
            if line.startswith("#") or not line.strip():
                continue
            parts = line.split()
            key, n, vals = parts[0], int(parts[1]), parts[2:]
            if key in ("names", "elements"):
                t[key] = vals
            elif key in ("n_spec", "n_reac"):
                t[key] = int(vals[0])
            elif key in ("label", "stoi_f", "stoi_b", "order", "elem_comp"):
                t[key] = np.array([int(v) for v in vals], dtype=np.int64)
            else:
                t[key] = np.array([float(v) for v in vals], dtype=np.float64)
            assert n == len(vals), key
    return t


def make_mech(t=None):
    if t is None:
        t = load_tables()
    m = Mech()
    ns, nr = t["n_spec"], t["n_reac"]
    m.n_spec, m.n_reac = ns, nr
    for key in ("mw", "t_low", "t_mid", "t_high", "LJ_potent_inv", "vis_coeff"):
        getattr(m, key)[:ns] = list(t[key])
    m.hi[:ns * 7] = list(t["high_coeff"])
    m.lo[:ns * 7] = list(t["low_coeff"])
    for key in ("WjDivWi_to_One4th", "sqrt_WiDivWjPl1Mul8", "binary_diffusivity_coeff", "kb_over_eps_jk"):
        getattr(m, key)[:ns * ns] = list(t[key])
    for key in ("label", "order"):
        getattr(m, key)[:nr] = [int(v) for v in t[key]]
    for key in ("stoi_f", "stoi_b"):
        getattr(m, key)[:nr * ns] = [int(v) for v in t[key]]
    for key in ("A", "b", "Ea", "A2", "b2", "Ea2", "troe_alpha", "troe_t3", "troe_t1", "troe_t2"):
        getattr(m, key)[:nr] = list(t[key])
    m.third_body_coeff[:nr * ns] = list(t["third_body_coeff"])
    return m


# ------------------------------------------------------------- point calls

def enthalpy(m, T):
    h = np.zeros(MAXS)
    lib().orc_enthalpy(T, C.byref(m), _p(h))
    return h[:m.n_spec].copy()


def enthalpy_cp(m, T):
    h, cp = np.zeros(MAXS), np.zeros(MAXS)
    lib().orc_enthalpy_and_cp(T, C.byref(m), _p(h), _p(cp))
    return h[:m.n_spec].copy(), cp[:m.n_spec].copy()


def cp(m, T):
    c = np.zeros(MAXS)
    lib().orc_cp(T, C.byref(m), _p(c))
    return c[:m.n_spec].copy()


def gibbs_rt(m, T):
    g = np.zeros(MAXS)
    lib().orc_gibbs_rt(T, C.byref(m), _p(g))
    return g[:m.n_spec].copy()


def sutherland(T):
    return lib().orc_sutherland(T)


def transport(m, T, p, rho, Y):
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    c = cp(m, T)
    mu, lam = C.c_double(), C.c_double()
    rhoD = np.zeros(MAXS)
    c_full = np.zeros(MAXS)
    c_full[:m.n_spec] = c
    lib().orc_transport(C.byref(m), T, p, rho, _p(Y), _p(c_full), C.byref(mu), C.byref(lam), _p(rhoD))
    return mu.value, lam.value, rhoD[:m.n_spec].copy()


def chem_source(m, rho, T, Y, want_jac=True):
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    om = np.zeros(MAXS)
    jac = np.zeros(MAXS * MAXS)
    jd = np.zeros(MAXS)
    lib().orc_chem_source_point(C.byref(m), rho, T, _p(Y), _p(om), _p(jac) if want_jac else None, _p(jd))
    ns = m.n_spec
    return om[:ns].copy(), jac[:ns * ns].reshape(ns, ns).copy(), jd[:ns].copy()


def rates(m, T, c):
    c = np.ascontiguousarray(c, dtype=np.float64)
    kf, kb = np.zeros(MAXR), np.zeros(MAXR)
    lib().orc_forward_rate(C.byref(m), T, _p(c), _p(kf))
    lib().orc_backward_rate(C.byref(m), T, _p(kf), _p(kb))
    return kf[:m.n_reac].copy(), kb[:m.n_reac].copy()


def progress(m, kf, kb, c):
    kf = np.ascontiguousarray(kf, dtype=np.float64)
    kb = np.ascontiguousarray(kb, dtype=np.float64)
    c = np.ascontiguousarray(c, dtype=np.float64)
    q, q1, q2 = np.zeros(MAXR), np.zeros(MAXR), np.zeros(MAXR)
    lib().orc_rate_of_progress(C.byref(m), _p(kf), _p(kb), _p(c), _p(q), _p(q1), _p(q2))
    nr = m.n_reac
    return q[:nr].copy(), q1[:nr].copy(), q2[:nr].copy()


def lu_solve(A, b):
    A = np.ascontiguousarray(A, dtype=np.float64).copy()
    b = np.ascontiguousarray(b, dtype=np.float64).copy()
    lib().orc_lu_solve(_p(A), _p(b), len(b))
    return b


def temperature_newton(m, rho, rhoE, vel3, Y, t_start):
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    v = np.ascontiguousarray(vel3, dtype=np.float64)
    p = C.c_double()
    it = C.c_int()
    t = lib().orc_temperature_newton(C.byref(m), rho, rhoE, _p(v), _p(Y), t_start, C.byref(p), C.byref(it))
    return t, p.value, it.value


def chem_batch(m, rho, T, Y, dt):
    """Y: (n_spec, n) SoA.  Returns dY (n_spec, n), dT (n)."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    T = np.ascontiguousarray(T, dtype=np.float64)
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    n = rho.shape[0]
    dY = np.zeros_like(Y)
    dT = np.zeros(n)
    lib().orc_chem_batch(C.byref(m), n, _p(rho), _p(T), _p(Y), dt, _p(dY), _p(dT))
    return dY, dT


# ------------------------------------------------------------- inflow

def inflow_state(mach, pressure, temperature, direction=(1.0, 0.0, 0.0), Y=None, m=None):
    """BoundCond.cpp:7-122 Inflow::Inflow for the (mach, pressure, temperature)
    combination: returns dict(rho,u,v,w,p,t,vel,sv)."""
    if Y is None:
        mw = MW_AIR
        gamma = GAMMA_AIR
        sv = np.zeros(0)
    else:
        sv = np.asarray(Y, dtype=np.float64)
        mw = 0.0
        for i in range(m.n_spec):
            mw += sv[i] / m.mw[i]
        mw = 1 / mw
        cpi = cp(m, temperature)
        cpt, cvt = 0.0, 0.0
        for i in range(m.n_spec):
            cpt += sv[i] * cpi[i]
            cvt += sv[i] * (cpi[i] - R_U / m.mw[i])
        gamma = cpt / cvt
    c = np.sqrt(gamma * R_U / mw * temperature)
    velocity = mach * c
    u, v, w = (d * velocity for d in direction)
    density = pressure * mw / (R_U * temperature)
    return dict(rho=density, u=u, v=v, w=w, p=pressure, t=temperature, vel=velocity, sv=sv, mw=mw, gamma=gamma)


# ------------------------------------------------------------- flow zone

FIELDS = dict(x=0, y=1, z=2, jac=3, metric=4, bv=5, sv=6, cv=7, dq=8, vel=9, acoustic=10, mach=11, mul=12, cond=13,
              rhoD=14, gamma=15, cp=16, dt_local=17, qn=18, bv_last=19, chem_jac=20, mut=21, turb_cond=22,
              wall_distance=23, turb_jac=24)


def make_cfg(case):
    """case: dict with scheme keys and 'bcs' list of dict(type, raw, state...)."""
    c = Cfg()
    c.mix = case.get("mix", 0)
    c.reaction = case.get("reaction", 0)
    c.chemSrcMethod = case.get("chemSrcMethod", 0)
    c.inviscid_scheme = case.get("inviscid_scheme", 3)
    c.reconstruction = case.get("reconstruction", 2)
    c.limiter = case.get("limiter", 0)
    c.viscous_order = case.get("viscous_order", 2)
    c.fixed_time_step = case.get("fixed_time_step", 0)
    c.output_screen = case.get("output_screen", 1)
    c.gradPInDiffusionFlux = case.get("gradPInDiffusionFlux", 0)
    c.implicit_method = case.get("implicit_method", 0)
    c.dplur_inner_step = case.get("DPLUR_inner_step", 3)
    c.turbulence = case.get("turbulence", 0)
    c.turb_implicit = case.get("turb_implicit", 1)
    c.compressibility_correction = case.get("compressibility_correction", 0)
    c.split_chem = case.get("split_chem", 0)
    c.steady_chem_dt = case.get("steady_chem_dt", 0.0)
    c.conservative_metric = int(case.get("conservative_metric", 0))
    c.cfl = case.get("cfl", 0.5)
    c.Pr = case.get("prandtl_number", 0.72)
    c.Sc = case.get("schmidt_number", 0.9)
    c.Prt = case.get("turbulent_prandtl_number", 0.9)
    c.Sct = case.get("turbulent_schmidt_number", 0.9)
    c.dt = case.get("dt", 1e-5)
    c.entropy_fix_factor = case.get("entropy_fix_factor", 0.125)
    c.n_bc = len(case["bcs"])
    for n, b in enumerate(case["bcs"]):
        B = c.bc[n]
        B.type = b["type"]
        B.raw[:] = list(b["raw"])
        st = b.get("state")
        if st is not None:
            for key in ("rho", "u", "v", "w", "p", "t", "vel", "a_inf", "gamma_inf", "entropy", "mw_inf", "p_total",
                        "t_total", "p_back", "mut"):
                if key in st:
                    setattr(B, key, st[key])
            if "sv" in st:
                B.sv[:len(st["sv"])] = list(st["sv"])
        B.t_wall = b.get("t_wall", 300.0)
        B.isothermal = 1 if b.get("isothermal", True) else 0
    return c


class Zone:
    """Flow oracle zone: reference DZone layout arrays in numpy views."""

    def __init__(self, mx, my, mz, ngg, case, x, y, z, mech=None):
        self.mx, self.my, self.mz, self.g = mx, my, mz, ngg
        self.cfg = make_cfg(case)
        self.mech = mech if mech is not None else Mech()
        self._x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        self._y = np.ascontiguousarray(y, dtype=np.float64).ravel()
        self._z = np.ascontiguousarray(z, dtype=np.float64).ravel()
        self.h = lib().orc_zone_create(mx, my, mz, ngg, C.byref(self.cfg), C.byref(self.mech), _p(self._x),
                                       _p(self._y), _p(self._z))

    def __del__(self):
        # a zone of a Mesh belongs to the mesh, which frees it
        if getattr(self, "h", None) and getattr(self, "_borrowed", None) is None:
            lib().orc_zone_destroy(self.h)
            self.h = None

    def field(self, name):
        n = C.c_int64()
        ptr = lib().orc_zone_field(self.h, FIELDS[name], C.byref(n))
        if n.value == 0:
            return np.zeros(0)
        # the array keeps this zone alive (its memory is the zone's), so a
        # view outliving the Python name of the zone never reads freed memory
        buf = (C.c_double * n.value).from_address(C.cast(ptr, C.c_void_p).value)
        buf._owner = self
        return np.frombuffer(buf, dtype=np.float64)

    def initialize(self):
        lib().orc_zone_initialize(self.h)

    def run_steady(self, n):
        res = np.zeros((n, 5))
        lib().orc_zone_run_steady(self.h, n, _p(res))
        return res

    def rhs(self):
        lib().orc_zone_rhs(self.h)

    def chem_step(self, dt):
        lib().orc_zone_chem_step(self.h, dt)

    def global_dt(self):
        return lib().orc_zone_global_dt(self.h)

    def source(self):
        lib().orc_zone_source(self.h)

    def store_last(self):
        lib().orc_zone_store_last(self.h)

    def residual(self, step):
        r = np.zeros(5)
        lib().orc_zone_residual(self.h, step, _p(r))
        return r

    def run(self, nsteps):
        dt = np.zeros(nsteps * 3)
        res = np.zeros(nsteps * 5)
        lib().orc_zone_run_rk3(self.h, nsteps, _p(dt), _p(res))
        return dt.reshape(nsteps, 3), res.reshape(nsteps, 5)

    def state(self):
        t, s = C.c_double(), C.c_int()
        lib().orc_zone_state(self.h, C.byref(t), C.byref(s))
        return t.value, s.value


class Mesh:
    """Multi-block oracle mesh: blocks joined by inner faces (lines as in the
    reference's inner-face file), each block an oracle zone."""

    def __init__(self, blocks, ngg, scheme, inner, mech=None):
        """blocks: list of dict(mx, my, mz, x, y, z, bcs); inner: list of
        (block, src7, tar7) in file order."""
        nb = len(blocks)
        self.mech = mech if mech is not None else Mech()
        self.cfgs = (Cfg * nb)()
        self._keep = []
        dims = np.zeros(3 * nb, dtype=np.int32)
        ptrs = (C.POINTER(C.c_double) * (3 * nb))()
        for b, blk in enumerate(blocks):
            self.cfgs[b] = make_cfg(dict(scheme, bcs=blk["bcs"]))
            dims[3 * b:3 * b + 3] = (blk["mx"], blk["my"], blk["mz"])
            for c, key in enumerate("xyz"):
                a = np.ascontiguousarray(blk[key], dtype=np.float64).ravel()
                self._keep.append(a)
                ptrs[3 * b + c] = _p(a)
        nface = np.zeros(nb, dtype=np.int32)
        lines = []
        for b in range(nb):
            for (fb, src, tar) in inner:
                if fb == b:
                    nface[b] += 1
                    lines.extend(list(src) + list(tar))
        lines = np.asarray(lines if lines else [0], dtype=np.int32)
        ip = C.POINTER(C.c_int)
        self.h = lib().orc_mesh_create(nb, dims.ctypes.data_as(ip), ngg, self.cfgs, C.byref(self.mech), ptrs,
                                       nface.ctypes.data_as(ip), lines.ctypes.data_as(ip))
        if not self.h:
            raise ValueError("unsupported inner face")
        self.nb = nb
        self.zones = []
        for b in range(nb):
            z = Zone.__new__(Zone)
            z.mx, z.my, z.mz, z.g = blocks[b]["mx"], blocks[b]["my"], blocks[b]["mz"], ngg
            z.h = lib().orc_mesh_zone(self.h, b)
            z._borrowed = self
            self.zones.append(z)

    def __del__(self):
        if getattr(self, "h", None):
            for z in self.zones:
                z.h = None
            lib().orc_mesh_destroy(self.h)
            self.h = None

    def initialize(self):
        lib().orc_mesh_initialize(self.h)

    def run(self, nsteps):
        dt = np.zeros(nsteps * 3)
        res = np.zeros(nsteps * 5)
        lib().orc_mesh_run_rk3(self.h, nsteps, _p(dt), _p(res))
        return dt.reshape(nsteps, 3), res.reshape(nsteps, 5)

    def run_steady(self, n):
        res = np.zeros((n, 5))
        lib().orc_mesh_run_steady(self.h, n, _p(res))
        return res

    def exchange(self, with_cv=False):
        lib().orc_mesh_exchange(self.h, 1 if with_cv else 0)


def inner_face_register(src7, tar7, ngg, dim):
    """The oracle's registration of one inner-face pair: dict of fields, or None."""
    f = InnerFace()
    a = np.ascontiguousarray(src7, dtype=np.int32)
    b = np.ascontiguousarray(tar7, dtype=np.int32)
    ip = C.POINTER(C.c_int)
    if lib().orc_inner_face_register(a.ctypes.data_as(ip), b.ctypes.data_as(ip), ngg, dim, C.byref(f)) != 0:
        return None
    return {k: (list(getattr(f, k)) if isinstance(getattr(f, k), C.Array) else getattr(f, k)) for k, _ in f._fields_}

