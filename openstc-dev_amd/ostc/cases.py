"""Seeded workload generators for the solver's benchmark and parity cases.

Each builder returns a plain dict describing a single-block case in the
reference's own input vocabulary: setup-file keys (Parameter.cpp:53-163
format), boundary-file lines (Mesh.cpp:686-706 format: i1 i2 j1 j2 k1 k2
label, 0-based), interior node coordinates (i fastest) and the seeded initial
primitive field.  write_case_dir() lays the case out as a reference case
directory (input_files/setup/*.txt, input_files/boundary_condition/,
input_files/chemistry/) so the solver reads it through its setup-file
interface.

Random numbers: splitmix64 -> uniform [0,1) (top 53 bits), Gaussian by
Box-Muller; identical streams feed the GPU run and the CPU baseline.
"""
import os
import shutil

import numpy as np

R_U = 8314.462618
GAMMA_AIR = 1.4
MW_AIR = 28.9647
PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHEM_DIR = os.path.join(PKG, "input_files", "chemistry")

_MASK = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014) vectorised over a counter."""

    def __init__(self, seed):
        self.state = np.uint64(seed & _MASK)

    def u64(self, n):
        with np.errstate(over="ignore"):
            ctr = np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
            z = self.state + ctr
            self.state = np.uint64(z[-1]) if n else self.state
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n):
        return (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n):
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        u1 = np.where(u1 <= 0.0, 1e-300, u1)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]


def sutherland(t):
    """Transport.cu:6-9."""
    return 1.716e-5 * (t / 273) ** 1.5 * (273 + 111) / (t + 111)


def air_inflow(mach, p, t):
    mw = MW_AIR
    c = np.sqrt(GAMMA_AIR * R_U / mw * t)
    return dict(rho=p * mw / (R_U * t), u=mach * c, v=0.0, w=0.0, p=p, t=t, vel=mach * c)


def ghost_shape(mx, my, mz, g):
    return (mz + 2 * g, my + 2 * g, mx + 2 * g)


def _setup_common(extra):
    base = {
        "int initial": 0, "int species": 0, "int reaction": 0, "int chemSrcMethod": 0, "bool turbulence": 0,
        "int turbulence_method": 1, "int RANS_model": 2, "int turb_implicit": 1, "int compressibility_correction": 0,
        "bool gradPInDiffusionFlux": 0, "bool if_collect_statistics": 0, "int start_collect_statistics_iter": 0,
        "string mechanism_file": "chemistry/combustion_mechanism.inp", "string therm_file": "chemistry/therm.dat",
        "string transport_file": "chemistry/tran.dat", "bool parallel": 0, "int output_file": 100000,
        "int output_screen": 1, "bool steady": 0, "int implicit_method": 0, "int DPLUR_inner_step": 3,
        "real convergence_criteria": 1e-7, "int temporal_scheme": 3, "bool fixed_time_step": 0, "real dt": 1e-5,
        "real total_simulation_time": 1e10, "int inviscid_scheme": 3, "int reconstruction": 2, "int limiter": 0,
        "real entropy_fix_factor": 0.125, "int viscous_order": 2, "int diffusivity_method": 0,
        "real schmidt_number": 0.9, "real prandtl_number": 0.72, "real turbulent_prandtl_number": 0.9,
        "real turbulent_schmidt_number": 0.9, "real c_chi": 1.0, "int groups_init": 1, "int if_monitor": 0,
        "real gridScale": 1.0, "int wall_distance": 0,
    }
    base.update(extra)
    return base


def c1_case(nx=200, ny=100, seed=1001, steps=100, inviscid_scheme=3):
    """2-D inviscid air, 200x100 Cartesian (spacing 1 mm), AUSM+ with
    MUSCL-minmod, SSP-RK3 at CFL 0.5.  Mach 2 supersonic inflow (101325 Pa,
    300 K) at i=0, supersonic outflow at i=max, symmetry at j=0 and j=max.
    Initial state: uniform inflow; interior density scaled by
    (1 + 0.01 N(0,1)) at fixed pressure, T from the equation of state."""
    dx = 1e-3
    xi = np.arange(nx) * dx
    yj = np.arange(ny) * dx
    X, Y = np.meshgrid(xi, yj, indexing="xy")  # (ny, nx)
    x = X[None, :, :].copy()
    y = Y[None, :, :].copy()
    z = np.zeros_like(x)
    inflow = air_inflow(2.0, 101325.0, 300.0)
    setup = _setup_common({
        "int dimension": 2, "real cfl": 0.5, "int inviscid_scheme": inviscid_scheme, "int reconstruction": 2,
        "int limiter": 0, "int viscous_order": 0, "int total_step": steps,
        "array string boundary_conditions": ["freestream", "outflow", "symmetry"],
        "string reference_state": "freestream", "string default_init": "freestream",
        "struct freestream": [("string", "type", "inflow"), ("int", "label", 5), ("real", "mach", 2.0),
                              ("real", "pressure", 101325.0), ("real", "temperature", 300.0), ("real", "u", 1.0),
                              ("real", "v", 0.0), ("real", "w", 0.0)],
        "struct outflow": [("string", "type", "outflow"), ("int", "label", 6)],
        "struct symmetry": [("string", "type", "symmetry"), ("int", "label", 3)],
    })
    bcs = [
        dict(type=5, raw=(0, 0, 0, ny - 1, 0, 0), label=5),
        dict(type=6, raw=(nx - 1, nx - 1, 0, ny - 1, 0, 0), label=6),
        dict(type=3, raw=(0, nx - 1, 0, 0, 0, 0), label=3),
        dict(type=3, raw=(0, nx - 1, ny - 1, ny - 1, 0, 0), label=3),
    ]
    scheme = dict(mix=0, inviscid_scheme=inviscid_scheme, reconstruction=2, limiter=0, viscous_order=0, cfl=0.5,
                  fixed_time_step=0, output_screen=1, entropy_fix_factor=0.125)
    return dict(name="c1", mx=nx, my=ny, mz=1, ngg=2, x=x, y=y, z=z, setup=setup, bcs=bcs, scheme=scheme,
                inflow=inflow, seed=seed, perturb="density", steps=steps)


def c2_grid(nx, ny, nz, lx=2.0, ly=1.0, lz=1.0, ratio=1.05, amp=0.05):
    """Curvilinear 3-D grid: wall-normal (j) geometric stretching by `ratio`,
    then x' = x + amp sin(2 pi y) sin(2 pi z)."""
    xi = np.linspace(0.0, lx, nx)
    steps = ratio ** np.arange(ny - 1)
    yj = np.concatenate([[0.0], np.cumsum(steps)])
    yj *= ly / yj[-1]
    zk = np.linspace(0.0, lz, nz)
    Z, Y, X = np.meshgrid(zk, yj, xi, indexing="ij")  # (nz, ny, nx)
    X = X + amp * np.sin(2 * np.pi * Y) * np.sin(2 * np.pi * Z)
    return X.copy(), Y.copy(), Z.copy()


def c2_case(nx=256, ny=128, nz=128, seed=1002, steps=20, mach=2.5, t_inf=300.0, re_unit=1e5, t_wall=300.0,
            reconstruction=4, inviscid_scheme=2, perturb=True):
    """3-D laminar Navier-Stokes in air on the sinusoidally mapped,
    wall-normal-stretched curvilinear grid; Roe + AWENO-5 requested
    (the reference runs AUSM+ for that pair), SSP-RK3 at CFL 0.5.  Mach 2.5,
    Re = 1e5 per metre at T = 300 K (Sutherland viscosity), isothermal no-slip
    wall at 300 K on j=0; supersonic inflow i=0, outflow i=max and j=max,
    symmetry k=0 and k=max."""
    x, y, z = c2_grid(nx, ny, nz)
    c = np.sqrt(GAMMA_AIR * R_U / MW_AIR * t_inf)
    u = mach * c
    rho = re_unit * sutherland(t_inf) / u
    p = rho * R_U / MW_AIR * t_inf
    inflow = air_inflow(mach, p, t_inf)
    setup = _setup_common({
        "int dimension": 3, "real cfl": 0.5, "int inviscid_scheme": inviscid_scheme,
        "int reconstruction": reconstruction, "int limiter": 0, "int viscous_order": 2, "int total_step": steps,
        "array string boundary_conditions": ["freestream", "outflow", "symmetry", "wall"],
        "string reference_state": "freestream", "string default_init": "freestream",
        "struct freestream": [("string", "type", "inflow"), ("int", "label", 5), ("real", "mach", mach),
                              ("real", "pressure", p), ("real", "temperature", t_inf), ("real", "u", 1.0),
                              ("real", "v", 0.0), ("real", "w", 0.0)],
        "struct outflow": [("string", "type", "outflow"), ("int", "label", 6)],
        "struct symmetry": [("string", "type", "symmetry"), ("int", "label", 3)],
        "struct wall": [("string", "type", "wall"), ("int", "label", 2), ("string", "thermal_type", "isothermal"),
                        ("real", "temperature", t_wall)],
    })
    bcs = [
        dict(type=5, raw=(0, 0, 0, ny - 1, 0, nz - 1), label=5),
        dict(type=6, raw=(nx - 1, nx - 1, 0, ny - 1, 0, nz - 1), label=6),
        dict(type=2, raw=(0, nx - 1, 0, 0, 0, nz - 1), label=2, t_wall=t_wall, isothermal=True),
        dict(type=6, raw=(0, nx - 1, ny - 1, ny - 1, 0, nz - 1), label=6),
        dict(type=3, raw=(0, nx - 1, 0, ny - 1, 0, 0), label=3),
        dict(type=3, raw=(0, nx - 1, 0, ny - 1, nz - 1, nz - 1), label=3),
    ]
    scheme = dict(mix=0, inviscid_scheme=inviscid_scheme, reconstruction=reconstruction, limiter=0, viscous_order=2,
                  cfl=0.5, fixed_time_step=0, output_screen=1)
    return dict(name="c2", mx=nx, my=ny, mz=nz, ngg=3 if reconstruction == 4 else 2, x=x, y=y, z=z, setup=setup,
                bcs=bcs, scheme=scheme, inflow=inflow, seed=seed, perturb="density" if perturb else None,
                steps=steps)


def wavy_grid(nx, ny, nz, lx=1.0, ly=0.8, lz=0.6, amp=0.03):
    """Curved grid displaced in all three directions (2-D when nz == 1):
    x' = x + amp sin(2 pi y/ly) sin(2 pi z/lz), y' = y + amp sin(2 pi x/lx)
    sin(2 pi z/lz), z' = z + amp sin(2 pi x/lx) sin(2 pi y/ly); in 2-D the
    z factors are dropped and z = 0."""
    xi = np.linspace(0.0, lx, nx)
    yj = np.linspace(0.0, ly, ny)
    zk = np.linspace(0.0, lz, nz) if nz > 1 else np.zeros(1)
    Z, Y, X = np.meshgrid(zk, yj, xi, indexing="ij")
    sx, sy = np.sin(2 * np.pi * X / lx), np.sin(2 * np.pi * Y / ly)
    sz = np.sin(2 * np.pi * Z / lz) if nz > 1 else np.ones_like(Z)
    x = X + amp * sy * sz
    y = Y + amp * sx * sz
    z = Z + amp * sx * sy if nz > 1 else Z
    return x.copy(), y.copy(), z.copy()


def uniform_flow_case(nx=14, ny=12, nz=10, reconstruction=2, inviscid_scheme=3, conservative_metric=1,
                      viscous_order=2, amp=0.03):
    """Free-stream preservation: a uniform Mach 2 air stream at 300 K,
    101325 Pa along (0.8, 0.48, 0.36) (2-D: (0.8, 0.6, 0)) on wavy_grid, every
    face a supersonic inflow of that state, no perturbation.  With the
    conservative metrics the right-hand side vanishes to rounding."""
    x, y, z = wavy_grid(nx, ny, nz, amp=amp)
    d = (0.8, 0.48, 0.36) if nz > 1 else (0.8, 0.6, 0.0)
    inflow = air_inflow(2.0, 101325.0, 300.0)
    vel = inflow["vel"]
    inflow = dict(inflow, u=vel * d[0], v=vel * d[1], w=vel * d[2])
    setup = _setup_common({
        "int dimension": 3 if nz > 1 else 2, "real cfl": 0.5, "int inviscid_scheme": inviscid_scheme,
        "int reconstruction": reconstruction, "int limiter": 0, "int viscous_order": viscous_order,
        "int total_step": 10, "bool conservative_metric": conservative_metric,
        "array string boundary_conditions": ["freestream"],
        "string reference_state": "freestream", "string default_init": "freestream",
        "struct freestream": [("string", "type", "inflow"), ("int", "label", 5), ("real", "mach", 2.0),
                              ("real", "pressure", 101325.0), ("real", "temperature", 300.0), ("real", "u", d[0]),
                              ("real", "v", d[1]), ("real", "w", d[2])],
    })
    faces = [(0, 0, 0, ny - 1, 0, nz - 1), (nx - 1, nx - 1, 0, ny - 1, 0, nz - 1),
             (0, nx - 1, 0, 0, 0, nz - 1), (0, nx - 1, ny - 1, ny - 1, 0, nz - 1)]
    if nz > 1:
        faces += [(0, nx - 1, 0, ny - 1, 0, 0), (0, nx - 1, 0, ny - 1, nz - 1, nz - 1)]
    bcs = [dict(type=5, raw=r, label=5) for r in faces]
    scheme = dict(mix=0, inviscid_scheme=inviscid_scheme, reconstruction=reconstruction, limiter=0,
                  viscous_order=viscous_order, cfl=0.5, fixed_time_step=0, output_screen=1,
                  conservative_metric=conservative_metric)
    return dict(name="uniform", mx=nx, my=ny, mz=nz, ngg=3 if reconstruction == 4 else 2, x=x, y=y, z=z,
                setup=setup, bcs=bcs, scheme=scheme, inflow=inflow, seed=1, perturb=None, steps=10)


def c3_case(nx=512, ny=256, seed=1003, steps=20, dx=1e-4, delta=5e-4, viscous_order=2, inviscid_scheme=4,
            reconstruction=4):
    """2-D reacting H2/air supersonic mixing layer with the 9-species,
    19-reaction H2-O2 mechanism (input_files/chemistry): upper stream air
    (Y_O2 = 0.233, Y_N2 = 0.767) at 1500 K and Mach 2, lower stream
    H2/N2 (0.3/0.7 by mass) at 300 K and Mach 1.2, both at 101325 Pa; a
    tanh profile of thickness `delta` across the centre line plus 1 % seeded
    velocity noise (see c3_fields).  HLLC with AWENO-5 (reconstruction 4,
    the multi-species characteristic interpolation, 3 ghost layers) or
    MUSCL-minmod (reconstruction 2, 2 ghost layers), second-order viscous
    fluxes with mixture-averaged transport, explicit finite-rate source
    inside SSP-RK3 at CFL 0.5.  Inflow i=0 (air above the centre line, fuel
    below, labels 5 and 51), outflow i=max, symmetry j=0 and j=max."""
    xi = np.arange(nx) * dx
    yj = np.arange(ny) * dx
    X, Y = np.meshgrid(xi, yj, indexing="xy")
    x = X[None, :, :].copy()
    y = Y[None, :, :].copy()
    z = np.zeros_like(x)
    air = [("real", "O2", 0.233), ("real", "N2", 0.767)]
    fuel = [("real", "H2", 0.3), ("real", "N2", 0.7)]
    common = [("real", "pressure", 101325.0), ("real", "u", 1.0), ("real", "v", 0.0), ("real", "w", 0.0)]
    setup = _setup_common({
        "int dimension": 2, "int species": 1, "int reaction": 1, "int chemSrcMethod": 0, "real cfl": 0.5,
        "int inviscid_scheme": inviscid_scheme, "int reconstruction": reconstruction, "int limiter": 0,
        "int viscous_order": viscous_order,
        "int total_step": steps,
        "array string boundary_conditions": ["air_stream", "fuel_stream", "outflow", "symmetry"],
        "string reference_state": "air_stream", "string default_init": "air_stream",
        "struct air_stream": [("string", "type", "inflow"), ("int", "label", 5), ("real", "mach", 2.0),
                              ("real", "temperature", 1500.0)] + common + air,
        "struct fuel_stream": [("string", "type", "inflow"), ("int", "label", 51), ("real", "mach", 1.2),
                               ("real", "temperature", 300.0)] + common + fuel,
        "struct outflow": [("string", "type", "outflow"), ("int", "label", 6)],
        "struct symmetry": [("string", "type", "symmetry"), ("int", "label", 3)],
    })
    half = ny // 2
    bcs = [
        dict(type=5, raw=(0, 0, half, ny - 1, 0, 0), label=5, stream="air_stream"),
        dict(type=5, raw=(0, 0, 0, half - 1, 0, 0), label=51, stream="fuel_stream"),
        dict(type=6, raw=(nx - 1, nx - 1, 0, ny - 1, 0, 0), label=6),
        dict(type=3, raw=(0, nx - 1, 0, 0, 0, 0), label=3),
        dict(type=3, raw=(0, nx - 1, ny - 1, ny - 1, 0, 0), label=3),
    ]
    scheme = dict(mix=1, reaction=1, chemSrcMethod=0, inviscid_scheme=inviscid_scheme, reconstruction=reconstruction,
                  limiter=0, viscous_order=viscous_order, cfl=0.5, fixed_time_step=0, output_screen=1,
                  entropy_fix_factor=0.125)
    return dict(name="c3", mx=nx, my=ny, mz=1, ngg=3 if reconstruction == 4 else 2, x=x, y=y, z=z, setup=setup,
                bcs=bcs, scheme=scheme,
                seed=seed, steps=steps, dx=dx, delta=delta, y0=0.5 * (ny - 1) * dx)


def mixing4_case(nx=48, ny=24, reconstruction=4, inviscid_scheme=4, steps=20):
    """The mixing layer of c3_case with the reference's 4-species mixing
    mechanism (chemistry/H2_mech_mixing.inp: H2, H2O, O2, N2) and no
    reactions: the same streams, profile, scheme and boundaries, species
    transport and mixture thermodynamics only."""
    case = c3_case(nx, ny, steps=steps, inviscid_scheme=inviscid_scheme, reconstruction=reconstruction)
    case["setup"].update({"int reaction": 0, "string mechanism_file": "chemistry/H2_mech_mixing.inp"})
    case["scheme"] = dict(case["scheme"], reaction=0)
    case["name"] = "mixing4"
    return case


def c3_fields(case, states, mw):
    """Initial bv (6, ...) and sv (9, ...) of the mixing layer: with
    f = (1 + tanh((y - y0) / delta)) / 2, u, T and Y blend from the lower
    (fuel) to the upper (air) stream state, p is uniform, rho follows from
    the equation of state, and interior u is scaled by (1 + 0.01 N(0,1))
    from the seeded generator.  `states` maps the two stream names to dicts
    with rho, u, p, t and sv (the inflow states the solver derives from the
    set-up structs); `mw` holds the species molecular weights."""
    mx, my, mz, g = case["mx"], case["my"], case["mz"], case["ngg"]
    shp = ghost_shape(mx, my, mz, g)
    up, lo = states["air_stream"], states["fuel_stream"]
    yc = (np.arange(my + 2 * g) - g) * case["dx"]
    f = 0.5 * (1.0 + np.tanh((yc - case["y0"]) / case["delta"]))[None, :, None]
    f = np.broadcast_to(f, shp)
    bv = np.empty((6,) + shp)
    bv[1] = lo["u"] + f * (up["u"] - lo["u"])
    bv[2] = 0.0
    bv[3] = 0.0
    bv[4] = up["p"]
    bv[5] = lo["t"] + f * (up["t"] - lo["t"])
    ns = len(mw)
    sv = np.empty((ns,) + shp)
    for l in range(ns):
        sv[l] = lo["sv"][l] + f * (up["sv"][l] - lo["sv"][l])
    mwinv = sum(sv[l] / mw[l] for l in range(ns))
    bv[0] = bv[4] / (R_U * bv[5] * mwinv)
    rng = SplitMix64(case["seed"])
    noise = 1.0 + 0.01 * rng.normal(mx * my * mz).reshape(mz, my, mx)
    sl = (slice(g, g + mz), slice(g, g + my), slice(g, g + mx))
    bv[1][sl] = bv[1][sl] * noise
    return bv, sv


def subsonic_case(nx=60, ny=30, seed=1006, mach=0.5, p=101325.0, t=300.0, pt_ratio=None, tt_ratio=None,
                  back_ratio=1.0):
    """2-D subsonic air channel (spacing 1 mm), AUSM+ with MUSCL-minmod,
    SSP-RK3 at CFL 0.5: subsonic inflow at i=0 (total pressure and
    temperature as ratios of the far state, default the isentropic totals of
    the Mach number), back pressure at i=max, farfield (Riemann invariants)
    at j=0 and j=max; the far state (Mach `mach`, p, T) is the reference and
    initial state, with the seeded 1 % density perturbation."""
    dx = 1e-3
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dx, indexing="xy")
    x, y = X[None].copy(), Y[None].copy()
    z = np.zeros_like(x)
    if pt_ratio is None:
        pt_ratio = (1 + 0.5 * (GAMMA_AIR - 1) * mach * mach) ** (GAMMA_AIR / (GAMMA_AIR - 1))
    if tt_ratio is None:
        tt_ratio = 1 + 0.5 * (GAMMA_AIR - 1) * mach * mach
    inflow = air_inflow(mach, p, t)
    setup = _setup_common({
        "int dimension": 2, "real cfl": 0.5, "int inviscid_scheme": 3, "int reconstruction": 2, "int limiter": 0,
        "int viscous_order": 0, "int total_step": 10,
        "array string boundary_conditions": ["farfield", "inlet", "outlet"],
        "string reference_state": "farfield", "string default_init": "farfield",
        "struct farfield": [("string", "type", "farfield"), ("int", "label", 4), ("real", "mach", mach),
                            ("real", "pressure", p), ("real", "temperature", t), ("real", "u", 1.0),
                            ("real", "v", 0.0), ("real", "w", 0.0)],
        "struct inlet": [("string", "type", "subsonic_inflow"), ("int", "label", 7),
                         ("real", "pt_pRef_ratio", pt_ratio), ("real", "Tt_TRef_ratio", tt_ratio),
                         ("real", "u", 1.0), ("real", "v", 0.0), ("real", "w", 0.0)],
        "struct outlet": [("string", "type", "back_pressure"), ("int", "label", 9),
                          ("real", "p_pRef_ratio", back_ratio)],
    })
    a = np.sqrt(GAMMA_AIR * R_U / MW_AIR * t)
    far = dict(inflow, a_inf=a, gamma_inf=GAMMA_AIR, entropy=p / inflow["rho"] ** GAMMA_AIR, mw_inf=MW_AIR, sv=[])
    bcs = [
        dict(type=7, raw=(0, 0, 0, ny - 1, 0, 0), label=7,
             state=dict(u=1.0, v=0.0, w=0.0, p_total=pt_ratio * p, t_total=tt_ratio * t, sv=[])),
        dict(type=9, raw=(nx - 1, nx - 1, 0, ny - 1, 0, 0), label=9, state=dict(p_back=back_ratio * p, sv=[])),
        dict(type=4, raw=(0, nx - 1, 0, 0, 0, 0), label=4, state=far),
        dict(type=4, raw=(0, nx - 1, ny - 1, ny - 1, 0, 0), label=4, state=far),
    ]
    scheme = dict(mix=0, inviscid_scheme=3, reconstruction=2, limiter=0, viscous_order=0, cfl=0.5,
                  fixed_time_step=0, output_screen=1)
    return dict(name="subsonic", mx=nx, my=ny, mz=1, ngg=2, x=x, y=y, z=z, setup=setup, bcs=bcs, scheme=scheme,
                inflow=inflow, seed=seed, perturb="density", steps=10)


def sst_state(st, mu, ratio, intensity):
    """BoundCond.cpp:83-95: k = 1.5 (|V| I)^2, mu_t = ratio mu,
    omega = rho k / mu_t, appended to the scalars of state `st`."""
    mut = ratio * mu
    k = 1.5 * st["vel"] ** 2 * intensity ** 2
    return dict(st, sv=list(st.get("sv", [])) + [k, st["rho"] * k / mut], mut=mut)


def stretched(n, length, first):
    """n nodes on [0, length] growing geometrically from spacing `first`."""
    lo, hi = 1.0, 2.0
    for _ in range(200):
        r = 0.5 * (lo + hi)
        total = first * (r ** (n - 1) - 1) / (r - 1)
        lo, hi = (r, hi) if total < length else (lo, r)
    r = 0.5 * (lo + hi)
    y = np.concatenate([[0.0], np.cumsum(first * r ** np.arange(n - 1))])
    return y * (length / y[-1])


def two_wall(n, length, first):
    """n nodes on [0, length] clustered symmetrically towards both ends,
    y = L/2 (1 + tanh(b (2s - 1)) / tanh b) on uniform s, with b chosen so
    that the first spacing equals `first`."""
    s_ = np.linspace(0.0, 1.0, n)
    lo, hi = 1e-6, 20.0
    for _ in range(200):
        b = 0.5 * (lo + hi)
        y1 = 0.5 * length * (1 + np.tanh(b * (2 * s_[1] - 1)) / np.tanh(b))
        lo, hi = (b, hi) if y1 > first else (lo, b)
    b = 0.5 * (lo + hi)
    y = 0.5 * length * (1 + np.tanh(b * (2 * s_ - 1)) / np.tanh(b))
    y[0], y[-1] = 0.0, length
    return y


def sst_plate_case(nx=64, ny=48, seed=1007, mach=2.0, p=20000.0, t=300.0, t_wall=300.0, lx=0.1, ly=0.02,
                   first=2e-5, le=8, intensity=0.01, ratio=10.0, compressibility=3, inviscid_scheme=3,
                   reconstruction=2):
    """2-D turbulent (k-omega SST) supersonic flat plate in air: Mach 2 inflow
    at i=0 carrying k, omega from a turbulence intensity and viscosity ratio,
    outflow at i=max and j=max, symmetry on j=0 ahead of the leading edge
    (node `le`) and an isothermal no-slip wall behind it; wall-normal
    geometric stretching from spacing `first`.  AUSM+ with MUSCL-minmod
    (reconstruction 2) or AWENO-5 carrying k and omega (reconstruction 4),
    second-order viscous fluxes with mu + mu_t, SSP-RK3 (or the steady path
    when the caller switches it) at CFL 0.5; Zeman compressibility
    correction by default."""
    xi = np.linspace(0.0, lx, nx)
    yj = stretched(ny, ly, first)
    X, Y = np.meshgrid(xi, yj, indexing="xy")
    x, y = X[None].copy(), Y[None].copy()
    z = np.zeros_like(x)
    inflow = sst_state(air_inflow(mach, p, t), sutherland(t), ratio, intensity)
    setup = _setup_common({
        "int dimension": 2, "real cfl": 0.5, "int inviscid_scheme": inviscid_scheme,
        "int reconstruction": reconstruction,
        "int limiter": 0, "int viscous_order": 2, "int total_step": 10, "bool turbulence": 1,
        "int turbulence_method": 1, "int RANS_model": 2, "int turb_implicit": 1,
        "int compressibility_correction": compressibility,
        "array string boundary_conditions": ["freestream", "outflow", "symmetry", "wall"],
        "string reference_state": "freestream", "string default_init": "freestream",
        "struct freestream": [("string", "type", "inflow"), ("int", "label", 5), ("real", "mach", mach),
                              ("real", "pressure", p), ("real", "temperature", t), ("real", "u", 1.0),
                              ("real", "v", 0.0), ("real", "w", 0.0), ("real", "turbulence_intensity", intensity),
                              ("real", "turb_viscosity_ratio", ratio)],
        "struct outflow": [("string", "type", "outflow"), ("int", "label", 6)],
        "struct symmetry": [("string", "type", "symmetry"), ("int", "label", 3)],
        "struct wall": [("string", "type", "wall"), ("int", "label", 2), ("string", "thermal_type", "isothermal"),
                        ("real", "temperature", t_wall)],
    })
    bcs = [
        dict(type=5, raw=(0, 0, 0, ny - 1, 0, 0), label=5),
        dict(type=6, raw=(nx - 1, nx - 1, 0, ny - 1, 0, 0), label=6),
        dict(type=6, raw=(0, nx - 1, ny - 1, ny - 1, 0, 0), label=6),
        dict(type=3, raw=(0, le - 1, 0, 0, 0, 0), label=3),
        dict(type=2, raw=(le, nx - 1, 0, 0, 0, 0), label=2, t_wall=t_wall, isothermal=True),
    ]
    scheme = dict(mix=0, inviscid_scheme=inviscid_scheme, reconstruction=reconstruction, limiter=0, viscous_order=2,
                  cfl=0.5, fixed_time_step=0, output_screen=1, turbulence=1, turb_implicit=1,
                  compressibility_correction=compressibility)
    return dict(name="sst_plate", mx=nx, my=ny, mz=1, ngg=3 if reconstruction == 4 else 2, x=x, y=y, z=z,
                setup=setup, bcs=bcs, scheme=scheme,
                inflow=inflow, seed=seed, perturb="density", steps=10, le=le, delta=0.2 * ly)


def sst_plate_fields(case):
    """Uniform inflow, a tanh boundary-layer profile u = U tanh(y / delta)
    over the plate (nodes i >= le) with T from the recovery-free
    Crocco-Busemann relation to the wall temperature, then the seeded 1 %
    density perturbation at fixed pressure; k and omega uniform."""
    bv, sv = initial_fields(dict(case, perturb=None))
    mx, my, mz, g = case["mx"], case["my"], case["mz"], case["ngg"]
    st = case["inflow"]
    yy = np.zeros(ghost_shape(mx, my, mz, g))
    yy[:, g:g + my, g:g + mx] = case["y"]
    f = np.tanh(np.maximum(yy, 0.0) / case["delta"])
    plate = np.zeros_like(yy, dtype=bool)
    plate[:, :, g + case["le"]:] = True
    tw = case["bcs"][-1]["t_wall"]
    u = np.where(plate, st["u"] * f, st["u"])
    t = np.where(plate, tw + (st["t"] - tw) * f, st["t"])
    bv[1] = u
    bv[5] = t
    rng = SplitMix64(case["seed"])
    noise = 1.0 + 0.01 * rng.normal(mx * my * mz).reshape(mz, my, mx)
    bv[0] = st["p"] * MW_AIR / (R_U * bv[5])
    sl = (slice(g, g + mz), slice(g, g + my), slice(g, g + mx))
    bv[0][sl] = bv[0][sl] * noise
    bv[5][sl] = bv[4][sl] * MW_AIR / (bv[0][sl] * R_U)
    return bv, sv


def c5_ports(nx, nz, lx, lz, x_port=0.25, size=2e-3, count=4):
    """Node ranges (i0, i1, k0, k1) of `count` square ports of side `size`
    centred at x = x_port * lx, spread evenly across z."""
    di = max(1, int(round(size / (lx / (nx - 1)))))
    dk = max(1, int(round(size / (lz / (nz - 1)))))
    ic = int(round(x_port * (nx - 1)))
    ports = []
    for c in range(count):
        kc = int(round((c + 0.5) / count * (nz - 1)))
        ports.append((ic - di // 2, ic - di // 2 + di - 1, kc - dk // 2, kc - dk // 2 + dk - 1))
    return ports


def c5_case(nx=384, ny=192, nz=128, seed=1005, lx=0.2, ly=0.05, lz=0.04, first=5e-6, p=101325.0, t_air=1200.0,
            mach=2.0, p_jet=202650.0, t_jet=300.0, t_wall=300.0, intensity=0.01, ratio=10.0, inner=3,
            chem=1, compressibility=3, reconstruction=4):
    """3-D reacting RANS (k-omega SST) supersonic H2/air channel with wall
    injection on one block: walls at j=0 and j=max with geometric stretching
    from spacing `first` towards both (y+ ~ 1), vitiated-air inflow at i=0
    (Y_O2 = 0.25, Y_H2O = 0.12, Y_N2 = 0.63; 1200 K, Mach 2), outflow at
    i=max, symmetry at k=0 and k=max, and sonic H2 jets at 300 K through
    four 2 mm square ports in the lower wall (inflow patches, normal
    injection).  9 species plus k and omega; HLLC with AWENO-5
    (reconstruction 4: the multi-species characteristic interpolation
    carrying k and omega, 3 ghost layers) or MUSCL-minmod (reconstruction
    2), viscous fluxes with mu + mu_t, steady pseudo-time with DPLUR (3
    inner sweeps) and the exact point-implicit chemistry, implicit SST
    source diagonal."""
    xi = np.linspace(0.0, lx, nx)
    yj = two_wall(ny, ly, first)
    zk = np.linspace(0.0, lz, nz)
    Z, Y, X = np.meshgrid(zk, yj, xi, indexing="ij")
    x, y, z = X.copy(), Y.copy(), Z.copy()
    common = [("real", "u", 1.0), ("real", "v", 0.0), ("real", "w", 0.0),
              ("real", "turbulence_intensity", intensity), ("real", "turb_viscosity_ratio", ratio)]
    setup = _setup_common({
        "int dimension": 3, "int species": 1, "int reaction": 1, "int chemSrcMethod": chem, "real cfl": 0.5,
        "int inviscid_scheme": 4, "int reconstruction": reconstruction, "int limiter": 0, "int viscous_order": 2,
        "int total_step": 10, "bool steady": 1, "int implicit_method": 1, "int DPLUR_inner_step": inner,
        "bool turbulence": 1, "int turbulence_method": 1, "int RANS_model": 2, "int turb_implicit": 1,
        "int compressibility_correction": compressibility,
        "array string boundary_conditions": ["vitiated_air", "h2_jet", "outflow", "symmetry", "wall"],
        "string reference_state": "vitiated_air", "string default_init": "vitiated_air",
        "struct vitiated_air": [("string", "type", "inflow"), ("int", "label", 5), ("real", "mach", mach),
                                ("real", "pressure", p), ("real", "temperature", t_air), ("real", "O2", 0.25),
                                ("real", "H2O", 0.12), ("real", "N2", 0.63)] + common,
        "struct h2_jet": [("string", "type", "inflow"), ("int", "label", 52), ("real", "mach", 1.0),
                          ("real", "pressure", p_jet), ("real", "temperature", t_jet), ("real", "H2", 1.0),
                          ("real", "u", 0.0), ("real", "v", 1.0), ("real", "w", 0.0),
                          ("real", "turbulence_intensity", intensity), ("real", "turb_viscosity_ratio", ratio)],
        "struct outflow": [("string", "type", "outflow"), ("int", "label", 6)],
        "struct symmetry": [("string", "type", "symmetry"), ("int", "label", 3)],
        "struct wall": [("string", "type", "wall"), ("int", "label", 2), ("string", "thermal_type", "isothermal"),
                        ("real", "temperature", t_wall)],
    })
    ports = c5_ports(nx, nz, lx, lz)
    # tensor partition of the lower wall face into wall and port rectangles
    ib = sorted({0, nx} | {a for p_ in ports for a in (p_[0], p_[1] + 1)})
    kb = sorted({0, nz} | {a for p_ in ports for a in (p_[2], p_[3] + 1)})
    bcs = [
        dict(type=5, raw=(0, 0, 0, ny - 1, 0, nz - 1), label=5, stream="vitiated_air"),
        dict(type=6, raw=(nx - 1, nx - 1, 0, ny - 1, 0, nz - 1), label=6),
        dict(type=3, raw=(0, nx - 1, 0, ny - 1, 0, 0), label=3),
        dict(type=3, raw=(0, nx - 1, 0, ny - 1, nz - 1, nz - 1), label=3),
        dict(type=2, raw=(0, nx - 1, ny - 1, ny - 1, 0, nz - 1), label=2, t_wall=t_wall, isothermal=True),
    ]
    for a in range(len(ib) - 1):
        for c in range(len(kb) - 1):
            i0, i1, k0, k1 = ib[a], ib[a + 1] - 1, kb[c], kb[c + 1] - 1
            jet = any(p_[0] <= i0 and i1 <= p_[1] and p_[2] <= k0 and k1 <= p_[3] for p_ in ports)
            if jet:
                bcs.append(dict(type=5, raw=(i0, i1, 0, 0, k0, k1), label=52, stream="h2_jet"))
            else:
                bcs.append(dict(type=2, raw=(i0, i1, 0, 0, k0, k1), label=2, t_wall=t_wall, isothermal=True))
    scheme = dict(mix=1, reaction=1, chemSrcMethod=chem, inviscid_scheme=4, reconstruction=reconstruction, limiter=0,
                  viscous_order=2, cfl=0.5, fixed_time_step=0, output_screen=1, implicit_method=1,
                  DPLUR_inner_step=inner, turbulence=1, turb_implicit=1, compressibility_correction=compressibility)
    return dict(name="c5", mx=nx, my=ny, mz=nz, ngg=3 if reconstruction == 4 else 2, x=x, y=y, z=z, setup=setup,
                bcs=bcs, scheme=scheme,
                seed=seed, steps=10, ports=ports)


def c5_fields(case, air):
    """Initial bv (6, ...) and sv (11, ...) of the channel: the vitiated-air
    inflow state everywhere (ghosts included) with the seeded 1 % density
    perturbation at fixed pressure in the interior.  `air` is the inflow
    state dict (rho, u, v, w, p, t, sv incl. k, omega, mw)."""
    mx, my, mz, g = case["mx"], case["my"], case["mz"], case["ngg"]
    bv, sv = initial_fields(dict(case, inflow=air, perturb="density"))
    return bv, sv


def initial_fields(case):
    """Uniform inflow everywhere (ghosts included), then the seeded interior
    density perturbation at fixed pressure.  Returns bv (6, ...) and sv
    (n_scalar, ...) in the ghosted reference layout."""
    mx, my, mz, g = case["mx"], case["my"], case["mz"], case["ngg"]
    shp = ghost_shape(mx, my, mz, g)
    st = case["inflow"]
    bv = np.empty((6,) + shp)
    for l, key in enumerate(("rho", "u", "v", "w", "p", "t")):
        bv[l] = st[key]
    nsc = len(st.get("sv", []))
    sv = np.empty((nsc,) + shp)
    for l in range(nsc):
        sv[l] = st["sv"][l]
    if case.get("perturb") == "density":
        rng = SplitMix64(case["seed"])
        n = mx * my * mz
        f = 1.0 + 0.01 * rng.normal(n).reshape(mz, my, mx)
        sl = (slice(g, g + mz), slice(g, g + my), slice(g, g + mx))
        rho = bv[0][sl] * f
        bv[0][sl] = rho
        mw = st.get("mw", MW_AIR)
        bv[5][sl] = bv[4][sl] * mw / (rho * R_U)
    return bv, sv


def chem_batch_inputs(n, seed=1004, species=("H2", "H", "O2", "O", "OH", "HO2", "H2O2", "H2O", "N2"), mw=None):
    """Chemistry-only batch: T ~ U(900, 2400) K, p ~ logU(0.5, 5) atm,
    equivalence ratio ln(phi) ~ N(ln 0.9487, 0.576) clipped to [0.3, 3],
    H2/O2/N2 from phi with air O2:N2 = 0.233:0.767 by mass, radicals
    H/O/OH/HO2 ~ logU(1e-6, 1e-3) mass fraction, H2O = H2O2 = 0, renormalised;
    rho from the equation of state.  Returns (rho, T, p, Y[ns, n])."""
    rng = SplitMix64(seed)
    t = 900.0 + 1500.0 * rng.uniform(n)
    p = 101325.0 * np.exp(np.log(0.5) + (np.log(5.0) - np.log(0.5)) * rng.uniform(n))
    phi = np.exp(np.log(0.9487) + 0.576 * rng.normal(n))
    phi = np.clip(phi, 0.3, 3.0)
    idx = {s: i for i, s in enumerate(species)}
    ns = len(species)
    Y = np.zeros((ns, n))
    # stoichiometric H2/O2 mass ratio: 2 * 2.016 / 31.998
    st = 2 * 2.016 / 31.998
    y_air = 1.0 / (1.0 + phi * st * (1.0 / 0.233))
    y_h2 = 1.0 - y_air
    Y[idx["H2"]] = y_h2
    Y[idx["O2"]] = 0.233 * y_air
    Y[idx["N2"]] = 0.767 * y_air
    for s in ("H", "O", "OH", "HO2"):
        Y[idx[s]] = np.exp(np.log(1e-6) + (np.log(1e-3) - np.log(1e-6)) * rng.uniform(n))
    Y /= Y.sum(axis=0, keepdims=True)
    if mw is None:
        mw = np.array([2.016, 1.008, 31.998, 15.999, 17.007, 33.006, 34.014, 18.015, 28.014])
    mwinv = (Y / mw[:, None]).sum(axis=0)
    rho = p / (R_U * t * mwinv)
    return rho, t, p, Y


def mixing_layer_states(n, seed=2003, species=("H2", "H", "O2", "O", "OH", "HO2", "H2O2", "H2O", "N2"), mw=None):
    """Chemistry training states spanning the reacting mixing layer: mixture
    fraction Z ~ U(0, 1) between the air stream (Y_O2 = 0.233, Y_N2 = 0.767)
    and the fuel stream (Y_H2 = 0.3, Y_N2 = 0.7), the frozen mixing
    temperature 1500 (1 - Z) + 300 Z plus U(-50, 600) K (floored at 250 K),
    p ~ logU(0.5, 2) atm, and H, O, OH, HO2, H2O2, H2O at 10^U(-16, -4)
    mass fraction (so that streams without radicals lie inside the range);
    renormalised, rho from the equation of state.  Returns (rho, T, p,
    Y[ns, n])."""
    rng = SplitMix64(seed)
    z = rng.uniform(n)
    t = 1500.0 * (1 - z) + 300.0 * z - 50.0 + 650.0 * rng.uniform(n)
    t = np.maximum(t, 250.0)
    p = 101325.0 * np.exp(np.log(0.5) + (np.log(2.0) - np.log(0.5)) * rng.uniform(n))
    idx = {s: i for i, s in enumerate(species)}
    Y = np.zeros((len(species), n))
    Y[idx["O2"]] = 0.233 * (1 - z)
    Y[idx["N2"]] = 0.767 * (1 - z) + 0.7 * z
    Y[idx["H2"]] = 0.3 * z
    for s in ("H", "O", "OH", "HO2", "H2O2", "H2O"):
        Y[idx[s]] = 10.0 ** (-16.0 + 12.0 * rng.uniform(n))
    Y /= Y.sum(axis=0, keepdims=True)
    if mw is None:
        mw = np.array([2.016, 1.008, 31.998, 15.999, 17.007, 33.006, 34.014, 18.015, 28.014])
    rho = p / (R_U * t * (Y / mw[:, None]).sum(axis=0))
    return rho, t, p, Y


def _fmt(v):
    if isinstance(v, (float, np.floating)):
        return repr(float(v))
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return str(v)


def write_case_dir(case, root):
    """Write a reference-format case directory for `case` under `root`."""
    setup_dir = os.path.join(root, "input_files", "setup")
    bc_dir = os.path.join(root, "input_files", "boundary_condition")
    os.makedirs(setup_dir, exist_ok=True)
    os.makedirs(bc_dir, exist_ok=True)
    chem = os.path.join(root, "input_files", "chemistry")
    if not os.path.exists(chem):
        shutil.copytree(CHEM_DIR, chem)
    lines = []
    for key, val in case["setup"].items():
        kind, name = key.split(" ", 1)
        if kind == "struct":
            lines.append(f"struct {name} {{")
            for t, k, v in val:
                lines.append(f"    {t} {k} = {_fmt(v)}")
            lines.append("}")
        elif kind == "array":
            t, nm = name.split(" ", 1)
            lines.append(f"array {t} {nm} {{")
            lines.append("    " + " ".join(_fmt(v) for v in val))
            lines.append("}")
        else:
            lines.append(f"{kind} {name} = {_fmt(val)}")
    with open(os.path.join(setup_dir, "default_settings.txt"), "w") as f:
        f.write("// generated case: " + case["name"] + "\n" + "\n".join(lines) + "\n")
    with open(os.path.join(bc_dir, "boundary   0.txt"), "w") as f:
        f.write(f"{len(case['bcs']):>12d}\n")
        for b in case["bcs"]:
            r = b["raw"]
            f.write("".join(f"{v:>6d}" for v in r) + f"{b['label']:>6d}\n")
    return root


# ------------------------------------------------------------ multi-block
def _rot(a, rotate):
    """Turn an array (..., k, j, i) by 180 degrees about k (i and j reversed)."""
    return a[..., ::-1, ::-1].copy() if rotate else a


def split_case(case, axis=0, at=None, rotate=False):
    """The single-block `case` cut into two blocks joined by one inner face:
    block 0 holds the index range [0, at] along `axis` (0 = i, 1 = j), block
    1 the range [at, m - 1] (the face layer belongs to both).  rotate: block
    1 is stored turned by 180 degrees about k (its i and j run backwards), so
    the inner face pairs a high face with a high face and reverses the
    tangential direction.  Boundary ranges are split with the blocks.
    Returns a multi-block case: blocks (mx, my, mz, x, y, z, bcs, lo, hi),
    inner (block, source line, target line) in the reference's 1-based
    inner-face format, and the single-block case under "full"."""
    if axis not in (0, 1):
        raise ValueError("split along i or j")
    m = (case["mx"], case["my"], case["mz"])
    at = m[axis] // 2 if at is None else at
    dim = 2 if case["mz"] == 1 else 3
    blocks = []
    for b, (lo, hi) in enumerate(((0, at), (at, m[axis] - 1))):
        rot = rotate and b == 1
        sl = [slice(None)] * 3
        sl[2 - axis] = slice(lo, hi + 1)
        x = _rot(case["x"][tuple(sl)], rot)
        y = _rot(case["y"][tuple(sl)], rot)
        z = _rot(case["z"][tuple(sl)], rot)
        ext = list(m)
        ext[axis] = hi - lo + 1
        bcs = []
        for bc in case["bcs"]:
            r = list(bc["raw"])
            a1, a2 = sorted((r[2 * axis], r[2 * axis + 1]))
            a1, a2 = max(a1, lo), min(a2, hi)
            if a1 > a2:
                continue
            # a face lying on the cut itself is not a boundary
            if r[2 * axis] == r[2 * axis + 1] and lo < r[2 * axis] < hi:
                continue
            r[2 * axis], r[2 * axis + 1] = a1 - lo, a2 - lo
            if rot:
                for d in (0, 1):
                    v1, v2 = ext[d] - 1 - r[2 * d], ext[d] - 1 - r[2 * d + 1]
                    r[2 * d], r[2 * d + 1] = min(v1, v2), max(v1, v2)
            bcs.append(dict(bc, raw=tuple(r)))
        blocks.append(dict(mx=ext[0], my=ext[1], mz=ext[2], x=x, y=y, z=z, bcs=bcs, lo=lo, hi=hi, rotated=rot))
    n1 = blocks[1]["mx"] if axis == 0 else blocks[1]["my"]
    tang = 1 - axis

    def line(b, face_idx, reverse, blk_id):
        v = [0] * 7
        v[2 * axis] = v[2 * axis + 1] = face_idx
        t1, t2 = (m[tang], 1) if reverse else (1, m[tang])
        if dim == 3:
            t1, t2 = -t1, -t2
        v[2 * tang], v[2 * tang + 1] = t1, t2
        v[4], v[5] = 1, m[2]
        v[6] = blk_id
        return v

    f0 = line(0, at + 1, False, 1)
    f1 = line(1, n1 if rotate else 1, rotate, 2)
    inner = [(0, f0, f1), (1, f1, f0)]
    mc = dict(case, name=case["name"] + "_split", blocks=blocks, inner=inner, full=case, split_axis=axis)
    for key in ("mx", "my", "mz", "x", "y", "z", "bcs"):
        mc.pop(key, None)
    return mc


def split_fields(mcase, bv, sv):
    """Per-block ghosted fields cut from the single-block ghosted fields."""
    g = mcase["ngg"]
    axis = mcase["split_axis"]
    out = []
    for blk in mcase["blocks"]:
        sl = [slice(None)] * 4
        sl[3 - axis] = slice(blk["lo"], blk["hi"] + 1 + 2 * g)
        fb = _rot(bv[tuple(sl)], blk["rotated"])
        fs = _rot(sv[tuple(sl)], blk["rotated"]) if sv.shape[0] else sv[tuple(sl)]
        out.append((np.ascontiguousarray(fb), np.ascontiguousarray(fs)))
    return out


def block_interior_of_full(mcase, b, arr, ncomp):
    """Block b's interior cut from a single-block ghosted array (ncomp, ...),
    in block b's index order."""
    full = mcase["full"]
    g = mcase["ngg"]
    a = np.asarray(arr).reshape((ncomp,) + ghost_shape(full["mx"], full["my"], full["mz"], g))
    a = a[:, g:g + full["mz"], g:g + full["my"], g:g + full["mx"]]
    blk = mcase["blocks"][b]
    sl = [slice(None)] * 4
    sl[3 - mcase["split_axis"]] = slice(blk["lo"], blk["hi"] + 1)
    return _rot(a[tuple(sl)], blk["rotated"])


def write_mesh_case_dir(mcase, root):
    """Reference-format case directory of a multi-block case: setup files,
    grid file, boundary file and inner-face file."""
    first = dict(mcase["full"], bcs=mcase["blocks"][0]["bcs"])
    write_case_dir(first, root)
    bc_dir = os.path.join(root, "input_files", "boundary_condition")
    with open(os.path.join(bc_dir, "boundary   0.txt"), "w") as f:
        for blk in mcase["blocks"]:
            f.write(f"{len(blk['bcs']):>12d}\n")
            for b in blk["bcs"]:
                f.write("".join(f"{v:>6d}" for v in b["raw"]) + f"{b['label']:>6d}\n")
    nb = len(mcase["blocks"])
    counts = [sum(1 for (fb, _, _) in mcase["inner"] if fb == b) for b in range(nb)]
    with open(os.path.join(bc_dir, "inner   0.txt"), "w") as f:
        f.write(f"{nb:>12d}\n")
        f.write("".join(f"{c:>12d}" for c in counts) + "\n")
        f.write(" imin  imax  jmin  jmax  kmin  kmax   id\n")
        for b in range(nb):
            for (fb, src, tar) in mcase["inner"]:
                if fb == b:
                    f.write("".join(f"{v:>9d}" for v in src) + "\n")
                    f.write("".join(f"{v:>9d}" for v in tar) + "\n")
    grid_dir = os.path.join(root, "input_files", "grid")
    os.makedirs(grid_dir, exist_ok=True)
    with open(os.path.join(grid_dir, "grid   0.grd"), "w") as f:
        f.write(f"{nb:>12d}\n")
        for blk in mcase["blocks"]:
            f.write(f"{blk['mx']:>12d}{blk['my']:>12d}{blk['mz']:>12d}\n")
        for blk in mcase["blocks"]:
            for key in ("x", "y", "z"):
                vals = np.asarray(blk[key], dtype=np.float64).ravel()
                for q in range(0, vals.size, 4):
                    f.write(" ".join(repr(float(v)) for v in vals[q:q + 4]) + "\n")
    return root
