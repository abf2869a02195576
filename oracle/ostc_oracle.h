/*
 * OpenSTC CPU oracle — TEST INFRASTRUCTURE ONLY.
 *
 * A serial FP64 restatement in plain C of the reference's per-cell routines on
 * the data-parallel hot path (thermo, transport, finite-rate chemistry,
 * reconstruction, Riemann fluxes, viscous flux, RK3 update, CFL dt, residual,
 * boundary conditions, grid metrics).  It is the checker the HIP path is
 * compared against and the timed CPU baseline; the product never links,
 * loads or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it.
 *
 * Every routine cites the reference file:line it follows (paths relative to
 * the reference's src/).  Arrays use the reference DZone layout
 * (gxl_lib/Array.hpp:141-155,174-194): component-major, column-major,
 *   val[(k+g)*disp1 + (j+g)*disp2 + (i+g) + l*sz],
 *   disp2 = mx+2g, disp1 = (my+2g)*disp2, sz = (mx+2g)(my+2g)(mz+2g).
 *
 * Parity pins (how each part is checked without trusting this file) are
 * listed in tests/test_oracle_*.py and tests/test_multiblock.py: closed
 * forms, published formulas, conservation and consistency properties,
 * finite-difference Jacobians, convergence orders and, for the multi-block
 * mesh, cut cases that must reproduce their single-block runs.  A routine
 * that no independent check reaches is marked "parity unpinned" in its
 * comment.
 */
#ifndef OSTC_ORACLE_H
#define OSTC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXS 20
#define ORC_MAXR 50
#define ORC_MAXBC 32

/* Species / reaction tables: the content of the reference's DParameter
 * (DParameter.cuh:13-92) for one mechanism.  Matrices are row-major,
 * packed with the live sizes (n_spec, n_reac). */
typedef struct {
  int n_spec, n_reac;
  double mw[ORC_MAXS], t_low[ORC_MAXS], t_mid[ORC_MAXS], t_high[ORC_MAXS];
  double hi[ORC_MAXS * 7], lo[ORC_MAXS * 7];
  double LJ_potent_inv[ORC_MAXS], vis_coeff[ORC_MAXS];
  double WjDivWi_to_One4th[ORC_MAXS * ORC_MAXS];
  double sqrt_WiDivWjPl1Mul8[ORC_MAXS * ORC_MAXS];
  double binary_diffusivity_coeff[ORC_MAXS * ORC_MAXS];
  double kb_over_eps_jk[ORC_MAXS * ORC_MAXS];
  int label[ORC_MAXR];
  int stoi_f[ORC_MAXR * ORC_MAXS], stoi_b[ORC_MAXR * ORC_MAXS];
  int order[ORC_MAXR];
  double A[ORC_MAXR], b[ORC_MAXR], Ea[ORC_MAXR];
  double A2[ORC_MAXR], b2[ORC_MAXR], Ea2[ORC_MAXR];
  double third_body_coeff[ORC_MAXR * ORC_MAXS];
  double troe_alpha[ORC_MAXR], troe_t3[ORC_MAXR], troe_t1[ORC_MAXR], troe_t2[ORC_MAXR];
} orc_mech;

/* ---- point thermo / transport / chemistry ---- */
void orc_enthalpy(double t, const orc_mech *m, double *h);
void orc_enthalpy_and_cp(double t, const orc_mech *m, double *h, double *cp);
void orc_cp(double t, const orc_mech *m, double *cp);
void orc_gibbs_rt(double t, const orc_mech *m, double *g);
double orc_sutherland(double t);
void orc_transport(const orc_mech *m, double t, double p, double rho, const double *Y, const double *cp_i,
                   double *mu, double *lambda, double *rhoD);
void orc_forward_rate(const orc_mech *m, double t, const double *c, double *kf);
void orc_backward_rate(const orc_mech *m, double t, const double *kf, double *kb);
void orc_rate_of_progress(const orc_mech *m, const double *kf, const double *kb, const double *c,
                          double *q, double *q1, double *q2);
void orc_chemical_source(const orc_mech *m, const double *q1, const double *q2, double *omega_d, double *omega);
/* omega [kg/m3/s]; jac (n_spec^2, may be NULL): d omega_m / d(rho Y_n), reference EPI form */
void orc_chem_source_point(const orc_mech *m, double rho, double t, const double *Y, double *omega, double *jac,
                           double *jac_diag);
void orc_lu_solve(double *lhs, double *rhs, int n);
/* Newton temperature from total energy (per volume) at fixed Y; returns T, writes p and iteration count */
double orc_temperature_newton(const orc_mech *m, double rho, double rhoE, const double *vel3, const double *Y,
                              double t_start, double *p_out, int *iters);
/* One dt point-implicit step of an isolated constant-volume cell: returns per-cell
 * dY (ns), dT.  Layout: Y and dY are SoA [l*n + i]. */
void orc_chem_batch(const orc_mech *m, int64_t n, const double *rho, const double *T, const double *Y,
                    double dt, double *dY, double *dT);

/* ---- flow solver ---- */
typedef struct {
  int type;                   /* 2 wall, 3 symmetry, 4 farfield, 5 inflow, 6 outflow, 7 subsonic inflow,
                                 9 back pressure */
  int raw[6];                 /* i1 i2 j1 j2 k1 k2 as in boundary file (0-based) */
  /* inflow state */
  double rho, u, v, w, p, t, vel;
  double sv[ORC_MAXS + 4];
  /* wall */
  double t_wall;
  int isothermal;
  /* farfield: a, gamma, entropy p / rho^gamma, mw of the far state (rho..vel, sv above) */
  double a_inf, gamma_inf, entropy, mw_inf;
  /* subsonic inflow: total pressure and temperature (u, v, w above: direction); back pressure */
  double p_total, t_total, p_back;
  /* turbulent viscosity of the inflow / farfield state (k, omega are in sv after the species) */
  double mut;
} orc_bc;

typedef struct {
  int mix;                    /* 0 Air (perfect gas), 1 multi-species */
  int reaction;               /* 0 none, 1 finite rate */
  int chemSrcMethod;          /* 0 explicit, 1 EPI, 2 DA */
  int inviscid_scheme;        /* 1 LF, 2 Roe, 3 AUSM+, 4 HLLC */
  int reconstruction;         /* 1 first order, 2 MUSCL, 3 NND2, 4 WENO5 (AWENO) */
  int limiter;                /* 0 minmod */
  int viscous_order;          /* 0 inviscid, 2 central */
  int fixed_time_step;
  int output_screen;
  int gradPInDiffusionFlux;
  int implicit_method;        /* steady path: 0 explicit (+ EPI/DA chemistry), 1 DPLUR */
  int dplur_inner_step;
  int turbulence;             /* 0 laminar, 1 k-omega SST (two scalars k, omega after the species) */
  int turb_implicit;          /* SST source Jacobian diagonal in the implicit treatment */
  int compressibility_correction; /* 0 none, 1 Wilcox, 2 Sarkar, 3 Zeman */
  int split_chem;             /* finite-rate source left out of the RHS (operator-split chemistry) */
  double cfl, Pr, Sc, Prt, Sct, dt, entropy_fix_factor;
  int n_bc;
  orc_bc bc[ORC_MAXBC];
  /* steady path with split_chem: one point-implicit chemistry step of this
   * dt per pseudo-iteration, after the flow update (0: none) */
  double steady_chem_dt;
  /* conservative symmetric metrics that keep a uniform flow exact (0: the
   * reference's metrics) */
  int conservative_metric;
} orc_cfg;

typedef struct orc_zone orc_zone;

orc_zone *orc_zone_create(int mx, int my, int mz, int ngg, const orc_cfg *cfg, const orc_mech *mech,
                          const double *x, const double *y, const double *z);
void orc_zone_destroy(orc_zone *zz);
/* field ids for orc_zone_field */
enum {
  ORC_X = 0, ORC_Y, ORC_Z, ORC_JAC, ORC_METRIC, ORC_BV, ORC_SV, ORC_CV, ORC_DQ, ORC_VEL, ORC_ACOUSTIC,
  ORC_MACH, ORC_MUL, ORC_COND, ORC_RHOD, ORC_GAMMA, ORC_CP, ORC_DT_LOCAL, ORC_QN, ORC_BV_LAST, ORC_CHEM_JAC,
  ORC_MUT, ORC_TURB_COND, ORC_WALL_DIST, ORC_TURB_JAC
};
double *orc_zone_field(orc_zone *zz, int id, int64_t *count);
/* initialise after bv/sv were filled: 2-D k-copy, BCs, |V|, properties, cv from bv (Driver.cu:40-98, RK.cuh:38-48) */
void orc_zone_initialize(orc_zone *zz);
/* dq = source + inviscid + viscous (RK.cuh:109-116) */
void orc_zone_rhs(orc_zone *zz);
/* dt_local + min (TimeAdvanceFunc.cuh:50-100, TimeAdvanceFunc.cu:155-201) */
double orc_zone_global_dt(orc_zone *zz);
/* nsteps full RK3 steps (RK.cuh:89-193); res_hist gets 4 residuals + err_max per step (may be NULL) */
void orc_zone_run_rk3(orc_zone *zz, int nsteps, double *dt_hist, double *res_hist);
void orc_zone_state(orc_zone *zz, double *t_phys, int *step);

/* nsteps steady iterations with local time steps (SteadySim.cuh:52-124):
 * res_hist gets 4 residuals + err_max per step (NaN where not computed) */
void orc_zone_run_steady(orc_zone *zz, int nsteps, double *res_hist);
/* the implicit treatment alone on the current dq (ImplicitTreatmentHPP.cuh:7-52);
 * dt_local / spectral radii must be current */
void orc_zone_implicit(orc_zone *zz);
/* operator-split chemistry: one point-implicit step dt of every interior cell
 * at its density and total energy (orc_chem_batch on the gathered state),
 * then boundary conditions with cv, the 2-D k-copy and the properties */
void orc_zone_chem_step(orc_zone *zz, double dt);

/* grid metrics on their own (Mesh.cpp:259-521, 790-944): x,y,z interior in, outputs ghosted */
void orc_grid_metrics(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z,
                      double *xg, double *yg, double *zg, double *jac, double *metric);

/* ---- multi-block meshes (one GPU: blocks exchange through inner faces) ---- */
/* an inner face after Mesh.cpp:34-187 InnerFace::register_boundary (0-based) */
typedef struct {
  int rs[3], re[3];           /* this block's range: face layer + ngg ghost layers, full tangential extent */
  int face, dir;              /* normal direction (0 i, 1 j, 2 k), -1 on the low face */
  int ts[3], te[3];           /* target block's start (and end) indices */
  int tblk, tface, tdir;      /* target block (0-based), its face normal and side */
  int src_tar[3];             /* target direction i takes its offset from this block's direction src_tar[i] */
  int loop[3], tloop[3];      /* index steps along each direction in this block and the target */
  int n[3];                   /* points per direction */
} orc_inner_face;
int orc_inner_face_register(const int *s7, const int *t7, int ngg, int dim, orc_inner_face *f);
typedef struct orc_mesh orc_mesh;
/* nb blocks: dims 3 per block, cfgs one per block (its boundaries), xyz 3
 * interior coordinate arrays per block, nface[b] inner faces of block b and
 * lines[14 per face] = source line (7 ints) then target line (7 ints) as in
 * the reference's inner-face file.  NULL for an unsupported face. */
orc_mesh *orc_mesh_create(int nb, const int *dims, int ngg, const orc_cfg *cfgs, const orc_mech *mech,
                          const double *const *xyz, const int *nface, const int *lines);
void orc_mesh_destroy(orc_mesh *M);
orc_zone *orc_mesh_zone(orc_mesh *M, int b);
const orc_inner_face *orc_mesh_face(orc_mesh *M, int q);
void orc_mesh_initialize(orc_mesh *M);
void orc_mesh_run_rk3(orc_mesh *M, int nsteps, double *dt_hist, double *res_hist);
void orc_mesh_run_steady(orc_mesh *M, int nsteps, double *res_hist);
void orc_mesh_exchange(orc_mesh *M, int with_cv);

/* point-level test hooks (see orc_flow.c) */
void orc_face_flux_air(int scheme, const double *pl5, const double *pr5, const double *met6, const double *jac2,
                       double *f5);
void orc_face_flux_roe_air(const double *pl5, const double *pr5, const double *met6, const double *jac2,
                           const double *delta2, double *f5);
void orc_weno5_scalar(const double *v6, double *vm, double *vp);
void orc_aweno_matrices(const double *cl7, const double *cr7, const double *met6, double *L25, double *R25);
void orc_ssprk3_coeffs(double *a, double *b, double *c);
void orc_face_flux(int scheme, const orc_mech *mech, const double *pl, const double *pr, const double *met6,
                   const double *jac2, const double *delta2, double *f);
void orc_jac_times_dq(const orc_mech *mech, int n_scalar, const double *b6, const double *sv, double gam,
                      const double *m3, double lam, const double *dq, double *out);
void orc_reconstruct_air(int reconstruction, const double *pv4, double *pl, double *pr);
void orc_aweno_mix_matrices(const orc_mech *mech, int nsc, const double *cl, const double *cr, const double *k0,
                            const double *k1, double *L, double *R);
int orc_aweno_mix_sides(const orc_mech *mech, int nsc, const double *c6, const double *k0, const double *k1,
                        double *pl, double *pr);
void orc_zone_store_last(orc_zone *zz);
void orc_zone_source(orc_zone *zz);
void orc_zone_residual(orc_zone *zz, int step, double *res5);
/* SST point hooks: compressibility corrections (SST.cu:24-42) and the
 * constants gamma1, gamma2 derived from the model constants (SST.cuh:13-34) */
double orc_sst_wilcox(double mt);
double orc_sst_zeman(double mt, double gamma_p1);
void orc_sst_constants(double *gamma1, double *gamma2);

#ifdef __cplusplus
}
#endif
#endif
