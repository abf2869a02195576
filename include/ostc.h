/*
 * OpenSTC-MI355X C ABI: the per-step data-parallel hot path of the
 * structured-grid reacting-flow solver (reconstruction + Riemann flux,
 * central viscous flux with thermo/transport, finite-rate chemistry with the
 * point-implicit source, SSP-RK3 update, CFL dt, residual, boundary
 * conditions) on one gfx950 GPU.
 *
 * Conventions: every entry point returns 0 on success and a nonzero status
 * otherwise (no exception crosses the ABI); ostc_last_error() gives the
 * calling thread's message.  One host thread drives one solver; all device
 * work is ordered on the solver's HIP stream and only calls returning host
 * scalars synchronise.  Field arrays use the reference's DZone layout
 * (src/gxl_lib/Array.hpp:141-155): component-major, i fastest, ngg ghost
 * layers per side.  The reference exposes these operations as C++ template
 * launchers over (Block&, DZone*, DParameter*); each entry below names the
 * one it replaces.
 */
#ifndef OSTC_H
#define OSTC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ostc_solver ostc_solver;

/* field identifiers for ostc_field_info / ostc_upload / ostc_download */
enum ostc_field {
  OSTC_BV = 0,       /* rho,u,v,w,p,T (6)            ghosted  Field.h:31 */
  OSTC_SV = 1,       /* Y_1..Y_ns, k, omega (n_scalar) ghosted Field.h:32 */
  OSTC_CV = 2,       /* rho, rho u, .., rho E, rho Y (n_var) ghosted Field.h:30 */
  OSTC_DQ = 3,       /* RHS * J (n_var)              no ghosts Field.h:53 */
  OSTC_VEL = 4,      /* |V|                          ghosted */
  OSTC_ACOUSTIC = 5, /* speed of sound               ghosted */
  OSTC_MACH = 6,
  OSTC_MUL = 7,      /* laminar viscosity */
  OSTC_COND = 8,     /* thermal conductivity */
  OSTC_RHOD = 9,     /* rho D_k (n_spec) */
  OSTC_GAMMA = 10,
  OSTC_CP = 11,
  OSTC_DT_LOCAL = 12,/* no ghosts */
  OSTC_QN = 13,      /* cv at t^n */
  OSTC_BV_LAST = 14, /* residual work array (4), no ghosts */
  OSTC_JAC = 15,     /* cell Jacobian J = det(dx/dxi), ghosted */
  OSTC_METRIC = 16,  /* 9 components d xi_r/d x_c, component-major, ghosted */
  OSTC_CHEM_JAC = 17,/* n_spec^2 EPI Jacobian, no ghosts */
  OSTC_MUT = 18,     /* SST turbulent viscosity, ghosted            Field.h:46 */
  OSTC_TURB_COND = 19, /* SST turbulent conductivity mu_t cp / Prt, ghosted  Field.h:47 */
  OSTC_WALL_DIST = 20, /* SST wall distance, ghosted                Field.h:23 */
  OSTC_TURB_JAC = 21 /* SST source Jacobian diagonal of k, omega (2), no ghosts */
};

/* Version of this ABI (bumped on any signature change). */
int ostc_abi_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char *ostc_last_error(void);

/* Reads <case_dir>/input_files/setup/*.txt in the reference order with its
 * grammar and derived keys, and (species != 0) the CHEMKIN, therm and tran
 * files named there; selects HIP device `device` (device < 0 opens the
 * solver host-only: setup files, tables and grid metrics, no device work).
 * Multi-species mechanisms of 4 or 9 species (the counts the kernels are
 * compiled for) with at most 24 reactions; other counts are an error.
 * Replaces: Parameter::Parameter (Parameter.cpp:8-52), Species/Reaction
 * constructors (ChemData.cpp:11-632), DParameter::DParameter (DParameter.cu:7-159). */
int ostc_init(const char *case_dir, int device, ostc_solver **out);
int ostc_destroy(ostc_solver *s);

/* Integer / real setup keys after derivation (e.g. "ngg", "n_var", "n_spec",
 * "n_reac", "inviscid_type"). */
int ostc_get_int(ostc_solver *s, const char *key, int *value);
int ostc_get_real(ostc_solver *s, const char *key, double *value);
/* Mechanism tables as parsed (names as in DParameter: "mw", "t_low",
 * "high_coeff", "stoi_f", "label", "A", "Ea2", "third_body_coeff", ...).
 * `out` holds `count` doubles (integer tables are converted). */
int ostc_get_table(ostc_solver *s, const char *name, double *out, int64_t count);

/* Adds one structured block from its interior node coordinates (i fastest,
 * length mx*my*mz); builds ghost nodes and metrics.
 * Replaces: Mesh::read_grid + init_ghost_grid + compute_jac_metric
 * (Mesh.cpp:599-686, 790-944, 259-521). Returns the block index in *blk. */
int ostc_add_block(ostc_solver *s, int mx, int my, int mz, const double *x, const double *y, const double *z,
                   int *blk);
/* One boundary-file line "i1 i2 j1 j2 k1 k2 label" (0-based).
 * Replaces: Mesh::read_boundary (Mesh.cpp:686-706). */
int ostc_add_boundary(ostc_solver *s, int blk, int i1, int i2, int j1, int j2, int k1, int k2, int label);
/* Reads <case_dir>/input_files/boundary_condition/boundary   0.txt for all blocks. */
int ostc_read_boundary_file(ostc_solver *s);
/* Reads <case_dir>/input_files/grid/grid   0.grd (block count, extents, then
 * per block all x, all y, all z with i fastest), scales it by gridScale and
 * adds every block.  Replaces: Mesh::read_grid + Mesh::scale
 * (Mesh.cpp:599-686, 571-578, 776-788). */
int ostc_read_grid(ostc_solver *s);
/* One inner face of block blk: the source and target lines of the reference's
 * inner-face file (7 integers each, 1-based; the target line ends with the
 * target block, 1-based; in 3-D one tangential pair is negative to state the
 * orientation).  2-D faces must join the same index direction on both sides.
 * Replaces: InnerFace::InnerFace + register_boundary (Mesh.cpp:34-187). */
int ostc_add_inner_face(ostc_solver *s, int blk, const int *src7, const int *tar7);
/* Reads <case_dir>/input_files/boundary_condition/inner   0.txt.
 * Replaces: Mesh::read_inner_interface (Mesh.cpp:710-742). */
int ostc_read_inner_file(ostc_solver *s);
/* Grid, boundary file and (when present) inner-face file of the case.
 * Replaces: the reading part of Mesh::Mesh (Mesh.cpp:550-568). */
int ostc_read_mesh(ostc_solver *s);
/* Number of blocks, and the interior extents (mx, my, mz) of one. */
int ostc_block_count(ostc_solver *s, int *n);
int ostc_block_dims(ostc_solver *s, int blk, int *dims);
/* Registered inner faces of a block: *count faces; out (31 ints per face, up
 * to cap faces): rs[3] re[3] face dir ts[3] te[3] target_block target_face
 * target_dir src_tar[3] loop[3] target_loop[3] n[3] 0 0 (0-based). */
int ostc_inner_faces(ostc_solver *s, int blk, int *count, int *out, int cap);
/* Allocates the device zones, builds the ghost nodes of every block (those
 * beyond an inner face from the neighbour block), computes the Jacobian and
 * metrics on the device and resolves boundary labels against the setup-file
 * structs.  Replaces: Field::setup_device_memory (Field.cu:133-251),
 * init_ghost_grid + init_inner_ghost_grid + compute_jac_metric
 * (Mesh.cpp:790-994, 259-521), DBoundCond::initialize_bc_on_GPU
 * (BoundCond.cu:167-384). */
int ostc_finalize(ostc_solver *s);

/* A block's Jacobian (ghosted) and 9 metric components (component-major):
 * before ostc_finalize and in host-only mode the host computation of the
 * block on its own (reflected ghosts); after ostc_finalize the device arrays
 * in use.  Either pointer may be NULL. */
int ostc_block_geometry(ostc_solver *s, int blk, double *jac, double *metric);

/* Field shape: element count, ghost layers, components. */
int ostc_field_info(ostc_solver *s, int blk, int field, int64_t *count, int *ghosts, int *ncomp);
int ostc_upload(ostc_solver *s, int blk, int field, const double *host);
int ostc_download(ostc_solver *s, int blk, int field, double *host);
/* Device address of a field (for callers that keep their data in HBM). */
int ostc_field_device_ptr(ostc_solver *s, int blk, int field, double **dptr);

/* A block's state in caller-owned device memory, DZone layout (Field.h:13-71):
 * ghosted arrays of (mx+2ngg)(my+2ngg)(mz+2ngg) points, component-major,
 * i fastest; dq without ghosts (mx*my*mz per component). */
typedef struct ostc_zone_view {
  int mx, my, mz, ngg, n_var, n_scalar;
  double *bv; /* 6 components: rho, u, v, w, p, T (required) */
  double *sv; /* n_scalar components: Y_1..Y_ns, k, omega (required when n_scalar > 0) */
  double *cv; /* n_var components: rho, rho u, rho v, rho w, rho E, rho Y.., rho k, rho omega (required) */
  double *qn; /* n_var components: cv at t^n on the interior cells, stored by RK stage 0 (NULL: the library keeps its own) */
  double *dq; /* n_var components, no ghosts (NULL: the library keeps its own) */
} ostc_zone_view;
/* Binds caller device arrays as block blk's state after ostc_finalize: the
 * kernels then read and write them in place.  The sizes must match the
 * block, the setup's ngg, n_var and n_scalar.  The library never frees
 * caller memory (it frees its own copy of each bound field here); the
 * arrays must stay valid until ostc_destroy or a new binding.
 * Replaces: the DZone device pointers a reference block carries
 * (Field.h:13-71, Field.cu:182-189), which ostc_bind_zone hands over. */
int ostc_bind_zone(ostc_solver *s, int blk, const ostc_zone_view *view);

/* Uniform initial field from the setup files' default_init inflow.
 * Replaces: initialize_from_start (Initialize.cu:5-47) for groups_init = 1. */
int ostc_initialize_uniform(ostc_solver *s, int blk);
/* Reference inflow state of struct `name`: out[0..6] = rho,u,v,w,p,T,|V|,
 * out[7..7+n_scalar) = scalars (species, then k and omega with SST), then
 * mu_t.  Replaces: Inflow::Inflow (BoundCond.cpp:7-122). */
int ostc_inflow_state(ostc_solver *s, const char *name, double *out, int n);
/* Start-of-run sequence on every block: 2-D k-copy, boundary conditions,
 * |V|, properties, last-step store, cv from bv on all points.
 * Replaces: Driver::initialize_computation (Driver.cu:40-98) + RK.cuh:38-48. */
int ostc_initialize(ostc_solver *s);

/* dq = source + inviscid + viscous for one block.
 * Replaces: compute_source (SourceTerm.cuh:11-37), compute_inviscid_flux
 * (SchemeSelector.cuh:30-47), compute_viscous_flux (SchemeSelector.cuh:49-69). */
int ostc_rhs(ostc_solver *s, int blk);
/* Local CFL dt of every interior cell of one block (the dt_local field and,
 * for the steady path, the spectral radii); stream-ordered, no host sync.
 * Replaces: local_time_step (TimeAdvanceFunc.cuh:50-100). */
int ostc_local_dt(ostc_solver *s, int blk);
/* Local CFL dt on every block and the global minimum, returned to the host.
 * Replaces: local_time_step (TimeAdvanceFunc.cuh:50-100) + global_time_step
 * (TimeAdvanceFunc.cu:155-180). */
int ostc_global_dt(ostc_solver *s, double *dt);
/* One SSP-RK3 stage update with a given dt.  Replaces: update_cv_and_bv_rk (RK.cuh:197-253). */
int ostc_rk3_stage(ostc_solver *s, int blk, double dt, int rk);
/* Boundary conditions in the reference priority order.
 * Replaces: DBoundCond::apply_boundary_conditions (BoundCond.cuh:673-822). */
int ostc_apply_bc(ostc_solver *s, int blk, int with_cv);
/* Inner-face exchange of every block in order: face layers averaged with
 * the neighbour, ghost layers from the neighbour's interior, cv refreshed
 * with with_cv.  Replaces: data_communication / inner_communication
 * (DataCommunication.cuh:27-110). */
int ostc_exchange(ostc_solver *s, int with_cv);
/* Stores rho, |V|, p, T of one block as the reference for the next residual.
 * Replaces: store_last_step (TimeAdvanceFunc.cu:6-17). */
int ostc_store_last(ostc_solver *s, int blk);
/* Residual over all blocks against the last ostc_store_last of each:
 * res[l] = sqrt(sum (q_l - q_l,last)^2 / N) for q = (rho, |V|, p, T), N the
 * points of all blocks, divided
 * by the values of the first call with step == 1 (floor 1e-20);
 * err_max = their maximum.  Synchronises; a NaN residual is an error.
 * Replaces: compute_residual (Residual.cuh:57-161, Residual.cu:49-89). */
int ostc_residual(ostc_solver *s, int step, double res[4], double *err_max);
/* Physical properties on [-1, m] (+ 2-D k-copy first when with_k_copy).
 * Replaces: update_physical_properties (FieldOperation.cuh:112-146). */
int ostc_update_properties(ostc_solver *s, int blk);
/* `nsteps` full SSP-RK3 steps with the dt recomputed on the device every
 * stage and the residual every step, no host synchronisation inside.
 * dt_hist (3*nsteps) and res_hist (5*nsteps: 4 residuals, err_max) may be NULL.
 * The device histories are kept between calls (reallocated only when a call
 * asks for more steps than any before).  After a failed call (nonzero
 * status) the block state is undefined: some stages may have run, and the
 * step counter and physical time are not advanced.
 * Replaces: RK3_bv (RK.cuh:19-195) without IO/monitors. */
int ostc_step_rk3(ostc_solver *s, int nsteps, double *dt_hist, double *res_hist);
/* The implicit treatment of the current dq of one block with the local dt
 * and spectral radii of the last ostc_global_dt: implicit_method 0 applies
 * the point-implicit chemistry (chemSrcMethod 1 EPI, 2 DA) in place;
 * implicit_method 1 runs DPLUR with DPLUR_inner_step sweeps and returns dq/dt.
 * Replaces: implicit_treatment (ImplicitTreatmentHPP.cuh:7-52), DPLUR
 * (DPLUR.cuh:259-314), EPI / DA (FiniteRateChem.cu:234-256, 420-432). */
int ostc_implicit(ostc_solver *s, int blk);
/* `nsteps` steady iterations with local time steps: RHS, local dt, implicit
 * treatment, update_bv, boundary conditions (no cv), properties, residual
 * (res_hist 5*nsteps, NaN where not computed).  As for ostc_step_rk3, the
 * block state is undefined after a failed call.
 * Replaces: steady_simulation (SteadySim.cuh:15-135) without IO, update_bv
 * (FieldOperation.cuh:221-296). */
int ostc_step_steady(ostc_solver *s, int nsteps, double *res_hist);
/* Physical time and step counter. */
int ostc_time(ostc_solver *s, double *t, int *step);
/* 1 = flux arithmetic in FP32 (state, thermo, Newton T and accumulation stay FP64). */
int ostc_set_flux_precision(ostc_solver *s, int fp32);

/* One point-implicit chemistry step dt of n isolated constant-volume cells:
 * rho[n], T[n], Y[n_spec*n] (component-major) in, dY[n_spec*n], dT[n] out.
 * All pointers are DEVICE pointers.  mode 0 = FP64 point-implicit integrator;
 * mode 1 = trained MLP surrogate with per-cell fallback to mode 0
 * (*n_fallback receives the fallback count; 9-species mechanism only).
 * Needs a reacting mechanism.  New entry point: the reference
 * reaches this arithmetic only inside compute_source + EPI
 * (FiniteRateChem.cu:7-63, 234-256). */
int ostc_chem_batch(ostc_solver *s, int64_t n, const double *rho, const double *T, const double *Y, double dt,
                    double *dY, double *dT, int mode, int64_t *n_fallback);
/* Allocation-free variant of ostc_chem_batch, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * arguments, kernels and results as ostc_chem_batch; mode 1 needs a fused
 * inference kernel (the default) and keeps the network outputs, fallback list
 * and fallback counter in one fixed device array per GPU (36 MiB, held until
 * the process ends), so n is limited to 1,048,576 cells (larger n is an
 * error).  Calls on one GPU share that array: calls that overlap on one GPU
 * (several solvers or threads) race.  Allocates nothing. */
int ostc_chem_batch_noalloc(ostc_solver *s, int64_t n, const double *rho, const double *T, const double *Y, double dt,
                            double *dY, double *dT, int mode, int64_t *n_fallback);
/* Output and restart files in the reference's format and places, under
 * <root>/output: field/flowfield.plt (Tecplot binary V112, one ordered zone
 * per block with the interior points; variables x, y, z, density, u, v, w,
 * pressure, temperature, mach, the species by name, tke, omega and mut for
 * SST), message/step.txt and, once a step-1 residual has set it,
 * message/residual_scale.txt.
 * Replaces: FieldIO::write_header / write_common_data_section / print_field
 * (FieldIO.h:58-600) and Residual.cuh:123-139. */
int ostc_write_flowfield(ostc_solver *s, const char *root);
/* Reads those files back into a finalized solver: variables are matched by
 * name (density/rho/roe, u, v, w, p/pressure, t/temperature, species names,
 * k/tke, omega; others ignored), the points the file holds are overwritten
 * (other points keep their values), species or turbulence variables the file
 * lacks take the default inflow state, and the next ostc_initialize keeps the
 * file's step and solution time; the residual scale is restored.
 * Replaces: read_flowfield + identify_variable_labels
 * (Initialize.cuh:64-313), Driver.cu:21-27. */
int ostc_read_flowfield(ostc_solver *s, const char *root);
/* Operator-split reacting steps.  ostc_set_chem_split(s, 1) leaves the
 * finite-rate source out of the right-hand side (the flow step runs
 * non-reacting); ostc_chem_step then advances every interior cell of a block
 * by one point-implicit chemistry step dt at its density and total energy
 * (mode 0 FP64 integrator, mode 1 surrogate + fallback, as ostc_chem_batch),
 * and re-applies the boundary conditions and properties.  No counterpart in
 * the reference, whose RK3 adds the explicit source in every stage
 * (RK.cuh:107-161); this is the split form of "RK3 + point-implicit". */
int ostc_set_chem_split(ostc_solver *s, int on);
/* Inline chemistry on the steady path: with dt > 0, ostc_step_steady runs
 * the flow pseudo-iteration without the chemical source (and without the
 * chemistry Jacobian in the implicit treatment), then advances every
 * interior cell by one chemistry step dt at its density and total energy
 * (mode 0 FP64 point-implicit integrator, mode 1 surrogate + fallback,
 * as ostc_chem_batch) before the boundary conditions.  dt <= 0 switches it
 * off (the reference's point-implicit source in the right-hand side).  No
 * counterpart in the reference, whose steady path keeps the source in the
 * right-hand side (SteadySim.cuh:73-83, ImplicitTreatmentHPP.cuh:7-52).
 * ostc_steady_chem_stats: cells integrated and fallback cells since the last
 * ostc_set_steady_chem. */
int ostc_set_steady_chem(ostc_solver *s, double dt, int mode);
/* The interior cells' rho, T and Y (component-major, n_spec x cells, i
 * fastest) of a block into caller DEVICE arrays of mx*my*mz cells: the
 * chemistry states of a running flow, e.g. to train the surrogate on. */
int ostc_chem_states(ostc_solver *s, int blk, double *rho, double *T, double *Y);
int ostc_steady_chem_stats(ostc_solver *s, int64_t *cells, int64_t *fallback);
int ostc_chem_step(ostc_solver *s, int blk, double dt, int mode, int64_t *n_fallback);

/* ---- MLP chemistry surrogate (no counterpart in the reference) ----
 * A 11 -> 256 -> 256 -> 256 -> 8 GELU network, BF16 MFMA with FP32
 * accumulation, mapping (log(Y+1e-12) of the 9 species, T, log p) of a cell
 * to its point-implicit increment over one dt: the 6 coordinates of dY in an
 * orthonormal basis of the null space of the element/mass conservation rows,
 * dT, and sigma_min(I - dt J) (the step's conditioning, for the gate), each
 * encoded as asinh(c / scale), standardised.  Parameters are FP32, one array; layer l is a row-major
 * [out_pad x ld] matrix with the bias in column kpad:
 *   l=0: 256 x 24 (kpad 16, 11 inputs)   l=1,2: 256 x 264 (kpad 256)
 *   l=3: 16 x 264 (kpad 256, 8 outputs); everything else is zero.
 * Network inputs/outputs of the forward/gradient hooks are standardised. */
typedef struct {
  int64_t epochs, batch;          /* batch: multiple of 8 */
  double lr, lr_min;              /* Adam, cosine decay lr -> lr_min */
  double beta1, beta2, eps;
  uint64_t seed;                  /* shuffling */
} ostc_surrogate_train_cfg;

typedef struct {
  double cons_tol;                /* fallback above this element/mass conservation error (1e-6) */
  double dy_max, dT_max;          /* fallback above these |dY_k|, |dT| [K] predicted */
  double margin;                  /* input range widening, fraction of the training range */
  double sig_min;                 /* fallback below this predicted sigma_min(I - dt J) */
} ostc_surrogate_gate;

/* Random initial parameters (uniform, variance 2/fan_in; 1/fan_in for the output layer). */
int ostc_surrogate_init(ostc_solver *s, uint64_t seed);
int ostc_surrogate_param_count(ostc_solver *s, int64_t *n);
int ostc_surrogate_get_params(ostc_solver *s, float *host);
int ostc_surrogate_set_params(ostc_solver *s, const float *host);
/* Forward pass of the training path (FP32 master weights): x [B x 11], y [B x 8], host arrays. */
int ostc_surrogate_forward(ostc_solver *s, int64_t B, const float *x, float *y);
/* One forward + backward pass of the training kernels on host arrays x [B x 11],
 * t [B x 8] (B multiple of 8): loss = mean (y - t)^2, grad = d loss / d params. */
int ostc_surrogate_gradient(ostc_solver *s, int64_t B, const float *x, const float *t, float *grad, double *loss);
/* Train on n device states (rho, T, Y component-major): targets from the FP64
 * point-implicit integrator with step dt, standardisation fitted to the set,
 * Adam on shuffled mini-batches.  loss_hist[epochs] may be NULL. */
int ostc_surrogate_train(ostc_solver *s, int64_t n, const double *rho, const double *T, const double *Y, double dt,
                         const ostc_surrogate_train_cfg *cfg, double *loss_hist);
int ostc_surrogate_set_gate(ostc_solver *s, const ostc_surrogate_gate *g);
int ostc_surrogate_get_gate(ostc_solver *s, ostc_surrogate_gate *g);
/* out[123]: x mean(11) std(11) min(11) max(11), u mean(8) std(8), dt, basis(9x6 row-major),
 * asinh scales(8); network output j decodes to c_j = scale_j sinh(mean_j + std_j out_j). */
int ostc_surrogate_norm(ostc_solver *s, double *out);
int ostc_surrogate_save(ostc_solver *s, const char *path);
int ostc_surrogate_load(ostc_solver *s, const char *path);
/* Inference kernel of ostc_chem_batch mode 1: 3 = fused network + gate, 128 cells per
 * workgroup, the two halves pipelined (matrix products of one half beside the GELU of
 * the other; default); 1 = the same without the pipeline (identical results);
 * 2 = fused, 64 cells; 0 = one GEMM launch per layer. */
int ostc_surrogate_set_kernel(ostc_solver *s, int kernel);
/* Standardised network outputs out[n x 8] (host) of n device states, without the gate. */
int ostc_surrogate_network_outputs(ostc_solver *s, int64_t n, const double *rho, const double *T, const double *Y,
                                   float *host_out);
/* Last surrogate chem batch: out[4] = fallback cells, network ms (features + MLP),
 * gate ms, FP64 fallback ms (HIP events on the solver stream). */
int ostc_surrogate_last_stats(ostc_solver *s, double *out);
/* Why the last inference's cells fell back: out[0] the fallback count,
 * out[r] the cells whose first failing gate condition was r: 1 input outside
 * the training range, 2 non-finite output, 3 |dY_k| > dy_max, 4 |dT| >
 * dT_max, 5 predicted sigma_min(I - dt J) < sig_min, 6 conservation error >
 * cons_tol, 7 temperature inversion failed, 8 dt other than the training dt
 * (not counted by ostc_chem_batch_noalloc). */
int ostc_surrogate_gate_reasons(ostc_solver *s, int64_t *out);

/* Device memory helpers for callers without their own allocator. */
int ostc_dev_alloc(ostc_solver *s, int64_t bytes, void **dptr);
int ostc_dev_free(ostc_solver *s, void *dptr);
int ostc_memcpy_h2d(ostc_solver *s, void *dst, const void *src, int64_t bytes);
int ostc_memcpy_d2h(ostc_solver *s, void *dst, const void *src, int64_t bytes);
int ostc_memcpy_d2d(ostc_solver *s, void *dst, const void *src, int64_t bytes);
int ostc_sync(ostc_solver *s);

/* Kernel timing with HIP events on the solver stream (0 off, 1 on) and a
 * text summary "name calls total_ms" per kernel family. */
int ostc_profile(ostc_solver *s, int enable);
int ostc_profile_report(ostc_solver *s, char *buf, int len);

#ifdef __cplusplus
}
#endif
#endif
