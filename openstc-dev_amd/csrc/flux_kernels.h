// Launch wrappers of the device kernels (host-callable, stream-ordered).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <vector>

#include "device_common.h"

namespace ostc {

// Species counts the mixture kernels are compiled for: the 9-species H2-O2
// mechanism and the 4-species H2 / O2 / H2O / N2 mixing mechanism; 0 is
// air.  with_species(ns, f) calls f(std::integral_constant<int, NS>{}) for
// a compiled count and returns hipErrorNotSupported otherwise;
// with_mixture(ns, f) the same without air.
inline bool species_supported(int ns) { return ns == 0 || ns == 4 || ns == 9; }

template <typename F>
hipError_t with_mixture(int ns, F &&f) {
  switch (ns) {
    case 4: return f(std::integral_constant<int, 4>{});
    case 9: return f(std::integral_constant<int, 9>{});
    default: return hipErrorNotSupported;
  }
}

template <typename F>
hipError_t with_species(int ns, F &&f) {
  if (ns == 0) return f(std::integral_constant<int, 0>{});
  return with_mixture(ns, f);
}

struct BCDesc {
  int type, face, dir;
  int rs[3], re[3];
  double rho, u, v, w, p, t, vel;
  double sv[kMS + 4];
  double t_wall;
  int isothermal;
  double a_inf, gamma_inf, entropy, mw_inf;  // farfield state (rho..vel, sv above)
  double p_total, t_total, p_back;           // subsonic inflow (u, v, w: direction), back pressure
  double mut;                                // turbulent viscosity of the inflow / farfield state
};

// per-step scratch kept on the device so a run needs no host round trip
struct StepScratch {
  double *partial;         // per-block partials (>= 4 * max_blocks)
  unsigned int *counter;   // last-block ticket
  double *dt;              // current stage dt
  double *dt_hist;         // 3 per step
  double *res_hist;        // 5 per step (4 residuals + err_max)
  double *res_scale;       // 4
  double *res_one;         // 5: the residual of one ostc_residual call
  double *blk_dt;          // one CFL minimum per block (multi-block cases)
  int max_blocks;
  int hist_cap;            // steps dt_hist / res_hist hold (grows, never shrinks)
};

// an inner face as the exchange kernel walks it (after InnerFace::register_boundary)
struct InnerFaceDesc {
  int rs[3], ts[3];            // start of this block's range and of the target's
  int face, face_at;           // normal direction and the index of the face layer
  int src_tar[3], loop[3], tloop[3], n[3];
};

// one direction (0 = i, 1 = j, 2 = k); init: the sweep writes dq instead of accumulating.
// dt_w (non-null, k-direction): where the sweep can, it also computes the
// stage's CFL time step (dt_local and the minimum into dt_w->dt, through
// dt_w->partial) and sets *dt_done; otherwise *dt_done is left unchanged.
hipError_t launch_flux_sweep(const ZoneView &z, const DParams &P, const DMech *m, int dir, bool init, bool fp32,
                             hipStream_t s, const StepScratch *dt_w = nullptr, bool *dt_done = nullptr);
// the same for the k-omega SST cases (n_turb == 2), flux_sst.hip
hipError_t launch_flux_sweep_sst(const ZoneView &z, const DParams &P, const DMech *m, int dir, bool init, bool fp32,
                                 hipStream_t s);
// finite-rate source; accumulate: add to dq (already initialised by launch_sst)
hipError_t launch_source(const ZoneView &z, const DParams &P, const DMech *m, bool accumulate, hipStream_t s);
hipError_t launch_local_dt(const ZoneView &z, const DParams &P, const StepScratch &w, double *dt_out,
                           hipStream_t s);
hipError_t launch_rk_update(const ZoneView &z, const DParams &P, const DMech *m, const double *dt, int rk,
                            hipStream_t s);
hipError_t launch_bc(const ZoneView &z, const DParams &P, const DMech *m, const BCDesc &bc, bool with_cv,
                     hipStream_t s);
hipError_t launch_eliminate_k(const ZoneView &z, const DParams &P, bool with_cv, hipStream_t s);
// k_planes_copied: a 2-D block whose k-ghost planes k_eliminate_k has just
// made copies of k = 0 (then the k = 0 values are computed once)
hipError_t launch_properties(const ZoneView &z, const DParams &P, const DMech *m, bool k_planes_copied,
                             hipStream_t s);
hipError_t launch_velocity(const ZoneView &z, hipStream_t s);
hipError_t launch_cv_from_bv(const ZoneView &z, const DParams &P, const DMech *m, hipStream_t s);
hipError_t launch_store_last(const ZoneView &z, hipStream_t s);
// residual of the current step into res_hist[slot*5 ..] (or into out when
// given); first != 0 sets the scale
hipError_t launch_residual(const ZoneView &z, const StepScratch &w, int slot, int first, hipStream_t s,
                           double *out = nullptr);
// the same over several blocks (sums of all blocks, divided by all their points)
hipError_t launch_residual_blocks(const std::vector<const ZoneView *> &zs, const StepScratch &w, int slot, int first,
                                  hipStream_t s, double *out = nullptr);
hipError_t launch_copy_dt(const double *src, double *dst, hipStream_t s);
// *out = min(in[0..n)) (n >= 1)
hipError_t launch_min_of(const double *in, int n, double *out, hipStream_t s);
// inner-face exchange of block z with target tz (DataCommunication.cuh:58-110)
hipError_t launch_inner_exchange(const ZoneView &z, const ZoneView &tz, const DParams &P, const DMech *m,
                                 const InnerFaceDesc &f, bool with_cv, hipStream_t s);
// J and metrics of a block from its (g + 1)-ghosted node coordinates, with the
// reference's ghost fall-back; bad: one byte per metric point (scratch);
// *interior_bad (device int, zeroed by the caller) set on a non-positive
// interior Jacobian.  conservative: then the conservative symmetric rows
// (AWENO-5 differences when aweno is set), with fsp_scratch 9 node arrays of
// the (g + 1)-ghosted layout.
hipError_t launch_metrics(int mx, int my, int mz, int g, const double *X, const double *Y, const double *Z,
                          double *jac, double *met, unsigned char *bad, int *interior_bad, hipStream_t s,
                          int conservative = 0, int aweno = 0, double *fsp_scratch = nullptr);
// one point-implicit chemistry step of n independent cells of an ns-species
// mechanism (4 or 9)
hipError_t launch_chem_batch(const DMech *m, int nr, int64_t n, const double *rho, const double *T, const double *Y,
                             double dt, double *dY, double *dT, hipStream_t s, int ns = 9);
// ---- steady path (implicit.hip)
// point-implicit chemistry on dq with the local dt (EPI or DA per chemSrcMethod)
hipError_t launch_chem_implicit(const ZoneView &z, const DParams &P, hipStream_t s);
// DPLUR: dq0, inner sweeps (wall dq zeroed before each), dq back to dq/dt
hipError_t launch_dplur(const ZoneView &z, const DParams &P, const DMech *m, int inner_steps,
                        const std::vector<BCDesc> &walls, hipStream_t s);
// local-dt update of bv / sv from dq (cv untouched)
hipError_t launch_update_bv(const ZoneView &z, const DParams &P, const DMech *m, hipStream_t s);
// operator-split chemistry on a zone's interior: gather (rho, T, Y SoA with
// stride mx*my*mz) and scatter (Y += dY, T += dT at fixed rho and rho E, p
// from the equation of state, rho Y rows of cv)
hipError_t launch_chem_gather(const ZoneView &z, const DParams &P, double *rho, double *T, double *Y, hipStream_t s);
hipError_t launch_chem_scatter(const ZoneView &z, const DParams &P, const DMech *m, const double *dY, const double *dT,
                               hipStream_t s);
// ---- k-omega SST (sst.hip)
// wall distance of every ghosted point: xyz the ghosted node coordinates
// (3 x sz, component-major), wall the nw wall points (3 x nw, component-major)
hipError_t launch_wall_distance(const ZoneView &z, const double *xyz, const double *wall, int nw, hipStream_t s);
// mu_t on [-1, m] from the current state (initialisation)
hipError_t launch_init_mut(const ZoneView &z, const DParams &P, const DMech *m, hipStream_t s);
// mu_t, lambda_t, the source-Jacobian diagonal, and dq = (0, ..., 0, S_k, S_omega) on the interior
hipError_t launch_sst(const ZoneView &z, const DParams &P, hipStream_t s);
// explicit steady step: dq_turb /= 1 - dt J_turb
hipError_t launch_turb_implicit(const ZoneView &z, const DParams &P, hipStream_t s);
// smallest singular value of I - dt J per cell (conditioning of the point-implicit step)
hipError_t launch_epi_sigma(const DMech *m, int64_t n, const double *rho, const double *T, const double *Y, double dt,
                            double *sig, hipStream_t s);
// the same for the cells idx[0..count) of arrays of stride n (idx values < n)
hipError_t launch_chem_batch_subset(const DMech *m, int nr, int64_t n, int64_t count, const int *idx, const double *rho,
                                    const double *T, const double *Y, double dt, double *dY, double *dT,
                                    hipStream_t s);

}  // namespace ostc
