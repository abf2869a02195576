// Device-side data views and per-point physics shared by the gfx950
// kernels.  All state is FP64 in HBM (reference DZone layout); the flux
// arithmetic is templated on its working precision.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ostc {

constexpr int kMS = 20;   // species capacity of the device tables
constexpr int kMR = 50;   // reaction capacity of the device tables

// DParameter.cuh:13-92 content, fixed capacity, row-major packed with the
// live n_spec / n_reac.  Lives in device memory; kernels read it with
// wave-uniform addresses.
struct DMech {
  int n_spec, n_reac;
  double mw[kMS], mw_inv[kMS], t_low[kMS], t_mid[kMS], hi[kMS * 7], lo[kMS * 7];
  double LJ_potent_inv[kMS], vis_coeff[kMS];
  double WjDivWi_to_One4th[kMS * kMS], sqrt_WiDivWjPl1Mul8[kMS * kMS];
  double binary_diffusivity_coeff[kMS * kMS], kb_over_eps_jk[kMS * kMS];
  int label[kMR], stoi_f[kMR * kMS], stoi_b[kMR * kMS], order[kMR];
  double A[kMR], b[kMR], Ea[kMR], A2[kMR], b2[kMR], Ea2[kMR];
  double third_body_coeff[kMR * kMS];
  double troe_alpha[kMR], troe_t3[kMR], troe_t1[kMR], troe_t2[kMR];
  // logarithms of LJ_potent_inv and kb_over_eps_jk: the collision-integral
  // powers T*^-0.145 become exp(-0.145 (log T + log k))
  double log_LJ_potent_inv[kMS], log_kb_over_eps_jk[kMS * kMS];
  // R_u / W_l, rounded as the quotient itself (so h_l, cp_l equal the
  // divided forms bit for bit)
  double r_mw[kMS];
};

// Scheme switches and constants (the scalar part of DParameter).
struct DParams {
  int mx, my, mz, g, dim;
  int nv, ns, nsc, nt;  // n_var, species, scalars (species + turbulence), turbulence scalars (0 or 2)
  int mix, reaction, chemSrcMethod;
  int inviscid_scheme, reconstruction, limiter, viscous, inviscid_type, gradP;
  double cfl, Pr, Sc, Prt, Sct;
  double efix;  // Roe entropy-fix factor
  int turb_implicit, compress;  // SST: implicit source diagonal; compressibility correction 0-3
  int nr;                       // reactions of the mechanism (0 without chemistry)
  int jface;                    // HLLC / LF face vector (J0 m0 + J1 m1) / 2 (conservative metrics)
};

// (a, b, c) of linear index t over an a-fastest na x nb x nc box, in 32-bit
// unsigned arithmetic: block point counts stay below 2^32 (checked when the
// blocks are built), and 64-bit division is a long instruction sequence
__device__ inline void unflatten3(int64_t t, int na, int nb, int &a, int &b, int &c) {
  const uint32_t u = (uint32_t)t, ua = (uint32_t)na, ub = (uint32_t)nb;
  const uint32_t r = u / ua;
  a = (int)(u - r * ua);
  c = (int)(r / ub);
  b = (int)(r - (uint32_t)c * ub);
}

// Raw device pointers of one block (DZone, Field.h:13-71).  bv/sv/cv/qn and
// the property arrays use the ngg-ghosted layout; dq/dt_local/bv_last use
// the ghost-free layout.  metric is 9 component-major arrays
// (index (r*3+c)*sz + p  =  d xi_r / d x_c).
struct ZoneView {
  int mx, my, mz, g;
  int64_t d2, d1, sz;      // ghosted
  int64_t e2, e1, sz0;     // ghost-free
  double *bv, *sv, *cv, *qn;
  double *vel, *acoustic, *mach, *mul, *cond, *rhoD, *gam, *cp;
  const double *jac, *metric;
  double *dq, *dt_local, *bv_last, *chem_jac;
  double *inv_sr;          // steady path: inviscid spectral radii, 3 x ghost-free (null otherwise)
  double *dq0, *dqk;       // DPLUR work arrays, n_var x ghost-free (null otherwise)
  double *dplur_h;         // DPLUR, mixtures: species enthalpies and mixture total enthalpy of each
                           // interior point, (n_spec + 1) x ghost-free (null otherwise)
  double *mut, *tcond;     // SST: turbulent viscosity and conductivity (ghosted; null when laminar)
  double *wdist;           // SST: wall distance (ghosted)
  double *turb_jac;        // SST: source Jacobian diagonal of k and omega, 2 x ghost-free

  __host__ __device__ int64_t at(int i, int j, int k) const {
    return (int64_t)(k + g) * d1 + (int64_t)(j + g) * d2 + (i + g);
  }
  __host__ __device__ int64_t at0(int i, int j, int k) const { return (int64_t)k * e1 + (int64_t)j * e2 + i; }
};

constexpr double kRuD = 8314.462618;
constexpr double kGam = 1.4;
constexpr double kMwAirD = 28.9647;

// ---------------------------------------------------------------- thermo
// Thermo.cu:5-79.  Below t_low: h extrapolated with cp(t_low), cp and g/RT
// frozen at t_low.

// The mechanism pointer made opaque to the optimiser: a property evaluation
// given fresh(m) loads its NASA-7 coefficients where it uses them
// (scalar-cache hits) instead of sharing loads hoisted once and held in
// registers for the rest of the kernel (the flux sweeps evaluate the
// properties at many temperatures per face).
__device__ inline const DMech *fresh(const DMech *m) {
  asm volatile("" : "+s"(m));
  return m;
}

template <int NS>
__device__ inline void d_enthalpy(double t, const DMech *__restrict__ m, double *h) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    if (t < m->t_low[i]) {
      const double tt = m->t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      const double *c = &m->lo[i * 7];
      double hh = c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 / 3 + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c[5];
      const double cp = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
      h[i] = hh + cp * (t - tt);
    } else {
      const double *c = t < m->t_mid[i] ? &m->lo[i * 7] : &m->hi[i * 7];
      h[i] = c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 / 3 + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5];
    }
    h[i] *= m->r_mw[i];
  }
}

template <int NS>
__device__ inline void d_enthalpy_cp(double t, const DMech *__restrict__ m, double *h, double *cp) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    if (t < m->t_low[i]) {
      const double tt = m->t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      const double *c = &m->lo[i * 7];
      const double hh = c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 / 3 + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c[5];
      cp[i] = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
      h[i] = hh + cp[i] * (t - tt);
    } else {
      const double *c = t < m->t_mid[i] ? &m->lo[i * 7] : &m->hi[i * 7];
      h[i] = c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 / 3 + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5];
      cp[i] = c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    h[i] *= m->r_mw[i];
    cp[i] *= m->r_mw[i];
  }
}

template <int NS>
__device__ inline void d_cp(double t, const DMech *__restrict__ m, double *cp) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t;
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    const double tt = t < m->t_low[i] ? m->t_low[i] : t;
    const double *c = (t < m->t_low[i] || t < m->t_mid[i]) ? &m->lo[i * 7] : &m->hi[i * 7];
    if (t < m->t_low[i]) {
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt;
      cp[i] = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
    } else {
      cp[i] = c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    cp[i] *= m->r_mw[i];
  }
}

template <int NS>
__device__ inline void d_gibbs_rt(double t, const DMech *__restrict__ m, double *g) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t_inv = 1 / t, log_t = log(t);
#pragma unroll
  for (int i = 0; i < NS; ++i) {
    if (t < m->t_low[i]) {
      const double tt = m->t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt_inv = 1 / tt, log_tt = log(tt);
      const double *c = &m->lo[i * 7];
      g[i] = c[0] * (1.0 - log_tt) - 0.5 * c[1] * tt - c[2] * tt2 / 6.0 - c[3] * tt3 / 12.0 - c[4] * tt4 * 0.05 +
             c[5] * tt_inv - c[6];
    } else {
      const double *c = t < m->t_mid[i] ? &m->lo[i * 7] : &m->hi[i * 7];
      g[i] = c[0] * (1.0 - log_t) - 0.5 * c[1] * t - c[2] * t2 / 6.0 - c[3] * t3 / 12.0 - c[4] * t4 * 0.05 +
             c[5] * t_inv - c[6];
    }
  }
}

// Transport.cu:6-9
// (T / 273)^1.5 as x sqrt(x): the same power within an ulp, without the
// log / exp of a general pow
__device__ inline double d_sutherland(double t) {
  const double x = t / 273;
  return 1.716e-5 * (x * sqrt(x)) * (273 + 111) / (t + 111);
}

// FieldOperation.cu:3-39: Newton on e(T) from the warm start; tolerance
// |1 - T1/T| <= 1e-3, at most 1000 iterations.
template <int NS>
__device__ inline double d_temperature(const DMech *__restrict__ m, double rho, double rhoE, double u, double v,
                                       double w, const double *Y, double t, double &p) {
  double mw = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) mw += Y[l] / m->mw[l];
  mw = 1 / mw;
  const double gas = kRuD / mw;
  const double e = rhoE / rho - 0.5 * (u * u + v * v + w * w);
  double err = 1;
  int iter = 0;
  while (err > 1e-3 && iter++ < 1000) {
    double h[NS > 0 ? NS : 1], cp[NS > 0 ? NS : 1];
    d_enthalpy_cp<NS>(t, m, h, cp);
    double cpt = 0, ht = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cpt += cp[l] * Y[l];
      ht += h[l] * Y[l];
    }
    const double et = ht - gas * t;
    const double cv = cpt - gas;
    const double t1 = t - (et - e) / cv;
    err = fabs(1 - t1 / t);
    t = t1;
  }
  p = rho * t * gas;
  return t;
}

// TimeAdvanceFunc.cuh:50-100: the CFL time step of the cell at point p
// (interior index q): the inviscid spectral radii of the three directions
// plus the viscous radius.  The inviscid radii are kept per direction when
// the zone stores them (z.inv_sr, for DPLUR).  One definition for the
// stand-alone pass and for the sweep that computes it on the way.
template <int MIX>
__device__ inline double cell_dt(const ZoneView &z, const DParams &P, int64_t p, int64_t q) {
  double mm[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) mm[c] = z.metric[c * z.sz + p];
  const double gxi = sqrt(mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]);
  const double geta = sqrt(mm[3] * mm[3] + mm[4] * mm[4] + mm[5] * mm[5]);
  const double gzeta = sqrt(mm[6] * mm[6] + mm[7] * mm[7] + mm[8] * mm[8]);
  const double u = z.bv[z.sz + p], v = z.bv[2 * z.sz + p], w = z.bv[3 * z.sz + p];
  const double U = u * mm[0] + v * mm[1] + w * mm[2];
  const double V = u * mm[3] + v * mm[4] + w * mm[5];
  const double W = u * mm[6] + v * mm[7] + w * mm[8];
  const double a = z.acoustic[p];
  const double r0 = fabs(U) + a * gxi, r1 = fabs(V) + a * geta;
  const double r2 = P.dim == 3 ? fabs(W) + a * gzeta : 0.0;
  const double sri = r0 + r1 + r2;
  if (z.inv_sr) {
    z.inv_sr[q] = r0;
    z.inv_sr[z.sz0 + q] = r1;
    z.inv_sr[2 * z.sz0 + q] = r2;
  }
  const double gamma = MIX ? z.gam[p] : kGam;
  const double c1 = fmax(gamma, 4.0 / 3.0) / z.bv[p];
  double c2 = z.mul[p] / P.Pr;
  if (z.mut) c2 += z.mut[p] / P.Prt;
  double srv = gxi * gxi + geta * geta;
  if (P.dim == 3) srv += gzeta * gzeta;
  srv *= c1 * c2;
  return P.cfl / (sri + srv);
}

}  // namespace ostc
