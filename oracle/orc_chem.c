/*
 * OpenSTC CPU oracle, part 1: NASA-7 thermo, mixture transport, finite-rate
 * chemistry, the point-implicit solve, the Newton temperature inversion and
 * the chemistry-batch integrator.  TEST INFRASTRUCTURE ONLY (see
 * ostc_oracle.h): nothing in the product links or calls this file.
 *
 * Serial FP64, compiled with -ffp-contract=off so every product/oracle
 * difference is attributable to the product.
 */
#include "ostc_oracle.h"

#include <math.h>
#include <string.h>

#define R_U 8314.462618        /* Constants.h:17 */
#define P_ATM 101325.0         /* Constants.h:22 */

/* Thermo.cu:5-25 compute_enthalpy: NASA-7 h [J/kg]; below t_low the enthalpy is
 * extrapolated linearly with cp(t_low). */
void orc_enthalpy(double t, const orc_mech *m, double *h) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  for (int i = 0; i < m->n_spec; ++i) {
    if (t < m->t_low[i]) {
      const double tt = m->t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      const double *c = &m->lo[i * 7];
      h[i] = c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 / 3 + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c[5];
      const double cp = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
      h[i] += cp * (t - tt);
    } else {
      const double *c = t < m->t_mid[i] ? &m->lo[i * 7] : &m->hi[i * 7];
      h[i] = c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 / 3 + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5];
    }
    h[i] *= R_U / m->mw[i];
  }
}

/* Thermo.cu:27-49 compute_enthalpy_and_cp */
void orc_enthalpy_and_cp(double t, const orc_mech *m, double *h, double *cp) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
  for (int i = 0; i < m->n_spec; ++i) {
    if (t < m->t_low[i]) {
      const double tt = m->t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt, tt5 = tt4 * tt;
      const double *c = &m->lo[i * 7];
      h[i] = c[0] * tt + 0.5 * c[1] * tt2 + c[2] * tt3 / 3 + 0.25 * c[3] * tt4 + 0.2 * c[4] * tt5 + c[5];
      cp[i] = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
      h[i] += cp[i] * (t - tt);
    } else {
      const double *c = t < m->t_mid[i] ? &m->lo[i * 7] : &m->hi[i * 7];
      h[i] = c[0] * t + 0.5 * c[1] * t2 + c[2] * t3 / 3 + 0.25 * c[3] * t4 + 0.2 * c[4] * t5 + c[5];
      cp[i] = c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    h[i] *= R_U / m->mw[i];
    cp[i] *= R_U / m->mw[i];
  }
}

/* Thermo.cu:51-65 compute_cp (frozen at t_low below the range) */
void orc_cp(double t, const orc_mech *m, double *cp) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t;
  for (int i = 0; i < m->n_spec; ++i) {
    if (t < m->t_low[i]) {
      const double tt = m->t_low[i];
      const double tt2 = tt * tt, tt3 = tt2 * tt, tt4 = tt3 * tt;
      const double *c = &m->lo[i * 7];
      cp[i] = c[0] + c[1] * tt + c[2] * tt2 + c[3] * tt3 + c[4] * tt4;
    } else {
      const double *c = t < m->t_mid[i] ? &m->lo[i * 7] : &m->hi[i * 7];
      cp[i] = c[0] + c[1] * t + c[2] * t2 + c[3] * t3 + c[4] * t4;
    }
    cp[i] *= R_U / m->mw[i];
  }
}

/* Thermo.cu:67-79 compute_gibbs_div_rt: g/RT, frozen at t_low below the range */
void orc_gibbs_rt(double t, const orc_mech *m, double *g) {
  const double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t_inv = 1 / t, log_t = log(t);
  for (int i = 0; i < m->n_spec; ++i) {
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

/* Transport.cu:6-9 */
double orc_sutherland(double t) { return 1.716e-5 * pow(t / 273, 1.5) * (273 + 111) / (t + 111); }

/* Transport.cu:143-145 */
static double omega_d(double t_red) { return 1.0 / pow(t_red, 0.145) + 1.0 / ((t_red + 0.5) * (t_red + 0.5)); }

/* Transport.cu:40-110 compute_transport_property: Wilke viscosity, the
 * 1.065/0.065 conductivity mixing rule and mixture-averaged rho*D_k. */
void orc_transport(const orc_mech *m, double t, double p, double rho, const double *Y, const double *cp,
                   double *mu_out, double *lambda_out, double *rhoD) {
  const int ns = m->n_spec;
  double mw_inv = 0;
  for (int l = 0; l < ns; ++l) mw_inv += Y[l] / m->mw[l];
  const double mw_total = 1 / mw_inv;
  double X[ORC_MAXS], vis[ORC_MAXS], d_ij[ORC_MAXS * ORC_MAXS];
  for (int l = 0; l < ns; ++l) {
    X[l] = Y[l] * mw_total / m->mw[l];
    const double t_dl = t * m->LJ_potent_inv[l];
    const double ci = 1.147 * pow(t_dl, -0.145) + pow(t_dl + 0.5, -2);
    vis[l] = m->vis_coeff[l] * sqrt(t) / ci;
  }
  double viscosity = 0, conductivity = 0;
  for (int a = 0; a < ns; ++a) {
    double vis_temp = 0;
    for (int b = 0; b < ns; ++b) {
      double pf = 1.0;
      if (a != b) {
        const double num = 1 + sqrt(vis[a] / vis[b]) * m->WjDivWi_to_One4th[a * ns + b];
        pf = num * num * m->sqrt_WiDivWjPl1Mul8[a * ns + b];
      }
      vis_temp += pf * X[b];
    }
    const double cond_temp = 1.065 * vis_temp - 0.065 * X[a];
    viscosity += vis[a] * X[a] / vis_temp;
    const double lam = vis[a] * (cp[a] + 1.25 * R_U / m->mw[a]);
    conductivity += lam * X[a] / cond_temp;
  }
  *mu_out = viscosity;
  *lambda_out = conductivity;
  const double t32p = t * sqrt(t) / p;
  for (int a = 0; a < ns; ++a)
    for (int b = a + 1; b < ns; ++b) {
      const double t_red = t * m->kb_over_eps_jk[a * ns + b];
      d_ij[a * ns + b] = m->binary_diffusivity_coeff[a * ns + b] * t32p / omega_d(t_red);
      d_ij[b * ns + a] = d_ij[a * ns + b];
    }
  const double eps = 1e-12;
  for (int l = 0; l < ns; ++l) {
    double num = 0, den = 0;
    for (int n = 0; n < ns; ++n) {
      if (l != n) {
        num += (X[n] + eps) * m->mw[n];
        den += (X[n] + eps) / d_ij[l * ns + n];
      }
    }
    rhoD[l] = rho * num / (den * mw_total);
  }
}

static double arrhenius(double t, double A, double b, double Ea) { return A * pow(t, b) * exp(-Ea / t); }

/* FiniteRateChem.cu:65-108 forward_reaction_rate: Arrhenius (+ duplicate),
 * third body, Lindemann and Troe fall-off (log10 form). */
void orc_forward_rate(const orc_mech *m, double t, const double *c, double *kf) {
  const int ns = m->n_spec;
  for (int i = 0; i < m->n_reac; ++i) {
    const int type = m->label[i];
    kf[i] = arrhenius(t, m->A[i], m->b[i], m->Ea[i]);
    if (type == 3) kf[i] += arrhenius(t, m->A2[i], m->b2[i], m->Ea2[i]);
    if (type > 3) {
      double cc = 0;
      for (int l = 0; l < ns; ++l) cc += c[l] * m->third_body_coeff[i * ns + l];
      if (type == 4) {
        kf[i] *= cc;
      } else {
        const double kf_low = arrhenius(t, m->A2[i], m->b2[i], m->Ea2[i]);
        const double kf_high = kf[i];
        const double pr = kf_low * cc / kf_high;
        double F = 1.0;
        if (type > 5) {
          double f_cent = (1 - m->troe_alpha[i]) * exp(-t / m->troe_t3[i]) + m->troe_alpha[i] * exp(-t / m->troe_t1[i]);
          if (type == 7) f_cent += exp(-m->troe_t2[i] / t);
          const double logFc = log10(f_cent);
          const double cc2 = -0.4 - 0.67 * logFc;
          const double nn = 0.75 - 1.27 * logFc;
          const double logPr = log10(pr);
          const double tempo = (logPr + cc2) / (nn - 0.14 * (logPr + cc2));
          const double pp = logFc / (1.0 + tempo * tempo);
          F = pow(10, pp);
        }
        kf[i] = kf_high * pr / (1.0 + pr) * F;
      }
    }
  }
}

/* FiniteRateChem.cu:114-149 backward_reaction_rate: kb = kf/Kc with
 * Kc = (p_atm/(R_u T)*1e-3)^order * exp(-dG/RT); REV gives kb directly. */
void orc_backward_rate(const orc_mech *m, double t, const double *kf, double *kb) {
  const int ns = m->n_spec;
  int n_gibbs = m->n_reac;
  for (int i = 0; i < m->n_reac; ++i) {
    if (m->label[i] == 0) {
      kb[i] = 0;
      --n_gibbs;
    } else if (m->label[i] == 2) {
      kb[i] = arrhenius(t, m->A2[i], m->b2[i], m->Ea2[i]);
      --n_gibbs;
    }
  }
  if (n_gibbs < 1) return;
  double g[ORC_MAXS];
  orc_gibbs_rt(t, m, g);
  const double temp_p = P_ATM / R_U * 1e-3;
  const double temp_t = temp_p / t;
  for (int i = 0; i < m->n_reac; ++i) {
    if (m->label[i] != 2 && m->label[i] != 0) {
      double dg = 0;
      for (int l = 0; l < ns; ++l) dg += g[l] * (m->stoi_b[i * ns + l] - m->stoi_f[i * ns + l]);
      const double kc = pow(temp_t, m->order[i]) * exp(-dg);
      kb[i] = kf[i] / kc;
    }
  }
}

/* FiniteRateChem.cu:151-176 rate_of_progress (pow over all species) */
void orc_rate_of_progress(const orc_mech *m, const double *kf, const double *kb, const double *c, double *q,
                          double *q1, double *q2) {
  const int ns = m->n_spec;
  for (int i = 0; i < m->n_reac; ++i) {
    if (m->label[i] != 0) {
      q1[i] = 1.0;
      q2[i] = 1.0;
      for (int j = 0; j < ns; ++j) {
        q1[i] *= pow(c[j], m->stoi_f[i * ns + j]);
        q2[i] *= pow(c[j], m->stoi_b[i * ns + j]);
      }
      q1[i] *= kf[i];
      q2[i] *= kb[i];
      q[i] = q1[i] - q2[i];
    } else {
      q1[i] = 1.0;
      q2[i] = 0.0;
      for (int j = 0; j < ns; ++j) q1[i] *= pow(c[j], m->stoi_f[i * ns + j]);
      q1[i] *= kf[i];
      q[i] = q1[i];
    }
  }
}

/* FiniteRateChem.cu:178-194 chemical_source [kg/m3/s] */
void orc_chemical_source(const orc_mech *m, const double *q1, const double *q2, double *omega_d, double *omega) {
  const int ns = m->n_spec, nr = m->n_reac;
  for (int i = 0; i < ns; ++i) {
    double creation = 0;
    omega_d[i] = 0;
    for (int j = 0; j < nr; ++j) {
      creation += q2[j] * m->stoi_f[j * ns + i] + q1[j] * m->stoi_b[j * ns + i];
      omega_d[i] += q1[j] * m->stoi_f[j * ns + i] + q2[j] * m->stoi_b[j * ns + i];
    }
    creation *= m->mw[i] * 1e+3;
    omega_d[i] *= m->mw[i] * 1e+3;
    omega[i] = creation - omega_d[i];
  }
}

/* FiniteRateChem.cu:7-63 finite_rate_chemistry (one point), with the EPI
 * Jacobian of FiniteRateChem.cu:196-217 and the diagonal of :219-232. */
void orc_chem_source_point(const orc_mech *m, double rho, double t, const double *Y, double *omega, double *jac,
                           double *jac_diag) {
  const int ns = m->n_spec, nr = m->n_reac;
  double c[ORC_MAXS] = {0}, kf[ORC_MAXR], kb[ORC_MAXR], q[ORC_MAXR], q1[ORC_MAXR], q2[ORC_MAXR], omega_d[ORC_MAXS];
  for (int l = 0; l < ns; ++l) c[l] = rho * Y[l] / m->mw[l] * 1e-3;
  memset(kb, 0, sizeof(kb));
  orc_forward_rate(m, t, c, kf);
  orc_backward_rate(m, t, kf, kb);
  orc_rate_of_progress(m, kf, kb, c, q, q1, q2);
  orc_chemical_source(m, q1, q2, omega_d, omega);
  if (jac) {
    for (int a = 0; a < ns; ++a)
      for (int b = 0; b < ns; ++b) {
        double zz = 0;
        if (Y[b] > 1e-25) {
          for (int r = 0; r < nr; ++r)
            zz += (m->stoi_b[r * ns + a] - m->stoi_f[r * ns + a]) *
                  (m->stoi_f[r * ns + b] * q1[r] - m->stoi_b[r * ns + b] * q2[r]);
          zz /= rho * Y[b];
        }
        jac[a * ns + b] = m->mw[a] * zz * 1e+3;
      }
  }
  if (jac_diag) {
    for (int l = 0; l < ns; ++l) {
      jac_diag[l] = 0;
      if (Y[l] > 1e-25) jac_diag[l] = -omega_d[l] / (Y[l] * rho);
    }
  }
}

/* FiniteRateChem.cu:364-418 solve_chem_system: in-place Crout-style LU with
 * partial (row) pivoting, first maximum wins, then forward/back substitution. */
void orc_lu_solve(double *lhs, double *rhs, int dim) {
  int ipiv[ORC_MAXS];
  for (int n = 0; n < dim; ++n) {
    int ik = n;
    for (int mm = n; mm < dim; ++mm) {
      for (int t = 0; t < n; ++t) lhs[mm * dim + n] -= lhs[mm * dim + t] * lhs[t * dim + n];
      if (fabs(lhs[mm * dim + n]) > fabs(lhs[ik * dim + n])) ik = mm;
    }
    ipiv[n] = ik;
    if (ik != n) {
      for (int t = 0; t < dim; ++t) {
        const double mid = lhs[ik * dim + t];
        lhs[ik * dim + t] = lhs[n * dim + t];
        lhs[n * dim + t] = mid;
      }
    }
    for (int p = n + 1; p < dim; ++p)
      for (int t = 0; t < n; ++t) lhs[n * dim + p] -= lhs[n * dim + t] * lhs[t * dim + p];
    for (int mm = n + 1; mm < dim; ++mm) lhs[mm * dim + n] /= lhs[n * dim + n];
  }
  for (int mm = 0; mm < dim; ++mm) {
    const int t = ipiv[mm];
    if (t != mm) {
      const double mid = rhs[t];
      rhs[t] = rhs[mm];
      rhs[mm] = mid;
    }
  }
  for (int mm = 1; mm < dim; ++mm)
    for (int t = 0; t < mm; ++t) rhs[mm] -= lhs[mm * dim + t] * rhs[t];
  rhs[dim - 1] /= lhs[dim * dim - 1];
  for (int mm = dim - 2; mm >= 0; --mm) {
    for (int t = mm + 1; t < dim; ++t) rhs[mm] -= lhs[mm * dim + t] * rhs[t];
    rhs[mm] /= lhs[mm * dim + mm];
  }
}

/* FieldOperation.cu:3-39 compute_temperature_and_pressure: Newton on
 * e(T) = sum Y h(T) - R T from the warm start t_start, stop when
 * |1 - T1/T| <= 1e-3 or after 1000 iterations; p = rho R T.
 * The iteration count and the loose tolerance make T path dependent; the
 * product must reproduce this exact iteration. */
double orc_temperature_newton(const orc_mech *m, double rho, double rhoE, const double *vel3, const double *Y,
                              double t_start, double *p_out, int *iters) {
  const int ns = m->n_spec;
  double mw = 0;
  for (int l = 0; l < ns; ++l) mw += Y[l] / m->mw[l];
  mw = 1 / mw;
  const double gas_const = R_U / mw;
  const double e = rhoE / rho - 0.5 * (vel3[0] * vel3[0] + vel3[1] * vel3[1] + vel3[2] * vel3[2]);
  double err = 1, t = t_start;
  int iter = 0;
  double h_i[ORC_MAXS], cp_i[ORC_MAXS];
  while (err > 1e-3 && iter++ < 1000) {
    orc_enthalpy_and_cp(t, m, h_i, cp_i);
    double cp_tot = 0, h = 0;
    for (int l = 0; l < ns; ++l) {
      cp_tot += cp_i[l] * Y[l];
      h += h_i[l] * Y[l];
    }
    const double e_t = h - gas_const * t;
    const double cv = cp_tot - gas_const;
    const double t1 = t - (e_t - e) / cv;
    err = fabs(1 - t1 / t);
    t = t1;
  }
  if (p_out) *p_out = rho * t * gas_const;
  if (iters) *iters = iter;
  return t;
}

/* Chemistry-only integration of isolated constant-volume cells.  The
 * reference has no such entry point; this composes its per-cell routines the
 * way its steady path does (SteadySim.cuh:76-90): finite_rate_chemistry at
 * frozen T with the exact-point-implicit Jacobian, the EPI solve
 * (I - dt J) x = omega (FiniteRateChem.cu:234-256), the conservative update
 * rho Y += dt x, then the Newton temperature inversion at the unchanged
 * internal energy (FieldOperation.cuh:221-296 with V = 0). */
void orc_chem_batch(const orc_mech *m, int64_t n, const double *rho_a, const double *T_a, const double *Y_a,
                    double dt, double *dY, double *dT) {
  const int ns = m->n_spec;
  /* cells are independent: an OpenMP build (-fopenmp) splits them over threads */
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const double rho = rho_a[i], t = T_a[i];
    double Y[ORC_MAXS], omega[ORC_MAXS], jac[ORC_MAXS * ORC_MAXS], lhs[ORC_MAXS * ORC_MAXS], h[ORC_MAXS];
    for (int l = 0; l < ns; ++l) Y[l] = Y_a[(int64_t)l * n + i];
    /* total energy of the old state (FieldOperation.cuh:233 compute_total_energy_1_point, V = 0) */
    double mw = 0;
    for (int l = 0; l < ns; ++l) mw += Y[l] / m->mw[l];
    mw = 1 / mw;
    const double p_old = rho * t * (R_U / mw);
    orc_enthalpy(t, m, h);
    double rhoE = 0;
    for (int l = 0; l < ns; ++l) rhoE += h[l] * Y[l];
    rhoE *= rho;
    rhoE -= p_old;
    orc_chem_source_point(m, rho, t, Y, omega, jac, NULL);
    for (int a = 0; a < ns; ++a)
      for (int b = 0; b < ns; ++b) lhs[a * ns + b] = (a == b ? 1.0 : 0.0) - dt * jac[a * ns + b];
    orc_lu_solve(lhs, omega, ns);
    const double rho_inv = 1.0 / rho;
    double Ynew[ORC_MAXS];
    for (int l = 0; l < ns; ++l) Ynew[l] = rho_inv * (rho * Y[l] + omega[l] * dt);
    const double v0[3] = {0, 0, 0};
    const double tn = orc_temperature_newton(m, rho, rhoE, v0, Ynew, t, NULL, NULL);
    for (int l = 0; l < ns; ++l) dY[(int64_t)l * n + i] = Ynew[l] - Y[l];
    dT[i] = tn - t;
  }
}
