// Per-point finite-rate chemistry, mixture transport and the point-implicit
// solve, FP64, templated on the species count so every per-lane array has a
// compile-time size and stays in registers.
#pragma once

#include "device_common.h"

namespace ostc {

__host__ __device__ constexpr double kPatmD() { return 101325.0; }

// A T^b exp(-Ea / T) as A exp(b ln T - Ea / T): one exponential per rate,
// with ln T and 1 / T (lt, it) computed once per cell
__device__ inline double d_arrhenius(double lt, double it, double A, double b, double Ea) {
  return A * exp(b * lt - Ea * it);
}

// x^n for the small integer exponents of stoichiometric coefficients and
// reaction orders, as products (x^0 = 1, x^1 = x and x^2 = x*x exactly as a
// correctly rounded pow gives them; x^3 and negative orders within one or
// two roundings); any other exponent goes to pow.  The exponent is uniform
// across a wavefront (it belongs to the reaction), so the switch is a
// scalar branch.
__device__ inline double d_ipow(double x, int n) {
  switch (n) {
    case 0: return 1.0;
    case 1: return x;
    case 2: return x * x;
    case 3: return x * x * x;
    case -1: return 1.0 / x;
    case -2: return 1.0 / (x * x);
    default: return pow(x, (double)n);
  }
}

// FiniteRateChem.cu:65-149 for reaction r: kf (Arrhenius + duplicate, third
// body, Lindemann, Troe) and kb = kf / Kc (or the REV set, 0 if irreversible).
template <int NS>
__device__ inline void d_rate(const DMech *__restrict__ m, int r, double t, double lt, double it, const double *c,
                              const double *g, double temp_t, double &kf, double &kb) {
  const int type = m->label[r];
  double k = d_arrhenius(lt, it, m->A[r], m->b[r], m->Ea[r]);
  if (type == 3) k += d_arrhenius(lt, it, m->A2[r], m->b2[r], m->Ea2[r]);
  if (type > 3) {
    double cc = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) cc += c[l] * m->third_body_coeff[r * NS + l];
    if (type == 4) {
      k *= cc;
    } else {
      const double klow = d_arrhenius(lt, it, m->A2[r], m->b2[r], m->Ea2[r]);
      const double khigh = k;
      const double pr = klow * cc / khigh;
      double F = 1.0;
      if (type > 5) {
        double fc = (1 - m->troe_alpha[r]) * exp(-t / m->troe_t3[r]) + m->troe_alpha[r] * exp(-t / m->troe_t1[r]);
        if (type == 7) fc += exp(-m->troe_t2[r] / t);
        const double lfc = log10(fc);
        const double c2 = -0.4 - 0.67 * lfc;
        const double nn = 0.75 - 1.27 * lfc;
        const double lpr = log10(pr);
        const double tempo = (lpr + c2) / (nn - 0.14 * (lpr + c2));
        const double pp = lfc / (1.0 + tempo * tempo);
        F = exp(2.302585092994046 * pp);  // 10^pp
      }
      k = khigh * pr / (1.0 + pr) * F;
    }
  }
  kf = k;
  if (type == 0) {
    kb = 0;
  } else if (type == 2) {
    kb = d_arrhenius(lt, it, m->A2[r], m->b2[r], m->Ea2[r]);
  } else {
    double dg = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) dg += g[l] * (m->stoi_b[r * NS + l] - m->stoi_f[r * NS + l]);
    const double kc = d_ipow(temp_t, m->order[r]) * exp(-dg);
    kb = k / kc;
  }
}

// FiniteRateChem.cu:7-63 + 151-217: omega [kg/m3/s] and, when jac != null,
// the exact-point-implicit Jacobian d omega_m / d(rho Y_n).  Rates, progress
// and source are fused per reaction so no per-reaction array is kept.
// omega_d (may be null): the destruction rates [kg/m3/s] of chemical_source.
template <int NS>
__device__ inline void d_chem_source(const DMech *__restrict__ m, double rho, double t, const double *Y, double *omega,
                                     double *jac, double *omega_d = nullptr) {
  double c[NS], g[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) c[l] = rho * Y[l] / m->mw[l] * 1e-3;
  d_gibbs_rt<NS>(t, m, g);
  const double temp_t = (kPatmD() / kRuD * 1e-3) / t;
  const double lt = log(t), it = 1.0 / t;
  const int nr = m->n_reac;
  double creation[NS], destr[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    creation[l] = 0;
    destr[l] = 0;
  }
  if (jac) {
#pragma unroll
    for (int a = 0; a < NS * NS; ++a) jac[a] = 0;
  }
  for (int r = 0; r < nr; ++r) {
    double kf, kb;
    d_rate<NS>(m, r, t, lt, it, c, g, temp_t, kf, kb);
    double q1 = 1.0, q2 = 1.0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      q1 *= d_ipow(c[l], m->stoi_f[r * NS + l]);
      q2 *= d_ipow(c[l], m->stoi_b[r * NS + l]);
    }
    q1 *= kf;
    q2 = m->label[r] != 0 ? q2 * kb : 0.0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const int nf = m->stoi_f[r * NS + l], nb = m->stoi_b[r * NS + l];
      creation[l] += q2 * nf + q1 * nb;
      destr[l] += q1 * nf + q2 * nb;
    }
    if (jac) {
#pragma unroll
      for (int n = 0; n < NS; ++n) {
        const double dn = m->stoi_f[r * NS + n] * q1 - m->stoi_b[r * NS + n] * q2;
#pragma unroll
        for (int a = 0; a < NS; ++a) jac[a * NS + n] += (m->stoi_b[r * NS + a] - m->stoi_f[r * NS + a]) * dn;
      }
    }
  }
#pragma unroll
  for (int l = 0; l < NS; ++l) omega[l] = creation[l] * (m->mw[l] * 1e+3) - destr[l] * (m->mw[l] * 1e+3);
  if (omega_d)
#pragma unroll
    for (int l = 0; l < NS; ++l) omega_d[l] = destr[l] * (m->mw[l] * 1e+3);
  if (jac) {
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const bool on = Y[n] > 1e-25;
#pragma unroll
      for (int a = 0; a < NS; ++a) jac[a * NS + n] = on ? m->mw[a] * (jac[a * NS + n] / (rho * Y[n])) * 1e+3 : 0.0;
    }
  }
}

// The same source and Jacobian in two passes over the reactions, for kernels
// that solve with I - dt J: the first computes omega and parks the rates of
// progress q1, q2 of reaction r at q[2r * stride], q[(2r + 1) * stride]
// (LDS, one column per thread), the second accumulates the Jacobian from
// them.  Keeping the 81 Jacobian sums out of the first pass (where the
// concentrations, free energies and creation / destruction sums are live)
// keeps the kernel in registers.  Sums run over r in the same order as in
// d_chem_source, so the results are the same.
template <int NS>
__device__ inline void d_chem_rates(const DMech *__restrict__ m, double rho, double t, const double *Y, double *omega,
                                    double *q, int stride) {
  double c[NS], g[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) c[l] = rho * Y[l] / m->mw[l] * 1e-3;
  d_gibbs_rt<NS>(t, m, g);
  const double temp_t = (kPatmD() / kRuD * 1e-3) / t;
  const double lt = log(t), it = 1.0 / t;
  const int nr = m->n_reac;
  double creation[NS], destr[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    creation[l] = 0;
    destr[l] = 0;
  }
  for (int r = 0; r < nr; ++r) {
    double kf, kb;
    d_rate<NS>(m, r, t, lt, it, c, g, temp_t, kf, kb);
    double q1 = 1.0, q2 = 1.0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      q1 *= d_ipow(c[l], m->stoi_f[r * NS + l]);
      q2 *= d_ipow(c[l], m->stoi_b[r * NS + l]);
    }
    q1 *= kf;
    q2 = m->label[r] != 0 ? q2 * kb : 0.0;
    q[2 * r * stride] = q1;
    q[(2 * r + 1) * stride] = q2;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const int nf = m->stoi_f[r * NS + l], nb = m->stoi_b[r * NS + l];
      creation[l] += q2 * nf + q1 * nb;
      destr[l] += q1 * nf + q2 * nb;
    }
  }
#pragma unroll
  for (int l = 0; l < NS; ++l) omega[l] = creation[l] * (m->mw[l] * 1e+3) - destr[l] * (m->mw[l] * 1e+3);
}

// The EPI Jacobian J from the parked rates of progress; with EPI, A = I - dt J
template <int NS, bool EPI = true>
__device__ inline void d_epi_matrix(const DMech *__restrict__ m, double rho, const double *Y, double dt,
                                    const double *q, int stride, double *A) {
#pragma unroll
  for (int a = 0; a < NS * NS; ++a) A[a] = 0;
  const int nr = m->n_reac;
  for (int r = 0; r < nr; ++r) {
    const double q1 = q[2 * r * stride], q2 = q[(2 * r + 1) * stride];
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      const double dn = m->stoi_f[r * NS + n] * q1 - m->stoi_b[r * NS + n] * q2;
#pragma unroll
      for (int a = 0; a < NS; ++a) A[a * NS + n] += (m->stoi_b[r * NS + a] - m->stoi_f[r * NS + a]) * dn;
    }
  }
#pragma unroll
  for (int n = 0; n < NS; ++n) {
    const bool on = Y[n] > 1e-25;
#pragma unroll
    for (int a = 0; a < NS; ++a) {
      const double J = on ? m->mw[a] * (A[a * NS + n] / (rho * Y[n])) * 1e+3 : 0.0;
      A[a * NS + n] = EPI ? (a == n ? 1.0 : 0.0) - dt * J : J;
    }
  }
}

// FiniteRateChem.cu:364-418 solve_chem_system: Crout-style LU with row
// pivoting (first maximum wins), then substitution; rhs overwritten.
template <int N>
__device__ inline void d_lu_solve(double *A, double *b) {
  int piv[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    // the running maximum is kept in a register (comparing against
    // A[ik][n] would index A dynamically and push it to scratch memory)
    int ik = n;
    double best = 0;
#pragma unroll
    for (int r = n; r < N; ++r) {
#pragma unroll
      for (int t = 0; t < n; ++t) A[r * N + n] -= A[r * N + t] * A[t * N + n];
      const double v = fabs(A[r * N + n]);
      if (r == n) {
        best = v;
      } else if (v > best) {
        best = v;
        ik = r;
      }
    }
    piv[n] = ik;
    // row swap n <-> ik as selects on static indices (a branch around the
    // swap keeps A out of registers)
#pragma unroll
    for (int r = n + 1; r < N; ++r) {
      const bool sw = r == ik;
#pragma unroll
      for (int t = 0; t < N; ++t) {
        const double x = A[r * N + t], y = A[n * N + t];
        A[r * N + t] = sw ? y : x;
        A[n * N + t] = sw ? x : y;
      }
    }
#pragma unroll
    for (int p = n + 1; p < N; ++p)
#pragma unroll
      for (int t = 0; t < n; ++t) A[n * N + p] -= A[n * N + t] * A[t * N + p];
#pragma unroll
    for (int r = n + 1; r < N; ++r) A[r * N + n] /= A[n * N + n];
  }
#pragma unroll
  for (int r = 0; r < N; ++r) {
#pragma unroll
    for (int q = r + 1; q < N; ++q) {
      const bool sw = q == piv[r];
      const double x = b[q], y = b[r];
      b[q] = sw ? y : x;
      b[r] = sw ? x : y;
    }
  }
#pragma unroll
  for (int r = 1; r < N; ++r)
#pragma unroll
    for (int t = 0; t < r; ++t) b[r] -= A[r * N + t] * b[t];
  b[N - 1] /= A[N * N - 1];
#pragma unroll
  for (int r = N - 2; r >= 0; --r) {
#pragma unroll
    for (int t = r + 1; t < N; ++t) b[r] -= A[r * N + t] * b[t];
    b[r] /= A[r * N + r];
  }
}

// Transport.cu:6-38 species viscosities of Chapman-Enskog form,
// vis = vis_coeff sqrt(T) / (1.147 T*^-0.145 + (T* + 0.5)^-2), T* = T / (eps/kb),
// with T*^-0.145 = exp(-0.145 (log T + log(kb/eps))); also sqrt(vis) and
// 1 / sqrt(vis), so that Wilke's sqrt(vis_a / vis_b) is a product
template <int NS>
__device__ inline void species_viscosities(const DMech *__restrict__ m, double t, double *vis, double *sv,
                                           double *isv) {
  const double st = sqrt(t), lt = log(t);
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    const double x5 = t * m->LJ_potent_inv[l] + 0.5;
    const double ci = 1.147 * exp(-0.145 * (lt + m->log_LJ_potent_inv[l])) + 1.0 / (x5 * x5);
    vis[l] = m->vis_coeff[l] * st / ci;
    sv[l] = sqrt(vis[l]);
    isv[l] = 1.0 / sv[l];
  }
}

// Transport.cu:113-140 compute_viscosity: Wilke's rule alone, at the mean
// molecular weight mw_total of composition Y
template <int NS>
__device__ inline double d_viscosity(const DMech *__restrict__ m, double t, double mw_total, const double *Y) {
  double X[NS], vis[NS], sv[NS], isv[NS];
  species_viscosities<NS>(m, t, vis, sv, isv);
#pragma unroll
  for (int l = 0; l < NS; ++l) X[l] = Y[l] * mw_total / m->mw[l];
  double mu = 0;
#pragma unroll
  for (int a = 0; a < NS; ++a) {
    double vt = 0;
#pragma unroll
    for (int b = 0; b < NS; ++b) {
      double pf = 1.0;
      if (a != b) {
        const double num = 1 + sv[a] * isv[b] * m->WjDivWi_to_One4th[a * NS + b];
        pf = num * num * m->sqrt_WiDivWjPl1Mul8[a * NS + b];
      }
      vt += pf * X[b];
    }
    mu += vis[a] * X[a] / vt;
  }
  return mu;
}

// Transport.cu:40-110 compute_transport_property
template <int NS>
__device__ inline void d_transport(const DMech *__restrict__ m, double t, double p, double rho, const double *Y,
                                   const double *cp, double &mu, double &lam, double *rhoD) {
  double mwi = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) mwi += Y[l] / m->mw[l];
  const double mw_total = 1 / mwi;
  double X[NS], vis[NS], sv[NS], isv[NS];
  const double st = sqrt(t);
  species_viscosities<NS>(m, t, vis, sv, isv);
#pragma unroll
  for (int l = 0; l < NS; ++l) X[l] = Y[l] * mw_total / m->mw[l];
  double viscosity = 0, conductivity = 0;
#pragma unroll
  for (int a = 0; a < NS; ++a) {
    double vt = 0;
#pragma unroll
    for (int b = 0; b < NS; ++b) {
      double pf = 1.0;
      if (a != b) {
        const double num = 1 + sv[a] * isv[b] * m->WjDivWi_to_One4th[a * NS + b];
        pf = num * num * m->sqrt_WiDivWjPl1Mul8[a * NS + b];
      }
      vt += pf * X[b];
    }
    const double ct = 1.065 * vt - 0.065 * X[a];
    viscosity += vis[a] * X[a] / vt;
    const double la = vis[a] * (cp[a] + 1.25 * kRuD / m->mw[a]);
    conductivity += la * X[a] / ct;
  }
  mu = viscosity;
  lam = conductivity;
  const double t32p = t * st / p, lt = log(t);
  // reciprocal binary diffusivities 1 / D_ab of the NS (NS - 1) / 2 pairs,
  // each once: T*^-0.145 = exp(-0.145 (log T + log k_ab))
  double idij[NS * (NS - 1) / 2 > 0 ? NS * (NS - 1) / 2 : 1];
#pragma unroll
  for (int a = 0; a < NS; ++a)
#pragma unroll
    for (int b = a + 1; b < NS; ++b) {
      const double tr = t * m->kb_over_eps_jk[a * NS + b];
      const double od = exp(-0.145 * (lt + m->log_kb_over_eps_jk[a * NS + b])) + 1.0 / ((tr + 0.5) * (tr + 0.5));
      idij[a * NS - a * (a + 1) / 2 + (b - a - 1)] = od / (m->binary_diffusivity_coeff[a * NS + b] * t32p);
    }
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    double num = 0, den = 0;
#pragma unroll
    for (int n = 0; n < NS; ++n) {
      if (l != n) {
        const int a = l < n ? l : n, b = l < n ? n : l;
        num += (X[n] + 1e-12) * m->mw[n];
        den += (X[n] + 1e-12) * idij[a * NS - a * (a + 1) / 2 + (b - a - 1)];
      }
    }
    rhoD[l] = rho * num / (den * mw_total);
  }
}

}  // namespace ostc
