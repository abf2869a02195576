// Steady-path kernels: the implicit treatment of the RHS (point-implicit
// chemistry, DPLUR) and the local-time-step update of the primitives.
//   ImplicitTreatmentHPP.cuh:7-52, FiniteRateChem.cu:234-256 / 420-432,
//   DPLUR.cuh:14-314, DPLUR.cu:5-52, FieldOperation.cuh:221-296.
// One thread per interior cell; dq, dq0, dqk, dt_local, inv_sr and chem_jac
// use the ghost-free component-major layout.
#include <hip/hip_runtime.h>

#include <utility>

#include "device_chem.h"
#include "flux_kernels.h"

namespace ostc {

namespace {

constexpr int kTB = 256;

__device__ inline bool cell_index(const ZoneView &z, int &i, int &j, int &k) {
  const int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x;
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  if (t >= n) return false;
  unflatten3(t, z.mx, z.my, i, j, k);
  return true;
}

// (diag I - dt J) x = b for the species rows, J the stored EPI Jacobian
template <int NS>
__device__ inline void epi_solve(const ZoneView &z, int64_t q, double diag, double dt, double *b) {
  double A[NS * NS];
#pragma unroll
  for (int a = 0; a < NS; ++a)
#pragma unroll
    for (int c = 0; c < NS; ++c)
      A[a * NS + c] = (a == c ? diag : 0.0) - dt * z.chem_jac[(int64_t)(a * NS + c) * z.sz0 + q];
  d_lu_solve<NS>(A, b);
}

// explicit steady step with a point-implicit chemical source:
// EPI (I - dt J) x = dq_species, or DA dq_l /= 1 - dt J_ll
template <int NS>
__global__ void __launch_bounds__(kTB) k_chem_implicit(ZoneView z, int method) {
  int i, j, k;
  if (!cell_index(z, i, j, k)) return;
  const int64_t q = z.at0(i, j, k);
  const double dt = z.dt_local[q];
  if (method == 1) {
    double b[NS];
#pragma unroll
    for (int l = 0; l < NS; ++l) b[l] = z.dq[(int64_t)(5 + l) * z.sz0 + q];
    epi_solve<NS>(z, q, 1.0, dt, b);
#pragma unroll
    for (int l = 0; l < NS; ++l) z.dq[(int64_t)(5 + l) * z.sz0 + q] = b[l];
  } else {
#pragma unroll
    for (int l = 0; l < NS; ++l) z.dq[(int64_t)(5 + l) * z.sz0 + q] /= 1 - dt * z.chem_jac[(int64_t)l * z.sz0 + q];
  }
}

// DPLUR.cuh:14-85 compute_DQ_0 (SST rows: SST.cu:249-258), then dq0 = dq
template <int NS, int NT>
__global__ void __launch_bounds__(kTB) k_dplur_dq0(ZoneView z, int fr, int chem, int timp) {
  constexpr int NV = 5 + NS + NT;
  int i, j, k;
  if (!cell_index(z, i, j, k)) return;
  const int64_t q = z.at0(i, j, k);
  const double dt = z.dt_local[q];
  const double diag = 1 + dt * (z.inv_sr[q] + z.inv_sr[z.sz0 + q] + z.inv_sr[2 * z.sz0 + q]);
  double d[NV];
#pragma unroll
  for (int l = 0; l < NV; ++l) d[l] = z.dq[(int64_t)l * z.sz0 + q] * dt;
#pragma unroll
  for (int l = 0; l < 5; ++l) d[l] /= diag;
  if constexpr (NS > 0) {
    if (fr && chem == 1) {
      epi_solve<NS>(z, q, diag, dt, d + 5);
    } else if (fr && chem == 2) {
#pragma unroll
      for (int l = 0; l < NS; ++l) d[5 + l] /= diag - dt * z.chem_jac[(int64_t)l * z.sz0 + q];
    } else {
#pragma unroll
      for (int l = 0; l < NS; ++l) d[5 + l] /= diag;
    }
  }
#pragma unroll
  for (int l = 0; l < NT; ++l) d[5 + NS + l] /= timp ? diag - dt * z.turb_jac[(int64_t)l * z.sz0 + q] : diag;
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    z.dq[(int64_t)l * z.sz0 + q] = d[l];
    z.dq0[(int64_t)l * z.sz0 + q] = d[l];
  }
}

// DPLUR.cuh:80-137: (A +- lambda I) dq at point (i, j, k) along dir
template <int NS, int NT>
__device__ inline void jac_times_dq(const ZoneView &z, const DMech *__restrict__ m, int i, int j, int k, int dir,
                                    double lam, double *out) {
  constexpr int NV = 5 + NS + NT;
  const int64_t p = z.at(i, j, k), q = z.at0(i, j, k);
  const double mx = z.metric[(dir * 3 + 0) * z.sz + p], my = z.metric[(dir * 3 + 1) * z.sz + p],
               mz = z.metric[(dir * 3 + 2) * z.sz + p];
  const double u = z.bv[z.sz + p], v = z.bv[2 * z.sz + p], w = z.bv[3 * z.sz + p];
  const double U = mx * u + my * v + mz * w;
  const double l1 = U + lam;
  const double e = 0.5 * z.vel[p] * z.vel[p];
  double d[NV];
#pragma unroll
  for (int l = 0; l < NV; ++l) d[l] = z.dq[(int64_t)l * z.sz0 + q];
  double gamma = kGam, b3 = 0, b4 = 0, h = 0;
  if constexpr (NS == 0) {
    h = gamma / (gamma - 1) * z.bv[4 * z.sz + p] / z.bv[p] + e;
  } else {
    // species enthalpies and h = sum Y h + e of this point, fixed during the
    // inner sweeps (k_dplur_enthalpy)
    const double t = z.bv[5 * z.sz + p];
    gamma = z.gam[p];
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      b3 += kRuD * t / m->mw[l] * d[5 + l];
      b4 += z.dplur_h[(int64_t)l * z.sz0 + q] * d[5 + l];
    }
    h = z.dplur_h[(int64_t)NS * z.sz0 + q];
    b3 *= gamma;
    b4 *= gamma - 1;
  }
  const double b1 = mx * d[1] + my * d[2] + mz * d[3] - U * d[0];
  const double b2 = (gamma - 1) * (e * d[0] - u * d[1] - v * d[2] - w * d[3] + d[4]);
  out[0] = b1 + l1 * d[0];
  out[1] = u * b1 + mx * b2 + l1 * d[1] + mx * (b3 - b4);
  out[2] = v * b1 + my * b2 + l1 * d[2] + my * (b3 - b4);
  out[3] = w * b1 + mz * b2 + l1 * d[3] + mz * (b3 - b4);
  out[4] = h * b1 + U * b2 + l1 * d[4] + U * (b3 - b4);
#pragma unroll
  for (int l = 0; l < NS + NT; ++l) out[5 + l] = l1 * d[5 + l] + z.sv[(int64_t)l * z.sz + p] * b1;
}

// the per-point part of DPLUR.cuh:80-137 for mixtures: species enthalpies
// at the point's temperature and h = sum_l Y_l h_l + |u|^2 / 2, summed in the
// order the Jacobian product uses
template <int NS>
__global__ void __launch_bounds__(kTB) k_dplur_enthalpy(ZoneView z, const DMech *__restrict__ m) {
  int i, j, k;
  if (!cell_index(z, i, j, k)) return;
  const int64_t p = z.at(i, j, k), q = z.at0(i, j, k);
  double hs[NS];
  d_enthalpy<NS>(z.bv[5 * z.sz + p], m, hs);
  double h = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    z.dplur_h[(int64_t)l * z.sz0 + q] = hs[l];
    h += z.sv[(int64_t)l * z.sz + p] * hs[l];
  }
  z.dplur_h[(int64_t)NS * z.sz0 + q] = h + 0.5 * z.vel[p] * z.vel[p];
}

// DPLUR.cuh:139-257, the end of one cell's inner sweep: dqk from dq0 and the
// neighbour sum tot (point-implicit species rows, implicit SST source)
template <int NS, int NT>
__device__ inline void dplur_finish(const ZoneView &z, int i, int j, int k, const double *tot, int fr, int chem,
                                    int timp) {
  constexpr int NV = 5 + NS + NT;
  const int64_t q = z.at0(i, j, k);
  const double dt = z.dt_local[q];
  const double diag = 1 + dt * (z.inv_sr[q] + z.inv_sr[z.sz0 + q] + z.inv_sr[2 * z.sz0 + q]);
#pragma unroll
  for (int l = 0; l < 5; ++l) z.dqk[(int64_t)l * z.sz0 + q] = z.dq0[(int64_t)l * z.sz0 + q] + dt * tot[l] / diag;
  if constexpr (NS > 0) {
    if (fr && chem == 1) {
      double b[NS];
#pragma unroll
      for (int l = 0; l < NS; ++l) b[l] = tot[5 + l] * dt;
      epi_solve<NS>(z, q, diag, dt, b);
#pragma unroll
      for (int l = 0; l < NS; ++l) z.dqk[(int64_t)(5 + l) * z.sz0 + q] = z.dq0[(int64_t)(5 + l) * z.sz0 + q] + b[l];
    } else if (fr && chem == 2) {
#pragma unroll
      for (int l = 0; l < NS; ++l)
        z.dqk[(int64_t)(5 + l) * z.sz0 + q] =
            z.dq0[(int64_t)(5 + l) * z.sz0 + q] + dt * tot[5 + l] / (diag - dt * z.chem_jac[(int64_t)l * z.sz0 + q]);
    } else {
#pragma unroll
      for (int l = 0; l < NS; ++l)
        z.dqk[(int64_t)(5 + l) * z.sz0 + q] = z.dq0[(int64_t)(5 + l) * z.sz0 + q] + dt * tot[5 + l] / diag;
    }
  }
  // SST.cu:260-273 implicit_treat_for_dqk
#pragma unroll
  for (int l = 0; l < NT; ++l) {
    const int c = 5 + NS + l;
    const double dg = timp ? diag - dt * z.turb_jac[(int64_t)l * z.sz0 + q] : diag;
    z.dqk[(int64_t)c * z.sz0 + q] = z.dq0[(int64_t)c * z.sz0 + q] + dt * tot[c] / dg;
  }
}

// the i- and j-neighbour term of one side: tot += s 0.5 (A +- lambda I) dq
template <int NS, int NT>
__device__ inline void dplur_add(const ZoneView &z, const DMech *__restrict__ m, int i, int j, int k, int dir,
                                 double sign, double *tot) {
  constexpr int NV = 5 + NS + NT;
  double jd[NV];
  jac_times_dq<NS, NT>(z, m, i, j, k, dir, sign * z.inv_sr[(int64_t)dir * z.sz0 + z.at0(i, j, k)], jd);
  if (sign > 0) {
#pragma unroll
    for (int l = 0; l < NV; ++l) tot[l] += 0.5 * jd[l];
  } else {
#pragma unroll
    for (int l = 0; l < NV; ++l) tot[l] -= 0.5 * jd[l];
  }
}

// DPLUR.cuh:139-257: dqk from the neighbours' current dq (interior only), one
// thread per cell
template <int NS, int NT>
__global__ void __launch_bounds__(kTB) k_dplur_inner(ZoneView z, const DMech *__restrict__ m, int fr, int chem,
                                                     int timp) {
  constexpr int NV = 5 + NS + NT;
  int i, j, k;
  if (!cell_index(z, i, j, k)) return;
  double tot[NV];
#pragma unroll
  for (int l = 0; l < NV; ++l) tot[l] = 0;
  const int c[3] = {i, j, k}, ext[3] = {z.mx, z.my, z.mz};
#pragma unroll
  for (int dd = 0; dd < 3; ++dd)
    if (c[dd] > 0) dplur_add<NS, NT>(z, m, i - (dd == 0), j - (dd == 1), k - (dd == 2), dd, 1.0, tot);
#pragma unroll
  for (int dd = 0; dd < 3; ++dd)
    if (c[dd] != ext[dd] - 1) dplur_add<NS, NT>(z, m, i + (dd == 0), j + (dd == 1), k + (dd == 2), dd, -1.0, tot);
  dplur_finish<NS, NT>(z, i, j, k, tot, fr, chem, timp);
}

// DPLUR.cu:20-52 set_dq_to_0 over one wall range (interior cells of the wall plane)
__global__ void k_zero_dq_range(ZoneView z, int nv, int i0, int i1, int j0, int j1, int k0, int k1) {
  const int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x;
  const int ni = i1 - i0 + 1, nj = j1 - j0 + 1, nk = k1 - k0 + 1;
  if (t >= (int64_t)ni * nj * nk) return;
  int i, j, k;
  unflatten3(t, ni, nj, i, j, k);
  i += i0;
  j += j0;
  k += k0;
  const int64_t q = z.at0(i, j, k);
  for (int l = 0; l < nv; ++l) z.dq[(int64_t)l * z.sz0 + q] = 0;
}

// DPLUR.cu:5-17 convert_dq_back_to_dqDt
// src: the array holding the last inner sweep's result (dq or dqk), so an
// odd sweep count needs no copy back into dq first.
__global__ void k_dq_back(ZoneView z, const double *__restrict__ src, int nv) {
  int i, j, k;
  if (!cell_index(z, i, j, k)) return;
  const int64_t q = z.at0(i, j, k);
  const double dt = z.dt_local[q];
  for (int l = 0; l < nv; ++l) z.dq[(int64_t)l * z.sz0 + q] = src[(int64_t)l * z.sz0 + q] / dt;
}

// FieldOperation.cuh:221-296 update_bv: rho, momentum, total energy and
// scalars advanced by dq dt_local / J from the old primitives; then T and p
// (Newton from the energy for mixtures).  cv is not touched.
template <int NS>
__global__ void __launch_bounds__(kTB) k_update_bv(ZoneView z, DParams P, const DMech *__restrict__ m) {
  int i, j, k;
  if (!cell_index(z, i, j, k)) return;
  const int64_t p = z.at(i, j, k), q = z.at0(i, j, k);
  const int64_t S = z.sz, S0 = z.sz0;
  const double rho_n = z.bv[p];
  const double u0 = z.bv[S + p], v0 = z.bv[2 * S + p], w0 = z.bv[3 * S + p];
  double En = 0.5 * (u0 * u0 + v0 * v0 + w0 * w0);
  double Y[NS > 0 ? NS : 1];
  if constexpr (NS > 0) {
    double h[NS];
    d_enthalpy<NS>(z.bv[5 * S + p], m, h);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      Y[l] = z.sv[(int64_t)l * S + p];
      En += h[l] * Y[l];
    }
    En = En * rho_n - z.bv[4 * S + p];
  } else {
    En = En * rho_n + z.bv[4 * S + p] / (kGam - 1);
  }
  const double dtj = z.dt_local[q] / z.jac[p];
  const double rho = rho_n + z.dq[q] * dtj;
  z.bv[p] = rho;
  const double ri = 1.0 / rho;
  const double u = ri * (rho_n * u0 + z.dq[S0 + q] * dtj);
  const double v = ri * (rho_n * v0 + z.dq[2 * S0 + q] * dtj);
  const double w = z.mz == 1 ? 0.0 : ri * (rho_n * w0 + z.dq[3 * S0 + q] * dtj);
  z.bv[S + p] = u;
  z.bv[2 * S + p] = v;
  z.bv[3 * S + p] = w;
  const double E = En + z.dq[4 * S0 + q] * dtj;
  for (int l = 0; l < P.nsc; ++l) {
    const double y = ri * (rho_n * z.sv[(int64_t)l * S + p] + z.dq[(int64_t)(5 + l) * S0 + q] * dtj);
    z.sv[(int64_t)l * S + p] = y;
    if constexpr (NS > 0)
      if (l < NS) Y[l] = y;
  }
  const double V2 = u * u + v * v + w * w;
  if (P.inviscid_type == 1) {
    // the AWENO path reconstructs conservative variables: keep cv the
    // updated conservative state (the reference's steady update leaves cv
    // at its initial values, FieldOperation.cuh:221-296)
    z.cv[p] = rho;
    z.cv[S + p] = rho * u;
    z.cv[2 * S + p] = rho * v;
    z.cv[3 * S + p] = rho * w;
    z.cv[4 * S + p] = E;
    for (int l = 0; l < P.nsc; ++l) z.cv[(int64_t)(5 + l) * S + p] = rho * z.sv[(int64_t)l * S + p];
  }
  if constexpr (NS > 0) {
    double pn;
    const double t = d_temperature<NS>(m, rho, E, u, v, w, Y, z.bv[5 * S + p], pn);
    z.bv[5 * S + p] = t;
    z.bv[4 * S + p] = pn;
  } else {
    const double pn = (kGam - 1) * (E - 0.5 * rho * V2);
    z.bv[4 * S + p] = pn;
    z.bv[5 * S + p] = pn * kMwAirD * ri / kRuD;
  }
  z.vel[p] = sqrt(V2);
}

inline unsigned nb(int64_t n) { return (unsigned)((n + kTB - 1) / kTB); }

}  // namespace

hipError_t launch_chem_implicit(const ZoneView &z, const DParams &P, hipStream_t s) {
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  if (!(P.mix && P.reaction == 1) || (P.chemSrcMethod != 1 && P.chemSrcMethod != 2)) return hipSuccess;
  return with_mixture(P.ns, [&](auto c) {
    hipLaunchKernelGGL(k_chem_implicit<decltype(c)::value>, dim3(nb(n)), dim3(kTB), 0, s, z, P.chemSrcMethod);
    return hipGetLastError();
  });
}

hipError_t launch_dplur(const ZoneView &z, const DParams &P, const DMech *m, int inner_steps,
                        const std::vector<BCDesc> &walls, hipStream_t s) {
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  const int nv = P.nv;
  const int fr = P.mix && P.reaction == 1;
  if (!z.inv_sr || !z.dq0 || !z.dqk) return hipErrorInvalidValue;
  if (!species_supported(P.ns) || (P.nt != 0 && P.nt != 2)) return hipErrorNotSupported;
  const int timp = P.nt && P.turb_implicit == 1;
  if (timp && !z.turb_jac) return hipErrorInvalidValue;
  const dim3 g(nb(n)), bk(kTB);
  hipError_t e = with_species(P.ns, [&](auto c) {
    constexpr int NS = decltype(c)::value;
    if (P.nt == 0)
      hipLaunchKernelGGL((k_dplur_dq0<NS, 0>), g, bk, 0, s, z, fr, P.chemSrcMethod, 0);
    else
      hipLaunchKernelGGL((k_dplur_dq0<NS, 2>), g, bk, 0, s, z, fr, P.chemSrcMethod, timp);
    if constexpr (NS > 0) {
      if (!z.dplur_h) return hipErrorInvalidValue;
      hipLaunchKernelGGL(k_dplur_enthalpy<NS>, g, bk, 0, s, z, m);
    }
    return hipGetLastError();
  });
  if (e != hipSuccess) return e;
  // The inner sweeps alternate between dq and dqk as input and output (the
  // reference copies dqk back into dq after each sweep); k_dq_back reads the
  // last sweep's output array (dqk for an odd count) and writes dq / dt, so
  // no copy is needed.
  for (int it = 0; it < inner_steps; ++it) {
    ZoneView zi = z;
    if (it & 1) std::swap(zi.dq, zi.dqk);
    for (const BCDesc &b : walls) {
      const int64_t cnt = (int64_t)(b.re[0] - b.rs[0] + 1) * (b.re[1] - b.rs[1] + 1) * (b.re[2] - b.rs[2] + 1);
      if (cnt <= 0) continue;
      if (b.rs[0] < 0 || b.rs[1] < 0 || b.rs[2] < 0 || b.re[0] >= z.mx || b.re[1] >= z.my || b.re[2] >= z.mz)
        return hipErrorInvalidValue;
      hipLaunchKernelGGL(k_zero_dq_range, dim3(nb(cnt)), dim3(kTB), 0, s, zi, nv, b.rs[0], b.re[0], b.rs[1], b.re[1],
                         b.rs[2], b.re[2]);
    }
    e = with_species(P.ns, [&](auto c) {
      constexpr int NS = decltype(c)::value;
      if (P.nt == 0)
        hipLaunchKernelGGL((k_dplur_inner<NS, 0>), g, bk, 0, s, zi, m, fr, P.chemSrcMethod, 0);
      else
        hipLaunchKernelGGL((k_dplur_inner<NS, 2>), g, bk, 0, s, zi, m, fr, P.chemSrcMethod, timp);
      return hipGetLastError();
    });
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_dq_back, dim3(nb(n)), dim3(kTB), 0, s, z, (const double *)((inner_steps & 1) ? z.dqk : z.dq),
                     nv);
  return hipGetLastError();
}

hipError_t launch_update_bv(const ZoneView &z, const DParams &P, const DMech *m, hipStream_t s) {
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  return with_species(P.ns, [&](auto c) {
    hipLaunchKernelGGL(k_update_bv<decltype(c)::value>, dim3(nb(n)), dim3(kTB), 0, s, z, P, m);
    return hipGetLastError();
  });
}

}  // namespace ostc
