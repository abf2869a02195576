// k-omega SST turbulence on gfx950: the wall distance, mu_t initialisation,
// the fused source kernel and the explicit-steady implicit diagonal.
//   SST.cu:8-247 (source, mu_t, implicit diagonal), SourceTerm.cuh:11-37,
//   FieldOperation.cuh:175-195 (initialize_mut), WallDistance.cuh:9-81 and
//   Driver.cu:100-121 (wall distance).
// k and omega are the scalars sv[ns], sv[ns + 1]; rho k, rho omega the
// conservative rows 5 + ns, 6 + ns.
#include "device_chem.h"
#include "device_common.h"
#include "flux_kernels.h"
#include "sst_device.h"

namespace ostc {

namespace {

constexpr int kTB = 256;

// Brute-force distance to the nearest wall point.  One thread per ghosted
// point; the workgroup stages the wall points through LDS in tiles of 256
// (every lane then reads the same LDS address: a broadcast).  Squared
// distances are minimised from 1e6 and the root taken at the end, as the
// reference does.
__global__ void __launch_bounds__(kTB) k_wall_distance(ZoneView z, const double *__restrict__ xyz,
                                                       const double *__restrict__ wall, int nw) {
  __shared__ double sx[kTB], sy[kTB], sz[kTB];
  const int64_t p = (int64_t)blockIdx.x * kTB + threadIdx.x;
  const bool live = p < z.sz;
  const double x = live ? xyz[p] : 0.0, y = live ? xyz[z.sz + p] : 0.0, zc = live ? xyz[2 * z.sz + p] : 0.0;
  double d2 = 1e+6;
  for (int base = 0; base < nw; base += kTB) {
    const int c = base + (int)threadIdx.x;
    if (c < nw) {
      sx[threadIdx.x] = wall[c];
      sy[threadIdx.x] = wall[nw + c];
      sz[threadIdx.x] = wall[2 * nw + c];
    }
    __syncthreads();
    const int n = nw - base < kTB ? nw - base : kTB;
    for (int l = 0; l < n; ++l) {
      const double dx = x - sx[l], dy = y - sy[l], dz = zc - sz[l];
      const double d = dx * dx + dy * dy + dz * dz;
      d2 = d < d2 ? d : d2;
    }
    __syncthreads();
  }
  if (live) z.wdist[p] = sqrt(d2);
}

// central point gradient 0.5 (xi_c dF/di + eta_c dF/dj + zeta_c dF/dk)
__device__ inline void point_grad(const ZoneView &z, const double *f, int64_t p, const double *mm, double *g) {
  const double di = f[p + 1] - f[p - 1], dj = f[p + z.d2] - f[p - z.d2], dk = f[p + z.d1] - f[p - z.d1];
#pragma unroll
  for (int c = 0; c < 3; ++c) g[c] = 0.5 * (mm[c] * di + mm[3 + c] * dj + mm[6 + c] * dk);
}

__device__ inline void load_metric(const ZoneView &z, int64_t p, double *mm) {
#pragma unroll
  for (int c = 0; c < 9; ++c) mm[c] = z.metric[(int64_t)c * z.sz + p];
}

// SST.cu:44-97 compute_mut
__device__ inline double sst_mut(const ZoneView &z, int ns, int64_t p, double mul) {
  double mm[9], gu[3], gv[3], gw[3];
  load_metric(z, p, mm);
  point_grad(z, z.bv + z.sz, p, mm, gu);
  point_grad(z, z.bv + 2 * z.sz, p, mm, gv);
  point_grad(z, z.bv + 3 * z.sz, p, mm, gw);
  const double rho = z.bv[p], tke = z.sv[(int64_t)ns * z.sz + p], omega = z.sv[(int64_t)(ns + 1) * z.sz + p];
  const double vort = sqrt((gv[0] - gu[1]) * (gv[0] - gu[1]) + (gw[0] - gu[2]) * (gw[0] - gu[2]) +
                           (gw[1] - gv[2]) * (gw[1] - gv[2]));
  double f2 = 1;
  const double dy = z.wdist[p];
  if (dy > 1e-25) {
    const double p1 = 2 * sqrt(tke) / (0.09 * omega * dy);
    const double p2 = 500 * mul / (rho * dy * dy * omega);
    const double a2 = fmax(p1, p2);
    f2 = tanh(a2 * a2);
  }
  const double den = fmax(sst::kA1 * omega, vort * f2);
  return den > 1e-25 ? sst::kA1 * rho * tke / den : 0.0;
}

// FieldOperation.cuh:175-195 initialize_mut on [-1, m]
template <int NS>
__global__ void __launch_bounds__(kTB) k_init_mut(ZoneView z, const DMech *__restrict__ m) {
  const int X = z.mx + 2, Y = z.my + 2, Z = z.mz + 2;
  const int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (t >= (int64_t)X * Y * Z) return;
  int i, j, k;
  unflatten3(t, X, Y, i, j, k);
  --i;
  --j;
  --k;
  const int64_t p = z.at(i, j, k);
  const double T = z.bv[5 * z.sz + p];
  double mul = d_sutherland(T);
  if constexpr (NS > 0) {
    double Yk[NS], mwi = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      Yk[l] = z.sv[(int64_t)l * z.sz + p];
      mwi += Yk[l] / m->mw[l];
    }
    mul = d_viscosity<NS>(m, T, 1 / mwi, Yk);
  }
  z.mut[p] = sst_mut(z, NS, p, mul);
}

// SST.cu:24-42
__device__ inline double wilcox_correction(double mt) {
  const double mt0 = 0.25;
  return mt - mt0 > 0 ? sst::kBetaStar * 2 * (mt * mt - mt0 * mt0) : 0.0;
}

__device__ inline double zeman_correction(double mt, double gamma_p1) {
  const double mt0 = 0.25 * sqrt(2.0 / gamma_p1);
  const double dmt = mt - mt0;
  if (dmt <= 0) return 0.0;
  constexpr double lambda2 = 0.66 * 0.66;
  return sst::kBetaStar * 0.75 * (1 - exp(-0.5 * gamma_p1 * dmt * dmt / lambda2));
}

// SST.cu:99-247 compute_source_and_mut + SourceTerm.cuh:19-28, one thread
// per interior cell: F1 / F2, mu_t, lambda_t = mu_t cp / Prt, the k and omega
// sources (production limited: above 10 mu_t |vorticity|^2 it is replaced by
// mu_t |vorticity|^2), the compressibility correction and the source
// Jacobian diagonal (-2 beta* omega, -2 beta omega).  dq is initialised here:
// zero on the flow and species rows, the source on the k and omega rows.
template <int NS>
__global__ void __launch_bounds__(kTB) k_sst(ZoneView z, DParams P) {
  const int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x;
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  if (t >= n) return;
  int i, j, k;
  unflatten3(t, z.mx, z.my, i, j, k);
  const int64_t p = z.at(i, j, k), q = z.at0(i, j, k);
  const double *ks = z.sv + (int64_t)NS * z.sz, *ws = z.sv + (int64_t)(NS + 1) * z.sz;
  double mm[9], gu[3], gv[3], gw[3], gk[3], go[3];
  load_metric(z, p, mm);
  point_grad(z, z.bv + z.sz, p, mm, gu);
  point_grad(z, z.bv + 2 * z.sz, p, mm, gv);
  point_grad(z, z.bv + 3 * z.sz, p, mm, gw);
  point_grad(z, ks, p, mm, gk);
  point_grad(z, ws, p, mm, go);
  const double u_x = gu[0], u_y = gu[1], u_z = gu[2], v_x = gv[0], v_y = gv[1], v_z = gv[2];
  const double w_x = gw[0], w_y = gw[1], w_z = gw[2];
  const double rho = z.bv[p], tke = ks[p], omega = ws[p];
  const double inter = 2 * rho * sst::kSigmaOmega2 / omega * (gk[0] * go[0] + gk[1] * go[1] + gk[2] * go[2]);
  const double rhoK = rho * tke;
  const double vort = sqrt((v_x - u_y) * (v_x - u_y) + (w_x - u_z) * (w_x - u_z) + (w_y - v_z) * (w_y - v_z));
  double f1 = 1, f2 = 1;
  const double dy = z.wdist[p];
  if (dy > 1e-25) {
    const double p1 = sqrt(tke) / (0.09 * omega * dy);
    const double d2 = dy * dy;
    const double p2 = 500 * z.mul[p] / (rho * d2 * omega);
    const double a2 = fmax(2 * p1, p2);
    f2 = tanh(a2 * a2);
    const double cd = fmax(1e-20, inter);
    const double p3 = 4 * rho * sst::kSigmaOmega2 * tke / (cd * d2);
    const double a1 = fmin(fmax(p1, p2), p3);
    f1 = tanh(a1 * a1 * a1 * a1);
  }
  double mut = 0;
  const double den = fmax(sst::kA1 * omega, vort * f2);
  if (den > 1e-25) mut = sst::kA1 * rhoK / den;
  z.mut[p] = mut;
  double beta = sst::kBeta2 + (sst::kBeta1 - sst::kBeta2) * f1;
  double beta_star = sst::kBetaStar;
  if (P.compress) {
    const double gam = NS > 0 ? z.gam[p] : kGam;
    const double mt = sqrt(2 * tke / (gam * z.bv[4 * z.sz + p] / rho));
    double c;
    if (P.compress == 1)
      c = wilcox_correction(mt);
    else if (P.compress == 2)
      c = sst::kBetaStar * mt * mt;
    else
      c = zeman_correction(mt, gam + 1.0);
    beta_star += c;
    beta -= c;
  }
  double sk = 0, sw = 0;
  if (mut > 1e-25) {
    const double divU = u_x + v_y + w_z;
    double prod_k = mut * (2 * (u_x * u_x + v_y * v_y + w_z * w_z) - 2.0 / 3 * divU * divU +
                           (u_y + v_x) * (u_y + v_x) + (u_z + w_x) * (u_z + w_x) + (v_z + w_y) * (v_z + w_y)) -
                    2.0 / 3 * rhoK * divU;
    const double prod_k2 = mut * vort * vort;
    if (prod_k > 10 * prod_k2) prod_k = prod_k2;
    const double J = z.jac[p];
    sk = J * (prod_k - beta_star * rhoK * omega);
    const double gamma = sst::kGamma2 + (sst::kGamma1 - sst::kGamma2) * f1;
    sw = J * (gamma * rho / mut * prod_k + (1 - f1) * inter - beta * rho * omega * omega);
  }
  for (int l = 0; l < 5 + NS; ++l) z.dq[(int64_t)l * z.sz0 + q] = 0.0;
  z.dq[(int64_t)(5 + NS) * z.sz0 + q] = sk;
  z.dq[(int64_t)(6 + NS) * z.sz0 + q] = sw;
  if (P.turb_implicit == 1) {
    z.turb_jac[q] = -2 * beta_star * omega;
    z.turb_jac[z.sz0 + q] = -2 * beta * omega;
  }
  const double cp = NS > 0 ? z.cp[p] : kGam * kRuD / kMwAirD / (kGam - 1);
  z.tcond[p] = mut * cp / P.Prt;
}

// SST.cu:8-22 implicit_treat_for_SST (explicit steady step)
__global__ void __launch_bounds__(kTB) k_turb_implicit(ZoneView z, int ns) {
  const int64_t q = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (q >= (int64_t)z.mx * z.my * z.mz) return;  // ghost-free layout: q is the cell index
  const double dt = z.dt_local[q];
  z.dq[(int64_t)(5 + ns) * z.sz0 + q] /= 1 - dt * z.turb_jac[q];
  z.dq[(int64_t)(6 + ns) * z.sz0 + q] /= 1 - dt * z.turb_jac[z.sz0 + q];
}

inline unsigned nblk(int64_t n, int tb) { return (unsigned)((n + tb - 1) / tb); }

}  // namespace

hipError_t launch_wall_distance(const ZoneView &z, const double *xyz, const double *wall, int nw, hipStream_t s) {
  if (!z.wdist || nw < 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_wall_distance, dim3(nblk(z.sz, kTB)), dim3(kTB), 0, s, z, xyz, wall, nw);
  return hipGetLastError();
}

hipError_t launch_init_mut(const ZoneView &z, const DParams &P, const DMech *m, hipStream_t s) {
  if (!z.mut || !z.wdist || P.nt != 2) return hipErrorInvalidValue;
  const int64_t n = (int64_t)(z.mx + 2) * (z.my + 2) * (z.mz + 2);
  return with_species(P.ns, [&](auto c) {
    hipLaunchKernelGGL(k_init_mut<decltype(c)::value>, dim3(nblk(n, kTB)), dim3(kTB), 0, s, z, m);
    return hipGetLastError();
  });
}

hipError_t launch_sst(const ZoneView &z, const DParams &P, hipStream_t s) {
  if (!z.mut || !z.tcond || !z.wdist || P.nt != 2 || (P.turb_implicit == 1 && !z.turb_jac))
    return hipErrorInvalidValue;
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  return with_species(P.ns, [&](auto c) {
    hipLaunchKernelGGL(k_sst<decltype(c)::value>, dim3(nblk(n, kTB)), dim3(kTB), 0, s, z, P);
    return hipGetLastError();
  });
}

hipError_t launch_turb_implicit(const ZoneView &z, const DParams &P, hipStream_t s) {
  if (!z.turb_jac || P.nt != 2) return hipErrorInvalidValue;
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  hipLaunchKernelGGL(k_turb_implicit, dim3(nblk(n, kTB)), dim3(kTB), 0, s, z, P.ns);
  return hipGetLastError();
}

}  // namespace ostc
