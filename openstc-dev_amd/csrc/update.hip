// Point-wise and reduction kernels of the RK3 stage: chemical source,
// CFL time step with an on-device global minimum, the SSP-RK3 update with
// the FP64 Newton temperature inversion, boundary conditions, 2-D k-copies,
// physical properties, the residual and the chemistry-batch integrator.
#include "device_chem.h"
#include "device_common.h"
#include "flux_kernels.h"
#include "sst_device.h"

#include <algorithm>

namespace ostc {

namespace {

constexpr int kTB = 256;

__device__ inline void cell_of(const ZoneView &z, int64_t t, int &i, int &j, int &k) {
  unflatten3(t, z.mx, z.my, i, j, k);
}

__device__ inline double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ------------------------------------------------------------- source
// SourceTerm.cuh:11-37 (finite-rate part): dq = 0 for the flow components,
// dq(5+l) = J omega_l, and the chemistry Jacobian (EPI) or its diagonal (DA).
// With accumulate (the SST kernel already initialised dq) the species rows
// are added to and the flow rows left alone.  The EPI Jacobian is built in a
// second pass over the rates of progress parked in LDS (one column per
// thread, no barrier).  The park is dynamic LDS of 2 n_reac rows of kSrcTB
// doubles, launched only with the EPI Jacobian: its size sets how many
// workgroups share a CU (4 of them for 19 reactions, 2 waves per SIMD).
constexpr int kSrcTB = 128;
constexpr int kParkMaxReac = 24;  // reactions per cell the LDS park holds

template <int NS>
__global__ void __launch_bounds__(kSrcTB) k_source(ZoneView z, DParams P, const DMech *__restrict__ m, int accumulate) {
  extern __shared__ double sq_park[];
  const int64_t t = (int64_t)blockIdx.x * kSrcTB + threadIdx.x;
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  if (t >= n) return;
  int i, j, k;
  cell_of(z, t, i, j, k);
  const int64_t p = z.at(i, j, k), q = z.at0(i, j, k);
  double Y[NS], om[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) Y[l] = z.sv[(int64_t)l * z.sz + p];
  const double rho = z.bv[p], T = z.bv[5 * z.sz + p], J = z.jac[p];
  if (P.chemSrcMethod == 1) {
    double jac[NS * NS];
    d_chem_rates<NS>(m, rho, T, Y, om, sq_park + threadIdx.x, kSrcTB);
    d_epi_matrix<NS, false>(m, rho, Y, 0.0, sq_park + threadIdx.x, kSrcTB, jac);
#pragma unroll
    for (int c = 0; c < NS * NS; ++c) z.chem_jac[(int64_t)c * z.sz0 + q] = jac[c];
  } else if (P.chemSrcMethod == 2) {
    // FiniteRateChem.cu:219-232 diagonal approximation: -omega_d / (rho Y)
    double od[NS];
    d_chem_source<NS>(m, rho, T, Y, om, nullptr, od);
#pragma unroll
    for (int l = 0; l < NS; ++l) z.chem_jac[(int64_t)l * z.sz0 + q] = Y[l] > 1e-25 ? -od[l] / (Y[l] * rho) : 0.0;
  } else {
    d_chem_source<NS>(m, rho, T, Y, om, nullptr);
  }
  if (accumulate) {
#pragma unroll
    for (int l = 0; l < NS; ++l) z.dq[(int64_t)(5 + l) * z.sz0 + q] += J * om[l];
    return;
  }
#pragma unroll
  for (int l = 0; l < 5; ++l) z.dq[(int64_t)l * z.sz0 + q] = 0.0;
#pragma unroll
  for (int l = 0; l < NS; ++l) z.dq[(int64_t)(5 + l) * z.sz0 + q] = J * om[l];
}

// ------------------------------------------------------------- dt
// TimeAdvanceFunc.cuh:50-100 local_time_step, then the global minimum
// (TimeAdvanceFunc.cu:155-201) in two passes on the device: a grid-stride
// pass writes one partial per block, a one-block pass reduces them.  No
// atomics, no host round trip; min is order-independent, hence exact.
constexpr int kRedBlocks = 1024;

template <int MIX>
__global__ void __launch_bounds__(kTB) k_local_dt(ZoneView z, DParams P, double *partial) {
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  double best = 1e+6;
  for (int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x; t < n; t += (int64_t)gridDim.x * kTB) {
    int i, j, k;
    cell_of(z, t, i, j, k);
    const int64_t p = z.at(i, j, k);
    const double dtl = cell_dt<MIX>(z, P, p, z.at0(i, j, k));
    z.dt_local[z.at0(i, j, k)] = dtl;
    best = fmin(best, dtl);
  }
  __shared__ double sm[kTB / 64];
  const double v = wave_min(best);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = sm[0];
#pragma unroll
    for (int w = 1; w < kTB / 64; ++w) r = fmin(r, sm[w]);
    partial[blockIdx.x] = r;
  }
}

__global__ void __launch_bounds__(kTB) k_min_final(const double *partial, int nb, double *out) {
  double r = 1e+6;
  for (int b = threadIdx.x; b < nb; b += kTB) r = fmin(r, partial[b]);
  __shared__ double sm[kTB / 64];
  r = wave_min(r);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = sm[0];
#pragma unroll
    for (int w = 1; w < kTB / 64; ++w) b = fmin(b, sm[w]);
    *out = b;
  }
}

// ------------------------------------------------------------- RK update
// RK.cuh:10-14 (SSP-RK3 a, b, c) and RK.cuh:197-253 update_cv_and_bv_rk.
template <int NS>
__global__ void __launch_bounds__(kTB) k_rk_update(ZoneView z, DParams P, const DMech *__restrict__ m,
                                                    const double *__restrict__ dtp, int rk) {
  const int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x;
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  if (t >= n) return;
  int i, j, k;
  cell_of(z, t, i, j, k);
  const double A = rk == 0 ? 1.0 : (rk == 1 ? 0.75 : 1.0 / 3.0);
  const double B = rk == 0 ? 0.0 : (rk == 1 ? 0.25 : 2.0 / 3.0);
  const double C = rk == 0 ? 1.0 : (rk == 1 ? 0.25 : 2.0 / 3.0);
  const int64_t p = z.at(i, j, k), q = z.at0(i, j, k);
  const double dtj = *dtp / z.jac[p];
  constexpr int NV = 5 + NS;
  double cv[NV];
  // stage 0 keeps the step's initial state in qn on the way (the stages
  // after it read qn only on interior cells): qn and cv hold the same value
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    const int64_t o = (int64_t)l * z.sz + p;
    const double c0 = z.cv[o];
    const double q0 = rk == 0 ? c0 : z.qn[o];
    if (rk == 0) z.qn[o] = c0;
    cv[l] = A * q0 + B * c0 + C * dtj * z.dq[(int64_t)l * z.sz0 + q];
  }
  if (z.mz == 1) cv[3] = 0;
#pragma unroll
  for (int l = 0; l < NV; ++l) z.cv[(int64_t)l * z.sz + p] = cv[l];
  const double ri = 1.0 / cv[0];
  // turbulence scalars (rho k, rho omega) after the species
  for (int l = NV; l < P.nv; ++l) {
    const int64_t o = (int64_t)l * z.sz + p;
    const double c0 = z.cv[o];
    const double q0 = rk == 0 ? c0 : z.qn[o];
    if (rk == 0) z.qn[o] = c0;
    const double c = A * q0 + B * c0 + C * dtj * z.dq[(int64_t)l * z.sz0 + q];
    z.cv[o] = c;
    z.sv[(int64_t)(l - 5) * z.sz + p] = c * ri;
  }
  const double u = cv[1] * ri, v = cv[2] * ri, w = cv[3] * ri;
  z.bv[p] = cv[0];
  z.bv[z.sz + p] = u;
  z.bv[2 * z.sz + p] = v;
  z.bv[3 * z.sz + p] = w;
  const double V2 = u * u + v * v + w * w;
  if constexpr (NS > 0) {
    double Y[NS];
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      Y[l] = cv[5 + l] * ri;
      z.sv[(int64_t)l * z.sz + p] = Y[l];
    }
    double pr;
    const double T = d_temperature<NS>(m, cv[0], cv[4], u, v, w, Y, z.bv[5 * z.sz + p], pr);
    z.bv[5 * z.sz + p] = T;
    z.bv[4 * z.sz + p] = pr;
  } else {
    const double pr = (kGam - 1) * (cv[4] - 0.5 * cv[0] * V2);
    z.bv[4 * z.sz + p] = pr;
    z.bv[5 * z.sz + p] = pr * kMwAirD * ri / kRuD;
  }
  z.vel[p] = sqrt(V2);
}

// FieldOperation.cuh:18-110 compute_total_energy / compute_cv_from_bv_1_point
template <int NS>
__device__ inline void cv_point(const ZoneView &z, const DMech *__restrict__ m, int64_t p, int nt) {
  const double rho = z.bv[p], u = z.bv[z.sz + p], v = z.bv[2 * z.sz + p], w = z.bv[3 * z.sz + p];
  for (int l = NS; l < NS + nt; ++l) z.cv[(int64_t)(5 + l) * z.sz + p] = rho * z.sv[(int64_t)l * z.sz + p];
  z.cv[p] = rho;
  z.cv[z.sz + p] = rho * u;
  z.cv[2 * z.sz + p] = rho * v;
  z.cv[3 * z.sz + p] = rho * w;
  const double V2 = u * u + v * v + w * w;
  double te = 0.5 * V2;
  if constexpr (NS > 0) {
    double h[NS];
    d_enthalpy<NS>(z.bv[5 * z.sz + p], m, h);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const double y = z.sv[(int64_t)l * z.sz + p];
      z.cv[(int64_t)(5 + l) * z.sz + p] = rho * y;
      te += h[l] * y;
    }
    te *= rho;
    te -= z.bv[4 * z.sz + p];
  } else {
    te *= rho;
    te += z.bv[4 * z.sz + p] / (kGam - 1);
  }
  z.cv[4 * z.sz + p] = te;
  z.vel[p] = sqrt(V2);
}

template <int NS>
__global__ void __launch_bounds__(kTB) k_cv_from_bv(ZoneView z, const DMech *__restrict__ m, int nt) {
  const int X = z.mx + 2 * z.g, Y = z.my + 2 * z.g, Z = z.mz + 2 * z.g;
  const int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (t >= (int64_t)X * Y * Z) return;
  cv_point<NS>(z, m, t, nt);
}

__global__ void __launch_bounds__(kTB) k_velocity(ZoneView z) {
  const int X = z.mx + 2 * z.g, Y = z.my + 2 * z.g, Z = z.mz + 2 * z.g;
  const int64_t p = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (p >= (int64_t)X * Y * Z) return;
  const double u = z.bv[z.sz + p], v = z.bv[2 * z.sz + p], w = z.bv[3 * z.sz + p];
  z.vel[p] = sqrt(u * u + v * v + w * w);
}

// ------------------------------------------------------- inner faces
// DataCommunication.cuh:58-110 inner_communication: one thread per point of
// an inner face's range in this block; the face layer is averaged with the
// target block's (both blocks written, |V| of both set), the ghost layers
// take the target's interior values.  Threads touch disjoint points: the
// ghost layers read target points off its face layer.
template <int NS, bool WITH_CV>
__global__ void __launch_bounds__(kTB) k_inner_exchange(ZoneView z, ZoneView tz, const DMech *__restrict__ m,
                                                        InnerFaceDesc f, int nsc, int nt) {
  const int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x;
  const int64_t n01 = (int64_t)f.n[0] * f.n[1];
  if (t >= n01 * f.n[2]) return;
  const int nn[3] = {(int)(t % f.n[0]), (int)((t / f.n[0]) % f.n[1]), (int)(t / n01)};
  int idx[3], it[3], d[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    d[i] = f.loop[i] * nn[i];
    idx[i] = f.rs[i] + d[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) it[i] = f.ts[i] + f.tloop[i] * d[f.src_tar[i]];
  const int64_t p = z.at(idx[0], idx[1], idx[2]), q = tz.at(it[0], it[1], it[2]);
  if (idx[f.face] == f.face_at) {
    for (int l = 0; l < 6; ++l) {
      const double ave = 0.5 * (tz.bv[l * tz.sz + q] + z.bv[l * z.sz + p]);
      z.bv[l * z.sz + p] = ave;
      tz.bv[l * tz.sz + q] = ave;
    }
    for (int l = 0; l < nsc; ++l) {
      const double ave = 0.5 * (tz.sv[l * tz.sz + q] + z.sv[l * z.sz + p]);
      z.sv[l * z.sz + p] = ave;
      tz.sv[l * tz.sz + q] = ave;
    }
    const double u = z.bv[z.sz + p], v = z.bv[2 * z.sz + p], w = z.bv[3 * z.sz + p];
    const double vel = sqrt(u * u + v * v + w * w);
    z.vel[p] = vel;
    tz.vel[q] = vel;
  } else {
    for (int l = 0; l < 6; ++l) z.bv[l * z.sz + p] = tz.bv[l * tz.sz + q];
    for (int l = 0; l < nsc; ++l) z.sv[l * z.sz + p] = tz.sv[l * tz.sz + q];
  }
  if (WITH_CV) cv_point<NS>(z, m, p, nt);
}

// ------------------------------------------------------------- BCs
// BoundCond.cuh:52-552: the face point of a registered range (where the
// reference writes it) and its ngg ghosts.  A ghost layer depends only on
// interior points, the untouched face point or constants, so each of the
// ngg + 1 layers of a point gets its own thread (layer 0: the face point).
// Farfield and subsonic inflow derive their ghosts from the face point's new
// values and keep one thread per point (layer 0 walks all); the SST wall's
// ghost omega is the face value, which each layer's thread recomputes from
// the first interior point exactly as the face thread does.
// Lanes run along the face's fastest index: the layer for i-faces (ghosts
// adjacent in memory), the point for j- and k-faces.
template <int NS, bool WITH_CV>
__global__ void __launch_bounds__(kTB) k_bc(ZoneView z, DParams P, const DMech *__restrict__ m, BCDesc b) {
  const int nx = b.re[0] - b.rs[0] + 1, ny = b.re[1] - b.rs[1] + 1, nz = b.re[2] - b.rs[2] + 1;
  const uint32_t npt = (uint32_t)nx * ny * nz, nl = z.g + 1;
  const uint32_t t = blockIdx.x * kTB + threadIdx.x;
  if (t >= npt * nl) return;
  uint32_t pt, gl;
  if (b.face == 0) {
    pt = t / nl;
    gl = t - pt * nl;
  } else {
    gl = t / npt;
    pt = t - gl * npt;
  }
  const bool par = !(b.type == 4 || b.type == 7);
  if (!par && gl != 0) return;
  const bool face = !par || gl == 0;                          // this thread writes the face point
  const int glo = par ? (int)gl : 1, ghi = par ? (int)gl : z.g;  // ghost layers of this thread
  int i, j, k;
  unflatten3(pt, nx, ny, i, j, k);
  i += b.rs[0];
  j += b.rs[1];
  k += b.rs[2];
  int d[3] = {0, 0, 0};
  d[b.face] = b.dir;
  const int64_t so[3] = {1, z.d2, z.d1};
  const int64_t sd = d[0] * so[0] + d[1] * so[1] + d[2] * so[2];  // outward step
  const int64_t p = z.at(i, j, k);
  double *bv = z.bv, *sv = z.sv;
  const int64_t S = z.sz;
  if (b.type == 3) {
    // symmetry (BoundCond.cuh:52-118); never refreshes ghost cv
    const int64_t pi = p - sd;
    double kx = z.metric[(b.face * 3 + 0) * S + p], ky = z.metric[(b.face * 3 + 1) * S + p],
           kz = z.metric[(b.face * 3 + 2) * S + p];
    const double km = sqrt(kx * kx + ky * ky + kz * kz);
    kx /= km;
    ky /= km;
    kz /= km;
    if (face) {
      const double u1 = bv[S + pi], v1 = bv[2 * S + pi], w1 = bv[3 * S + pi];
      const double uk = kx * u1 + ky * v1 + kz * w1;
      const double ut = u1 - kx * uk, vt = v1 - ky * uk, wt = w1 - kz * uk;
      bv[S + p] = ut;
      bv[2 * S + p] = vt;
      bv[3 * S + p] = wt;
      z.vel[p] = sqrt(ut * ut + vt * vt + wt * wt);
      bv[p] = bv[pi];
      bv[4 * S + p] = bv[4 * S + pi];
      bv[5 * S + p] = bv[5 * S + pi];
      for (int l = 0; l < P.nsc; ++l) sv[l * S + p] = sv[l * S + pi];
    }
    for (int g = max(glo, 1); g <= ghi; ++g) {
      const int64_t pg = p + g * sd, pm = p - g * sd;
      bv[pg] = bv[pm];
      const double u = bv[S + pm], v = bv[2 * S + pm], w = bv[3 * S + pm];
      const double uk = kx * u + ky * v + kz * w;
      const double ug = u - 2 * uk * kx, vg = v - 2 * uk * ky, wg = w - 2 * uk * kz;
      bv[S + pg] = ug;
      bv[2 * S + pg] = vg;
      bv[3 * S + pg] = wg;
      z.vel[pg] = sqrt(ug * ug + vg * vg + wg * wg);
      bv[4 * S + pg] = bv[4 * S + pm];
      bv[5 * S + pg] = bv[5 * S + pm];
      for (int l = 0; l < P.nsc; ++l) sv[l * S + pg] = sv[l * S + pm];
      if (z.mut) z.mut[pg] = z.mut[pm];
    }
  } else if (b.type == 6) {
    // outflow (BoundCond.cuh:120-148)
    for (int g = max(glo, 1); g <= ghi; ++g) {
      const int64_t pg = p + g * sd;
#pragma unroll
      for (int l = 0; l < 6; ++l) bv[l * S + pg] = bv[l * S + p];
      for (int l = 0; l < P.nsc; ++l) sv[l * S + pg] = sv[l * S + p];
      if (z.mut) z.mut[pg] = z.mut[p];
      if (WITH_CV) cv_point<NS>(z, m, pg, P.nt);
    }
  } else if (b.type == 5) {
    // inflow (BoundCond.cuh:150-212)
    for (int g = par ? glo : 0; g <= ghi; ++g) {
      const int64_t pg = p + g * sd;
      bv[pg] = b.rho;
      bv[S + pg] = b.u;
      bv[2 * S + pg] = b.v;
      bv[3 * S + pg] = b.w;
      bv[4 * S + pg] = b.p;
      bv[5 * S + pg] = b.t;
      for (int l = 0; l < P.nsc; ++l) sv[l * S + pg] = b.sv[l];
      if (z.mut) z.mut[pg] = b.mut;
      if (g == 0) z.vel[pg] = b.vel;
      if (WITH_CV) cv_point<NS>(z, m, pg, P.nt);
    }
  } else if (b.type == 4) {
    // farfield (BoundCond.cuh:213-424); never refreshes ghost cv
    double nx = z.metric[(b.face * 3 + 0) * S + p], ny = z.metric[(b.face * 3 + 1) * S + p],
           nz = z.metric[(b.face * 3 + 2) * S + p];
    const double sc = b.dir / sqrt(nx * nx + ny * ny + nz * nz);
    nx *= sc;
    ny *= sc;
    nz *= sc;
    const double ub = bv[S + p], vb = bv[2 * S + p], wb = bv[3 * S + p];
    const double uf = nx * ub + ny * vb + nz * wb;
    double gb = kGam, mw = kMwAirD;
    double svb[NS > 0 ? NS : 1];
    if constexpr (NS > 0) {
      double mi = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        svb[l] = sv[l * S + p];
        mi += svb[l] / m->mw[l];
      }
      gb = z.gam[p];
      mw = 1 / mi;
    }
    const double pb = bv[4 * S + p], rb = bv[p];
    const double ab = sqrt(gb * pb / rb);
    const double mb = uf / ab;
    if (mb >= 1) {
      for (int g = 1; g <= z.g; ++g) {
        const int64_t pg = p + g * sd;
#pragma unroll
        for (int l = 0; l < 6; ++l) bv[l * S + pg] = bv[l * S + p];
        for (int l = 0; l < P.nsc; ++l) sv[l * S + pg] = sv[l * S + p];
        if (z.mut) z.mut[pg] = z.mut[p];
      }
      return;
    }
    double st[6];
    bool far_sv = false;
    if (mb <= -1) {
      st[0] = b.rho;
      st[1] = b.u;
      st[2] = b.v;
      st[3] = b.w;
      st[4] = b.p;
      st[5] = b.t;
      far_sv = true;
    } else {
      const double rp = uf + 2 * ab / (gb - 1);
      const double uinf = b.u * nx + b.v * ny + b.w * nz;
      const double rn = uinf - 2 * b.a_inf / (b.gamma_inf - 1);
      const double Un = 0.5 * (rp + rn);
      double rho, pr, t, u, v, w;
      if constexpr (NS == 0) {
        const double cb = 0.25 * (kGam - 1) * (rp - rn);
        double sb;
        if (mb <= 0) {
          sb = b.entropy;
          u = b.u + (Un - uinf) * nx;
          v = b.v + (Un - uinf) * ny;
          w = b.w + (Un - uinf) * nz;
          far_sv = true;
        } else {
          sb = pb / pow(rb, kGam);
          u = ub + (Un - uf) * nx;
          v = vb + (Un - uf) * ny;
          w = wb + (Un - uf) * nz;
        }
        rho = pow(cb * cb / (kGam * sb), 1 / (kGam - 1));
        pr = rho * cb * cb / kGam;
        t = pr * kMwAirD / (rho * kRuD);
      } else {
        if (mb < 0) {
          u = b.u + (Un - uinf) * nx;
          v = b.v + (Un - uinf) * ny;
          w = b.w + (Un - uinf) * nz;
#pragma unroll
          for (int l = 0; l < NS; ++l) svb[l] = b.sv[l];
          far_sv = true;
          mw = b.mw_inf;
        } else {
          u = ub + (Un - uf) * nx;
          v = vb + (Un - uf) * ny;
          w = wb + (Un - uf) * nz;
        }
        double gamma = kGam, err = 1;
        while (err > 1e-4) {
          const double glast = gamma;
          const double cb = 0.25 * (gamma - 1) * (rp - rn);
          const double sb = mb <= 0 ? b.entropy : pb / pow(rb, gamma);
          rho = pow(cb * cb / (gamma * sb), 1 / (gamma - 1));
          pr = rho * cb * cb / gamma;
          t = pr * mw / (rho * kRuD);
          double cpi[NS], cpt = 0;
          d_cp<NS>(t, m, cpi);
#pragma unroll
          for (int l = 0; l < NS; ++l) cpt += cpi[l] * svb[l];
          gamma = cpt / (cpt - kRuD / mw);
          err = fabs(1 - gamma / glast);
        }
      }
      st[0] = rho;
      st[1] = u;
      st[2] = v;
      st[3] = w;
      st[4] = pr;
      st[5] = t;
    }
    const double mutb = z.mut ? (far_sv ? b.mut : z.mut[p]) : 0.0;
    for (int g = 0; g <= z.g; ++g) {
      const int64_t pg = p + g * sd;
#pragma unroll
      for (int l = 0; l < 6; ++l) bv[l * S + pg] = st[l];
      for (int l = 0; l < P.nsc; ++l) {
        double y = far_sv ? b.sv[l] : sv[l * S + p];
        if constexpr (NS > 0)
          if (!far_sv && l < NS) y = svb[l];
        sv[l * S + pg] = y;
      }
      if (z.mut) z.mut[pg] = mutb;
    }
    z.vel[p] = sqrt(st[1] * st[1] + st[2] * st[2] + st[3] * st[3]);
  } else if (b.type == 7) {
    // subsonic inflow, air (BoundCond.cuh:554-636): ghosts only
    double nx = z.metric[(b.face * 3 + 0) * S + p], ny = z.metric[(b.face * 3 + 1) * S + p],
           nz = z.metric[(b.face * 3 + 2) * S + p];
    const double sc = b.dir / sqrt(nx * nx + ny * ny + nz * nz);
    nx *= sc;
    ny *= sc;
    nz *= sc;
    const double ub = bv[S + p], vb = bv[2 * S + p], wb = bv[3 * S + p];
    const double uf = nx * ub + ny * vb + nz * wb;
    const double g1 = kGam - 1;
    const double a = sqrt(kGam * bv[4 * S + p] / bv[p]);
    const double rn = fabs(uf) - 2 * a / g1;
    const double hti = bv[4 * S + p] / bv[p] * kGam / g1 + 0.5 * (ub * ub + vb * vb + wb * wb);
    const double qa = 1 + 2.0 / g1, qb = 2 * rn, qc = g1 * (0.5 * rn * rn - hti);
    const double dl = qb * qb - 4 * qa * qc;
    double an = a;
    if (dl >= 0) {
      const double ap = (-qb + sqrt(dl)) / (2 * qa), am = (-qb - sqrt(dl)) / (2 * qa);
      an = ap > am ? ap : am;
    }
    const double un = rn + 2 * an / g1;
    const double M = un / an;
    const double pr = b.p_total * pow(1 + 0.5 * g1 * M * M, -kGam / g1);
    const double t = b.t_total * pow(pr / b.p_total, g1 / kGam);
    const double rho = pr * kMwAirD / (t * kRuD);
    for (int g = 1; g <= z.g; ++g) {
      const int64_t pg = p + g * sd;
      bv[pg] = rho;
      bv[S + pg] = un * b.u;
      bv[2 * S + pg] = un * b.v;
      bv[3 * S + pg] = un * b.w;
      bv[4 * S + pg] = pr;
      bv[5 * S + pg] = t;
      const double ubar = bv[S + pg] * nx + bv[2 * S + pg] * ny + bv[3 * S + pg] * nz;
      for (int l = 0; l < P.nsc; ++l) sv[l * S + pg] = ubar > 0 ? sv[l * S + p] : b.sv[l];
      if (z.mut) z.mut[pg] = z.mut[p];
      if (WITH_CV) cv_point<NS>(z, m, pg, P.nt);
    }
  } else if (b.type == 9) {
    // back pressure (BoundCond.cuh:638-670)
    for (int g = max(glo, 1); g <= ghi; ++g) {
      const int64_t pg = p + g * sd;
#pragma unroll
      for (int l = 0; l < 4; ++l) bv[l * S + pg] = bv[l * S + p];
      bv[4 * S + pg] = b.p_back;
      bv[5 * S + pg] = bv[5 * S + p];
      for (int l = 0; l < P.nsc; ++l) sv[l * S + pg] = sv[l * S + p];
      if (z.mut) z.mut[pg] = z.mut[p];
      if (WITH_CV) cv_point<NS>(z, m, pg, P.nt);
    }
  } else if (b.type == 2) {
    // no-slip wall (BoundCond.cuh:424-552)
    const int64_t pi = p - sd;
    const double tw = b.isothermal ? b.t_wall : bv[5 * S + pi];
    const double pw = bv[4 * S + pi];
    double mw = kMwAirD;
    double Yw[NS > 0 ? NS : 1];
    if constexpr (NS > 0) {
      double s = 0;
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        Yw[l] = sv[l * S + pi];
        s += Yw[l] / m->mw[l];
      }
      mw = 1 / s;
    }
    const double rw = pw * mw / (tw * kRuD);
    // SST wall values: k = 0, omega = 60 mu_w / (rho_w beta1 d^2) with the
    // wall distance of the first interior point
    double om_w = 0;
    if (P.nt) {
      double mu_w = d_sutherland(tw);
      if constexpr (NS > 0) mu_w = d_viscosity<NS>(m, tw, mw, Yw);
      const double dy = z.wdist[pi];
      om_w = dy > 1e-25 ? 60 * mu_w / (rw * sst::kBeta1 * dy * dy) : sv[(NS + 1) * S + pi];
    }
    if (face) {
      if constexpr (NS > 0) {
#pragma unroll
        for (int l = 0; l < NS; ++l) sv[l * S + p] = Yw[l];
      }
      bv[p] = rw;
      bv[S + p] = 0;
      bv[2 * S + p] = 0;
      bv[3 * S + p] = 0;
      bv[4 * S + p] = pw;
      bv[5 * S + p] = tw;
      z.vel[p] = 0;
      if (P.nt) {
        sv[NS * S + p] = 0;
        sv[(NS + 1) * S + p] = om_w;
      }
      if (WITH_CV) cv_point<NS>(z, m, p, P.nt);
    }
    for (int g = max(glo, 1); g <= ghi; ++g) {
      const int64_t pa = p - g * sd, pg = p + g * sd;
      const double ui = bv[S + pa], vi = bv[2 * S + pa], wi = bv[3 * S + pa], pii = bv[4 * S + pa],
                   ti = bv[5 * S + pa];
      double tg = ti;
      if (b.isothermal) {
        tg = 2 * tw - ti;
        if (tg <= 0.1 * tw) tg = tw;
      }
      if constexpr (NS > 0) {
        double s = 0;
#pragma unroll
        for (int l = 0; l < NS; ++l) {
          const double y = sv[l * S + pa];
          sv[l * S + pg] = y;
          s += y / m->mw[l];
        }
        mw = 1 / s;
      }
      bv[pg] = pii * mw / (tg * kRuD);
      bv[S + pg] = -ui;
      bv[2 * S + pg] = -vi;
      bv[3 * S + pg] = -wi;
      bv[4 * S + pg] = pii;
      bv[5 * S + pg] = tg;
      if (P.nt) {
        sv[NS * S + pg] = 0;
        sv[(NS + 1) * S + pg] = om_w;
        z.mut[pg] = 0;
      }
      if (WITH_CV) cv_point<NS>(z, m, pg, P.nt);
    }
  }
}

// FieldOperation.cuh:372-400 eliminate_k_gradient (2-D runs)
template <bool WITH_CV>
__global__ void __launch_bounds__(kTB) k_eliminate_k(ZoneView z, DParams P) {
  const int X = z.mx + 2 * z.g, Y = z.my + 2 * z.g;
  const int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (t >= (int64_t)X * Y) return;
  const uint32_t jj = (uint32_t)t / (uint32_t)X;
  const int i = (int)((uint32_t)t - jj * (uint32_t)X) - z.g, j = (int)jj - z.g;
  const int64_t p0 = z.at(i, j, 0);
  for (int k = 1; k <= z.g; ++k) {
    const int64_t pu = p0 + k * z.d1, pd = p0 - k * z.d1;
#pragma unroll
    for (int l = 0; l < 6; ++l) {
      const double v = z.bv[l * z.sz + p0];
      z.bv[l * z.sz + pu] = v;
      z.bv[l * z.sz + pd] = v;
    }
    for (int l = 0; l < P.nsc; ++l) {
      const double v = z.sv[l * z.sz + p0];
      z.sv[l * z.sz + pu] = v;
      z.sv[l * z.sz + pd] = v;
    }
    if (WITH_CV)
      for (int l = 0; l < P.nv; ++l) {
        const double v = z.cv[l * z.sz + p0];
        z.cv[l * z.sz + pu] = v;
        z.cv[l * z.sz + pd] = v;
      }
  }
}

// FieldOperation.cuh:112-146 update_physical_properties on [-1, m].  PLANE:
// a 2-D block, whose k-ghost planes the preceding k_eliminate_k made copies
// of k = 0 (bv and every scalar): the k = 0 values are computed once and
// stored on k = -1, 0, 1 (bit-identical to computing each plane), the Mach
// number from each plane's own |u|.
template <int NS, bool PLANE>
__global__ void __launch_bounds__(kTB) k_properties(ZoneView z, DParams P, const DMech *__restrict__ m) {
  const int X = z.mx + 2, Y = z.my + 2, Z = PLANE ? 1 : z.mz + 2;
  const int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (t >= (int64_t)X * Y * Z) return;
  int i, j, k;
  unflatten3(t, X, Y, i, j, k);
  --i;
  --j;
  k = PLANE ? 0 : k - 1;
  const int64_t p = z.at(i, j, k);
  const double T = z.bv[5 * z.sz + p];
  double a, mu, lam, cpt = 0, gam = 0, rd[NS > 0 ? NS : 1];
  if constexpr (NS > 0) {
    double Yk[NS], cp[NS];
    d_cp<NS>(T, m, cp);
    double mwi = 0, cvt = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      Yk[l] = z.sv[(int64_t)l * z.sz + p];
      mwi += Yk[l] / m->mw[l];
      cpt += Yk[l] * cp[l];
      cvt += Yk[l] * (cp[l] - m->r_mw[l]);
    }
    const double mw = 1 / mwi;
    gam = cpt / cvt;
    a = sqrt(gam * kRuD * T / mw);
    d_transport<NS>(m, T, z.bv[4 * z.sz + p], z.bv[p], Yk, cp, mu, lam, rd);
  } else {
    const double c_temp = kGam * kRuD / kMwAirD;
    const double cp = kGam * kRuD / kMwAirD / (kGam - 1);
    a = sqrt(c_temp * T);
    mu = d_sutherland(T);
    lam = mu * cp / P.Pr;
  }
  for (int kk = PLANE ? -1 : 0; kk <= (PLANE ? 1 : 0); ++kk) {
    const int64_t q = p + kk * z.d1;
    if constexpr (NS > 0) {
      z.cp[q] = cpt;
      z.gam[q] = gam;
#pragma unroll
      for (int l = 0; l < NS; ++l) z.rhoD[(int64_t)l * z.sz + q] = rd[l];
    }
    z.acoustic[q] = a;
    z.mul[q] = mu;
    z.cond[q] = lam;
    z.mach[q] = z.vel[q] / a;
  }
}

// TimeAdvanceFunc.cu:6-17 store_last_step
__global__ void __launch_bounds__(kTB) k_store_last(ZoneView z) {
  const int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x;
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  if (t >= n) return;
  int i, j, k;
  cell_of(z, t, i, j, k);
  const int64_t p = z.at(i, j, k), q = z.at0(i, j, k);
  z.bv_last[q] = z.bv[p];
  z.bv_last[z.sz0 + q] = z.vel[p];
  z.bv_last[2 * z.sz0 + q] = z.bv[4 * z.sz + p];
  z.bv_last[3 * z.sz0 + q] = z.bv[5 * z.sz + p];
}

// Residual.cuh:57-161 + TimeAdvanceFunc.cu:19-33: squared change of
// (rho, |V|, p, T) summed per block (grid-stride), then one block finishes
// the RMS, the step-1 scale (floor 1e-20) and err_max on the device.
__global__ void __launch_bounds__(kTB) k_residual(ZoneView z, double *partial, int ld, int off) {
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  double s[4] = {0, 0, 0, 0};
  for (int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x; t < n; t += (int64_t)gridDim.x * kTB) {
    int i, j, k;
    cell_of(z, t, i, j, k);
    const int64_t p = z.at(i, j, k), q = z.at0(i, j, k);
    const double cur[4] = {z.bv[p], z.vel[p], z.bv[4 * z.sz + p], z.bv[5 * z.sz + p]};
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const double d = cur[l] - z.bv_last[l * z.sz0 + q];
      z.bv_last[l * z.sz0 + q] = d * d;
      s[l] += d * d;
    }
  }
  __shared__ double sm[4][kTB / 64];
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const double v = wave_sum(s[l]);
    if ((threadIdx.x & 63) == 0) sm[l][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      double b = 0;
#pragma unroll
      for (int w = 0; w < kTB / 64; ++w) b += sm[l][w];
      partial[(size_t)l * ld + off + blockIdx.x] = b;
    }
  }
}

__global__ void __launch_bounds__(kTB) k_residual_final(const double *partial, int nb, int64_t n, double *res_scale,
                                                         double *out, int first) {
  double r[4] = {0, 0, 0, 0};
  for (int b = threadIdx.x; b < nb; b += kTB)
#pragma unroll
    for (int l = 0; l < 4; ++l) r[l] += partial[(size_t)l * nb + b];
  __shared__ double sm[4][kTB / 64];
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const double v = wave_sum(r[l]);
    if ((threadIdx.x & 63) == 0) sm[l][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double res[4];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      double b = 0;
#pragma unroll
      for (int w = 0; w < kTB / 64; ++w) b += sm[l][w];
      res[l] = sqrt(b / (double)n);
      if (first) res_scale[l] = res[l] < 1e-20 ? 1e-20 : res[l];
      res[l] /= res_scale[l];
      out[l] = res[l];
    }
    double e = res[0];
#pragma unroll
    for (int l = 1; l < 4; ++l)
      if (fabs(res[l]) > e) e = res[l];
    out[4] = e;
  }
}

__global__ void k_copy1(const double *src, double *dst) { *dst = *src; }

// ------------------------------------------------------------- chemistry batch
// One point-implicit step of isolated constant-volume cells: omega and the
// exact Jacobian at frozen T, (I - dt J) x = omega, rho Y += dt x, then the
// Newton temperature at the unchanged internal energy (the composition of
// FiniteRateChem.cu:7-63/234-256 and FieldOperation.cuh:221-296 at V = 0).
// With idx, thread t handles cell idx[t] for t < count; otherwise cell t < n.
// The rates of progress are parked in LDS between the source pass and the
// Jacobian pass (d_chem_rates / d_epi_matrix), one column per thread: no
// thread reads another's column, so the kernel has no barrier.  The park is
// dynamic LDS of 2 n_reac rows of kChemTB doubles.
constexpr int kChemTB = 128;

template <int NS>
__global__ void __launch_bounds__(kChemTB) k_chem_batch(const DMech *__restrict__ m, int64_t n, int64_t count,
                                                        const int *__restrict__ idx, const double *__restrict__ rho_a,
                                                        const double *__restrict__ T_a, const double *__restrict__ Y_a,
                                                        double dt, double *__restrict__ dY, double *__restrict__ dT) {
  extern __shared__ double sq_park[];
  const int64_t t = (int64_t)blockIdx.x * kChemTB + threadIdx.x;
  if (t >= count) return;
  const int64_t c = idx ? (int64_t)idx[t] : t;
  const double rho = rho_a[c], T = T_a[c];
  double Y[NS], om[NS], J[NS * NS], h[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) Y[l] = Y_a[(int64_t)l * n + c];
  double mwi = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) mwi += Y[l] / m->mw[l];
  const double mw = 1 / mwi;
  const double p_old = rho * T * (kRuD / mw);
  d_enthalpy<NS>(T, m, h);
  double rhoE = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) rhoE += h[l] * Y[l];
  rhoE *= rho;
  rhoE -= p_old;
  d_chem_rates<NS>(m, rho, T, Y, om, sq_park + threadIdx.x, kChemTB);
  d_epi_matrix<NS>(m, rho, Y, dt, sq_park + threadIdx.x, kChemTB, J);
  d_lu_solve<NS>(J, om);
  const double ri = 1.0 / rho;
  double Yn[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) Yn[l] = ri * (rho * Y[l] + om[l] * dt);
  double pn;
  const double Tn = d_temperature<NS>(m, rho, rhoE, 0.0, 0.0, 0.0, Yn, T, pn);
#pragma unroll
  for (int l = 0; l < NS; ++l) dY[(int64_t)l * n + c] = Yn[l] - Y[l];
  dT[c] = Tn - T;
}

// Smallest singular value of the point-implicit matrix I - dt J of each cell
// (J the exact-point-implicit Jacobian of the chemical source at the cell's
// state), by one-sided Jacobi rotations of its columns in FP64: the
// conditioning of the point-implicit step, used as a training target.
template <int NS>
__global__ void __launch_bounds__(128) k_epi_sigma(const DMech *__restrict__ m, int64_t n,
                                                   const double *__restrict__ rho_a, const double *__restrict__ T_a,
                                                   const double *__restrict__ Y_a, double dt,
                                                   double *__restrict__ sig) {
  const int64_t c = (int64_t)blockIdx.x * 128 + threadIdx.x;
  if (c >= n) return;
  double Y[NS], om[NS], A[NS * NS];
#pragma unroll
  for (int l = 0; l < NS; ++l) Y[l] = Y_a[(int64_t)l * n + c];
  d_chem_source<NS>(m, rho_a[c], T_a[c], Y, om, A);
#pragma unroll
  for (int a = 0; a < NS; ++a)
#pragma unroll
    for (int b = 0; b < NS; ++b) A[a * NS + b] = (a == b ? 1.0 : 0.0) - dt * A[a * NS + b];
  for (int sweep = 0; sweep < 12; ++sweep) {
    double off = 0;
#pragma unroll
    for (int p = 0; p < NS - 1; ++p)
#pragma unroll
      for (int q = p + 1; q < NS; ++q) {
        double al = 0, be = 0, ga = 0;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          al += A[k * NS + p] * A[k * NS + p];
          be += A[k * NS + q] * A[k * NS + q];
          ga += A[k * NS + p] * A[k * NS + q];
        }
        const double rel = fabs(ga) / sqrt(al * be + 1e-300);
        off = fmax(off, rel);
        if (rel > 1e-15) {
          const double zeta = (be - al) / (2 * ga);
          const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
          const double cs = 1 / sqrt(1 + t * t), sn = cs * t;
#pragma unroll
          for (int k = 0; k < NS; ++k) {
            const double x = A[k * NS + p], y = A[k * NS + q];
            A[k * NS + p] = cs * x - sn * y;
            A[k * NS + q] = sn * x + cs * y;
          }
        }
      }
    if (off < 1e-14) break;
  }
  double mn = 1e300;
#pragma unroll
  for (int p = 0; p < NS; ++p) {
    double s2 = 0;
#pragma unroll
    for (int k = 0; k < NS; ++k) s2 += A[k * NS + p] * A[k * NS + p];
    mn = fmin(mn, s2);
  }
  sig[c] = sqrt(mn);
}

// Operator-split chemistry on a zone: the interior state is gathered into
// batch arrays (rho, T, Y SoA with stride n, cells i fastest), integrated by
// the batch kernels (or the surrogate), and scattered back: Y += dY,
// T += dT at the unchanged density and total energy, p from the equation of
// state, the rho Y rows of cv.
template <int NS>
__global__ void __launch_bounds__(kTB) k_chem_gather(ZoneView z, double *__restrict__ rho, double *__restrict__ T,
                                                     double *__restrict__ Y) {
  const int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x;
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  if (t >= n) return;
  int i, j, k;
  cell_of(z, t, i, j, k);
  const int64_t p = z.at(i, j, k);
  rho[t] = z.bv[p];
  T[t] = z.bv[5 * z.sz + p];
#pragma unroll
  for (int l = 0; l < NS; ++l) Y[(int64_t)l * n + t] = z.sv[(int64_t)l * z.sz + p];
}

template <int NS>
__global__ void __launch_bounds__(kTB) k_chem_scatter(ZoneView z, const DMech *__restrict__ m,
                                                      const double *__restrict__ dY, const double *__restrict__ dT) {
  const int64_t t = (int64_t)blockIdx.x * kTB + threadIdx.x;
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  if (t >= n) return;
  int i, j, k;
  cell_of(z, t, i, j, k);
  const int64_t p = z.at(i, j, k);
  const double rho = z.bv[p];
  double mwi = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) {
    const double y = z.sv[(int64_t)l * z.sz + p] + dY[(int64_t)l * n + t];
    z.sv[(int64_t)l * z.sz + p] = y;
    z.cv[(int64_t)(5 + l) * z.sz + p] = rho * y;
    mwi += y / m->mw[l];
  }
  const double T = z.bv[5 * z.sz + p] + dT[t];
  z.bv[5 * z.sz + p] = T;
  z.bv[4 * z.sz + p] = rho * T * kRuD * mwi;
}

inline unsigned nblk(int64_t n, int tb) { return (unsigned)((n + tb - 1) / tb); }

}  // namespace

hipError_t launch_chem_gather(const ZoneView &z, const DParams &P, double *rho, double *T, double *Y, hipStream_t s) {
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  return with_mixture(P.ns, [&](auto c) {
    hipLaunchKernelGGL(k_chem_gather<decltype(c)::value>, dim3(nblk(n, kTB)), dim3(kTB), 0, s, z, rho, T, Y);
    return hipGetLastError();
  });
}

hipError_t launch_chem_scatter(const ZoneView &z, const DParams &P, const DMech *m, const double *dY, const double *dT,
                               hipStream_t s) {
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  return with_mixture(P.ns, [&](auto c) {
    hipLaunchKernelGGL(k_chem_scatter<decltype(c)::value>, dim3(nblk(n, kTB)), dim3(kTB), 0, s, z, m, dY, dT);
    return hipGetLastError();
  });
}

hipError_t launch_source(const ZoneView &z, const DParams &P, const DMech *m, bool accumulate, hipStream_t s) {
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  if (P.chemSrcMethod == 1 && (P.nr <= 0 || P.nr > kParkMaxReac)) return hipErrorInvalidValue;
  const size_t park = P.chemSrcMethod == 1 ? (size_t)2 * P.nr * kSrcTB * sizeof(double) : 0;
  return with_mixture(P.ns, [&](auto c) {
    hipLaunchKernelGGL(k_source<decltype(c)::value>, dim3(nblk(n, kSrcTB)), dim3(kSrcTB), park, s, z, P, m,
                       accumulate ? 1 : 0);
    return hipGetLastError();
  });
}

hipError_t launch_local_dt(const ZoneView &z, const DParams &P, const StepScratch &w, double *dt_out, hipStream_t s) {
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  const int nb = (int)std::min<int64_t>(kRedBlocks, nblk(n, kTB));
  if (nb > w.max_blocks) return hipErrorInvalidValue;
  if (P.mix)
    hipLaunchKernelGGL(k_local_dt<1>, dim3(nb), dim3(kTB), 0, s, z, P, w.partial);
  else
    hipLaunchKernelGGL(k_local_dt<0>, dim3(nb), dim3(kTB), 0, s, z, P, w.partial);
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(kTB), 0, s, (const double *)w.partial, nb, dt_out);
  return hipGetLastError();
}

hipError_t launch_rk_update(const ZoneView &z, const DParams &P, const DMech *m, const double *dt, int rk,
                            hipStream_t s) {
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  return with_species(P.ns, [&](auto c) {
    hipLaunchKernelGGL(k_rk_update<decltype(c)::value>, dim3(nblk(n, kTB)), dim3(kTB), 0, s, z, P, m, dt, rk);
    return hipGetLastError();
  });
}

hipError_t launch_bc(const ZoneView &z, const DParams &P, const DMech *m, const BCDesc &bc, bool with_cv,
                     hipStream_t s) {
  const int64_t n = (int64_t)(bc.re[0] - bc.rs[0] + 1) * (bc.re[1] - bc.rs[1] + 1) * (bc.re[2] - bc.rs[2] + 1);
  if (n <= 0) return hipSuccess;
  if (n * (z.g + 1) > ((int64_t)1 << 32) - kTB) return hipErrorInvalidValue;  // 32-bit thread index
  const dim3 g(nblk(n * (z.g + 1), kTB)), b(kTB);
  return with_species(P.ns, [&](auto c) {
    constexpr int NS = decltype(c)::value;
    if (with_cv)
      hipLaunchKernelGGL((k_bc<NS, true>), g, b, 0, s, z, P, m, bc);
    else
      hipLaunchKernelGGL((k_bc<NS, false>), g, b, 0, s, z, P, m, bc);
    return hipGetLastError();
  });
}

hipError_t launch_eliminate_k(const ZoneView &z, const DParams &P, bool with_cv, hipStream_t s) {
  const int64_t n = (int64_t)(z.mx + 2 * z.g) * (z.my + 2 * z.g);
  if (with_cv)
    hipLaunchKernelGGL(k_eliminate_k<true>, dim3(nblk(n, kTB)), dim3(kTB), 0, s, z, P);
  else
    hipLaunchKernelGGL(k_eliminate_k<false>, dim3(nblk(n, kTB)), dim3(kTB), 0, s, z, P);
  return hipGetLastError();
}

hipError_t launch_properties(const ZoneView &z, const DParams &P, const DMech *m, bool k_planes_copied,
                             hipStream_t s) {
  const bool plane = k_planes_copied && P.dim == 2 && z.mz == 1;
  const int64_t n = (int64_t)(z.mx + 2) * (z.my + 2) * (plane ? 1 : z.mz + 2);
  const dim3 g(nblk(n, kTB)), b(kTB);
  return with_species(P.ns, [&](auto c) {
    constexpr int NS = decltype(c)::value;
    if (plane)
      hipLaunchKernelGGL((k_properties<NS, true>), g, b, 0, s, z, P, m);
    else
      hipLaunchKernelGGL((k_properties<NS, false>), g, b, 0, s, z, P, m);
    return hipGetLastError();
  });
}

hipError_t launch_velocity(const ZoneView &z, hipStream_t s) {
  const int64_t n = (int64_t)(z.mx + 2 * z.g) * (z.my + 2 * z.g) * (z.mz + 2 * z.g);
  hipLaunchKernelGGL(k_velocity, dim3(nblk(n, kTB)), dim3(kTB), 0, s, z);
  return hipGetLastError();
}

hipError_t launch_cv_from_bv(const ZoneView &z, const DParams &P, const DMech *m, hipStream_t s) {
  const int64_t n = (int64_t)(z.mx + 2 * z.g) * (z.my + 2 * z.g) * (z.mz + 2 * z.g);
  return with_species(P.ns, [&](auto c) {
    hipLaunchKernelGGL(k_cv_from_bv<decltype(c)::value>, dim3(nblk(n, kTB)), dim3(kTB), 0, s, z, m, P.nt);
    return hipGetLastError();
  });
}

hipError_t launch_store_last(const ZoneView &z, hipStream_t s) {
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  hipLaunchKernelGGL(k_store_last, dim3(nblk(n, kTB)), dim3(kTB), 0, s, z);
  return hipGetLastError();
}

hipError_t launch_residual(const ZoneView &z, const StepScratch &w, int slot, int first, hipStream_t s,
                           double *out) {
  const int64_t n = (int64_t)z.mx * z.my * z.mz;
  const int nb = (int)std::min<int64_t>(kRedBlocks, nblk(n, kTB));
  if (nb > w.max_blocks) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_residual, dim3(nb), dim3(kTB), 0, s, z, w.partial, nb, 0);
  hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(kTB), 0, s, (const double *)w.partial, nb, n, w.res_scale,
                     out ? out : w.res_hist + (size_t)slot * 5, first);
  return hipGetLastError();
}

hipError_t launch_residual_blocks(const std::vector<const ZoneView *> &zs, const StepScratch &w, int slot, int first,
                                  hipStream_t s, double *out) {
  if (zs.size() == 1) return launch_residual(*zs[0], w, slot, first, s, out);
  int ld = 0;
  int64_t n = 0;
  for (const ZoneView *z : zs) {
    const int64_t c = (int64_t)z->mx * z->my * z->mz;
    ld += (int)std::min<int64_t>(kRedBlocks, nblk(c, kTB));
    n += c;
  }
  if (ld > w.max_blocks) return hipErrorInvalidValue;
  int off = 0;
  for (const ZoneView *z : zs) {
    const int nb = (int)std::min<int64_t>(kRedBlocks, nblk((int64_t)z->mx * z->my * z->mz, kTB));
    hipLaunchKernelGGL(k_residual, dim3(nb), dim3(kTB), 0, s, *z, w.partial, ld, off);
    off += nb;
  }
  hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(kTB), 0, s, (const double *)w.partial, ld, n, w.res_scale,
                     out ? out : w.res_hist + (size_t)slot * 5, first);
  return hipGetLastError();
}

hipError_t launch_inner_exchange(const ZoneView &z, const ZoneView &tz, const DParams &P, const DMech *m,
                                 const InnerFaceDesc &f, bool with_cv, hipStream_t s) {
  const int64_t n = (int64_t)f.n[0] * f.n[1] * f.n[2];
  if (n <= 0) return hipSuccess;
  if (n > ((int64_t)1 << 31)) return hipErrorInvalidValue;
  const dim3 g(nblk(n, kTB)), b(kTB);
  return with_species(P.ns, [&](auto c) {
    constexpr int NS = decltype(c)::value;
    if (with_cv)
      hipLaunchKernelGGL((k_inner_exchange<NS, true>), g, b, 0, s, z, tz, m, f, P.nsc, P.nt);
    else
      hipLaunchKernelGGL((k_inner_exchange<NS, false>), g, b, 0, s, z, tz, m, f, P.nsc, P.nt);
    return hipGetLastError();
  });
}

hipError_t launch_min_of(const double *in, int n, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_min_final, dim3(1), dim3(kTB), 0, s, in, n, out);
  return hipGetLastError();
}

hipError_t launch_copy_dt(const double *src, double *dst, hipStream_t s) {
  hipLaunchKernelGGL(k_copy1, dim3(1), dim3(1), 0, s, src, dst);
  return hipGetLastError();
}

hipError_t launch_chem_batch(const DMech *m, int nr, int64_t n, const double *rho, const double *T, const double *Y,
                             double dt, double *dY, double *dT, hipStream_t s, int ns) {
  if (n <= 0) return hipSuccess;
  if (nr <= 0 || nr > kParkMaxReac) return hipErrorInvalidValue;
  const size_t park = (size_t)2 * nr * kChemTB * sizeof(double);
  return with_mixture(ns, [&](auto c) {
    hipLaunchKernelGGL(k_chem_batch<decltype(c)::value>, dim3(nblk(n, kChemTB)), dim3(kChemTB), park, s, m, n, n,
                       (const int *)nullptr, rho, T, Y, dt, dY, dT);
    return hipGetLastError();
  });
}

hipError_t launch_epi_sigma(const DMech *m, int64_t n, const double *rho, const double *T, const double *Y, double dt,
                            double *sig, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_epi_sigma<9>, dim3(nblk(n, 128)), dim3(128), 0, s, m, n, rho, T, Y, dt, sig);
  return hipGetLastError();
}

hipError_t launch_chem_batch_subset(const DMech *m, int nr, int64_t n, int64_t count, const int *idx,
                                    const double *rho, const double *T, const double *Y, double dt, double *dY,
                                    double *dT, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  if (count > n || nr <= 0 || nr > kParkMaxReac) return hipErrorInvalidValue;
  const size_t park = (size_t)2 * nr * kChemTB * sizeof(double);
  hipLaunchKernelGGL(k_chem_batch<9>, dim3(nblk(count, kChemTB)), dim3(kChemTB), park, s, m, n, count, idx, rho, T, Y,
                     dt, dY, dT);
  return hipGetLastError();
}

}  // namespace ostc
