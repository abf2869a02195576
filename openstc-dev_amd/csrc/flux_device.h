// Fused per-direction flux sweeps for gfx950.
//
// One kernel per grid direction computes, for every face i+1/2 of a line,
//   F = F_inviscid(reconstruction + Riemann solver) - F_viscous
// and the AWENO high-order central correction G in flux form, then updates
//   dq(i) -= (F(i+1/2) - F(i-1/2)) + [cds(i)] (G(i+1/2) - G(i-1/2))
// in the same pass: no face-flux buffer in HBM and no separate viscous or
// CDS launches (the reference uses up to 3 launches per direction:
// InviscidScheme.cu:10-46/160-205, AWENO.cu:213-264, SchemeSelector.cuh:49-157).
//
// Thread mapping (wave64):
//  * i-sweep: a wavefront owns 64 consecutive faces of the row-after-row
//    face sequence of a k-plane (63 new cells); the left face of a cell
//    comes from the neighbouring lane through a cross-lane shuffle.  Loads
//    along i are coalesced.
//  * j- and k-sweeps: lanes run along i (coalesced); a workgroup marches
//    along a line segment, its wavefronts computing consecutive faces and
//    passing them to the next row through LDS.
//
// The G form of the CDS term: with node fluxes f, the reference's
//   c1 (f[+3]-f[-3]) + c2 (f[+2]-f[-2]) + c3 (f[+1]-f[-1])
// equals G(i+1/2) - G(i-1/2) for
//   G(i+1/2) = c1 (f[-2]+f[+3]) + (c1+c2) (f[-1]+f[+2]) + (c1+c2+c3) (f[0]+f[+1]).
// The reference applies it to cells [0, n-2*ngg) only (AWENO.cu:234-258);
// that range is kept.
//
// This header holds the device code and the launch template; flux.hip
// instantiates the laminar sweeps and flux_sst.hip the k-omega SST ones (a
// translation unit of its own, built with its own optimisation level).
#pragma once

#include <algorithm>
#include <climits>

#include "device_common.h"
#include "flux_kernels.h"
#include "sst_device.h"

namespace ostc {

namespace {

template <typename R>
__device__ inline R minmod(R a, R b) {
  const R sa = a < R(0) ? R(-1) : R(1), sb = b < R(0) ? R(-1) : R(1);
  return R(0.5) * (sa + sb) * fmin(fabs(a), fabs(b));
}

// ------------------------------------------------------ reconstruction
// Reconstruction.cuh:20-72, Reconstruction.cu:5-46 on primitives
// (rho,u,v,w,p,Y...) of the 4 points p0-st .. p0+2st (recon 2 MUSCL, 3 NND2,
// otherwise first order: the two adjacent points only).  Appends rhoE at
// [nv] and gamma at [nv+1].
template <int NS, int NT, typename R>
__device__ inline void recon_pv(const ZoneView &z, int recon, const DMech *__restrict__ m, int64_t p0, int64_t st,
                                R *pl, R *pr) {
  constexpr int NV = 5 + NS + NT;
  const R kappa = R(1.0 / 3.0);
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    const double *src = l < 5 ? z.bv + (int64_t)l * z.sz : z.sv + (int64_t)(l - 5) * z.sz;
    const R q1 = (R)src[p0], q2 = (R)src[p0 + st];
    if (recon == 2) {
      const R q0 = (R)src[p0 - st], q3 = (R)src[p0 + 2 * st];
      const R di = q1 - q0, di1 = q2 - q1, di2 = q3 - q2;
      const R dnl = minmod(di, di1), dpl = minmod(di1, di), dnr = minmod(di1, di2), dpr = minmod(di2, di1);
      pl[l] = q1 + R(0.25) * ((R(1) - kappa) * dnl + (R(1) + kappa) * dpl);
      pr[l] = q2 - R(0.25) * ((R(1) - kappa) * dpr + (R(1) + kappa) * dnr);
    } else if (recon == 3) {
      const R q0 = (R)src[p0 - st], q3 = (R)src[p0 + 2 * st];
      const R di = q1 - q0, di1 = q2 - q1, di2 = q3 - q2;
      pl[l] = q1 + R(0.5) * minmod(di, di1);
      pr[l] = q2 - R(0.5) * minmod(di1, di2);
    } else {
      pl[l] = q1;
      pr[l] = q2;
    }
  }
  if constexpr (NS > 0) {
    double el = 0.5 * ((double)pl[1] * pl[1] + (double)pl[2] * pl[2] + (double)pl[3] * pl[3]);
    double er = 0.5 * ((double)pr[1] * pr[1] + (double)pr[2] * pr[2] + (double)pr[3] * pr[3]);
    double mil = 0, mir = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      mil += (double)pl[5 + l] / m->mw[l];
      mir += (double)pr[5 + l] / m->mw[l];
    }
    const double tl = (double)pl[4] / ((double)pl[0] * kRuD * mil), tr = (double)pr[4] / ((double)pr[0] * kRuD * mir);
    double h[NS], c[NS];
    double cpl = 0, cvl = 0, cpr = 0, cvr = 0;
    d_enthalpy_cp<NS>(tl, fresh(m), h, c);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cpl += c[l] * (double)pl[l + 5];
      cvl += (double)pl[l + 5] * (c[l] - m->r_mw[l]);
      el += h[l] * (double)pl[l + 5];
    }
    d_enthalpy_cp<NS>(tr, fresh(m), h, c);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cpr += c[l] * (double)pr[l + 5];
      cvr += (double)pr[l + 5] * (c[l] - m->r_mw[l]);
      er += h[l] * (double)pr[l + 5];
    }
    pl[NV] = (R)((double)pl[0] * el - (double)pl[4]);
    pr[NV] = (R)((double)pr[0] * er - (double)pr[4]);
    pl[NV + 1] = (R)(cpl / cvl);
    pr[NV + 1] = (R)(cpr / cvr);
  } else {
    const R el = R(0.5) * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
    const R er = R(0.5) * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
    pl[NV] = el * pl[0] + pl[4] / R(kGam - 1);
    pr[NV] = er * pr[0] + pr[4] / R(kGam - 1);
    pl[NV + 1] = R(kGam);
    pr[NV + 1] = R(kGam);
  }
}

// ------------------------------------------------------ reciprocals
// 1/x for finite, normal x: the hardware estimate refined by two
// Newton-Raphson steps r <- r + r (1 - x r), each squaring the relative
// error, so the result is within an ulp of 1/x.  A quotient a * rcp(b)
// then differs from a / b by rounding only; shared denominators are
// inverted once.
__device__ inline double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// ------------------------------------------------------ AWENO (air)
// AWENO.cu:151-206 WENO5 on the six characteristic values v[s] of the
// stencil points s = 0..5 (i-2..i+3): v-(i+1/2) and v+(i+1/2).
// The nonlinear weights a_k = c_k (1 + tau / (eps + b_k)^2) of each side
// share one reciprocal of the product of their three denominators
// (each denominator lies in [1e-12, 1e14] for characteristic variables in
// density units, far from the double range), and the two normalisations
// share one reciprocal of the product of the weight sums.
template <typename R>
__device__ inline void weno5_vals(const R *v, R &vm, R &vp) {
  const R t12 = R(13.0 / 12), eps = R(1e-6);
  const R o16 = R(1.0 / 16), f8 = R(5.0 / 8), f16 = R(5.0 / 16);
  // left-biased candidates and smoothness indicators
  const R u0 = R(0.125) * (R(3) * v[0] - R(10) * v[1] + R(15) * v[2]);
  const R u1 = R(0.125) * (-v[1] + R(6) * v[2] + R(3) * v[3]);
  const R u2 = R(0.125) * (R(3) * v[2] + R(6) * v[3] - v[4]);
  const R bl0 = t12 * (v[0] + v[2] - R(2) * v[1]) * (v[0] + v[2] - R(2) * v[1]) +
                R(0.25) * (v[0] + R(3) * v[2] - R(4) * v[1]) * (v[0] + R(3) * v[2] - R(4) * v[1]);
  const R bl1 = t12 * (v[1] + v[3] - R(2) * v[2]) * (v[1] + v[3] - R(2) * v[2]) + R(0.25) * (v[3] - v[1]) * (v[3] - v[1]);
  const R bl2 = t12 * (v[2] + v[4] - R(2) * v[3]) * (v[2] + v[4] - R(2) * v[3]) +
                R(0.25) * (R(3) * v[2] + v[4] - R(4) * v[3]) * (R(3) * v[2] + v[4] - R(4) * v[3]);
  // right-biased: candidates u1, u2, u3
  const R u3 = R(0.125) * (R(15) * v[3] - R(10) * v[4] + R(3) * v[5]);
  const R br0 = t12 * (v[1] + v[3] - R(2) * v[2]) * (v[1] + v[3] - R(2) * v[2]) +
                R(0.25) * (v[1] + R(3) * v[3] - R(4) * v[2]) * (v[1] + R(3) * v[3] - R(4) * v[2]);
  const R br1 = t12 * (v[2] + v[4] - R(2) * v[3]) * (v[2] + v[4] - R(2) * v[3]) + R(0.25) * (v[4] - v[2]) * (v[4] - v[2]);
  const R br2 = t12 * (v[3] + v[5] - R(2) * v[4]) * (v[3] + v[5] - R(2) * v[4]) +
                R(0.25) * (R(3) * v[3] + v[5] - R(4) * v[4]) * (R(3) * v[3] + v[5] - R(4) * v[4]);
  const R taul = (bl0 - bl2) * (bl0 - bl2), taur = (br0 - br2) * (br0 - br2);
  const R dl0 = (eps + bl0) * (eps + bl0), dl1 = (eps + bl1) * (eps + bl1), dl2 = (eps + bl2) * (eps + bl2);
  const R dr0 = (eps + br0) * (eps + br0), dr1 = (eps + br1) * (eps + br1), dr2 = (eps + br2) * (eps + br2);
  const R il = taul * rcp(dl0 * dl1 * dl2), ir = taur * rcp(dr0 * dr1 * dr2);
  const R al0 = o16 + o16 * (il * (dl1 * dl2)), al1 = f8 + f8 * (il * (dl0 * dl2)), al2 = f16 + f16 * (il * (dl0 * dl1));
  const R ar0 = f16 + f16 * (ir * (dr1 * dr2)), ar1 = f8 + f8 * (ir * (dr0 * dr2)), ar2 = o16 + o16 * (ir * (dr0 * dr1));
  const R sl = al0 + al1 + al2, sr = ar0 + ar1 + ar2;
  const R is = rcp(sl * sr);
  vm = (al0 * u0 + al1 * u1 + al2 * u2) * (sr * is);
  vp = (ar0 * u1 + ar1 * u2 + ar2 * u3) * (sl * is);
}

// one characteristic row: v(s) = L . cvb[s][0..4], then WENO5
template <typename R>
__device__ inline void weno5_row(const R *L, const R (*cvb)[6], R &vm, R &vp) {
  R v[6];
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    R a = R(0);
#pragma unroll
    for (int c = 0; c < 5; ++c) a += L[c] * cvb[s][c];
    v[s] = a;
  }
  weno5_vals(v, vm, vp);
}

// AWENO.cu:17-149 AWENO_interpolation<Air>.  mk0/mk1: metric row of the
// sweep direction at the two points adjacent to the face.
template <typename R>
__device__ inline void aweno_air(const R (*cvb)[6], const R *mk0, const R *mk1, R *pl, R *pr) {
  const R *cl = cvb[2], *cr = cvb[3];
  const R sl = sqrt(cl[0]), sr = sqrt(cr[0]);
  const R isum = rcp(sl + sr), icl = rcp(cl[0]), icr = rcp(cr[0]);
  const R rlc = sl * isum, rrc = sr * isum;
  const R um = rlc * cl[1] * icl + rrc * cr[1] * icr;
  const R vm = rlc * cl[2] * icl + rrc * cr[2] * icr;
  const R wm = rlc * cl[3] * icl + rrc * cr[3] * icr;
  const R ekm = R(0.5) * (um * um + vm * vm + wm * wm);
  const R gm1 = R(kGam - 1);
  const R hl = (cl[4] + cl[5]) * icl, hr = (cr[4] + cr[5]) * icr;
  const R hm = rlc * hl + rrc * hr;
  const R cm2 = gm1 * (hm - ekm);
  const R cm = sqrt(cm2);
  const R ic2 = rcp(cm2), ic = cm * ic2, hc2 = R(0.5) * ic2;
  R kx = R(0.5) * (mk0[0] + mk1[0]), ky = R(0.5) * (mk0[1] + mk1[1]), kz = R(0.5) * (mk0[2] + mk1[2]);
  const R igK = rcp(sqrt(kx * kx + ky * ky + kz * kz));
  kx *= igK;
  ky *= igK;
  kz *= igK;
  const R Uk = kx * um + ky * vm + kz * wm;
  const R alpha = gm1 * ekm;
  const R ac = R(1) - alpha * ic2, g2 = gm1 * ic2;
  R L[25];
  L[0] = (alpha + Uk * cm) * hc2;
  L[1] = -(gm1 * um + kx * cm) * hc2;
  L[2] = -(gm1 * vm + ky * cm) * hc2;
  L[3] = -(gm1 * wm + kz * cm) * hc2;
  L[4] = gm1 * hc2;
  L[5] = kx * ac - (kz * vm - ky * wm) * ic;
  L[6] = kx * g2 * um;
  L[7] = (kx * gm1 * vm + kz * cm) * ic2;
  L[8] = (kx * gm1 * wm - ky * cm) * ic2;
  L[9] = -kx * g2;
  L[10] = ky * ac - (kx * wm - kz * um) * ic;
  L[11] = (ky * gm1 * um - kz * cm) * ic2;
  L[12] = ky * g2 * vm;
  L[13] = (ky * gm1 * wm + kx * cm) * ic2;
  L[14] = -ky * g2;
  L[15] = kz * ac - (ky * um - kx * vm) * ic;
  L[16] = (kz * gm1 * um + ky * cm) * ic2;
  L[17] = (kz * gm1 * vm - kx * cm) * ic2;
  L[18] = kz * g2 * wm;
  L[19] = -kz * g2;
  L[20] = (alpha - Uk * cm) * hc2;
  L[21] = -(gm1 * um - kx * cm) * hc2;
  L[22] = -(gm1 * vm - ky * cm) * hc2;
  L[23] = -(gm1 * wm - kz * cm) * hc2;
  L[24] = gm1 * hc2;
  R vmn[5], vpl[5];
#pragma unroll
  for (int r = 0; r < 5; ++r) weno5_row(&L[r * 5], cvb, vmn[r], vpl[r]);
  // right eigenvectors (AWENO.cu:97-122), applied row by row
  const R ag = alpha * R(1.0 / (kGam - 1));
  const R Rm[25] = {R(1),
                    kx,
                    ky,
                    kz,
                    R(1),
                    um - kx * cm,
                    kx * um,
                    ky * um - kz * cm,
                    kz * um + ky * cm,
                    um + kx * cm,
                    vm - ky * cm,
                    kx * vm + kz * cm,
                    ky * vm,
                    kz * vm - kx * cm,
                    vm + ky * cm,
                    wm - kz * cm,
                    kx * wm - ky * cm,
                    ky * wm + kx * cm,
                    kz * wm,
                    wm + kz * cm,
                    hm - Uk * cm,
                    kx * ag + (kz * vm - ky * wm) * cm,
                    ky * ag + (kx * wm - kz * um) * cm,
                    kz * ag + (ky * um - kx * vm) * cm,
                    hm + Uk * cm};
  R ql[5], qr[5];
#pragma unroll
  for (int a = 0; a < 5; ++a) {
    R sa = R(0), sb = R(0);
#pragma unroll
    for (int b = 0; b < 5; ++b) {
      sa += Rm[a * 5 + b] * vmn[b];
      sb += Rm[a * 5 + b] * vpl[b];
    }
    ql[a] = sa;
    qr[a] = sb;
  }
  const R iql = rcp(ql[0]), iqr = rcp(qr[0]);
  pl[0] = ql[0];
  pl[1] = ql[1] * iql;
  pl[2] = ql[2] * iql;
  pl[3] = ql[3] * iql;
  pl[4] = gm1 * (ql[4] - R(0.5) * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]) * ql[0]);
  pl[5] = ql[4];
  pl[6] = R(kGam);
  pr[0] = qr[0];
  pr[1] = qr[1] * iqr;
  pr[2] = qr[2] * iqr;
  pr[3] = qr[3] * iqr;
  pr[4] = gm1 * (qr[4] - R(0.5) * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]) * qr[0]);
  pr[5] = qr[4];
  pr[6] = R(kGam);
}

// ------------------------------------------------------ AWENO (mixtures, scalars)
// The reference's mixture interpolation is an empty stub (AWENO.cu:10-15) and
// its air interpolation carries no transported scalars (AWENO.cu:53-54).  The
// build's definition, shared with the oracle: the air algorithm
// (AWENO.cu:17-149) on the 5 + n_scalar conservative variables, projected on
// the characteristic fields of the frozen multi-species Euler Jacobian at the
// face's Roe average (the eigen-system of the mixture Roe flux, Roe.cu:186-327):
//  * gamma and the species enthalpies h_l at the Roe temperature
//    T~ = <p / rho> / (R sum Y~_l / W_l) (sqrt(rho)-weighted averages), the
//    sound speed c~ = sqrt(gamma R T~ sum Y~/W), and the total enthalpy
//    H~ = ek~ + sum Y~_l h_l(T~) (which makes L and R exact inverses);
//  * acoustic waves from c1 = dp / c~^2, the entropy / shear combinations of
//    air, and one wave Q_{5+l} - Y~_l Q_0 per scalar (species, k, omega);
//  * each reconstructed side is turned into primitives by a Newton inversion
//    of its internal energy to 1e-12 relative (FP64, at most 30 iterations,
//    from T~), with gamma of the side at its own temperature.
// Laminar air with k and omega (NS = 0) keeps gamma = 1.4 and the
// sqrt(rho)-averaged H.
template <int NS, int NT, typename R>
struct MixEig {
  static constexpr int NSC = NS + NT;
  R u, v, w, h, ek, c, ic, c2, gm1, nx, ny, nz, U;
  double t;
  R y[NSC > 0 ? NSC : 1];
  R a[NS > 0 ? NS : 1];  // left rows:  W~/W_l - h_l (gamma - 1) / c~^2
  R b[NS > 0 ? NS : 1];  // right rows: h_l - W~/W_l c~^2 / (gamma - 1)
};

// Roe state of the face between cl and cr = (Q[NV], p); k0/k1 the metric
// rows of the sweep direction at the two points
template <int NS, int NT, typename R>
__device__ inline void mix_roe_state(const DMech *__restrict__ m, const R *cl, const R *cr, const R *k0, const R *k1,
                                     MixEig<NS, NT, R> &E) {
  constexpr int NV = 5 + NS + NT;
  const R dl = sqrt(cl[0]), dr = sqrt(cr[0]), id = rcp(dl + dr);
  const R il = rcp(cl[0]), ir = rcp(cr[0]);
  const R wl = dl * il, wr = dr * ir;  // sqrt(rho) / rho of each side
  E.u = (wl * cl[1] + wr * cr[1]) * id;
  E.v = (wl * cl[2] + wr * cr[2]) * id;
  E.w = (wl * cl[3] + wr * cr[3]) * id;
  E.ek = R(0.5) * (E.u * E.u + E.v * E.v + E.w * E.w);
#pragma unroll
  for (int l = 0; l < NS + NT; ++l) E.y[l] = (wl * cl[5 + l] + wr * cr[5 + l]) * id;
  R gm1;
  if constexpr (NS > 0) {
    double mwi = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) mwi += (double)E.y[l] / m->mw[l];
    const double t = (double)((dl * (cl[NV] * il) + dr * (cr[NV] * ir)) * id) / (kRuD * mwi);
    double hi[NS], cpi[NS];
    d_enthalpy_cp<NS>(t, fresh(m), hi, cpi);
    double cp = 0, cv = 0, hs = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cp += cpi[l] * (double)E.y[l];
      cv += (double)E.y[l] * (cpi[l] - m->r_mw[l]);
      hs += (double)E.y[l] * hi[l];
    }
    const double gamma = cp / cv;
    const double c = sqrt(gamma * kRuD * mwi * t);
    const double mw = 1 / mwi;
    E.t = t;
    E.c = (R)c;
    E.h = E.ek + (R)hs;
    gm1 = (R)(gamma - 1);
    const double c2 = c * c, g2 = (gamma - 1) / c2, cg = c2 / (gamma - 1);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      E.a[l] = (R)(mw / m->mw[l] - hi[l] * g2);
      E.b[l] = (R)(hi[l] - mw / m->mw[l] * cg);
    }
  } else {
    E.t = 0;
    E.h = (wl * (cl[4] + cl[NV]) + wr * (cr[4] + cr[NV])) * id;
    gm1 = R(kGam - 1);
    E.c = sqrt(gm1 * (E.h - E.ek));
  }
  E.gm1 = gm1;
  E.c2 = E.c * E.c;
  E.ic = rcp(E.c);
  const R kx = R(0.5) * (k0[0] + k1[0]), ky = R(0.5) * (k0[1] + k1[1]), kz = R(0.5) * (k0[2] + k1[2]);
  const R igk = rcp(sqrt(kx * kx + ky * ky + kz * kz));
  E.nx = kx * igk;
  E.ny = ky * igk;
  E.nz = kz * igk;
  E.U = E.nx * E.u + E.ny * E.v + E.nz * E.w;
}

// characteristic amplitudes w = L Q of a conservative vector Q
template <int NS, int NT, typename R>
__device__ inline void mix_left(const MixEig<NS, NT, R> &E, const R *Q, R *wv) {
  constexpr int NV = 5 + NS + NT;
  R c1 = E.gm1 * (E.ek * Q[0] - E.u * Q[1] - E.v * Q[2] - E.w * Q[3] + Q[4]) * rcp(E.c2);
#pragma unroll
  for (int l = 0; l < NS; ++l) c1 += E.a[l] * Q[5 + l];
  const R c2v = (E.nx * Q[1] + E.ny * Q[2] + E.nz * Q[3] - E.U * Q[0]) * E.ic;
  const R c3 = Q[0] - c1;
  wv[0] = R(0.5) * (c1 - c2v);
  wv[1] = E.nx * c3 - ((E.nz * E.v - E.ny * E.w) * Q[0] - E.nz * Q[2] + E.ny * Q[3]) * E.ic;
  wv[2] = E.ny * c3 - ((E.nx * E.w - E.nz * E.u) * Q[0] - E.nx * Q[3] + E.nz * Q[1]) * E.ic;
  wv[3] = E.nz * c3 - ((E.ny * E.u - E.nx * E.v) * Q[0] - E.ny * Q[1] + E.nx * Q[2]) * E.ic;
  wv[4] = R(0.5) * (c1 + c2v);
#pragma unroll
  for (int l = 5; l < NV; ++l) wv[l] = Q[l] - E.y[l - 5] * Q[0];
}

// conservative vector Q = R w of characteristic amplitudes w
template <int NS, int NT, typename R>
__device__ inline void mix_right(const MixEig<NS, NT, R> &E, const R *wv, R *Q) {
  constexpr int NV = 5 + NS + NT;
  const R c = E.c;
  const R c0 = E.nx * wv[1] + E.ny * wv[2] + E.nz * wv[3];
  const R C1 = c0 + wv[0] + wv[4], C2 = c * (wv[4] - wv[0]);
  R C3 = R(0);
#pragma unroll
  for (int l = 0; l < NS; ++l) C3 += E.b[l] * wv[5 + l];
  Q[0] = C1;
  Q[1] = E.u * C1 + E.nx * C2 - c * (E.nz * wv[2] - E.ny * wv[3]);
  Q[2] = E.v * C1 + E.ny * C2 - c * (E.nx * wv[3] - E.nz * wv[1]);
  Q[3] = E.w * C1 + E.nz * C2 - c * (E.ny * wv[1] - E.nx * wv[2]);
  Q[4] = E.h * C1 + E.U * C2 - E.c2 * c0 * rcp(E.gm1) +
         c * ((E.nz * E.v - E.ny * E.w) * wv[1] + (E.nx * E.w - E.nz * E.u) * wv[2] +
              (E.ny * E.u - E.nx * E.v) * wv[3]) +
         C3;
#pragma unroll
  for (int l = 5; l < NV; ++l) Q[l] = wv[l] + E.y[l - 5] * C1;
}

// Temperature of a mixture of mass fractions Y at internal energy e per unit
// mass: Newton on sum Y h(T) - R T sum Y/W = e from t0, until |dT| <= 1e-12 T
// (at most 30 iterations).  FP64 in every precision mode.
template <int NS>
__device__ inline double d_energy_temperature(const DMech *__restrict__ m, const double *Y, double e, double t0,
                                              double &mwi) {
  mwi = 0;
#pragma unroll
  for (int l = 0; l < NS; ++l) mwi += Y[l] / m->mw[l];
  double t = t0;
  for (int it = 0; it < 30; ++it) {
    double h[NS], cp[NS];
    d_enthalpy_cp<NS>(t, fresh(m), h, cp);
    double f = 0, fp = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      f += Y[l] * h[l];
      fp += Y[l] * cp[l];
    }
    f -= kRuD * t * mwi + e;
    fp -= kRuD * mwi;
    const double dt = f / fp;
    t -= dt;
    if (fabs(dt) <= 1e-12 * t) break;
  }
  return t;
}

// primitives (rho, u, v, w, p, scalars; rho E at [NV], gamma at [NV+1]) of a
// reconstructed conservative side Q
template <int NS, int NT, typename R>
__device__ inline void mix_side(const DMech *__restrict__ m, const MixEig<NS, NT, R> &E, const R *Q, R *pv) {
  constexpr int NV = 5 + NS + NT;
  const R rho = Q[0], ri = rcp(rho);
  pv[0] = rho;
  pv[1] = Q[1] * ri;
  pv[2] = Q[2] * ri;
  pv[3] = Q[3] * ri;
#pragma unroll
  for (int l = 5; l < NV; ++l) pv[l] = Q[l] * ri;
  const R ek = R(0.5) * (pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3]);
  pv[NV] = Q[4];
  if constexpr (NS > 0) {
    double Y[NS], mwi;
#pragma unroll
    for (int l = 0; l < NS; ++l) Y[l] = (double)pv[5 + l];
    const double t = d_energy_temperature<NS>(m, Y, (double)(Q[4] * ri - ek), E.t, mwi);
    pv[4] = (R)((double)rho * kRuD * t * mwi);
    double cpi[NS];
    d_cp<NS>(t, fresh(m), cpi);
    double cp = 0, cv = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cp += cpi[l] * Y[l];
      cv += Y[l] * (cpi[l] - m->r_mw[l]);
    }
    pv[NV + 1] = (R)(cp / cv);
  } else {
    pv[4] = R(kGam - 1) * (Q[4] - ek * rho);
    pv[NV + 1] = R(kGam);
  }
}

// A reconstructed side is admissible when its density and pressure are
// positive, its gamma exceeds 1, everything is finite, and each mass
// fraction lies in [-0.01, 1.01].
template <int NS, int NT, typename R>
__device__ inline bool mix_side_ok(const R *pv) {
  constexpr int NV = 5 + NS + NT;
  bool ok = pv[0] > R(0) && pv[4] > R(0) && pv[NV + 1] > R(1) && isfinite(pv[4]) && isfinite(pv[NV]) &&
            isfinite(pv[NV + 1]);
#pragma unroll
  for (int l = 5; l < NV; ++l) {
    ok = ok && isfinite(pv[l]);
    if (l < 5 + NS) ok = ok && pv[l] >= R(-0.01) && pv[l] <= R(1.01);
  }
  return ok;
}

// AWENO interpolation of one face from the stencil points p0 + (s - 2) st,
// s = 0..5, every value read from memory (conservative variables and p).
// Where either reconstructed side is not admissible (WENO overshoot at a
// jump no candidate stencil avoids, e.g. a jet port two cells wide), both
// sides take the first-order values of the two adjacent points.
// Register use: the scalar waves Q_{5+l} - Y~_l Q_0 are formed and
// interpolated one scalar at a time (their stencil is read when needed), and
// the acoustic rows keep only the 6 x 5 flow rows of the stencil, so the
// 6 x NV amplitude block of the whole stencil is never live at once; the
// arithmetic and its order are those of mix_left / weno5 per component.
template <int NS, int NT, typename R>
__device__ inline void aweno_mix(const ZoneView &z, const DMech *__restrict__ m, int64_t p0, int64_t st,
                                 const R *mk0, const R *mk1, R *pl, R *pr) {
  constexpr int NV = 5 + NS + NT;
  MixEig<NS, NT, R> E;
  {
    R cl[NV + 1], cr[NV + 1];
#pragma unroll
    for (int l = 0; l < NV; ++l) {
      cl[l] = (R)z.cv[(int64_t)l * z.sz + p0];
      cr[l] = (R)z.cv[(int64_t)l * z.sz + p0 + st];
    }
    cl[NV] = (R)z.bv[4 * z.sz + p0];
    cr[NV] = (R)z.bv[4 * z.sz + p0 + st];
    mix_roe_state<NS, NT, R>(m, cl, cr, mk0, mk1, E);
  }
  // flow rows of the stencil, then (in place) their five wave amplitudes
  R q[6][5];
  R c1[6];
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int64_t p = p0 + (s - 2) * st;
#pragma unroll
    for (int c = 0; c < 5; ++c) q[s][c] = (R)z.cv[(int64_t)c * z.sz + p];
    c1[s] = E.gm1 * (E.ek * q[s][0] - E.u * q[s][1] - E.v * q[s][2] - E.w * q[s][3] + q[s][4]) * rcp(E.c2);
  }
  R wm[NV], wp[NV];
  // scalar waves, one scalar at a time; the species also complete c1
#pragma unroll
  for (int l = 0; l < NS + NT; ++l) {
    R v[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const R ql = (R)z.cv[(int64_t)(5 + l) * z.sz + p0 + (s - 2) * st];
      if (l < NS) c1[s] += E.a[l] * ql;
      v[s] = ql - E.y[l] * q[s][0];
    }
    weno5_vals(v, wm[5 + l], wp[5 + l]);
  }
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const R *Q = q[s];
    const R c2v = (E.nx * Q[1] + E.ny * Q[2] + E.nz * Q[3] - E.U * Q[0]) * E.ic;
    const R c3 = Q[0] - c1[s];
    const R w0 = R(0.5) * (c1[s] - c2v);
    const R w1 = E.nx * c3 - ((E.nz * E.v - E.ny * E.w) * Q[0] - E.nz * Q[2] + E.ny * Q[3]) * E.ic;
    const R w2 = E.ny * c3 - ((E.nx * E.w - E.nz * E.u) * Q[0] - E.nx * Q[3] + E.nz * Q[1]) * E.ic;
    const R w3 = E.nz * c3 - ((E.ny * E.u - E.nx * E.v) * Q[0] - E.ny * Q[1] + E.nx * Q[2]) * E.ic;
    const R w4 = R(0.5) * (c1[s] + c2v);
    q[s][0] = w0;
    q[s][1] = w1;
    q[s][2] = w2;
    q[s][3] = w3;
    q[s][4] = w4;
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    const R v[6] = {q[0][c], q[1][c], q[2][c], q[3][c], q[4][c], q[5][c]};
    weno5_vals(v, wm[c], wp[c]);
  }
  R qs[NV];
  mix_right<NS, NT, R>(E, wm, qs);
  mix_side<NS, NT, R>(m, E, qs, pl);
  mix_right<NS, NT, R>(E, wp, qs);
  mix_side<NS, NT, R>(m, E, qs, pr);
  if (!(mix_side_ok<NS, NT, R>(pl) && mix_side_ok<NS, NT, R>(pr))) recon_pv<NS, NT, R>(z, 1, m, p0, st, pl, pr);
}

// ------------------------------------------------------ Riemann solvers
// AusmP.cu:6-94.  pl/pr: (rho,u,v,w,p,Y..., rhoE, gamma); mk0/mk1 metric
// rows and j0/j1 Jacobians of the two adjacent points.
template <int NS, int NT, typename R>
__device__ inline void ausm_plus(const R *pl, const R *pr, const R *mk0, const R *mk1, R j0, R j1, R *f) {
  constexpr int NV = 5 + NS + NT;
  const R k1 = R(0.5) * (j0 * mk0[0] + j1 * mk1[0]);
  const R k2 = R(0.5) * (j0 * mk0[1] + j1 * mk1[1]);
  const R k3 = R(0.5) * (j0 * mk0[2] + j1 * mk1[2]);
  const R gk = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
  const R igk = rcp(gk);
  const R ul = (k1 * pl[1] + k2 * pl[2] + k3 * pl[3]) * igk;
  const R ur = (k1 * pr[1] + k2 * pr[2] + k3 * pr[3]) * igk;
  const R c = R(0.5) * (sqrt(pl[NV + 1] * pl[4] * rcp(pl[0])) + sqrt(pr[NV + 1] * pr[4] * rcp(pr[0])));
  const R ic = rcp(c);
  const R ml = ul * ic, mr = ur * ic;
  const R alpha = R(3 / 16.0);
  R mlp, mrn, plp, prn;
  if (fabs(ml) > R(1)) {
    mlp = R(0.5) * (ml + fabs(ml));
    plp = mlp / ml;
  } else {
    const R a = (ml + R(1)) * (ml + R(1)) * R(0.25);
    const R b = (ml * ml - R(1)) * (ml * ml - R(1));
    mlp = a + R(0.125) * b;
    plp = a * (R(2) - ml) + alpha * ml * b;
  }
  if (fabs(mr) > R(1)) {
    mrn = R(0.5) * (mr - fabs(mr));
    prn = mrn / mr;
  } else {
    const R a = (mr - R(1)) * (mr - R(1)) * R(0.25);
    const R b = (mr * mr - R(1)) * (mr * mr - R(1));
    mrn = -a - R(0.125) * b;
    prn = a * (R(2) + mr) - alpha * mr * b;
  }
  const R pc = plp * pl[4] + prn * pr[4];
  const R mh = mlp + mrn;
  const R mpos = R(0.5) * (mh + fabs(mh)), mneg = R(0.5) * (mh - fabs(mh));
  const R mass = c * (pl[0] * mpos + pr[0] * mneg);
  const R coeff = mass * gk;
  // upwind state, selected element by element (a selected pointer would
  // force both state arrays out of registers)
  const bool up = mass >= R(0);
  f[0] = coeff;
  f[1] = coeff * (up ? pl[1] : pr[1]) + pc * k1;
  f[2] = coeff * (up ? pl[2] : pr[2]) + pc * k2;
  f[3] = coeff * (up ? pl[3] : pr[3]) + pc * k3;
  f[4] = coeff * ((up ? pl[NV] : pr[NV]) + (up ? pl[4] : pr[4])) / (up ? pl[0] : pr[0]);
  if (up) {
#pragma unroll
    for (int l = 5; l < NV; ++l) f[l] = coeff * pl[l];
  } else {
#pragma unroll
    for (int l = 5; l < NV; ++l) f[l] = coeff * pr[l];
  }
}

// Face metric and Jacobian of HLLC and LF: the reference's (m0 + m1) / 2 and
// (J0 + J1) / 2 (HLLC.cu:64-71, LaxFriedrichs.cu:22-35); with conservative
// metrics (jface) the face vector (J0 m0 + J1 m1) / 2 and the factor 1, as
// AUSM+ and Roe take it, so that the face vectors of a cell close.
template <typename R>
__device__ inline void face_vector(const R *mk0, const R *mk1, R j0, R j1, bool jface, R &kx, R &ky, R &kz, R &ja) {
  if (jface) {
    kx = R(0.5) * (j0 * mk0[0] + j1 * mk1[0]);
    ky = R(0.5) * (j0 * mk0[1] + j1 * mk1[1]);
    kz = R(0.5) * (j0 * mk0[2] + j1 * mk1[2]);
    ja = R(1);
  } else {
    kx = R(0.5) * (mk0[0] + mk1[0]);
    ky = R(0.5) * (mk0[1] + mk1[1]);
    kz = R(0.5) * (mk0[2] + mk1[2]);
    ja = R(0.5) * (j0 + j1);
  }
}

// HLLC.cu:10-143
template <int NS, int NT, typename R>
__device__ inline void hllc(const DMech *__restrict__ m, const R *pl, const R *pr, const R *mk0, const R *mk1, R j0,
                            R j1, bool jface, R *f) {
  constexpr int NV = 5 + NS + NT;
  const R sl0 = sqrt(pl[0]), sr0 = sqrt(pr[0]);
  const R isum = rcp(sl0 + sr0), ipl = rcp(pl[0]), ipr = rcp(pr[0]);
  const R rlc = sl0 * isum, rrc = sr0 * isum;
  const R ut = pl[1] * rlc + pr[1] * rrc, vt = pl[2] * rlc + pr[2] * rrc, wt = pl[3] * rlc + pr[3] * rrc;
  const R hl = (pl[NV] + pl[4]) * ipl, hr = (pr[NV] + pr[4]) * ipr;
  const R ht = hl * rlc + hr * rrc;
  const R ekt = R(0.5) * (ut * ut + vt * vt + wt * wt);
  R ct;
  if constexpr (NS == 0) {
    ct = sqrt(R(kGam - 1) * (ht - ekt));
  } else {
    double svm[NS];
    double mwi = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      svm[l] = (double)(rlc * pl[l + 5] + rrc * pr[l + 5]);
      mwi += svm[l] / m->mw[l];
    }
    const double t = (double)(rlc * (pl[4] * ipl) + rrc * (pr[4] * ipr)) / (kRuD * mwi);
    double h[NS], c[NS], cpp = 0, cvv = 0;
    d_enthalpy_cp<NS>(t, fresh(m), h, c);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cpp += c[l] * svm[l];
      cvv += svm[l] * (c[l] - m->r_mw[l]);
    }
    ct = (R)sqrt(cpp / cvv * kRuD * mwi * t);
  }
  R kx, ky, kz, ja;
  face_vector(mk0, mk1, j0, j1, jface, kx, ky, kz, ja);
  const R gK = sqrt(kx * kx + ky * ky + kz * kz);
  const R igK = rcp(gK);
  const R Utb = (kx * ut + ky * vt + kz * wt) * igK;
  const R Ul = kx * pl[1] + ky * pl[2] + kz * pl[3];
  const R Ulg = Ul * igK;
  const R cl = sqrt(pl[NV + 1] * pl[4] * ipl);
  const R sl = fmin(Ulg - cl, Utb - ct);
  if (sl >= R(0)) {
    const R rU = pl[0] * Ul;
    f[0] = ja * rU;
    f[1] = ja * (rU * pl[1] + pl[4] * kx);
    f[2] = ja * (rU * pl[2] + pl[4] * ky);
    f[3] = ja * (rU * pl[3] + pl[4] * kz);
    f[4] = ja * ((pl[NV] + pl[4]) * Ul);
#pragma unroll
    for (int l = 5; l < NV; ++l) f[l] = ja * rU * pl[l];
    return;
  }
  const R Ur = kx * pr[1] + ky * pr[2] + kz * pr[3];
  const R Urg = Ur * igK;
  const R cr = sqrt(pr[NV + 1] * pr[4] * ipr);
  const R sr = fmax(Urg + cr, Utb + ct);
  if (sr < R(0)) {
    const R rU = pr[0] * Ur;
    f[0] = ja * rU;
    f[1] = ja * (rU * pr[1] + pr[4] * kx);
    f[2] = ja * (rU * pr[2] + pr[4] * ky);
    f[3] = ja * (rU * pr[3] + pr[4] * kz);
    f[4] = ja * ((pr[NV] + pr[4]) * Ur);
#pragma unroll
    for (int l = 5; l < NV; ++l) f[l] = ja * rU * pr[l];
    return;
  }
  const R sm = ((pr[0] * Ur * (sr - Urg) - pl[0] * Ul * (sl - Ulg)) * igK + pl[4] - pr[4]) /
               (pr[0] * (sr - Urg) - pl[0] * (sl - Ulg));
  const R pm = pl[0] * (sl - Ulg) * (sm - Ulg) + pl[4];
  // star state on the side of the contact, selected element by element
  const bool left = sm >= R(0);
  const R s = left ? sl : sr;
  const R Uq = left ? Ul : Ur;
  const R iq = left ? ipl : ipr;
  const R q0 = left ? pl[0] : pr[0], q4 = left ? pl[4] : pr[4];
  const R pC = rcp(s - sm);
  const R Q = ja * pC * sm * (s * gK - Uq) * q0;
  const R dP = (s * pm - sm * q4) * pC * ja;
  f[0] = Q;
  f[1] = Q * (left ? pl[1] : pr[1]) + dP * kx;
  f[2] = Q * (left ? pl[2] : pr[2]) + dP * ky;
  f[3] = Q * (left ? pl[3] : pr[3]) + dP * kz;
  f[4] = Q * (left ? pl[NV] : pr[NV]) * iq + pC * ja * (s * pm * sm * gK - sm * q4 * Uq);
  if (left) {
#pragma unroll
    for (int l = 5; l < NV; ++l) f[l] = Q * pl[l];
  } else {
#pragma unroll
    for (int l = 5; l < NV; ++l) f[l] = Q * pr[l];
  }
}

// Roe.cu:60-87: entropy-fix delta of one point, eps (|U| + |V| (+|W|) +
// a * mean |grad xi|); 2-D drops the third direction.
template <typename R>
__device__ inline R efix_point(const ZoneView &z, const DParams &P, int64_t p) {
  R mm[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) mm[c] = (R)z.metric[c * z.sz + p];
  const R u = (R)z.bv[1 * z.sz + p], v = (R)z.bv[2 * z.sz + p], w = (R)z.bv[3 * z.sz + p];
  const R U = fabs(u * mm[0] + v * mm[1] + w * mm[2]);
  const R V = fabs(u * mm[3] + v * mm[4] + w * mm[5]);
  const R W = fabs(u * mm[6] + v * mm[7] + w * mm[8]);
  const R kx = sqrt(mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2]);
  const R ky = sqrt(mm[3] * mm[3] + mm[4] * mm[4] + mm[5] * mm[5]);
  const R kz = sqrt(mm[6] * mm[6] + mm[7] * mm[7] + mm[8] * mm[8]);
  const R a = (R)z.acoustic[p];
  if (P.dim == 2) return (R)P.efix * (U + V + a * R(0.5) * (kx + ky));
  return (R)P.efix * (U + V + W + a * (kx + ky + kz) / R(3));
}

// Roe.cu:186-375: half sum of the left and right fluxes, each with its own
// J * metric, minus 1/2 |A| dQ from the characteristic decomposition at the
// Roe average (species rows included; the mixture's gamma and c from the
// Roe-averaged temperature, FP64), Harten's entropy fix with the face
// average of the point deltas d0, d1.
template <int NS, int NT, typename R>
__device__ inline void roe(const DMech *__restrict__ m, const R *pl, const R *pr, const R *mk0, const R *mk1, R j0,
                           R j1, R d0, R d1, R *f) {
  constexpr int NV = 5 + NS + NT;
#pragma unroll
  for (int l = 0; l < NV; ++l) f[l] = R(0);
  // half of each side's physical flux (called once per side so that no
  // pointer to either state array is selected at run time)
  auto half_flux = [&](const R *q, const R *mk, const R J) {
    const R Jx = J * mk[0], Jy = J * mk[1], Jz = J * mk[2];
    const R Uk = q[1] * Jx + q[2] * Jy + q[3] * Jz;
    const R co = Uk * q[0];
    f[0] += R(0.5) * co;
    f[1] += R(0.5) * (co * q[1] + q[4] * Jx);
    f[2] += R(0.5) * (co * q[2] + q[4] * Jy);
    f[3] += R(0.5) * (co * q[3] + q[4] * Jz);
    f[4] += R(0.5) * Uk * (q[4] + q[NV]);
#pragma unroll
    for (int l = 5; l < NV; ++l) f[l] += R(0.5) * co * q[l];
  };
  half_flux(pl, mk0, j0);
  half_flux(pr, mk1, j1);
  const R dl = sqrt(pl[0]), dr = sqrt(pr[0]), id = R(1) / (dl + dr);
  const R u = (dl * pl[1] + dr * pr[1]) * id, v = (dl * pl[2] + dr * pr[2]) * id, w = (dl * pl[3] + dr * pr[3]) * id;
  const R ek = R(0.5) * (u * u + v * v + w * w);
  const R h = (dl * (pl[NV] + pl[4]) / pl[0] + dr * (pr[NV] + pr[4]) / pr[0]) * id;
  R gamma = R(kGam), c, mw = R(kMwAirD);
  R svm[NV > 5 ? NV - 5 : 1], hi[NS > 0 ? NS : 1];
#pragma unroll
  for (int l = 0; l < NV - 5; ++l) svm[l] = (dl * pl[l + 5] + dr * pr[l + 5]) * id;
  if constexpr (NS == 0) {
    c = sqrt((gamma - R(1)) * (h - ek));
  } else {
    double mwi = 0;
#pragma unroll
    for (int l = 0; l < NS; ++l) mwi += (double)svm[l] / m->mw[l];
    const double t = (double)((dl * (pl[4] / pl[0]) + dr * (pr[4] / pr[0])) * id) / (kRuD * mwi);
    double hh[NS], cc[NS], cp = 0, cv = 0;
    d_enthalpy_cp<NS>(t, fresh(m), hh, cc);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cp += cc[l] * (double)svm[l];
      cv += (double)svm[l] * (cc[l] - m->r_mw[l]);
      hi[l] = (R)hh[l];
    }
    gamma = (R)(cp / cv);
    c = (R)sqrt(cp / cv * kRuD * mwi * t);
    mw = (R)(1.0 / mwi);
  }
  R kx = R(0.5) * (mk0[0] + mk1[0]), ky = R(0.5) * (mk0[1] + mk1[1]), kz = R(0.5) * (mk0[2] + mk1[2]);
  const R gk = sqrt(kx * kx + ky * ky + kz * kz);
  R Uk = kx * u + ky * v + kz * w;
  R lam[3] = {Uk - gk * c, Uk, Uk + gk * c};
  const R d = R(0.5) * (d0 + d1);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lam[a] = fabs(lam[a]);
    if (lam[a] < d) lam[a] = R(0.5) * (lam[a] * lam[a] / d + d);
  }
  kx /= gk;
  ky /= gk;
  kz /= gk;
  Uk /= gk;
  const R ja = R(0.5) * (j0 + j1);
  R dq[NV];
  dq[0] = ja * (pr[0] - pl[0]);
#pragma unroll
  for (int l = 1; l < NV; ++l) dq[l] = ja * (pr[0] * pr[l] - pl[0] * pl[l]);
  dq[4] = ja * (pr[NV] - pl[NV]);
  const R g1 = gamma - R(1), ic2 = R(1) / (c * c);
  R c1 = g1 * (ek * dq[0] - u * dq[1] - v * dq[2] - w * dq[3] + dq[4]) * ic2;
  R c2 = (kx * dq[1] + ky * dq[2] + kz * dq[3] - Uk * dq[0]) / c;
#pragma unroll
  for (int l = 0; l < NS; ++l) c1 += (mw / (R)m->mw[l] - hi[l] * g1 * ic2) * dq[l + 5];
  R c3 = dq[0] - c1;
  R b[NV];
  b[0] = -lam[0] * (R(0.5) * (c1 - c2));
  b[1] = -lam[1] * (kx * c3 - ((kz * v - ky * w) * dq[0] - kz * dq[2] + ky * dq[3]) / c);
  b[2] = -lam[1] * (ky * c3 - ((kx * w - kz * u) * dq[0] - kx * dq[3] + kz * dq[1]) / c);
  b[3] = -lam[1] * (kz * c3 - ((ky * u - kx * v) * dq[0] - ky * dq[1] + kx * dq[2]) / c);
  b[4] = -lam[2] * (R(0.5) * (c1 + c2));
#pragma unroll
  for (int l = 0; l < NV - 5; ++l) b[l + 5] = -lam[1] * (dq[l + 5] - svm[l] * dq[0]);
  const R c0 = kx * b[1] + ky * b[2] + kz * b[3];
  c1 = c0 + b[0] + b[4];
  c2 = c * (b[4] - b[0]);
  c3 = R(0);
#pragma unroll
  for (int l = 0; l < NS; ++l) c3 += (hi[l] - mw / (R)m->mw[l] * c * c / g1) * b[l + 5];
  f[0] += R(0.5) * c1;
  f[1] += R(0.5) * (u * c1 + kx * c2 - c * (kz * b[2] - ky * b[3]));
  f[2] += R(0.5) * (v * c1 + ky * c2 - c * (kx * b[3] - kz * b[1]));
  f[3] += R(0.5) * (w * c1 + kz * c2 - c * (ky * b[1] - kx * b[2]));
  f[4] += R(0.5) * (h * c1 + Uk * c2 - c * c * c0 / g1 +
                    c * ((kz * v - ky * w) * b[1] + (kx * w - kz * u) * b[2] + (ky * u - kx * v) * b[3]) + c3);
#pragma unroll
  for (int l = 0; l < NV - 5; ++l) f[5 + l] += R(0.5) * (b[l + 5] + svm[l] * c1);
}

// LaxFriedrichs.cu:14-49 (air only; the reference's generic version is empty)
template <int NS, int NT, typename R>
__device__ inline void lax_friedrichs(const R *pl, const R *pr, const R *mk0, const R *mk1, R j0, R j1, bool jface,
                                      R *f) {
  constexpr int NV = 5 + NS + NT;
  R kx, ky, kz, ja;
  face_vector(mk0, mk1, j0, j1, jface, kx, ky, kz, ja);
  const R gK = sqrt(kx * kx + ky * ky + kz * kz);
  const R Ukl = pl[1] * kx + pl[2] * ky + pl[3] * kz, Ukr = pr[1] * kx + pr[2] * ky + pr[3] * kz;
  const R cl = sqrt(R(kGam) * pl[4] / pl[0]), cr = sqrt(R(kGam) * pr[4] / pr[0]);
  const R sr = fmax(fabs(Ukl) + cl * gK, fabs(Ukr) + cr * gK);
  const R hj = R(0.5) * ja;
  const R rUl = pl[0] * Ukl, rUr = pr[0] * Ukr;
  f[0] = (rUl + rUr - sr * (pr[0] - pl[0])) * hj;
  f[1] = (rUl * pl[1] + rUr * pr[1] + kx * (pl[4] + pr[4]) - sr * (pr[1] * pr[0] - pl[1] * pl[0])) * hj;
  f[2] = (rUl * pl[2] + rUr * pr[2] + ky * (pl[4] + pr[4]) - sr * (pr[2] * pr[0] - pl[2] * pl[0])) * hj;
  f[3] = (rUl * pl[3] + rUr * pr[3] + kz * (pl[4] + pr[4]) - sr * (pr[3] * pr[0] - pl[3] * pl[0])) * hj;
  f[4] = ((pl[NV] + pl[4]) * Ukl + (pr[NV] + pr[4]) * Ukr - sr * (pr[NV] - pl[NV])) * hj;
#pragma unroll
  for (int l = 5; l < NV; ++l) f[l] = R(0);
}

// ------------------------------------------------------ viscous face
// ViscousScheme.cuh:22-821: central face gradients (normal difference plus
// 4-point cross averages), chain rule with face-averaged metrics, Stokes
// stress, Fourier flux, mixture-averaged diffusion with the correction
// velocity; the flux is projected on the J-weighted face metric.
template <int NS, int NT, typename R>
__device__ inline void viscous_face(const ZoneView &z, const DParams &P, const DMech *__restrict__ m, int dir,
                                    int64_t p0, int64_t st, R *fv) {
  const int64_t p1 = p0 + st;
  const int64_t so[3] = {1, z.d2, z.d1};
  R mm[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) mm[c] = R(0.5) * ((R)z.metric[c * z.sz + p0] + (R)z.metric[c * z.sz + p1]);
  // derivatives along the three computational directions of field src
  auto grad = [&](const double *src, R *g) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d == dir) {
        g[d] = (R)src[p1] - (R)src[p0];
      } else if (d < 2 || z.mz > 1) {
        const int64_t o = so[d];
        g[d] = R(0.25) * ((R)src[p0 + o] - (R)src[p0 - o] + (R)src[p1 + o] - (R)src[p1 - o]);
      } else {
        g[d] = R(0);
      }
    }
  };
  auto chain = [&](const R *g, R *gx) {
#pragma unroll
    for (int c = 0; c < 3; ++c) gx[c] = g[0] * mm[c] + g[1] * mm[3 + c] + g[2] * mm[6 + c];
  };
  R g[3], ux[3], vx[3], wx[3], tx[3];
  grad(z.bv + 1 * z.sz, g);
  chain(g, ux);
  grad(z.bv + 2 * z.sz, g);
  chain(g, vx);
  grad(z.bv + 3 * z.sz, g);
  chain(g, wx);
  grad(z.bv + 5 * z.sz, g);
  chain(g, tx);
  const R mul = R(0.5) * ((R)z.mul[p0] + (R)z.mul[p1]);
  R mut = R(0);
  if constexpr (NT > 0) mut = R(0.5) * ((R)z.mut[p0] + (R)z.mut[p1]);
  const R vis = mul + mut;
  const R v3 = vis * R(1.0 / 3.0);
  R txx = v3 * (R(4) * ux[0] - R(2) * vx[1] - R(2) * wx[2]);
  R tyy = v3 * (R(4) * vx[1] - R(2) * ux[0] - R(2) * wx[2]);
  R tzz = v3 * (R(4) * wx[2] - R(2) * ux[0] - R(2) * vx[1]);
  const R txy = vis * (ux[1] + vx[0]), txz = vis * (ux[2] + wx[0]), tyz = vis * (vx[2] + wx[1]);
  if constexpr (NT > 0) {
    // -2/3 rho k on the normal stresses, faded out within 1e-3 of a wall
    const double *ks = z.sv + (int64_t)NS * z.sz;
    const R dyd = fmin(R(0.5) * ((R)z.wdist[p0] + (R)z.wdist[p1]) / R(1e-3), R(1));
    const R rk = R(-2.0 / 3 * 0.5) * ((R)z.bv[p0] * (R)ks[p0] + (R)z.bv[p1] * (R)ks[p1]) * dyd;
    txx += rk;
    tyy += rk;
    tzz += rk;
  }
  const R J0 = (R)z.jac[p0], J1 = (R)z.jac[p1];
  const int rb = dir * 3;
  const R kx = R(0.5) * ((R)z.metric[(rb + 0) * z.sz + p0] * J0 + (R)z.metric[(rb + 0) * z.sz + p1] * J1);
  const R ky = R(0.5) * ((R)z.metric[(rb + 1) * z.sz + p0] * J0 + (R)z.metric[(rb + 1) * z.sz + p1] * J1);
  const R kz = R(0.5) * ((R)z.metric[(rb + 2) * z.sz + p0] * J0 + (R)z.metric[(rb + 2) * z.sz + p1] * J1);
  fv[0] = R(0);
  fv[1] = kx * txx + ky * txy + kz * txz;
  fv[2] = kx * txy + ky * tyy + kz * tyz;
  fv[3] = kx * txz + ky * tyz + kz * tzz;
  const R um = R(0.5) * ((R)z.bv[1 * z.sz + p0] + (R)z.bv[1 * z.sz + p1]);
  const R vm = R(0.5) * ((R)z.bv[2 * z.sz + p0] + (R)z.bv[2 * z.sz + p1]);
  const R wm = R(0.5) * ((R)z.bv[3 * z.sz + p0] + (R)z.bv[3 * z.sz + p1]);
  R cond = R(0.5) * ((R)z.cond[p0] + (R)z.cond[p1]);
  if constexpr (NT > 0) cond += R(0.5) * ((R)z.tcond[p0] + (R)z.tcond[p1]);
  fv[4] = um * fv[1] + vm * fv[2] + wm * fv[3] + cond * (kx * tx[0] + ky * tx[1] + kz * tx[2]);
  if constexpr (NS > 0) {
    R dif[NS], yk[NS], ddf[NS];
    R sum_gy_w = R(0), sum_rdy = R(0), corr = R(0), mw_tot = R(0);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const double *ys = z.sv + (int64_t)l * z.sz;
      yk[l] = R(0.5) * ((R)ys[p0] + (R)ys[p1]);
      dif[l] = R(0.5) * ((R)z.rhoD[(int64_t)l * z.sz + p0] + (R)z.rhoD[(int64_t)l * z.sz + p1]);
      if constexpr (NT > 0) dif[l] += mut / (R)P.Sct;
      R gy[3], yx[3];
      grad(ys, gy);
      chain(gy, yx);
      const R gg = kx * yx[0] + ky * yx[1] + kz * yx[2];
      ddf[l] = gg;
      corr += dif[l] * gg;
      sum_gy_w += gg / (R)m->mw[l];
      mw_tot += yk[l] / (R)m->mw[l];
      sum_rdy += dif[l] * yk[l];
    }
    mw_tot = R(1.0) / mw_tot;
    corr -= mw_tot * sum_rdy * sum_gy_w;
    if (P.gradP) {
      R gp[3], px[3];
      grad(z.bv + 4 * z.sz, gp);
      chain(gp, px);
      const R gpp = (kx * px[0] + ky * px[1] + kz * px[2]) /
                    (R(0.5) * ((R)z.bv[4 * z.sz + p1] + (R)z.bv[4 * z.sz + p0]));
#pragma unroll
      for (int l = 0; l < NS; ++l) {
        ddf[l] += (mw_tot / (R)m->mw[l] - R(1)) * yk[l] * gpp;
        corr += (mw_tot / (R)m->mw[l] - R(1)) * yk[l] * gpp * dif[l];
      }
    }
    double h[NS];
    d_enthalpy<NS>(0.5 * (z.bv[5 * z.sz + p0] + z.bv[5 * z.sz + p1]), fresh(m), h);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      const R df = dif[l] * (ddf[l] - mw_tot * yk[l] * sum_gy_w) - yk[l] * corr;
      fv[5 + l] = df;
      fv[4] += (R)h[l] * df;
    }
  }
  if constexpr (NT > 0) {
    // SST diffusion of k and omega with sigma blended by F1 at the face
    const double *ks = z.sv + (int64_t)NS * z.sz, *ws = z.sv + (int64_t)(NS + 1) * z.sz;
    R gk[3], kxv[3], go[3], oxv[3];
    grad(ks, gk);
    chain(gk, kxv);
    grad(ws, go);
    chain(go, oxv);
    const R wd = R(0.5) * ((R)z.wdist[p0] + (R)z.wdist[p1]);
    R f1 = R(1);
    if (wd > R(1e-25)) {
      const R km = R(0.5) * ((R)ks[p0] + (R)ks[p1]);
      const R om = R(0.5) * ((R)ws[p0] + (R)ws[p1]);
      const R p1v = sqrt(km) / (R(0.09) * om * wd);
      const R rhom = R(0.5) * ((R)z.bv[p0] + (R)z.bv[p1]);
      const R d2 = wd * wd;
      const R p2v = R(500) * mul / (rhom * d2 * om);
      const R cd = fmax(R(1e-20), R(2) * rhom * R(sst::kSigmaOmega2) / om *
                                      (kxv[0] * oxv[0] + kxv[1] * oxv[1] + kxv[2] * oxv[2]));
      const R p3v = R(4) * rhom * R(sst::kSigmaOmega2) * km / (cd * d2);
      const R a1 = fmin(fmax(p1v, p2v), p3v);
      f1 = tanh(a1 * a1 * a1 * a1);
    }
    const R sk = R(sst::kSigmaK2) + R(sst::kSigmaK1 - sst::kSigmaK2) * f1;
    const R sw = R(sst::kSigmaOmega2) + R(sst::kSigmaOmega1 - sst::kSigmaOmega2) * f1;
    fv[5 + NS] = (mul + mut * sk) * (kx * kxv[0] + ky * kxv[1] + kz * kxv[2]);
    fv[6 + NS] = (mul + mut * sw) * (kx * oxv[0] + ky * oxv[1] + kz * oxv[2]);
  }
}

// ------------------------------------------------------ face assembly
// AWENO.cu:233-242: node flux of the CDS correction at point q (metric row
// rb of the sweep direction), J (rho U, rho U u + kx p, .., U (rho E + p)),
// and (mixtures and transported scalars, the build's definition) J U rho Y_l
// for the components 5 .. NG-1.
template <typename R, int NG = 5>
__device__ inline void node_flux(const ZoneView &z, int rb, int64_t q, R *fn) {
  const R kx = (R)z.metric[(rb + 0) * z.sz + q], ky = (R)z.metric[(rb + 1) * z.sz + q],
          kz = (R)z.metric[(rb + 2) * z.sz + q];
  const R J = (R)z.jac[q];
  const R u = (R)z.bv[1 * z.sz + q], v = (R)z.bv[2 * z.sz + q], w = (R)z.bv[3 * z.sz + q];
  const R Uk = kx * u + ky * v + kz * w;
  const R rU = (R)z.bv[q] * Uk;
  const R p = (R)z.bv[4 * z.sz + q];
  fn[0] = rU * J;
  fn[1] = J * (rU * u + kx * p);
  fn[2] = J * (rU * v + ky * p);
  fn[3] = J * (rU * w + kz * p);
  fn[4] = J * Uk * ((R)z.cv[4 * z.sz + q] + p);
#pragma unroll
  for (int l = 5; l < NG; ++l) fn[l] = J * Uk * (R)z.cv[(int64_t)l * z.sz + q];
}

// components of the CDS flux G: all conservative variables for the mixture /
// scalar AWENO, the 5 flow rows for air, none (one unused slot) without AWENO
template <int NS, int NT, int SCHEME>
struct CdsWidth {
  static constexpr int value = SCHEME == 1 ? 5 + NS + NT : 1;
};

// CDS weights of the flux form G(i+1/2) = sum_s w_s f(i-2+s), s = 0..5
template <typename R>
__device__ inline R cds_weight(int s) {
  const R c1 = R(19.0 / 3840), c2 = R(-13.0 / 320), c3 = R(17.0 / 256);
  return (s == 0 || s == 5) ? c1 : (s == 1 || s == 4) ? c1 + c2 : c1 + c2 + c3;
}

// Riemann solver on the reconstructed states plus the viscous flux:
// F(face) = inviscid - viscous.
template <int NS, int NT, int SCHEME, typename R>
__device__ inline void riemann_viscous(const ZoneView &z, const DParams &P, const DMech *__restrict__ m, int dir,
                                       int64_t p0, int64_t st, const R *pl, const R *pr, const R *mk0, const R *mk1,
                                       R j0, R j1, R *F) {
  constexpr int NV = 5 + NS + NT;
  // inviscid_scheme 2 with AWENO (inviscid_type 1) runs AUSM+, as the
  // reference's AWENO switch has no Roe case (InviscidScheme.cu:290-303)
  switch (P.inviscid_scheme) {
    case 1: lax_friedrichs<NS, NT, R>(pl, pr, mk0, mk1, j0, j1, P.jface != 0, F); break;
    case 4: hllc<NS, NT, R>(m, pl, pr, mk0, mk1, j0, j1, P.jface != 0, F); break;
    case 2:
      if constexpr (SCHEME == 0) {
        roe<NS, NT, R>(m, pl, pr, mk0, mk1, j0, j1, efix_point<R>(z, P, p0), efix_point<R>(z, P, p0 + st), F);
        break;
      }
      [[fallthrough]];
    default: ausm_plus<NS, NT, R>(pl, pr, mk0, mk1, j0, j1, F); break;
  }
  if (P.viscous) {
    R fv[NV];
    viscous_face<NS, NT, R>(z, P, m, dir, p0, st, fv);
#pragma unroll
    for (int l = 0; l < NV; ++l) F[l] -= fv[l];
  }
}

// F(face) = inviscid - viscous; G(face) = CDS flux-form term (AWENO only),
// every stencil value read from memory.  Always inlined: as an outlined call
// (several kernels use it) its arrays and the kernel arguments went to the
// stack.
template <int NS, int NT, int SCHEME, typename R>
__device__ __attribute__((always_inline)) inline void face_flux(const ZoneView &z, const DParams &P, const DMech *__restrict__ m, int dir, int64_t p0,
                                 int64_t st, bool want_g, R *F, R *G) {
  constexpr int NV = 5 + NS + NT;
  constexpr int NG = CdsWidth<NS, NT, SCHEME>::value;
  const int rb = dir * 3;
  R mk0[3], mk1[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    mk0[c] = (R)z.metric[(rb + c) * z.sz + p0];
    mk1[c] = (R)z.metric[(rb + c) * z.sz + p0 + st];
  }
  const R j0 = (R)z.jac[p0], j1 = (R)z.jac[p0 + st];
  R pl[NV + 2], pr[NV + 2];
  if constexpr (SCHEME == 1) {
    if constexpr (NS + NT > 0) {
      aweno_mix<NS, NT, R>(z, m, p0, st, mk0, mk1, pl, pr);
    } else {
      R cvb[6][6];
#pragma unroll
      for (int s = 0; s < 6; ++s) {
        const int64_t q = p0 + (s - 2) * st;
#pragma unroll
        for (int c = 0; c < 5; ++c) cvb[s][c] = (R)z.cv[(int64_t)c * z.sz + q];
        cvb[s][5] = (R)z.bv[4 * z.sz + q];
      }
      aweno_air(cvb, mk0, mk1, pl, pr);
    }
    if (want_g) {
#pragma unroll
      for (int l = 0; l < NG; ++l) G[l] = R(0);
#pragma unroll
      for (int s = 0; s < 6; ++s) {
        R fn[NG];
        node_flux<R, NG>(z, rb, p0 + (s - 2) * st, fn);
#pragma unroll
        for (int l = 0; l < NG; ++l) G[l] += cds_weight<R>(s) * fn[l];
      }
    }
  } else {
    recon_pv<NS, NT, R>(z, P.reconstruction, m, p0, st, pl, pr);
  }
  riemann_viscous<NS, NT, SCHEME, R>(z, P, m, dir, p0, st, pl, pr, mk0, mk1, j0, j1, F);
}

template <int NS, int NT, typename R, int NG = 5>
__device__ inline void add_dq(const ZoneView &z, int64_t q0, bool init, const R *Fc, const R *Fp, bool cds, const R *Gc,
                              const R *Gp) {
  constexpr int NV = 5 + NS + NT;
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    double d = (double)Fc[l] - (double)Fp[l];
    if (l < NG && cds) d += (double)Gc[l] - (double)Gp[l];
    double *dst = z.dq + (int64_t)l * z.sz0 + q0;
    *dst = init ? -d : *dst - d;
  }
}

// Workgroups are handed to the 8 XCDs round-robin by linear id; this maps
// the hardware id to a logical id so that each XCD receives one contiguous
// range of logical blocks (neighbouring lines and planes then share that
// XCD's L2).  A bijection on [0, nb) for any nb.
__device__ inline int xcd_logical_block(int hw, int nb) {
  constexpr int kXcd = 8;
  const int q = nb / kXcd, rem = nb % kXcd;
  const int x = hw % kXcd, t = hw / kXcd;
  return x * q + (x < rem ? x : rem) + t;
}

// i-direction sweep.  The faces of one k-plane are numbered row after row,
// n + 1 per row (face f + 1/2 for f = -1 .. n-1), and a wavefront takes 63
// new faces of that sequence plus the one before them, so no lane idles at
// the end of a row; lane l's cell takes its left face from lane l - 1
// through a cross-lane shuffle.
template <int NS, int NT, int SCHEME, typename R>
__global__ void __launch_bounds__(256) sweep_i(ZoneView z, DParams P, const DMech *__restrict__ m, int init, int bpp) {
  constexpr int NV = 5 + NS + NT;
  const int lane = threadIdx.x;
  const int n = z.mx;
  const int L = xcd_logical_block(blockIdx.x, gridDim.x);
  const int k = L / bpp;
  const unsigned fo = (unsigned)((L - k * bpp) * 4 + threadIdx.y) * 63u + lane;
  const int j = (int)(fo / (unsigned)(n + 1));
  const int i = (int)(fo - (unsigned)j * (n + 1)) - 1;
  const bool live = j < z.my;
  const int ncds = SCHEME == 1 ? n - 2 * z.g : 0;
  constexpr int NG = CdsWidth<NS, NT, SCHEME>::value;
  R F[NV], G[NG];
#pragma unroll
  for (int l = 0; l < NV; ++l) F[l] = R(0);
#pragma unroll
  for (int l = 0; l < NG; ++l) G[l] = R(0);
  if (live) face_flux<NS, NT, SCHEME, R>(z, P, m, 0, z.at(i, j, k), 1, i <= ncds - 1, F, G);
  R Fp[NV], Gp[NG];
#pragma unroll
  for (int l = 0; l < NV; ++l) Fp[l] = __shfl_up(F[l], 1, 64);
#pragma unroll
  for (int l = 0; l < NG; ++l) Gp[l] = __shfl_up(G[l], 1, 64);
  if (live && lane > 0 && i >= 0) add_dq<NS, NT, R, NG>(z, z.at0(i, j, k), init != 0, F, Fp, i < ncds, G, Gp);
}

// AWENO i-sweep (air) with the stencil staged in LDS.  Points of a k-plane
// are numbered row after row, n + 6 per row (i = -3 .. n+2); face f + 1/2
// of row j (number F = j (n+1) + f + 1 in the face sequence) has its six
// stencil points at P .. P + 5, P = F + 5 j.  A workgroup's 253 consecutive
// faces therefore need one contiguous run of points (the rows join without
// a gap), which its threads stage once -- conservative variables, pressure
// and the CDS node flux of each point -- before every face reads its stencil
// and node fluxes from LDS.  Used when rows have at least 63 faces (the run
// then fits kISlots points).
constexpr int kISlots = 288;

template <typename R>
__global__ void __launch_bounds__(256) sweep_i_aweno(ZoneView z, DParams P, const DMech *__restrict__ m, int init,
                                                     int bpp) {
  __shared__ R sQ[6][kISlots];  // cv 0..4, p
  __shared__ R sN[5][kISlots];  // node flux
  const int lane = threadIdx.x, tid = threadIdx.y * 64 + lane;
  const int n = z.mx;
  const unsigned rowf = (unsigned)(n + 1), rowp = (unsigned)(n + 6);
  const int L = xcd_logical_block(blockIdx.x, gridDim.x);
  const int k = L / bpp, b = L - k * bpp;
  const unsigned nf = (unsigned)z.my * rowf;  // faces of the plane
  const unsigned fo = (unsigned)(b * 4 + threadIdx.y) * 63u + lane;
  const int j = (int)(fo / rowf);
  const int i = (int)(fo - (unsigned)j * rowf) - 1;
  const bool live = fo < nf;
  // the workgroup's faces [fo0, fo1] and their points [P0, P0 + S)
  const unsigned fo0 = (unsigned)b * 252u, fo1 = min(fo0 + 252u, nf - 1u);
  const unsigned P0 = fo0 + 5u * (fo0 / rowf);
  const unsigned S = fo1 + 5u * (fo1 / rowf) + 6u - P0;
  for (unsigned q = tid; q < S; q += 256) {
    const unsigned Pq = P0 + q, jq = Pq / rowp;
    const int64_t p = z.at((int)(Pq - jq * rowp) - 3, (int)jq, k);
#pragma unroll
    for (int c = 0; c < 5; ++c) sQ[c][q] = (R)z.cv[(int64_t)c * z.sz + p];
    sQ[5][q] = (R)z.bv[4 * z.sz + p];
    R fn[5];
    node_flux<R>(z, 0, p, fn);
#pragma unroll
    for (int c = 0; c < 5; ++c) sN[c][q] = fn[c];
  }
  __syncthreads();
  const int ncds = n - 2 * z.g;
  R F[5], G[5];
#pragma unroll
  for (int l = 0; l < 5; ++l) F[l] = G[l] = R(0);
  if (live) {
    const unsigned t0 = fo + 5u * (unsigned)j - P0;  // slot of the face's first stencil point
    const int64_t p0 = z.at(i, j, k);
    R mk0[3], mk1[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      mk0[c] = (R)z.metric[c * z.sz + p0];
      mk1[c] = (R)z.metric[c * z.sz + p0 + 1];
    }
    const R j0 = (R)z.jac[p0], j1 = (R)z.jac[p0 + 1];
    R cvb[6][6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
#pragma unroll
      for (int c = 0; c < 6; ++c) cvb[s][c] = sQ[c][t0 + s];
#pragma unroll
      for (int c = 0; c < 5; ++c) G[c] += cds_weight<R>(s) * sN[c][t0 + s];
    }
    R pl[7], pr[7];
    aweno_air(cvb, mk0, mk1, pl, pr);
    riemann_viscous<0, 0, 1, R>(z, P, m, 0, p0, 1, pl, pr, mk0, mk1, j0, j1, F);
  }
  R Fp[5], Gp[5];
#pragma unroll
  for (int l = 0; l < 5; ++l) {
    Fp[l] = __shfl_up(F[l], 1, 64);
    Gp[l] = __shfl_up(G[l], 1, 64);
  }
  if (live && lane > 0 && i >= 0) add_dq<0, 0, R>(z, z.at0(i, j, k), init != 0, F, Fp, i < ncds, G, Gp);
}

// j- and k-direction sweeps: lanes along i (coalesced); a workgroup of kRows
// wavefronts marches along one line segment [c0, c1) of the sweep direction
// in batches of kRows consecutive faces, one face per thread.  A cell's left
// face is the previous row's, read through LDS; the first row of a batch
// reads the last face of the batch before, which that batch's last row
// leaves in one of two carry slots (alternating, so that one barrier
// between the flux stores and the cell updates suffices).  Each face is
// computed once (plus one face per segment).
template <int NS, typename R>
struct JkRows {
  // 8 rows (512 threads, <= 256 VGPRs) for laminar air; 4 with transported
  // scalars, whose face arithmetic needs more registers than a 512-thread
  // block allows
  static constexpr int value = NS > 0 ? 4 : 8;
};

// line-segment decomposition shared by the j/k sweeps
struct JkLine {
  int i, o, c0, c1, nbatch;
  int64_t st;
};

template <int DIR, int kRows>
__device__ inline JkLine jk_line(const ZoneView &z, int gx, int nseg) {
  JkLine L;
  const int b = xcd_logical_block(blockIdx.x, gridDim.x);
  const int bx = b % gx, rest = b / gx;
  const int seg = rest % nseg;
  L.o = rest / nseg;  // k for DIR 1, j for DIR 2
  L.i = bx * 64 + threadIdx.x;
  const int n = DIR == 1 ? z.my : z.mz;
  L.c0 = (int)((int64_t)seg * n / nseg);
  L.c1 = (int)((int64_t)(seg + 1) * n / nseg);
  L.nbatch = (L.c1 - L.c0 + kRows) / kRows;  // faces c0 - 1 .. c1 - 1
  L.st = DIR == 1 ? z.d2 : z.d1;
  return L;
}

template <int DIR, int NS, int NT, int SCHEME, typename R>
__global__ void __launch_bounds__((64 * JkRows<NS + NT, R>::value)) sweep_jk(ZoneView z, DParams P,
                                                                           const DMech *__restrict__ m, int init,
                                                                           int gx, int nseg) {
  constexpr int NV = 5 + NS + NT;
  constexpr int NG = CdsWidth<NS, NT, SCHEME>::value;  // the CDS flux G exists for AWENO only
  constexpr int kRows = JkRows<NS + NT, R>::value;
  __shared__ R sF[kRows + 2][NV][64];  // slots 0, 1: carry; 2 + r: row r
  __shared__ R sG[kRows + 2][NG][64];
  const int lane = threadIdx.x, r = threadIdx.y;
  const JkLine L = jk_line<DIR, kRows>(z, gx, nseg);
  const int n = DIR == 1 ? z.my : z.mz;
  const int ncds = SCHEME == 1 ? n - 2 * z.g : 0;
  for (int b = 0; b < L.nbatch; ++b) {
    const int f = L.c0 - 1 + b * kRows + r;  // this thread's face is f + 1/2
    const bool live = L.i < z.mx && f < L.c1;
    R F[NV], G[NG];
#pragma unroll
    for (int l = 0; l < NV; ++l) F[l] = R(0);
#pragma unroll
    for (int l = 0; l < NG; ++l) G[l] = R(0);
    if (live) {
      const int64_t p0 = DIR == 1 ? z.at(L.i, f, L.o) : z.at(L.i, L.o, f);
      face_flux<NS, NT, SCHEME, R>(z, P, m, DIR, p0, L.st, f <= ncds - 1, F, G);
    }
#pragma unroll
    for (int l = 0; l < NV; ++l) sF[r + 2][l][lane] = F[l];
#pragma unroll
    for (int l = 0; l < NG; ++l) sG[r + 2][l][lane] = G[l];
    if (r == kRows - 1) {
      const int cs = (b + 1) & 1;
#pragma unroll
      for (int l = 0; l < NV; ++l) sF[cs][l][lane] = F[l];
#pragma unroll
      for (int l = 0; l < NG; ++l) sG[cs][l][lane] = G[l];
    }
    __syncthreads();
    if (live && f >= L.c0) {
      const int ls = r > 0 ? r + 1 : (b & 1);
      R Fp[NV], Gp[NG];
#pragma unroll
      for (int l = 0; l < NV; ++l) Fp[l] = sF[ls][l][lane];
#pragma unroll
      for (int l = 0; l < NG; ++l) Gp[l] = sG[ls][l][lane];
      const int64_t q = DIR == 1 ? z.at0(L.i, f, L.o) : z.at0(L.i, L.o, f);
      add_dq<NS, NT, R, NG>(z, q, init != 0, F, Fp, f < ncds, G, Gp);
    }
    __syncthreads();  // the next batch's flux stores overwrite the slots read here
  }
}

// AWENO j/k sweep (air): the rows of the line pass through an LDS ring of
// kRows + 5 rows, each row staged once by one wavefront -- its conservative
// variables and pressure (the characteristic stencil) and its CDS node flux
// -- so a face reads its six stencil rows from LDS instead of six wavefronts
// loading every row from memory.  Face fluxes pass to the next row as in
// sweep_jk.  DT: the cell updates also compute the CFL time step of their
// cell (cell_dt, written to dt_local) and each workgroup leaves the minimum
// over its cells in dt_partial[blockIdx.x] -- the stand-alone time-step pass
// folded into the last sweep of the right-hand side.
template <int DIR, int kRows, typename R, bool DT = false>
__global__ void __launch_bounds__(64 * kRows) sweep_jk_aweno(ZoneView z, DParams P, const DMech *__restrict__ m,
                                                             int init, int gx, int nseg, double *dt_partial) {
  constexpr int kRing = kRows + 5;
  __shared__ R sQ[kRing][6][64];     // cv 0..4, p
  __shared__ R sN[kRing][5][64];     // node flux
  __shared__ R sF[kRows + 2][5][64];  // slots 0, 1: carry; 2 + r: row r
  __shared__ R sG[kRows + 2][5][64];
  const int lane = threadIdx.x, r = threadIdx.y;
  const JkLine L = jk_line<DIR, kRows>(z, gx, nseg);
  const int n = DIR == 1 ? z.my : z.mz;
  const int ncds = n - 2 * z.g;
  const int rb = DIR * 3;
  const bool col = L.i < z.mx;
  double best = 1e+6;  // DT: smallest time step of this thread's cells
  // row q of the line (q >= -3) into ring slot (q + 3) % kRing
  auto stage = [&](int q) {
    const int64_t p = DIR == 1 ? z.at(L.i, q, L.o) : z.at(L.i, L.o, q);
    const int sl = (q + 3) % kRing;
#pragma unroll
    for (int c = 0; c < 5; ++c) sQ[sl][c][lane] = (R)z.cv[(int64_t)c * z.sz + p];
    sQ[sl][5][lane] = (R)z.bv[4 * z.sz + p];
    R fn[5];
    node_flux<R>(z, rb, p, fn);
#pragma unroll
    for (int c = 0; c < 5; ++c) sN[sl][c][lane] = fn[c];
  };
  // the five rows before the first batch's leading rows: c0 - 3 .. c0 + 1
  // (strided over the wavefronts, so any kRows covers them)
  if (col)
    for (int q = r; q < 5; q += kRows) stage(L.c0 - 3 + q);
  for (int b = 0; b < L.nbatch; ++b) {
    const int f = L.c0 - 1 + b * kRows + r;  // this thread's face is f + 1/2
    const bool live = col && f < L.c1;
    if (live) stage(f + 3);
    __syncthreads();
    R F[5], G[5];
#pragma unroll
    for (int l = 0; l < 5; ++l) F[l] = G[l] = R(0);
    if (live) {
      const int64_t p0 = DIR == 1 ? z.at(L.i, f, L.o) : z.at(L.i, L.o, f);
      R mk0[3], mk1[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        mk0[c] = (R)z.metric[(rb + c) * z.sz + p0];
        mk1[c] = (R)z.metric[(rb + c) * z.sz + p0 + L.st];
      }
      const R j0 = (R)z.jac[p0], j1 = (R)z.jac[p0 + L.st];
      R cvb[6][6];
#pragma unroll
      for (int s = 0; s < 6; ++s) {
        const int sl = (f + 1 + s) % kRing;  // row f - 2 + s
#pragma unroll
        for (int c = 0; c < 6; ++c) cvb[s][c] = sQ[sl][c][lane];
#pragma unroll
        for (int c = 0; c < 5; ++c) G[c] += cds_weight<R>(s) * sN[sl][c][lane];
      }
      R pl[7], pr[7];
      aweno_air(cvb, mk0, mk1, pl, pr);
      riemann_viscous<0, 0, 1, R>(z, P, m, DIR, p0, L.st, pl, pr, mk0, mk1, j0, j1, F);
    }
#pragma unroll
    for (int l = 0; l < 5; ++l) {
      sF[r + 2][l][lane] = F[l];
      sG[r + 2][l][lane] = G[l];
    }
    if (r == kRows - 1) {
      const int cs = (b + 1) & 1;
#pragma unroll
      for (int l = 0; l < 5; ++l) {
        sF[cs][l][lane] = F[l];
        sG[cs][l][lane] = G[l];
      }
    }
    __syncthreads();
    if (live && f >= L.c0) {
      const int ls = r > 0 ? r + 1 : (b & 1);
      R Fp[5], Gp[5];
#pragma unroll
      for (int l = 0; l < 5; ++l) {
        Fp[l] = sF[ls][l][lane];
        Gp[l] = sG[ls][l][lane];
      }
      const int64_t q = DIR == 1 ? z.at0(L.i, f, L.o) : z.at0(L.i, L.o, f);
      add_dq<0, 0, R>(z, q, init != 0, F, Fp, f < ncds, G, Gp);
      if constexpr (DT) {
        const double dtl = cell_dt<0>(z, P, DIR == 1 ? z.at(L.i, f, L.o) : z.at(L.i, L.o, f), q);
        z.dt_local[q] = dtl;
        best = fmin(best, dtl);
      }
    }
    // no barrier here: the next batch's staging and flux stores come after
    // its first barrier, which every thread reaches only once done here
    // (ring rows overwritten next were last read before this batch's
    // second barrier)
  }
  if constexpr (DT) {
    __shared__ double sm[kRows];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) best = fmin(best, __shfl_xor(best, o, 64));
    if (lane == 0) sm[r] = best;
    __syncthreads();
    if (r == 0 && lane == 0) {
      double v = sm[0];
#pragma unroll
      for (int w = 1; w < kRows; ++w) v = fmin(v, sm[w]);
      dt_partial[blockIdx.x] = v;
    }
  }
}

}  // namespace

// wavefronts (face rows) per workgroup of the staged AWENO j/k sweep: 4
// rows take 81 KB of LDS in FP64, so two independent workgroups share a CU
// and one computes while the other waits at a barrier (8 rows, one
// 124 KB workgroup per CU, ran the sweeps 1.3-1.4x slower)
constexpr int kAwenoRows = 4;

// Segments per line of the j/k sweeps.  With B workgroups of the kernel
// resident per CU (its occupancy), the sweep takes about
// ceil(workgroups / (B CUs)) rounds of (faces per segment / rows) batches;
// the fewest segments that minimise that product are taken.
static int sweep_segments(int per_cu, int cross, int n, int rows) {
  static int ncu_of[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  if (!ncu_of[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
    ncu_of[dev] = v;
  }
  const int64_t slots = (int64_t)ncu_of[dev] * std::max(per_cu, 1);
  int best = 1;
  int64_t best_cost = INT64_MAX;
  for (int ns = 1; ns <= std::min(n, 4096); ++ns) {
    const int64_t rounds = ((int64_t)cross * ns + slots - 1) / slots;
    const int len = (n + ns - 1) / ns;
    const int64_t cost = rounds * ((len + rows) / rows);
    if (cost < best_cost) {
      best_cost = cost;
      best = ns;
    }
  }
  return best;
}

// dt_partial / dt_out (non-null): the k-sweep of the staged air path
// also computes the stage's CFL time step (sweep_jk_aweno<2, ., ., true>)
// and reduces it into *dt_out; *dt_done tells the caller whether it did.
template <int NS, int NT, int SCHEME, typename R>
static hipError_t launch_sweep_t(const ZoneView &z, const DParams &P, const DMech *m, int dir, bool init,
                                 hipStream_t s, double *dt_partial = nullptr, double *dt_out = nullptr,
                                 int max_partials = 0, bool *dt_done = nullptr) {
  if (dir == 0) {
    const int64_t waves = ((int64_t)z.my * (z.mx + 1) + 62) / 63;
    const int bpp = (int)((waves + 3) / 4);
    const dim3 gr((unsigned)((int64_t)bpp * z.mz));
    // the staged AWENO i-sweep needs rows of at least 63 faces (its point
    // run then fits kISlots: 253 faces + 6 + 5 per row boundary, at most 5)
    if constexpr (SCHEME == 1 && NS == 0 && NT == 0) {
      if (z.mx + 1 >= 64) {
        hipLaunchKernelGGL((sweep_i_aweno<R>), gr, dim3(64, 4, 1), 0, s, z, P, m, init ? 1 : 0, bpp);
        return hipGetLastError();
      }
    }
    hipLaunchKernelGGL((sweep_i<NS, NT, SCHEME, R>), gr, dim3(64, 4, 1), 0, s, z, P, m, init ? 1 : 0, bpp);
    return hipGetLastError();
  }
  constexpr bool kStaged = SCHEME == 1 && NS == 0 && NT == 0;
  constexpr int kRows = kStaged ? kAwenoRows : JkRows<NS + NT, R>::value;
  const int gx = (z.mx + 63) / 64, gy = dir == 1 ? z.mz : z.my, n = dir == 1 ? z.my : z.mz;
  const int cross = gx * gy;
  // workgroups resident per CU (registers and LDS of the kernel), queried
  // once per kernel
  auto occupancy = [](const void *fn) {
    int v = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, fn, 64 * kRows, 0) == hipSuccess && v > 0 ? v : 1;
  };
  auto kernel = [](int d) -> const void * {
    if constexpr (kStaged)
      return d == 1 ? (const void *)sweep_jk_aweno<1, kRows, R> : (const void *)sweep_jk_aweno<2, kRows, R, true>;
    else
      return d == 1 ? (const void *)sweep_jk<1, NS, NT, SCHEME, R> : (const void *)sweep_jk<2, NS, NT, SCHEME, R>;
  };
  static const int per_cu_j = occupancy(kernel(1));
  static const int per_cu_k = occupancy(kernel(2));
  const int nseg = sweep_segments(dir == 1 ? per_cu_j : per_cu_k, cross, n, kRows);
  const dim3 gr((unsigned)((int64_t)cross * nseg));
  const int in = init ? 1 : 0;
  if constexpr (kStaged) {
    const bool with_dt = dir == 2 && dt_partial && dt_out && !z.inv_sr && P.dim == 3 &&
                         (int64_t)gr.x <= (int64_t)max_partials;
    if (dir == 1) {
      hipLaunchKernelGGL((sweep_jk_aweno<1, kRows, R>), gr, dim3(64, kRows, 1), 0, s, z, P, m, in, gx, nseg,
                         nullptr);
    } else if (with_dt) {
      hipLaunchKernelGGL((sweep_jk_aweno<2, kRows, R, true>), gr, dim3(64, kRows, 1), 0, s, z, P, m, in, gx, nseg,
                         dt_partial);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      e = launch_min_of(dt_partial, (int)gr.x, dt_out, s);
      if (e == hipSuccess && dt_done) *dt_done = true;
      return e;
    } else {
      hipLaunchKernelGGL((sweep_jk_aweno<2, kRows, R>), gr, dim3(64, kRows, 1), 0, s, z, P, m, in, gx, nseg,
                         nullptr);
    }
  } else {
    if (dir == 1)
      hipLaunchKernelGGL((sweep_jk<1, NS, NT, SCHEME, R>), gr, dim3(64, kRows, 1), 0, s, z, P, m, in, gx, nseg);
    else
      hipLaunchKernelGGL((sweep_jk<2, NS, NT, SCHEME, R>), gr, dim3(64, kRows, 1), 0, s, z, P, m, in, gx, nseg);
  }
  return hipGetLastError();
}

}  // namespace ostc
