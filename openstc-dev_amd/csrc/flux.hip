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
//  * i-sweep: a wavefront owns 64 consecutive faces (the 63 cells between
//    them); the left face of a cell comes from the neighbouring lane through
//    a cross-lane shuffle.  Loads along i are coalesced.
//  * j- and k-sweeps: lanes run along i (coalesced); the 8 wavefronts of a
//    workgroup own 8 consecutive faces along the sweep direction and pass
//    face fluxes to the next row through LDS (7 cells per 8 faces).
//
// The G form of the CDS term: with node fluxes f, the reference's
//   c1 (f[+3]-f[-3]) + c2 (f[+2]-f[-2]) + c3 (f[+1]-f[-1])
// equals G(i+1/2) - G(i-1/2) for
//   G(i+1/2) = c1 (f[-2]+f[+3]) + (c1+c2) (f[-1]+f[+2]) + (c1+c2+c3) (f[0]+f[+1]).
// The reference applies it to cells [0, n-2*ngg) only (AWENO.cu:234-258);
// that range is kept.
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
// (rho,u,v,w,p,Y...) of the 4 points p0-st .. p0+2st.  Appends rhoE at [nv]
// and gamma at [nv+1].
template <int NS, int NT, typename R>
__device__ inline void recon_pv(const ZoneView &z, const DParams &P, const DMech *__restrict__ m, int64_t p0,
                                int64_t st, R *pl, R *pr) {
  constexpr int NV = 5 + NS + NT;
  const R kappa = R(1.0 / 3.0);
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    const double *src = l < 5 ? z.bv + (int64_t)l * z.sz : z.sv + (int64_t)(l - 5) * z.sz;
    const R q0 = (R)src[p0 - st], q1 = (R)src[p0], q2 = (R)src[p0 + st], q3 = (R)src[p0 + 2 * st];
    if (P.reconstruction == 2) {
      const R di = q1 - q0, di1 = q2 - q1, di2 = q3 - q2;
      const R dnl = minmod(di, di1), dpl = minmod(di1, di), dnr = minmod(di1, di2), dpr = minmod(di2, di1);
      pl[l] = q1 + R(0.25) * ((R(1) - kappa) * dnl + (R(1) + kappa) * dpl);
      pr[l] = q2 - R(0.25) * ((R(1) - kappa) * dpr + (R(1) + kappa) * dnr);
    } else if (P.reconstruction == 3) {
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
    d_enthalpy_cp<NS>(tl, m, h, c);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cpl += c[l] * (double)pl[l + 5];
      cvl += (double)pl[l + 5] * (c[l] - kRuD / m->mw[l]);
      el += h[l] * (double)pl[l + 5];
    }
    d_enthalpy_cp<NS>(tr, m, h, c);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cpr += c[l] * (double)pr[l + 5];
      cvr += (double)pr[l + 5] * (c[l] - kRuD / m->mw[l]);
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

// ------------------------------------------------------ AWENO (air)
// AWENO.cu:151-206 WENO5 on one characteristic row; cvb[s][0..5] =
// (rho, rho u, rho v, rho w, rho E, p) at stencil points s = 0..5 (i-2..i+3).
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
  const R t12 = R(13.0 / 12), eps = R(1e-6);
  const R o16 = R(1.0 / 16), f8 = R(5.0 / 8), f16 = R(5.0 / 16);
  R v0 = R(0.125) * (R(3) * v[0] - R(10) * v[1] + R(15) * v[2]);
  R v1 = R(0.125) * (-v[1] + R(6) * v[2] + R(3) * v[3]);
  R v2 = R(0.125) * (R(3) * v[2] + R(6) * v[3] - v[4]);
  R b0 = t12 * (v[0] + v[2] - R(2) * v[1]) * (v[0] + v[2] - R(2) * v[1]) +
         R(0.25) * (v[0] + R(3) * v[2] - R(4) * v[1]) * (v[0] + R(3) * v[2] - R(4) * v[1]);
  R b1 = t12 * (v[1] + v[3] - R(2) * v[2]) * (v[1] + v[3] - R(2) * v[2]) + R(0.25) * (v[3] - v[1]) * (v[3] - v[1]);
  R b2 = t12 * (v[2] + v[4] - R(2) * v[3]) * (v[2] + v[4] - R(2) * v[3]) +
         R(0.25) * (R(3) * v[2] + v[4] - R(4) * v[3]) * (R(3) * v[2] + v[4] - R(4) * v[3]);
  R tau = (b0 - b2) * (b0 - b2);
  R a0 = o16 + o16 * tau / ((eps + b0) * (eps + b0));
  R a1 = f8 + f8 * tau / ((eps + b1) * (eps + b1));
  R a2 = f16 + f16 * tau / ((eps + b2) * (eps + b2));
  vm = (a0 * v0 + a1 * v1 + a2 * v2) / (a0 + a1 + a2);
  v0 = v1;
  v1 = v2;
  v2 = R(0.125) * (R(15) * v[3] - R(10) * v[4] + R(3) * v[5]);
  b0 = t12 * (v[1] + v[3] - R(2) * v[2]) * (v[1] + v[3] - R(2) * v[2]) +
       R(0.25) * (v[1] + R(3) * v[3] - R(4) * v[2]) * (v[1] + R(3) * v[3] - R(4) * v[2]);
  b1 = t12 * (v[2] + v[4] - R(2) * v[3]) * (v[2] + v[4] - R(2) * v[3]) + R(0.25) * (v[4] - v[2]) * (v[4] - v[2]);
  b2 = t12 * (v[3] + v[5] - R(2) * v[4]) * (v[3] + v[5] - R(2) * v[4]) +
       R(0.25) * (R(3) * v[3] + v[5] - R(4) * v[4]) * (R(3) * v[3] + v[5] - R(4) * v[4]);
  tau = (b0 - b2) * (b0 - b2);
  a0 = f16 + f16 * tau / ((eps + b0) * (eps + b0));
  a1 = f8 + f8 * tau / ((eps + b1) * (eps + b1));
  a2 = o16 + o16 * tau / ((eps + b2) * (eps + b2));
  vp = (a0 * v0 + a1 * v1 + a2 * v2) / (a0 + a1 + a2);
}

// AWENO.cu:17-149 AWENO_interpolation<Air>.  mk0/mk1: metric row of the
// sweep direction at the two points adjacent to the face.
template <typename R>
__device__ inline void aweno_air(const R (*cvb)[6], const R *mk0, const R *mk1, R *pl, R *pr) {
  const R *cl = cvb[2], *cr = cvb[3];
  const R sl = sqrt(cl[0]), sr = sqrt(cr[0]);
  const R rlc = sl / (sl + sr), rrc = sr / (sl + sr);
  const R um = rlc * cl[1] / cl[0] + rrc * cr[1] / cr[0];
  const R vm = rlc * cl[2] / cl[0] + rrc * cr[2] / cr[0];
  const R wm = rlc * cl[3] / cl[0] + rrc * cr[3] / cr[0];
  const R ekm = R(0.5) * (um * um + vm * vm + wm * wm);
  const R gm1 = R(kGam - 1);
  const R hl = (cl[4] + cl[5]) / cl[0], hr = (cr[4] + cr[5]) / cr[0];
  const R hm = rlc * hl + rrc * hr;
  const R cm2 = gm1 * (hm - ekm);
  const R cm = sqrt(cm2);
  R kx = R(0.5) * (mk0[0] + mk1[0]), ky = R(0.5) * (mk0[1] + mk1[1]), kz = R(0.5) * (mk0[2] + mk1[2]);
  const R gK = sqrt(kx * kx + ky * ky + kz * kz);
  kx /= gK;
  ky /= gK;
  kz /= gK;
  const R Uk = kx * um + ky * vm + kz * wm;
  const R alpha = gm1 * ekm;
  R L[25];
  L[0] = (alpha + Uk * cm) / cm2 * R(0.5);
  L[1] = -(gm1 * um + kx * cm) / cm2 * R(0.5);
  L[2] = -(gm1 * vm + ky * cm) / cm2 * R(0.5);
  L[3] = -(gm1 * wm + kz * cm) / cm2 * R(0.5);
  L[4] = gm1 / cm2 * R(0.5);
  L[5] = kx * (R(1) - alpha / cm2) - (kz * vm - ky * wm) / cm;
  L[6] = kx * gm1 * um / cm2;
  L[7] = (kx * gm1 * vm + kz * cm) / cm2;
  L[8] = (kx * gm1 * wm - ky * cm) / cm2;
  L[9] = -kx * gm1 / cm2;
  L[10] = ky * (R(1) - alpha / cm2) - (kx * wm - kz * um) / cm;
  L[11] = (ky * gm1 * um - kz * cm) / cm2;
  L[12] = ky * gm1 * vm / cm2;
  L[13] = (ky * gm1 * wm + kx * cm) / cm2;
  L[14] = -ky * gm1 / cm2;
  L[15] = kz * (R(1) - alpha / cm2) - (ky * um - kx * vm) / cm;
  L[16] = (kz * gm1 * um + ky * cm) / cm2;
  L[17] = (kz * gm1 * vm - kx * cm) / cm2;
  L[18] = kz * gm1 * wm / cm2;
  L[19] = -kz * gm1 / cm2;
  L[20] = (alpha - Uk * cm) / cm2 * R(0.5);
  L[21] = -(gm1 * um - kx * cm) / cm2 * R(0.5);
  L[22] = -(gm1 * vm - ky * cm) / cm2 * R(0.5);
  L[23] = -(gm1 * wm - kz * cm) / cm2 * R(0.5);
  L[24] = gm1 / cm2 * R(0.5);
  R vmn[5], vpl[5];
#pragma unroll
  for (int r = 0; r < 5; ++r) weno5_row(&L[r * 5], cvb, vmn[r], vpl[r]);
  // right eigenvectors (AWENO.cu:97-122), applied row by row
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
                    kx * alpha / gm1 + (kz * vm - ky * wm) * cm,
                    ky * alpha / gm1 + (kx * wm - kz * um) * cm,
                    kz * alpha / gm1 + (ky * um - kx * vm) * cm,
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
  pl[0] = ql[0];
  pl[1] = ql[1] / ql[0];
  pl[2] = ql[2] / ql[0];
  pl[3] = ql[3] / ql[0];
  pl[4] = gm1 * (ql[4] - R(0.5) * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]) * ql[0]);
  pl[5] = ql[4];
  pl[6] = R(kGam);
  pr[0] = qr[0];
  pr[1] = qr[1] / qr[0];
  pr[2] = qr[2] / qr[0];
  pr[3] = qr[3] / qr[0];
  pr[4] = gm1 * (qr[4] - R(0.5) * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]) * qr[0]);
  pr[5] = qr[4];
  pr[6] = R(kGam);
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
  const R ul = (k1 * pl[1] + k2 * pl[2] + k3 * pl[3]) / gk;
  const R ur = (k1 * pr[1] + k2 * pr[2] + k3 * pr[3]) / gk;
  const R c = R(0.5) * (sqrt(pl[NV + 1] * pl[4] / pl[0]) + sqrt(pr[NV + 1] * pr[4] / pr[0]));
  const R ml = ul / c, mr = ur / c;
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
  const R *s = mass >= R(0) ? pl : pr;
  f[0] = coeff;
  f[1] = coeff * s[1] + pc * k1;
  f[2] = coeff * s[2] + pc * k2;
  f[3] = coeff * s[3] + pc * k3;
  f[4] = coeff * (s[NV] + s[4]) / s[0];
#pragma unroll
  for (int l = 5; l < NV; ++l) f[l] = coeff * s[l];
}

// HLLC.cu:10-143
template <int NS, int NT, typename R>
__device__ inline void hllc(const DMech *__restrict__ m, const R *pl, const R *pr, const R *mk0, const R *mk1, R j0,
                            R j1, R *f) {
  constexpr int NV = 5 + NS + NT;
  const R sl0 = sqrt(pl[0]), sr0 = sqrt(pr[0]);
  const R rlc = sl0 / (sl0 + sr0), rrc = sr0 / (sl0 + sr0);
  const R ut = pl[1] * rlc + pr[1] * rrc, vt = pl[2] * rlc + pr[2] * rrc, wt = pl[3] * rlc + pr[3] * rrc;
  const R hl = (pl[NV] + pl[4]) / pl[0], hr = (pr[NV] + pr[4]) / pr[0];
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
    const double t = (double)(rlc * (pl[4] / pl[0]) + rrc * (pr[4] / pr[0])) / (kRuD * mwi);
    double h[NS], c[NS], cpp = 0, cvv = 0;
    d_enthalpy_cp<NS>(t, m, h, c);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cpp += c[l] * svm[l];
      cvv += svm[l] * (c[l] - kRuD / m->mw[l]);
    }
    ct = (R)sqrt(cpp / cvv * kRuD * mwi * t);
  }
  const R kx = R(0.5) * (mk0[0] + mk1[0]), ky = R(0.5) * (mk0[1] + mk1[1]), kz = R(0.5) * (mk0[2] + mk1[2]);
  const R gK = sqrt(kx * kx + ky * ky + kz * kz);
  const R Utb = (kx * ut + ky * vt + kz * wt) / gK;
  const R ja = R(0.5) * (j0 + j1);
  const R Ul = kx * pl[1] + ky * pl[2] + kz * pl[3];
  const R cl = sqrt(pl[NV + 1] * pl[4] / pl[0]);
  const R sl = fmin(Ul / gK - cl, Utb - ct);
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
  const R cr = sqrt(pr[NV + 1] * pr[4] / pr[0]);
  const R sr = fmax(Ur / gK + cr, Utb + ct);
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
  const R sm = ((pr[0] * Ur * (sr - Ur / gK) - pl[0] * Ul * (sl - Ul / gK)) / gK + pl[4] - pr[4]) /
               (pr[0] * (sr - Ur / gK) - pl[0] * (sl - Ul / gK));
  const R pm = pl[0] * (sl - Ul / gK) * (sm - Ul / gK) + pl[4];
  const bool left = sm >= R(0);
  const R *q = left ? pl : pr;
  const R s = left ? sl : sr;
  const R Uq = left ? Ul : Ur;
  const R pC = R(1) / (s - sm);
  const R Q = ja * pC * sm * (s * gK - Uq) * q[0];
  const R dP = (s * pm - sm * q[4]) * pC * ja;
  f[0] = Q;
  f[1] = Q * q[1] + dP * kx;
  f[2] = Q * q[2] + dP * ky;
  f[3] = Q * q[3] + dP * kz;
  f[4] = Q * q[NV] / q[0] + pC * ja * (s * pm * sm * gK - sm * q[4] * Uq);
#pragma unroll
  for (int l = 5; l < NV; ++l) f[l] = Q * q[l];
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
#pragma unroll
  for (int sd = 0; sd < 2; ++sd) {
    const R *q = sd ? pr : pl;
    const R *mk = sd ? mk1 : mk0;
    const R J = sd ? j1 : j0;
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
  }
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
    d_enthalpy_cp<NS>(t, m, hh, cc);
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      cp += cc[l] * (double)svm[l];
      cv += (double)svm[l] * (cc[l] - kRuD / m->mw[l]);
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
__device__ inline void lax_friedrichs(const R *pl, const R *pr, const R *mk0, const R *mk1, R j0, R j1, R *f) {
  constexpr int NV = 5 + NS + NT;
  const R kx = R(0.5) * (mk0[0] + mk1[0]), ky = R(0.5) * (mk0[1] + mk1[1]), kz = R(0.5) * (mk0[2] + mk1[2]);
  const R gK = sqrt(kx * kx + ky * ky + kz * kz);
  const R Ukl = pl[1] * kx + pl[2] * ky + pl[3] * kz, Ukr = pr[1] * kx + pr[2] * ky + pr[3] * kz;
  const R cl = sqrt(R(kGam) * pl[4] / pl[0]), cr = sqrt(R(kGam) * pr[4] / pr[0]);
  const R sr = fmax(fabs(Ukl) + cl * gK, fabs(Ukr) + cr * gK);
  const R hj = R(0.5) * R(0.5) * (j0 + j1);
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
  R txx = vis * (R(4) * ux[0] - R(2) * vx[1] - R(2) * wx[2]) / R(3.0);
  R tyy = vis * (R(4) * vx[1] - R(2) * ux[0] - R(2) * wx[2]) / R(3.0);
  R tzz = vis * (R(4) * wx[2] - R(2) * ux[0] - R(2) * vx[1]) / R(3.0);
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
    d_enthalpy<NS>(0.5 * (z.bv[5 * z.sz + p0] + z.bv[5 * z.sz + p1]), m, h);
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
// F(face) = inviscid - viscous; G(face) = CDS flux-form term (AWENO only).
template <int NS, int NT, int SCHEME, typename R>
__device__ inline void face_flux(const ZoneView &z, const DParams &P, const DMech *__restrict__ m, int dir, int64_t p0,
                                 int64_t st, bool want_g, R *F, R *G) {
  constexpr int NV = 5 + NS + NT;
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
    R cvb[6][6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const int64_t q = p0 + (s - 2) * st;
#pragma unroll
      for (int c = 0; c < 5; ++c) cvb[s][c] = (R)z.cv[(int64_t)c * z.sz + q];
      cvb[s][5] = (R)z.bv[4 * z.sz + q];
    }
    aweno_air(cvb, mk0, mk1, pl, pr);
    if (want_g) {
      // node fluxes (AWENO.cu:233-242) at the 6 stencil points
      const R c1 = R(19.0 / 3840), c2 = R(-13.0 / 320), c3 = R(17.0 / 256);
      const R wgt[6] = {c1, c1 + c2, c1 + c2 + c3, c1 + c2 + c3, c1 + c2, c1};
#pragma unroll
      for (int l = 0; l < 5; ++l) G[l] = R(0);
#pragma unroll
      for (int s = 0; s < 6; ++s) {
        const int64_t q = p0 + (s - 2) * st;
        const R kx = (R)z.metric[(rb + 0) * z.sz + q], ky = (R)z.metric[(rb + 1) * z.sz + q],
                kz = (R)z.metric[(rb + 2) * z.sz + q];
        const R J = (R)z.jac[q];
        const R u = (R)z.bv[1 * z.sz + q], v = (R)z.bv[2 * z.sz + q], w = (R)z.bv[3 * z.sz + q];
        const R Uk = kx * u + ky * v + kz * w;
        const R rU = (R)z.bv[q] * Uk;
        const R p = (R)z.bv[4 * z.sz + q];
        G[0] += wgt[s] * (rU * J);
        G[1] += wgt[s] * (J * (rU * u + kx * p));
        G[2] += wgt[s] * (J * (rU * v + ky * p));
        G[3] += wgt[s] * (J * (rU * w + kz * p));
        G[4] += wgt[s] * (J * Uk * ((R)z.cv[4 * z.sz + q] + p));
      }
    }
  } else {
    recon_pv<NS, NT, R>(z, P, m, p0, st, pl, pr);
  }
  // inviscid_scheme 2 with AWENO (inviscid_type 1) runs AUSM+, as the
  // reference's AWENO switch has no Roe case (InviscidScheme.cu:290-303)
  switch (P.inviscid_scheme) {
    case 1: lax_friedrichs<NS, NT, R>(pl, pr, mk0, mk1, j0, j1, F); break;
    case 4: hllc<NS, NT, R>(m, pl, pr, mk0, mk1, j0, j1, F); break;
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

template <int NS, int NT, typename R>
__device__ inline void add_dq(const ZoneView &z, int64_t q0, bool init, const R *Fc, const R *Fp, bool cds, const R *Gc,
                              const R *Gp) {
  constexpr int NV = 5 + NS + NT;
#pragma unroll
  for (int l = 0; l < NV; ++l) {
    double d = (double)Fc[l] - (double)Fp[l];
    if (l < 5 && cds) d += (double)Gc[l] - (double)Gp[l];
    double *dst = z.dq + (int64_t)l * z.sz0 + q0;
    *dst = init ? -d : *dst - d;
  }
}

// i-direction sweep: lane l of a wavefront owns face (i0 + l - 1) + 1/2.
template <int NS, int NT, int SCHEME, typename R>
__global__ void __launch_bounds__(256) sweep_i(ZoneView z, DParams P, const DMech *__restrict__ m, int init) {
  constexpr int NV = 5 + NS + NT;
  const int lane = threadIdx.x;
  const int i = blockIdx.x * 63 + lane - 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  const int n = z.mx;
  const bool live = j < z.my && i < n;
  const int ncds = SCHEME == 1 ? n - 2 * z.g : 0;
  R F[NV], G[5];
#pragma unroll
  for (int l = 0; l < NV; ++l) F[l] = R(0);
#pragma unroll
  for (int l = 0; l < 5; ++l) G[l] = R(0);
  if (live) face_flux<NS, NT, SCHEME, R>(z, P, m, 0, z.at(i, j, k), 1, i <= ncds - 1, F, G);
  R Fp[NV], Gp[5];
#pragma unroll
  for (int l = 0; l < NV; ++l) Fp[l] = __shfl_up(F[l], 1, 64);
#pragma unroll
  for (int l = 0; l < 5; ++l) Gp[l] = __shfl_up(G[l], 1, 64);
  if (live && lane > 0) add_dq<NS, NT, R>(z, z.at0(i, j, k), init != 0, F, Fp, i < ncds, G, Gp);
}

// j- and k-direction sweeps: lanes along i (coalesced), kRows wavefronts
// per workgroup own kRows consecutive faces along the sweep direction; each
// thread computes one face and the left face of its cell comes from the
// previous row through LDS (one redundant face per kRows - 1 cells).
template <int NS>
struct JkRows {
  // 8 rows (512 threads, <= 256 VGPRs) for laminar air; 4 with transported
  // scalars, whose face arithmetic needs more registers than a 512-thread
  // block allows
  static constexpr int value = NS > 0 ? 4 : 8;
};

template <int DIR, int NS, int NT, int SCHEME, typename R>
__global__ void __launch_bounds__(64 * JkRows<NS + NT>::value) sweep_jk(ZoneView z, DParams P,
                                                                        const DMech *__restrict__ m, int init) {
  constexpr int NV = 5 + NS + NT;
  constexpr int kRows = JkRows<NS + NT>::value;
  __shared__ R sF[kRows][NV][64];
  __shared__ R sG[kRows][5][64];
  const int lane = threadIdx.x, r = threadIdx.y;
  const int i = blockIdx.x * 64 + lane;
  const int o = blockIdx.y;  // k for DIR 1, j for DIR 2
  const int n = DIR == 1 ? z.my : z.mz;
  const int s = blockIdx.z * (kRows - 1) + r - 1;  // this thread's face is s + 1/2
  const bool live = i < z.mx && s < n;
  const int64_t st = DIR == 1 ? z.d2 : z.d1;
  const int ncds = SCHEME == 1 ? n - 2 * z.g : 0;
  R F[NV], G[5];
#pragma unroll
  for (int l = 0; l < NV; ++l) F[l] = R(0);
#pragma unroll
  for (int l = 0; l < 5; ++l) G[l] = R(0);
  if (live) {
    const int64_t p0 = DIR == 1 ? z.at(i, s, o) : z.at(i, o, s);
    face_flux<NS, NT, SCHEME, R>(z, P, m, DIR, p0, st, s <= ncds - 1, F, G);
  }
#pragma unroll
  for (int l = 0; l < NV; ++l) sF[r][l][lane] = F[l];
#pragma unroll
  for (int l = 0; l < 5; ++l) sG[r][l][lane] = G[l];
  __syncthreads();
  if (live && r > 0) {
    R Fp[NV], Gp[5];
#pragma unroll
    for (int l = 0; l < NV; ++l) Fp[l] = sF[r - 1][l][lane];
#pragma unroll
    for (int l = 0; l < 5; ++l) Gp[l] = sG[r - 1][l][lane];
    const int64_t q = DIR == 1 ? z.at0(i, s, o) : z.at0(i, o, s);
    add_dq<NS, NT, R>(z, q, init != 0, F, Fp, s < ncds, G, Gp);
  }
}

}  // namespace

template <int NS, int NT, int SCHEME, typename R>
hipError_t launch_sweep_t(const ZoneView &z, const DParams &P, const DMech *m, int dir, bool init, hipStream_t s) {
  if (dir == 0) {
    const dim3 gr((z.mx + 62) / 63, (z.my + 3) / 4, z.mz);
    hipLaunchKernelGGL((sweep_i<NS, NT, SCHEME, R>), gr, dim3(64, 4, 1), 0, s, z, P, m, init ? 1 : 0);
  } else if (dir == 1) {
    constexpr int kRows = JkRows<NS + NT>::value;
    const dim3 gr((z.mx + 63) / 64, z.mz, (z.my + kRows - 2) / (kRows - 1));
    hipLaunchKernelGGL((sweep_jk<1, NS, NT, SCHEME, R>), gr, dim3(64, kRows, 1), 0, s, z, P, m, init ? 1 : 0);
  } else {
    constexpr int kRows = JkRows<NS + NT>::value;
    const dim3 gr((z.mx + 63) / 64, z.my, (z.mz + kRows - 2) / (kRows - 1));
    hipLaunchKernelGGL((sweep_jk<2, NS, NT, SCHEME, R>), gr, dim3(64, kRows, 1), 0, s, z, P, m, init ? 1 : 0);
  }
  return hipGetLastError();
}

hipError_t launch_flux_sweep(const ZoneView &z, const DParams &P, const DMech *m, int dir, bool init, bool fp32,
                             hipStream_t s) {
  const bool aweno = P.inviscid_type == 1;
  if (P.nt == 0) {
    if (P.ns == 0) {
      if (aweno)
        return fp32 ? launch_sweep_t<0, 0, 1, float>(z, P, m, dir, init, s)
                    : launch_sweep_t<0, 0, 1, double>(z, P, m, dir, init, s);
      return fp32 ? launch_sweep_t<0, 0, 0, float>(z, P, m, dir, init, s)
                  : launch_sweep_t<0, 0, 0, double>(z, P, m, dir, init, s);
    }
    if (P.ns == 9 && !aweno)
      return fp32 ? launch_sweep_t<9, 0, 0, float>(z, P, m, dir, init, s)
                  : launch_sweep_t<9, 0, 0, double>(z, P, m, dir, init, s);
    return hipErrorNotSupported;
  }
  // SST: k and omega are transported after the species; the characteristic
  // AWENO reconstruction does not carry them (AWENO.cu:53-54)
  if (P.nt != 2 || aweno) return hipErrorNotSupported;
  if (P.ns == 0)
    return fp32 ? launch_sweep_t<0, 2, 0, float>(z, P, m, dir, init, s)
                : launch_sweep_t<0, 2, 0, double>(z, P, m, dir, init, s);
  if (P.ns == 9)
    return fp32 ? launch_sweep_t<9, 2, 0, float>(z, P, m, dir, init, s)
                : launch_sweep_t<9, 2, 0, double>(z, P, m, dir, init, s);
  return hipErrorNotSupported;
}

}  // namespace ostc
