/*
 * OpenSTC CPU oracle, part 2: grid metrics, boundary conditions, inviscid
 * (MUSCL/NND/1st + AUSM+/HLLC/LF, AWENO-5 + CDS) and viscous fluxes, the
 * CFL time step, the SSP-RK3 update and the residual.  TEST INFRASTRUCTURE
 * ONLY (see ostc_oracle.h).
 */
#include "ostc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define R_U 8314.462618     /* Constants.h:17 */
#define GAMMA_AIR 1.4       /* Constants.h:20 */
#define MW_AIR 28.9647      /* Constants.h:25 */

typedef struct {
  int g;
  int64_t d2, d1, sz;
} lay;

static lay mklay(int mx, int my, int mz, int g) {
  lay L;
  L.g = g;
  L.d2 = mx + 2 * g;
  L.d1 = (int64_t)(my + 2 * g) * L.d2;
  L.sz = L.d1 * (mz + 2 * g);
  return L;
}

static inline int64_t at(lay L, int i, int j, int k) {
  return (int64_t)(k + L.g) * L.d1 + (int64_t)(j + L.g) * L.d2 + (i + L.g);
}

struct orc_zone {
  int mx, my, mz, g, nv, ns, nt, nsc, dim;
  orc_cfg cfg;
  orc_mech mech;
  lay Lc, Lg, L0;
  double *x, *y, *z, *jac, *metric;
  double *bv, *sv, *cv, *qn, *vel, *acoustic, *mach, *mul, *cond, *rhoD, *gam, *cp;
  double *dq, *dt_local, *bv_last, *chem_jac, *inv_sr, *visc_sr;
  /* SST: turbulent viscosity and conductivity, wall distance (ghosted), source Jacobian diagonal (2 x L0) */
  double *mut, *tcond, *wdist, *turb_jac;
  /* registered boundaries (range after Boundary::register_boundary + wall trim) */
  int nbc;
  int bface[ORC_MAXBC], bdir[ORC_MAXBC], rs[ORC_MAXBC][3], re[ORC_MAXBC][3];
  int order[ORC_MAXBC];
  double dt, t_phys;
  int step;
  double res[4], res_scale[4];
  int inviscid_type;
};

#define M9(zz, p, r, c) ((zz)->metric[(p) * 9 + ((r) - 1) * 3 + ((c) - 1)])

/* ------------------------------------------------------------------ grid */

/* Mesh.cpp:790-858 init_ghost_grid, first part: reflected ghost nodes on
 * the six faces (ngg + 1 layers), 2-D k-ghosts at z0+k. */
static void ghost_faces(int mx, int my, int mz, int ngg, int dim, lay L, double *X, double *Y, double *Z) {
  double *C[3] = {X, Y, Z};
  for (int c = 0; c < 3; ++c) {
    double *A = C[c];
#define A_(i, j, k) A[at(L, i, j, k)]
    for (int k = 0; k < mz; ++k)
      for (int j = 0; j < my; ++j) {
        for (int i = -1; i >= -ngg - 1; --i) A_(i, j, k) = 2 * A_(0, j, k) - A_(-i, j, k);
        for (int i = mx; i < mx + ngg + 1; ++i) A_(i, j, k) = 2 * A_(mx - 1, j, k) - A_(2 * (mx - 1) - i, j, k);
      }
    for (int k = 0; k < mz; ++k)
      for (int i = 0; i < mx; ++i) {
        for (int j = -1; j >= -ngg - 1; --j) A_(i, j, k) = 2 * A_(i, 0, k) - A_(i, -j, k);
        for (int j = my; j < my + ngg + 1; ++j) A_(i, j, k) = 2 * A_(i, my - 1, k) - A_(i, 2 * (my - 1) - j, k);
      }
    if (dim == 3) {
      for (int j = 0; j < my; ++j)
        for (int i = 0; i < mx; ++i) {
          for (int k = -1; k >= -ngg - 1; --k) A_(i, j, k) = 2 * A_(i, j, 0) - A_(i, j, -k);
          for (int k = mz; k < mz + ngg + 1; ++k) A_(i, j, k) = 2 * A_(i, j, mz - 1) - A_(i, j, 2 * (mz - 1) - k);
        }
    } else {
      for (int j = 0; j < my; ++j)
        for (int i = 0; i < mx; ++i) {
          for (int k = -1; k >= -ngg - 1; --k) A_(i, j, k) = A_(i, j, 0) + (c == 2 ? k : 0);
          for (int k = mz; k < mz + ngg + 1; ++k) A_(i, j, k) = A_(i, j, mz - 1) + (c == 2 ? k - mz + 1 : 0);
        }
    }
#undef A_
  }
}

/* Mesh.cpp:861-937 init_ghost_grid, last part: parallelogram edges and
 * corners from the face ghosts (after the inner faces' ghosts were taken
 * from the neighbour block). */
static void ghost_edges(int mx, int my, int mz, int ngg, lay L, double *X, double *Y, double *Z) {
  double *C[3] = {X, Y, Z};
  for (int c = 0; c < 3; ++c) {
    double *A = C[c];
#define A_(i, j, k) A[at(L, i, j, k)]
    for (int i = 1; i <= ngg + 1; ++i)
      for (int j = 1; j <= ngg + 1; ++j)
        for (int k = 0; k < mz; ++k) {
          const int i1 = -i, j1 = -j, i2 = mx - 1 + i, j2 = my - 1 + j;
          A_(i1, j1, k) = A_(i1, 0, k) + A_(0, j1, k) - A_(0, 0, k);
          A_(i2, j2, k) = A_(i2, my - 1, k) + A_(mx - 1, j2, k) - A_(mx - 1, my - 1, k);
          A_(i1, j2, k) = A_(i1, my - 1, k) + A_(0, j2, k) - A_(0, my - 1, k);
          A_(i2, j1, k) = A_(i2, 0, k) + A_(mx - 1, j1, k) - A_(mx - 1, 0, k);
        }
    for (int i = 1; i <= ngg + 1; ++i)
      for (int j = 0; j < my; ++j)
        for (int k = 1; k <= ngg + 1; ++k) {
          const int i1 = -i, k1 = -k, i2 = mx - 1 + i, k2 = mz - 1 + k;
          A_(i1, j, k1) = A_(i1, j, 0) + A_(0, j, k1) - A_(0, j, 0);
          A_(i2, j, k2) = A_(i2, j, mz - 1) + A_(mx - 1, j, k2) - A_(mx - 1, j, mz - 1);
          A_(i1, j, k2) = A_(i1, j, mz - 1) + A_(0, j, k2) - A_(0, j, mz - 1);
          A_(i2, j, k1) = A_(i2, j, 0) + A_(mx - 1, j, k1) - A_(mx - 1, j, 0);
        }
    for (int i = -ngg - 1; i <= mx + ngg; ++i)
      for (int j = 1; j <= ngg + 1; ++j)
        for (int k = 1; k <= ngg + 1; ++k) {
          const int j1 = -j, k1 = -k, j2 = my - 1 + j, k2 = mz - 1 + k;
          A_(i, j1, k1) = A_(i, j1, 0) + A_(i, 0, k1) - A_(i, 0, 0);
          A_(i, j2, k2) = A_(i, j2, mz - 1) + A_(i, my - 1, k2) - A_(i, my - 1, mz - 1);
          A_(i, j1, k2) = A_(i, j1, mz - 1) + A_(i, 0, k2) - A_(i, 0, mz - 1);
          A_(i, j2, k1) = A_(i, j2, 0) + A_(i, my - 1, k1) - A_(i, my - 1, 0);
        }
#undef A_
  }
}

/* Mesh.cpp:790-944 init_ghost_grid of a single block (no inner faces) */
static void ghost_grid(int mx, int my, int mz, int ngg, int dim, lay L, double *X, double *Y, double *Z) {
  ghost_faces(mx, my, mz, ngg, dim, L, X, Y, Z);
  ghost_edges(mx, my, mz, ngg, L, X, Y, Z);
}

/* one point of Mesh.cpp:259-521: central differences of x,y,z; J and the
 * metrics d(xi_r)/d(x_c).  Returns 0 when J <= 0. */
static int metric_point(lay Lc, const double *X, const double *Y, const double *Z, int i, int j, int k, double *jac,
                        double *m) {
#define P(A, a, b, c) A[at(Lc, a, b, c)]
  const double dxd1 = (P(X, i + 1, j, k) - P(X, i - 1, j, k)) * 0.5, dyd1 = (P(Y, i + 1, j, k) - P(Y, i - 1, j, k)) * 0.5,
               dzd1 = (P(Z, i + 1, j, k) - P(Z, i - 1, j, k)) * 0.5;
  const double dxd2 = (P(X, i, j + 1, k) - P(X, i, j - 1, k)) * 0.5, dyd2 = (P(Y, i, j + 1, k) - P(Y, i, j - 1, k)) * 0.5,
               dzd2 = (P(Z, i, j + 1, k) - P(Z, i, j - 1, k)) * 0.5;
  const double dxd3 = (P(X, i, j, k + 1) - P(X, i, j, k - 1)) * 0.5, dyd3 = (P(Y, i, j, k + 1) - P(Y, i, j, k - 1)) * 0.5,
               dzd3 = (P(Z, i, j, k + 1) - P(Z, i, j, k - 1)) * 0.5;
#undef P
  const double J = dxd1 * dyd2 * dzd3 + dxd2 * dyd3 * dzd1 + dxd3 * dyd1 * dzd2 - dxd1 * dyd3 * dzd2 -
                   dxd2 * dyd1 * dzd3 - dxd3 * dyd2 * dzd1;
  if (J <= 0) return 0;
  *jac = J;
  m[0] = (dyd2 * dzd3 - dyd3 * dzd2) / J;
  m[1] = (dxd3 * dzd2 - dxd2 * dzd3) / J;
  m[2] = (dxd2 * dyd3 - dxd3 * dyd2) / J;
  m[3] = (dyd3 * dzd1 - dyd1 * dzd3) / J;
  m[4] = (dxd1 * dzd3 - dxd3 * dzd1) / J;
  m[5] = (dxd3 * dyd1 - dxd1 * dyd3) / J;
  m[6] = (dyd1 * dzd2 - dyd2 * dzd1) / J;
  m[7] = (dxd2 * dzd1 - dxd1 * dzd2) / J;
  m[8] = (dxd1 * dyd2 - dxd2 * dyd1) / J;
  return 1;
}

/* Mesh.cpp:259-521 compute_jac_metric: interior, then the six ghost slabs in
 * the reference's order; a non-positive J copies the inner neighbour. */
static int metrics(int mx, int my, int mz, int ngg, lay Lc, lay L, const double *X, const double *Y, const double *Z,
                   double *jac, double *met) {
#define DO(i, j, k, di, dj, dk)                                                              \
  do {                                                                                       \
    const int64_t p_ = at(L, i, j, k);                                                       \
    if (!metric_point(Lc, X, Y, Z, i, j, k, &jac[p_], &met[p_ * 9])) {                       \
      const int64_t q_ = at(L, (i) + (di), (j) + (dj), (k) + (dk));                          \
      jac[p_] = jac[q_];                                                                     \
      memcpy(&met[p_ * 9], &met[q_ * 9], 9 * sizeof(double));                                \
    }                                                                                        \
  } while (0)
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j) {
      int kmin = 1, kmax = mz - 1;
      if (mz == 1) {
        kmin = 0;
        kmax = 1;
      }
      for (int k = kmin; k < kmax; ++k) {
        const int64_t p = at(L, i, j, k);
        if (!metric_point(Lc, X, Y, Z, i, j, k, &jac[p], &met[p * 9])) return -1;
      }
    }
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j)
      for (int k = 0; k >= -ngg; --k) DO(i, j, k, 0, 0, 1);
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j)
      for (int k = mz - 1; k < mz + ngg; ++k) DO(i, j, k, 0, 0, -1);
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 0; j >= -ngg; --j)
      for (int k = -ngg; k < mz + ngg; ++k) DO(i, j, k, 0, 1, 0);
  for (int i = 1; i < mx - 1; ++i)
    for (int j = my - 1; j < my + ngg; ++j)
      for (int k = -ngg; k < mz + ngg; ++k) DO(i, j, k, 0, -1, 0);
  for (int i = 0; i >= -ngg; --i)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int k = -ngg; k < mz + ngg; ++k) DO(i, j, k, 1, 0, 0);
  for (int i = mx - 1; i < mx + ngg; ++i)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int k = -ngg; k < mz + ngg; ++k) DO(i, j, k, -1, 0, 0);
#undef DO
  return 0;
}

/* Conservative symmetric metrics (build option conservative_metric; the
 * reference's metrics are the non-conservative differences above and do not
 * preserve a uniform flow on curved grids).  Row r of J d(xi_r)/d(x_c):
 *   J xi_c   = 1/2 [D_zeta(A^c_eta) - D_eta(A^c_zeta)]
 *   J eta_c  = 1/2 [D_xi(A^c_zeta)  - D_zeta(A^c_xi)]
 *   J zeta_c = 1/2 [D_eta(A^c_xi)   - D_xi(A^c_eta)]
 * with A^x_d = y_d z - y z_d, A^y_d = z_d x - z x_d, A^z_d = x_d y - x y_d
 * (Thomas & Lombard 1979; Vinokur & Yee 2002).  D_d, also for the inner
 * derivatives x_d, is the difference the flux scheme applies along d at that
 * index: the face average of J * metric differenced, (f[+1] - f[-1]) / 2,
 * plus, for AWENO-5 on cells [0, n - 2 ngg) where the CDS correction acts,
 * c3 (f[+1] - f[-1]) + c2 (f[+2] - f[-2]) + c1 (f[+3] - f[-3]).  Operators
 * along different directions commute, so the discrete metric identities
 * sum_r D_r(J xi_r,c) = 0 hold to rounding and a uniform flow keeps a zero
 * right-hand side.  J stays the determinant of metric_point; the stored
 * metric is the row divided by it.  A row whose stencil leaves the node
 * array keeps the reference row (such nodes are never read by the fluxes). */
static const double kFspW1 = 145.0 / 256, kFspW2 = -13.0 / 320, kFspW3 = 19.0 / 3840;

/* D_d of F at (i, j, k) on the (ngg + 1)-ghosted node layout; NaN when the
 * stencil leaves it */
static double fsp_diff(lay Lc, const double *F, int d, int i, int j, int k, int n, int ngg, int aweno) {
  const int c = d == 0 ? i : (d == 1 ? j : k);
  const int wide = aweno && c >= 0 && c < n - 2 * ngg;
  const int w = wide ? 3 : 1;
  if (c - w < -(ngg + 1) || c + w > n + ngg) return NAN;
  const int di = d == 0, dj = d == 1, dk = d == 2;
#define F_(s) F[at(Lc, i + (s) * di, j + (s) * dj, k + (s) * dk)]
  if (!wide) return (F_(1) - F_(-1)) * 0.5;
  return (F_(1) - F_(-1)) * kFspW1 + (F_(2) - F_(-2)) * kFspW2 + (F_(3) - F_(-3)) * kFspW3;
#undef F_
}

static void fsp_metrics(int mx, int my, int mz, int ngg, int aweno, lay Lc, lay L, const double *X, const double *Y,
                        const double *Z, const double *jac, double *met) {
  const int n[3] = {mx, my, mz};
  const double *C[3] = {X, Y, Z};
  /* A[c * 3 + d] on the node layout */
  double *A[9];
  for (int q = 0; q < 9; ++q) A[q] = (double *)malloc(sizeof(double) * Lc.sz);
  for (int k = -ngg - 1; k <= mz + ngg; ++k)
    for (int j = -ngg - 1; j <= my + ngg; ++j)
      for (int i = -ngg - 1; i <= mx + ngg; ++i) {
        const int64_t p = at(Lc, i, j, k);
        for (int d = 0; d < 3; ++d) {
          double der[3];
          for (int c = 0; c < 3; ++c) der[c] = fsp_diff(Lc, C[c], d, i, j, k, n[d], ngg, aweno);
          /* column c uses the two other coordinates a = c+1, b = c+2 (mod 3): A = a_d b - a b_d */
          for (int c = 0; c < 3; ++c) {
            const int a = (c + 1) % 3, b = (c + 2) % 3;
            A[c * 3 + d][p] = der[a] * C[b][p] - C[a][p] * der[b];
          }
        }
      }
  for (int k = -ngg; k < mz + ngg; ++k)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int i = -ngg; i < mx + ngg; ++i) {
        const int64_t p = at(L, i, j, k);
        for (int r = 0; r < 3; ++r) {
          /* J xi_r,c = 1/2 [D_{r+2}(A^c_{r+1}) - D_{r+1}(A^c_{r+2})] */
          const int d1 = (r + 1) % 3, d2 = (r + 2) % 3;
          double row[3];
          int ok = 1;
          for (int c = 0; c < 3; ++c) {
            row[c] = (fsp_diff(Lc, A[c * 3 + d1], d2, i, j, k, n[d2], ngg, aweno) -
                      fsp_diff(Lc, A[c * 3 + d2], d1, i, j, k, n[d1], ngg, aweno)) *
                     0.5;
            if (isnan(row[c])) ok = 0;
          }
          if (!ok) continue;
          for (int c = 0; c < 3; ++c) met[p * 9 + r * 3 + c] = row[c] / jac[p];
        }
      }
  for (int q = 0; q < 9; ++q) free(A[q]);
}

void orc_grid_metrics(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z, double *xg,
                      double *yg, double *zg, double *jac, double *metric) {
  lay Lc = mklay(mx, my, mz, ngg + 1), L = mklay(mx, my, mz, ngg);
  const int dim = mz == 1 ? 2 : 3;
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        const int64_t s = ((int64_t)k * my + j) * mx + i;
        xg[at(Lc, i, j, k)] = x[s];
        yg[at(Lc, i, j, k)] = y[s];
        zg[at(Lc, i, j, k)] = z[s];
      }
  ghost_grid(mx, my, mz, ngg, dim, Lc, xg, yg, zg);
  metrics(mx, my, mz, ngg, Lc, L, xg, yg, zg, jac, metric);
}

/* --------------------------------------------------------------- zone */

static void *cal(int64_t n) { return calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

/* WallDistance.cuh:9-81 + Driver.cu:100-121: the wall points are the nodes
 * of every wall boundary's registered (trimmed) range; every ghosted point
 * gets the distance to the nearest of them by brute force (squared distances
 * minimised from 1e6, square root at the end). */
static void wall_distance(orc_zone *zz) {
  int64_t nw = 0;
  for (int b = 0; b < zz->cfg.n_bc; ++b)
    if (zz->cfg.bc[b].type == 2)
      nw += (int64_t)(zz->re[b][0] - zz->rs[b][0] + 1) * (zz->re[b][1] - zz->rs[b][1] + 1) *
            (zz->re[b][2] - zz->rs[b][2] + 1);
  double *wp = malloc(sizeof(double) * 3 * (nw > 0 ? nw : 1));
  int64_t c = 0;
  for (int b = 0; b < zz->cfg.n_bc; ++b) {
    if (zz->cfg.bc[b].type != 2) continue;
    for (int k = zz->rs[b][2]; k <= zz->re[b][2]; ++k)
      for (int j = zz->rs[b][1]; j <= zz->re[b][1]; ++j)
        for (int i = zz->rs[b][0]; i <= zz->re[b][0]; ++i) {
          const int64_t p = at(zz->Lc, i, j, k);
          wp[3 * c] = zz->x[p];
          wp[3 * c + 1] = zz->y[p];
          wp[3 * c + 2] = zz->z[p];
          ++c;
        }
  }
  const int g = zz->g;
  for (int k = -g; k < zz->mz + g; ++k)
    for (int j = -g; j < zz->my + g; ++j)
      for (int i = -g; i < zz->mx + g; ++i) {
        const int64_t p = at(zz->Lc, i, j, k);
        const double x = zz->x[p], y = zz->y[p], z = zz->z[p];
        double d2 = 1e+6;
        for (int64_t l = 0; l < nw; ++l) {
          const double d = (x - wp[3 * l]) * (x - wp[3 * l]) + (y - wp[3 * l + 1]) * (y - wp[3 * l + 1]) +
                           (z - wp[3 * l + 2]) * (z - wp[3 * l + 2]);
          if (d2 > d) d2 = d;
        }
        zz->wdist[at(zz->Lg, i, j, k)] = sqrt(d2);
      }
  free(wp);
}

static void zone_register(orc_zone *zz);

/* x == NULL leaves the grid, metrics and boundaries to the caller (the
 * multi-block mesh sets them after the inner faces are known) */
orc_zone *orc_zone_create(int mx, int my, int mz, int ngg, const orc_cfg *cfg, const orc_mech *mech, const double *x,
                          const double *y, const double *z) {
  orc_zone *zz = (orc_zone *)calloc(1, sizeof(orc_zone));
  zz->mx = mx;
  zz->my = my;
  zz->mz = mz;
  zz->g = ngg;
  zz->dim = mz == 1 ? 2 : 3;
  zz->cfg = *cfg;
  if (mech) zz->mech = *mech;
  zz->ns = cfg->mix ? zz->mech.n_spec : 0;
  zz->nt = cfg->turbulence ? 2 : 0;
  zz->nsc = zz->ns + zz->nt;
  zz->nv = 5 + zz->nsc;
  zz->Lc = mklay(mx, my, mz, ngg + 1);
  zz->Lg = mklay(mx, my, mz, ngg);
  zz->L0 = mklay(mx, my, mz, 0);
  const int64_t sg = zz->Lg.sz, s0 = zz->L0.sz, sc = zz->Lc.sz;
  zz->x = cal(sc);
  zz->y = cal(sc);
  zz->z = cal(sc);
  zz->jac = cal(sg);
  zz->metric = cal(sg * 9);
  zz->bv = cal(sg * 6);
  zz->sv = cal(sg * (zz->nsc ? zz->nsc : 1));
  zz->cv = cal(sg * zz->nv);
  zz->qn = cal(sg * zz->nv);
  zz->vel = cal(sg);
  zz->acoustic = cal(sg);
  zz->mach = cal(sg);
  zz->mul = cal(sg);
  zz->cond = cal(sg);
  zz->rhoD = cal(sg * (zz->ns ? zz->ns : 1));
  zz->gam = cal(sg);
  zz->cp = cal(sg);
  zz->dq = cal(s0 * zz->nv);
  zz->dt_local = cal(s0);
  zz->bv_last = cal(s0 * 4);
  zz->chem_jac = cal(s0 * (zz->ns ? zz->ns * zz->ns : 1));
  zz->inv_sr = cal(s0 * 3);
  zz->visc_sr = cal(s0);
  zz->mut = cal(sg);
  zz->tcond = cal(sg);
  zz->wdist = cal(sg);
  zz->turb_jac = cal(s0 * 2);
  if (x) {
    orc_grid_metrics(mx, my, mz, ngg, x, y, z, zz->x, zz->y, zz->z, zz->jac, zz->metric);
    zone_register(zz);
  }
  return zz;
}

/* boundaries and the SST wall distance of a zone whose grid is set */
static void zone_register(orc_zone *zz) {
  const orc_cfg *cfg = &zz->cfg;
  const int ngg = zz->g;
  /* inviscid_type: Parameter.cpp:165-219 deduce_known_info */
  zz->inviscid_type = 0;
  if (cfg->inviscid_scheme == 2) zz->inviscid_type = 2;
  if (cfg->reconstruction == 4 || cfg->reconstruction == 5) zz->inviscid_type = 1;
  if (cfg->conservative_metric)
    fsp_metrics(zz->mx, zz->my, zz->mz, ngg, zz->inviscid_type == 1, zz->Lc, zz->Lg, zz->x, zz->y, zz->z, zz->jac,
                zz->metric);
  /* boundaries: Mesh.cpp:14-31 register_boundary, BoundCond.cu:369-384 wall trim */
  zz->nbc = cfg->n_bc;
  for (int b = 0; b < cfg->n_bc; ++b) {
    const orc_bc *B = &cfg->bc[b];
    int s[3] = {B->raw[0], B->raw[2], B->raw[4]}, e[3] = {B->raw[1], B->raw[3], B->raw[5]};
    int face = 0, dir = 1;
    for (int d = 0; d < zz->dim; ++d)
      if (s[d] == e[d]) {
        face = d;
        if (s[d] == 0) dir = -1;
        break;
      }
    for (int d = 0; d < zz->dim; ++d) {
      if (d == face) continue;
      if (e[d] < s[d]) {
        int t = e[d];
        e[d] = s[d];
        s[d] = t;
      }
      s[d] -= ngg;
      e[d] += ngg;
    }
    if (zz->dim == 2) {
      s[2] = -ngg;
      e[2] = ngg;
    }
    if (B->type == 2)
      for (int q = 0; q < 3; ++q) {
        if (q == face) continue;
        s[q] += ngg;
        e[q] -= ngg;
      }
    zz->bface[b] = face;
    zz->bdir[b] = dir;
    for (int d = 0; d < 3; ++d) {
      zz->rs[b][d] = s[d];
      zz->re[b][d] = e[d];
    }
  }
  /* application order: BoundCond.cuh:673-822 (farfield, outflow, back pressure,
   * inflow, subsonic inflow, symmetry, wall; file order within one type) */
  const int prio[7] = {4, 6, 9, 5, 7, 3, 2};
  int n = 0;
  for (int p = 0; p < 7; ++p)
    for (int b = 0; b < cfg->n_bc; ++b)
      if (cfg->bc[b].type == prio[p]) zz->order[n++] = b;
  zz->nbc = n;
  if (zz->nt) wall_distance(zz);
}

void orc_zone_destroy(orc_zone *zz) {
  if (!zz) return;
  double *ptrs[] = {zz->x, zz->y, zz->z, zz->jac, zz->metric, zz->bv, zz->sv, zz->cv, zz->qn, zz->vel, zz->acoustic,
                    zz->mach, zz->mul, zz->cond, zz->rhoD, zz->gam, zz->cp, zz->dq, zz->dt_local, zz->bv_last,
                    zz->chem_jac, zz->inv_sr, zz->visc_sr, zz->mut, zz->tcond, zz->wdist, zz->turb_jac};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
  free(zz);
}

double *orc_zone_field(orc_zone *zz, int id, int64_t *count) {
  const int64_t sg = zz->Lg.sz, s0 = zz->L0.sz, sc = zz->Lc.sz;
  switch (id) {
    case ORC_X: *count = sc; return zz->x;
    case ORC_Y: *count = sc; return zz->y;
    case ORC_Z: *count = sc; return zz->z;
    case ORC_JAC: *count = sg; return zz->jac;
    case ORC_METRIC: *count = sg * 9; return zz->metric;
    case ORC_BV: *count = sg * 6; return zz->bv;
    case ORC_SV: *count = sg * zz->nsc; return zz->sv;
    case ORC_CV: *count = sg * zz->nv; return zz->cv;
    case ORC_DQ: *count = s0 * zz->nv; return zz->dq;
    case ORC_VEL: *count = sg; return zz->vel;
    case ORC_ACOUSTIC: *count = sg; return zz->acoustic;
    case ORC_MACH: *count = sg; return zz->mach;
    case ORC_MUL: *count = sg; return zz->mul;
    case ORC_COND: *count = sg; return zz->cond;
    case ORC_RHOD: *count = sg * zz->ns; return zz->rhoD;
    case ORC_GAMMA: *count = sg; return zz->gam;
    case ORC_CP: *count = sg; return zz->cp;
    case ORC_DT_LOCAL: *count = s0; return zz->dt_local;
    case ORC_QN: *count = sg * zz->nv; return zz->qn;
    case ORC_BV_LAST: *count = s0 * 4; return zz->bv_last;
    case ORC_CHEM_JAC: *count = s0 * zz->ns * zz->ns; return zz->chem_jac;
    case ORC_MUT: *count = sg; return zz->mut;
    case ORC_TURB_COND: *count = sg; return zz->tcond;
    case ORC_WALL_DIST: *count = sg; return zz->wdist;
    case ORC_TURB_JAC: *count = s0 * 2; return zz->turb_jac;
  }
  *count = 0;
  return NULL;
}

void orc_zone_state(orc_zone *zz, double *t_phys, int *step) {
  *t_phys = zz->t_phys;
  *step = zz->step;
}

#define BV(l, i, j, k) zz->bv[(l) * zz->Lg.sz + at(zz->Lg, i, j, k)]
#define SV(l, i, j, k) zz->sv[(l) * zz->Lg.sz + at(zz->Lg, i, j, k)]
#define CV(l, i, j, k) zz->cv[(l) * zz->Lg.sz + at(zz->Lg, i, j, k)]
#define QN(l, i, j, k) zz->qn[(l) * zz->Lg.sz + at(zz->Lg, i, j, k)]
#define DQ(l, i, j, k) zz->dq[(l) * zz->L0.sz + at(zz->L0, i, j, k)]
#define S1(a, i, j, k) zz->a[at(zz->Lg, i, j, k)]

/* FieldOperation.cuh:18-41 compute_total_energy (also sets |V|) */
static void total_energy(orc_zone *zz, int i, int j, int k) {
  const double u = BV(1, i, j, k), v = BV(2, i, j, k), w = BV(3, i, j, k);
  const double V2 = u * u + v * v + w * w;
  double te = 0.5 * V2;
  if (zz->cfg.mix) {
    double h[ORC_MAXS];
    orc_enthalpy(BV(5, i, j, k), &zz->mech, h);
    for (int l = 0; l < zz->ns; ++l) te += h[l] * SV(l, i, j, k);
    te *= BV(0, i, j, k);
    te -= BV(4, i, j, k);
  } else {
    te *= BV(0, i, j, k);
    te += BV(4, i, j, k) / (GAMMA_AIR - 1);
  }
  CV(4, i, j, k) = te;
  S1(vel, i, j, k) = sqrt(V2);
}

/* FieldOperation.cuh:83-110 compute_cv_from_bv_1_point */
static void cv_from_bv_point(orc_zone *zz, int i, int j, int k) {
  const double rho = BV(0, i, j, k);
  CV(0, i, j, k) = rho;
  CV(1, i, j, k) = rho * BV(1, i, j, k);
  CV(2, i, j, k) = rho * BV(2, i, j, k);
  CV(3, i, j, k) = rho * BV(3, i, j, k);
  for (int l = 0; l < zz->nsc; ++l) CV(5 + l, i, j, k) = rho * SV(l, i, j, k);
  total_energy(zz, i, j, k);
}

/* ------------------------------------------------------------ boundaries */

/* BoundCond.cuh:52-118 apply_symmetry (never refreshes ghost cv) */
static void bc_symmetry(orc_zone *zz, int b, int i, int j, int k) {
  const int face = zz->bface[b];
  int dir[3] = {0, 0, 0};
  dir[face] = zz->bdir[b];
  const int ii0 = i - dir[0], ij0 = j - dir[1], ik0 = k - dir[2];
  const double *m = &zz->metric[at(zz->Lg, i, j, k) * 9];
  double kx = m[face * 3 + 0], ky = m[face * 3 + 1], kz = m[face * 3 + 2];
  const double km = sqrt(kx * kx + ky * ky + kz * kz);
  kx /= km;
  ky /= km;
  kz /= km;
  const double u1 = BV(1, ii0, ij0, ik0), v1 = BV(2, ii0, ij0, ik0), w1 = BV(3, ii0, ij0, ik0);
  double u_k = kx * u1 + ky * v1 + kz * w1;
  const double ut = u1 - kx * u_k, vt = v1 - ky * u_k, wt = w1 - kz * u_k;
  BV(1, i, j, k) = ut;
  BV(2, i, j, k) = vt;
  BV(3, i, j, k) = wt;
  S1(vel, i, j, k) = sqrt(ut * ut + vt * vt + wt * wt);
  BV(0, i, j, k) = BV(0, ii0, ij0, ik0);
  BV(4, i, j, k) = BV(4, ii0, ij0, ik0);
  BV(5, i, j, k) = BV(5, ii0, ij0, ik0);
  for (int l = 0; l < zz->nsc; ++l) SV(l, i, j, k) = SV(l, ii0, ij0, ik0);
  for (int g = 1; g <= zz->g; ++g) {
    const int gi = i + g * dir[0], gj = j + g * dir[1], gk = k + g * dir[2];
    const int ii = i - g * dir[0], ij = j - g * dir[1], ik = k - g * dir[2];
    BV(0, gi, gj, gk) = BV(0, ii, ij, ik);
    const double u = BV(1, ii, ij, ik), v = BV(2, ii, ij, ik), w = BV(3, ii, ij, ik);
    u_k = kx * u + ky * v + kz * w;
    BV(1, gi, gj, gk) = u - 2 * u_k * kx;
    BV(2, gi, gj, gk) = v - 2 * u_k * ky;
    BV(3, gi, gj, gk) = w - 2 * u_k * kz;
    S1(vel, gi, gj, gk) = sqrt(BV(1, gi, gj, gk) * BV(1, gi, gj, gk) + BV(2, gi, gj, gk) * BV(2, gi, gj, gk) +
                               BV(3, gi, gj, gk) * BV(3, gi, gj, gk));
    BV(4, gi, gj, gk) = BV(4, ii, ij, ik);
    BV(5, gi, gj, gk) = BV(5, ii, ij, ik);
    for (int l = 0; l < zz->nsc; ++l) SV(l, gi, gj, gk) = SV(l, ii, ij, ik);
    if (zz->nt) S1(mut, gi, gj, gk) = S1(mut, ii, ij, ik);
  }
}

/* BoundCond.cuh:120-148 apply_outflow */
static void bc_outflow(orc_zone *zz, int b, int i, int j, int k, int with_cv) {
  int dir[3] = {0, 0, 0};
  dir[zz->bface[b]] = zz->bdir[b];
  for (int g = 1; g <= zz->g; ++g) {
    const int gi = i + g * dir[0], gj = j + g * dir[1], gk = k + g * dir[2];
    for (int l = 0; l < 6; ++l) BV(l, gi, gj, gk) = BV(l, i, j, k);
    for (int l = 0; l < zz->nsc; ++l) SV(l, gi, gj, gk) = SV(l, i, j, k);
    if (zz->nt) S1(mut, gi, gj, gk) = S1(mut, i, j, k);
    if (with_cv) cv_from_bv_point(zz, gi, gj, gk);
  }
}

/* BoundCond.cuh:150-212 apply_inflow */
static void bc_inflow(orc_zone *zz, const orc_bc *B, int b, int i, int j, int k, int with_cv) {
  int dir[3] = {0, 0, 0};
  dir[zz->bface[b]] = zz->bdir[b];
  BV(0, i, j, k) = B->rho;
  BV(1, i, j, k) = B->u;
  BV(2, i, j, k) = B->v;
  BV(3, i, j, k) = B->w;
  BV(4, i, j, k) = B->p;
  BV(5, i, j, k) = B->t;
  for (int l = 0; l < zz->nsc; ++l) SV(l, i, j, k) = B->sv[l];
  if (zz->nt) S1(mut, i, j, k) = B->mut;
  S1(vel, i, j, k) = B->vel;
  if (with_cv) cv_from_bv_point(zz, i, j, k);
  for (int g = 1; g <= zz->g; ++g) {
    const int gi = i + g * dir[0], gj = j + g * dir[1], gk = k + g * dir[2];
    BV(0, gi, gj, gk) = B->rho;
    BV(1, gi, gj, gk) = B->u;
    BV(2, gi, gj, gk) = B->v;
    BV(3, gi, gj, gk) = B->w;
    BV(4, gi, gj, gk) = B->p;
    BV(5, gi, gj, gk) = B->t;
    for (int l = 0; l < zz->nsc; ++l) SV(l, gi, gj, gk) = B->sv[l];
    if (zz->nt) S1(mut, gi, gj, gk) = B->mut;
    if (with_cv) cv_from_bv_point(zz, gi, gj, gk);
  }
}

/* BoundCond.cuh:424-552 apply_wall (no-slip; isothermal ghost T = 2Tw - Ti,
 * clipped at 0.1 Tw; rho from p and T) */
/* outward unit normal of a boundary face at (i,j,k): metric row of the face
 * direction, scaled by the face direction */
static void face_normal(const orc_zone *zz, int b, int i, int j, int k, double *n) {
  const int64_t p = at(zz->Lg, i, j, k);
  const double *m = &zz->metric[p * 9 + zz->bface[b] * 3];
  const double s = zz->bdir[b] / sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  n[0] = m[0] * s;
  n[1] = m[1] * s;
  n[2] = m[2] * s;
}

static void set_point(orc_zone *zz, int i, int j, int k, const double *b6, const double *svv, double mut) {
  for (int l = 0; l < 6; ++l) BV(l, i, j, k) = b6[l];
  for (int l = 0; l < zz->nsc; ++l) SV(l, i, j, k) = svv[l];
  if (zz->nt) S1(mut, i, j, k) = mut;
}

/* BoundCond.cuh:213-424 apply_farfield: supersonic in/outflow by the normal
 * Mach number of the boundary point, otherwise Riemann invariants with the
 * entropy of the far state (inflow) or of the boundary point (outflow);
 * mixtures iterate gamma(T) to 1e-4.  Ghost cv is never refreshed. */
static void bc_farfield(orc_zone *zz, const orc_bc *B, int b, int i, int j, int k) {
  int dir[3] = {0, 0, 0};
  dir[zz->bface[b]] = zz->bdir[b];
  double n[3];
  face_normal(zz, b, i, j, k, n);
  const double ub = BV(1, i, j, k), vb = BV(2, i, j, k), wb = BV(3, i, j, k);
  const double uf = n[0] * ub + n[1] * vb + n[2] * wb;
  const int ns = zz->ns, nsc = zz->nsc;
  double gb = GAMMA_AIR, mw = MW_AIR, svb[ORC_MAXS + 4];
  if (zz->cfg.mix) {
    for (int l = 0; l < nsc; ++l) svb[l] = SV(l, i, j, k);
    gb = S1(gam, i, j, k);
    double mi = 0;
    for (int l = 0; l < ns; ++l) mi += svb[l] / zz->mech.mw[l];
    mw = 1.0 / mi;
  }
  const double pb = BV(4, i, j, k), rb = BV(0, i, j, k);
  const double ab = sqrt(gb * pb / rb);
  const double mb = uf / ab;
  double st[6];
  if (mb <= -1) {
    const double far[6] = {B->rho, B->u, B->v, B->w, B->p, B->t};
    set_point(zz, i, j, k, far, B->sv, B->mut);
    S1(vel, i, j, k) = sqrt(B->u * B->u + B->v * B->v + B->w * B->w);
    for (int g = 1; g <= zz->g; ++g)
      set_point(zz, i + g * dir[0], j + g * dir[1], k + g * dir[2], far, B->sv, B->mut);
    return;
  }
  if (mb >= 1) {
    for (int g = 1; g <= zz->g; ++g) {
      const int gi = i + g * dir[0], gj = j + g * dir[1], gk = k + g * dir[2];
      for (int l = 0; l < 6; ++l) BV(l, gi, gj, gk) = BV(l, i, j, k);
      for (int l = 0; l < nsc; ++l) SV(l, gi, gj, gk) = SV(l, i, j, k);
      if (zz->nt) S1(mut, gi, gj, gk) = S1(mut, i, j, k);
    }
    return;
  }
  const double rp = uf + 2 * ab / (gb - 1);
  const double uinf = B->u * n[0] + B->v * n[1] + B->w * n[2];
  const double rn = uinf - 2 * B->a_inf / (B->gamma_inf - 1);
  const double Un = 0.5 * (rp + rn);
  double sb, rho, p, t, u, v, w;
  double mut = zz->nt ? S1(mut, i, j, k) : 0.0;
  if (!zz->cfg.mix) {
    const double cb = 0.25 * (GAMMA_AIR - 1) * (rp - rn);
    if (mb <= 0) {
      sb = B->entropy;
      u = B->u + (Un - uinf) * n[0];
      v = B->v + (Un - uinf) * n[1];
      w = B->w + (Un - uinf) * n[2];
      for (int l = 0; l < nsc; ++l) svb[l] = B->sv[l];
      mut = B->mut;
    } else {
      sb = pb / pow(rb, GAMMA_AIR);
      u = ub + (Un - uf) * n[0];
      v = vb + (Un - uf) * n[1];
      w = wb + (Un - uf) * n[2];
    }
    rho = pow(cb * cb / (GAMMA_AIR * sb), 1 / (GAMMA_AIR - 1));
    p = rho * cb * cb / GAMMA_AIR;
    t = p * MW_AIR / (rho * R_U);
  } else {
    if (mb < 0) {
      u = B->u + (Un - uinf) * n[0];
      v = B->v + (Un - uinf) * n[1];
      w = B->w + (Un - uinf) * n[2];
      for (int l = 0; l < nsc; ++l) svb[l] = B->sv[l];
      mw = B->mw_inf;
      mut = B->mut;
    } else {
      u = ub + (Un - uf) * n[0];
      v = vb + (Un - uf) * n[1];
      w = wb + (Un - uf) * n[2];
    }
    double gamma = GAMMA_AIR, err = 1, glast;
    while (err > 1e-4) {
      glast = gamma;
      const double cb = 0.25 * (gamma - 1) * (rp - rn);
      sb = mb <= 0 ? B->entropy : pb / pow(rb, gamma);
      rho = pow(cb * cb / (gamma * sb), 1 / (gamma - 1));
      p = rho * cb * cb / gamma;
      t = p * mw / (rho * R_U);
      double cpi[ORC_MAXS], cpt = 0;
      orc_cp(t, &zz->mech, cpi);
      for (int l = 0; l < ns; ++l) cpt += cpi[l] * svb[l];
      gamma = cpt / (cpt - R_U / mw);
      err = fabs(1 - gamma / glast);
    }
  }
  st[0] = rho;
  st[1] = u;
  st[2] = v;
  st[3] = w;
  st[4] = p;
  st[5] = t;
  set_point(zz, i, j, k, st, svb, mut);
  S1(vel, i, j, k) = sqrt(u * u + v * v + w * w);
  for (int g = 1; g <= zz->g; ++g) set_point(zz, i + g * dir[0], j + g * dir[1], k + g * dir[2], st, svb, mut);
}

/* BoundCond.cuh:554-636 apply_subsonic_inflow (air): the outgoing Riemann
 * invariant and total enthalpy of the boundary point fix the speed of sound
 * (larger root), total pressure / temperature the state; ghosts only. */
static void bc_subsonic(orc_zone *zz, const orc_bc *B, int b, int i, int j, int k, int with_cv) {
  int dir[3] = {0, 0, 0};
  dir[zz->bface[b]] = zz->bdir[b];
  double n[3];
  face_normal(zz, b, i, j, k, n);
  const double ub = BV(1, i, j, k), vb = BV(2, i, j, k), wb = BV(3, i, j, k);
  const double uf = n[0] * ub + n[1] * vb + n[2] * wb;
  const double g1 = GAMMA_AIR - 1;
  const double a = sqrt(GAMMA_AIR * BV(4, i, j, k) / BV(0, i, j, k));
  const double rn = fabs(uf) - 2 * a / g1;
  const double hti = BV(4, i, j, k) / BV(0, i, j, k) * GAMMA_AIR / g1 + 0.5 * (ub * ub + vb * vb + wb * wb);
  const double qa = 1 + 2.0 / g1, qb = 2 * rn, qc = g1 * (0.5 * rn * rn - hti);
  const double dl = qb * qb - 4 * qa * qc;
  double an = a;
  if (dl >= 0) {
    const double ap = (-qb + sqrt(dl)) / (2 * qa), am = (-qb - sqrt(dl)) / (2 * qa);
    an = ap > am ? ap : am;
  }
  const double un = rn + 2 * an / g1;
  const double M = un / an;
  const double p = B->p_total * pow(1 + 0.5 * g1 * M * M, -GAMMA_AIR / g1);
  const double t = B->t_total * pow(p / B->p_total, g1 / GAMMA_AIR);
  const double rho = p * MW_AIR / (t * R_U);
  for (int g = 1; g <= zz->g; ++g) {
    const int gi = i + g * dir[0], gj = j + g * dir[1], gk = k + g * dir[2];
    BV(0, gi, gj, gk) = rho;
    BV(1, gi, gj, gk) = un * B->u;
    BV(2, gi, gj, gk) = un * B->v;
    BV(3, gi, gj, gk) = un * B->w;
    BV(4, gi, gj, gk) = p;
    BV(5, gi, gj, gk) = t;
    const double ubar = BV(1, gi, gj, gk) * n[0] + BV(2, gi, gj, gk) * n[1] + BV(3, gi, gj, gk) * n[2];
    for (int l = 0; l < zz->nsc; ++l) SV(l, gi, gj, gk) = ubar > 0 ? SV(l, i, j, k) : B->sv[l];
    if (zz->nt) S1(mut, gi, gj, gk) = S1(mut, i, j, k);
    if (with_cv) cv_from_bv_point(zz, gi, gj, gk);
  }
}

/* BoundCond.cuh:638-670 apply_back_pressure: extrapolation with p fixed */
static void bc_back_pressure(orc_zone *zz, const orc_bc *B, int b, int i, int j, int k, int with_cv) {
  int dir[3] = {0, 0, 0};
  dir[zz->bface[b]] = zz->bdir[b];
  for (int g = 1; g <= zz->g; ++g) {
    const int gi = i + g * dir[0], gj = j + g * dir[1], gk = k + g * dir[2];
    for (int l = 0; l < 4; ++l) BV(l, gi, gj, gk) = BV(l, i, j, k);
    BV(4, gi, gj, gk) = B->p_back;
    BV(5, gi, gj, gk) = BV(5, i, j, k);
    for (int l = 0; l < zz->nsc; ++l) SV(l, gi, gj, gk) = SV(l, i, j, k);
    if (zz->nt) S1(mut, gi, gj, gk) = S1(mut, i, j, k);
    if (with_cv) cv_from_bv_point(zz, gi, gj, gk);
  }
}

/* SST.cuh:13-34 model constants (Menter's SST: inner set 1, k-epsilon set 2) */
#define SST_BETA_STAR 0.09
#define SST_SQRT_BETA_STAR 0.3
#define SST_KAPPA 0.41
#define SST_SIGMA_K1 0.85
#define SST_SIGMA_OMEGA1 0.5
#define SST_BETA1 0.0750
#define SST_A1 0.31
#define SST_SIGMA_K2 1.0
#define SST_SIGMA_OMEGA2 0.856
#define SST_BETA2 0.0828
static const double SST_GAMMA1 = SST_BETA1 / SST_BETA_STAR - SST_SIGMA_OMEGA1 * SST_KAPPA * SST_KAPPA / SST_SQRT_BETA_STAR;
static const double SST_GAMMA2 = SST_BETA2 / SST_BETA_STAR - SST_SIGMA_OMEGA2 * SST_KAPPA * SST_KAPPA / SST_SQRT_BETA_STAR;

/* Transport.cu:11-38 / 113-140 compute_viscosity: Wilke's mixing rule over
 * the species viscosities of the collision-integral fit */
static double wilke_viscosity(const orc_mech *m, double t, double mw_total, const double *Y) {
  const int ns = m->n_spec;
  double x[ORC_MAXS], vis[ORC_MAXS];
  for (int i = 0; i < ns; ++i) {
    x[i] = Y[i] * mw_total / m->mw[i];
    const double tdl = t * m->LJ_potent_inv[i];
    const double ci = 1.147 * pow(tdl, -0.145) + pow(tdl + 0.5, -2);
    vis[i] = m->vis_coeff[i] * sqrt(t) / ci;
  }
  double mu = 0;
  for (int i = 0; i < ns; ++i) {
    double vt = 0;
    for (int j = 0; j < ns; ++j) {
      double phi = 1.0;
      if (i != j) {
        const double num = 1 + sqrt(vis[i] / vis[j]) * m->WjDivWi_to_One4th[i * ns + j];
        phi = num * num * m->sqrt_WiDivWjPl1Mul8[i * ns + j];
      }
      vt += phi * x[j];
    }
    mu += vis[i] * x[i] / vt;
  }
  return mu;
}

static void bc_wall(orc_zone *zz, const orc_bc *B, int b, int i, int j, int k, int with_cv) {
  int dir[3] = {0, 0, 0};
  dir[zz->bface[b]] = zz->bdir[b];
  double t_wall = B->t_wall;
  const int i0 = i - dir[0], j0 = j - dir[1], k0 = k - dir[2];
  if (!B->isothermal) t_wall = BV(5, i0, j0, k0);
  const double p = BV(4, i0, j0, k0);
  double mw = MW_AIR;
  if (zz->cfg.mix) {
    mw = 0;
    for (int l = 0; l < zz->ns; ++l) {
      SV(l, i, j, k) = SV(l, i0, j0, k0);
      mw += SV(l, i, j, k) / zz->mech.mw[l];
    }
    mw = 1 / mw;
  }
  const double rho_wall = p * mw / (t_wall * R_U);
  BV(0, i, j, k) = rho_wall;
  BV(1, i, j, k) = 0;
  BV(2, i, j, k) = 0;
  BV(3, i, j, k) = 0;
  BV(4, i, j, k) = p;
  BV(5, i, j, k) = t_wall;
  S1(vel, i, j, k) = 0;
  if (zz->nt) {
    /* SST wall values: k = 0, omega = 60 mu_w / (rho_w beta1 d^2) with the
     * wall distance of the first interior point */
    double Yw[ORC_MAXS];
    for (int l = 0; l < zz->ns; ++l) Yw[l] = SV(l, i, j, k);
    const double mu_wall = zz->cfg.mix ? wilke_viscosity(&zz->mech, t_wall, mw, Yw) : orc_sutherland(t_wall);
    const double dy = S1(wdist, i0, j0, k0);
    SV(zz->ns, i, j, k) = 0;
    if (dy > 1e-25)
      SV(zz->ns + 1, i, j, k) = 60 * mu_wall / (rho_wall * SST_BETA1 * dy * dy);
    else
      SV(zz->ns + 1, i, j, k) = SV(zz->ns + 1, i0, j0, k0);
  }
  if (with_cv) cv_from_bv_point(zz, i, j, k);
  for (int g = 1; g <= zz->g; ++g) {
    const int ai = i - g * dir[0], aj = j - g * dir[1], ak = k - g * dir[2];
    const int gi = i + g * dir[0], gj = j + g * dir[1], gk = k + g * dir[2];
    const double ui = BV(1, ai, aj, ak), vi = BV(2, ai, aj, ak), wi = BV(3, ai, aj, ak);
    const double pi = BV(4, ai, aj, ak), ti = BV(5, ai, aj, ak);
    double tg = ti;
    if (B->isothermal) {
      tg = 2 * t_wall - ti;
      if (tg <= 0.1 * t_wall) tg = t_wall;
    }
    if (zz->cfg.mix) {
      mw = 0;
      for (int l = 0; l < zz->ns; ++l) {
        SV(l, gi, gj, gk) = SV(l, ai, aj, ak);
        mw += SV(l, gi, gj, gk) / zz->mech.mw[l];
      }
      mw = 1 / mw;
    }
    const double rho_g = pi * mw / (tg * R_U);
    BV(0, gi, gj, gk) = rho_g;
    BV(1, gi, gj, gk) = -ui;
    BV(2, gi, gj, gk) = -vi;
    BV(3, gi, gj, gk) = -wi;
    BV(4, gi, gj, gk) = pi;
    BV(5, gi, gj, gk) = tg;
    if (zz->nt) {
      SV(zz->ns, gi, gj, gk) = 0;
      SV(zz->ns + 1, gi, gj, gk) = SV(zz->ns + 1, i, j, k);
      S1(mut, gi, gj, gk) = 0;
    }
    if (with_cv) cv_from_bv_point(zz, gi, gj, gk);
  }
}

/* BoundCond.cuh:673-822 apply_boundary_conditions: one type after another
 * in priority order (a later one overwrites shared corner points). */
static void apply_bcs(orc_zone *zz, int with_cv) {
  for (int n = 0; n < zz->nbc; ++n) {
    const int b = zz->order[n];
    const orc_bc *B = &zz->cfg.bc[b];
    for (int k = zz->rs[b][2]; k <= zz->re[b][2]; ++k)
      for (int j = zz->rs[b][1]; j <= zz->re[b][1]; ++j)
        for (int i = zz->rs[b][0]; i <= zz->re[b][0]; ++i) {
          switch (B->type) {
            case 3: bc_symmetry(zz, b, i, j, k); break;
            case 5: bc_inflow(zz, B, b, i, j, k, with_cv); break;
            case 6: bc_outflow(zz, b, i, j, k, with_cv); break;
            case 2: bc_wall(zz, B, b, i, j, k, with_cv); break;
            case 4: bc_farfield(zz, B, b, i, j, k); break;
            case 7: bc_subsonic(zz, B, b, i, j, k, with_cv); break;
            case 9: bc_back_pressure(zz, B, b, i, j, k, with_cv); break;
            default: break;
          }
        }
  }
}

/* FieldOperation.cuh:372-400 eliminate_k_gradient (2-D) */
static void eliminate_k_gradient(orc_zone *zz, int with_cv) {
  const int g = zz->g;
  for (int j = -g; j < zz->my + g; ++j)
    for (int i = -g; i < zz->mx + g; ++i)
      for (int k = 1; k <= g; ++k) {
        for (int l = 0; l < 6; ++l) {
          BV(l, i, j, k) = BV(l, i, j, 0);
          BV(l, i, j, -k) = BV(l, i, j, 0);
        }
        for (int l = 0; l < zz->nsc; ++l) {
          SV(l, i, j, k) = SV(l, i, j, 0);
          SV(l, i, j, -k) = SV(l, i, j, 0);
        }
        if (with_cv)
          for (int l = 0; l < zz->nv; ++l) {
            CV(l, i, j, k) = CV(l, i, j, 0);
            CV(l, i, j, -k) = CV(l, i, j, 0);
          }
      }
}

/* FieldOperation.cuh:112-146 update_physical_properties on [-1, m] */
static void update_properties(orc_zone *zz) {
  for (int k = -1; k <= zz->mz; ++k)
    for (int j = -1; j <= zz->my; ++j)
      for (int i = -1; i <= zz->mx; ++i) {
        const double t = BV(5, i, j, k);
        if (zz->cfg.mix) {
          double cp[ORC_MAXS], Y[ORC_MAXS];
          orc_cp(t, &zz->mech, cp);
          double mw = 0, cp_tot = 0, cvv = 0;
          for (int l = 0; l < zz->ns; ++l) {
            Y[l] = SV(l, i, j, k);
            mw += Y[l] / zz->mech.mw[l];
            cp_tot += Y[l] * cp[l];
            cvv += Y[l] * (cp[l] - R_U / zz->mech.mw[l]);
          }
          mw = 1 / mw;
          S1(cp, i, j, k) = cp_tot;
          S1(gam, i, j, k) = cp_tot / cvv;
          S1(acoustic, i, j, k) = sqrt(S1(gam, i, j, k) * R_U * t / mw);
          double rhoD[ORC_MAXS];
          orc_transport(&zz->mech, t, BV(4, i, j, k), BV(0, i, j, k), Y, cp, &S1(mul, i, j, k), &S1(cond, i, j, k),
                        rhoD);
          for (int l = 0; l < zz->ns; ++l) zz->rhoD[l * zz->Lg.sz + at(zz->Lg, i, j, k)] = rhoD[l];
        } else {
          const double c_temp = GAMMA_AIR * R_U / MW_AIR;
          const double cp = GAMMA_AIR * R_U / MW_AIR / (GAMMA_AIR - 1);
          S1(acoustic, i, j, k) = sqrt(c_temp * t);
          S1(mul, i, j, k) = orc_sutherland(t);
          S1(cond, i, j, k) = S1(mul, i, j, k) * cp / zz->cfg.Pr;
        }
        S1(mach, i, j, k) = S1(vel, i, j, k) / S1(acoustic, i, j, k);
      }
}

static void initialize_mut(orc_zone *zz);

/* Driver.cu:40-98 initialize_computation followed by RK.cuh:38-48
 * (store_last_step, compute_cv_from_bv on every point). */
void orc_zone_initialize(orc_zone *zz) {
  const int g = zz->g;
  if (zz->dim == 2) eliminate_k_gradient(zz, 0);
  /* the mixture farfield reads gamma of its boundary point; the reference
   * applies the first BCs before any properties exist (Driver.cu:53-92,
   * reading unset memory), here they are computed from the initial field */
  if (zz->cfg.mix) update_properties(zz);
  apply_bcs(zz, 0);
  for (int k = -g; k < zz->mz + g; ++k)
    for (int j = -g; j < zz->my + g; ++j)
      for (int i = -g; i < zz->mx + g; ++i)
        S1(vel, i, j, k) = sqrt(BV(1, i, j, k) * BV(1, i, j, k) + BV(2, i, j, k) * BV(2, i, j, k) +
                                BV(3, i, j, k) * BV(3, i, j, k));
  if (zz->nt) initialize_mut(zz);
  update_properties(zz);
  for (int k = 0; k < zz->mz; ++k)
    for (int j = 0; j < zz->my; ++j)
      for (int i = 0; i < zz->mx; ++i) {
        const int64_t p = at(zz->L0, i, j, k);
        zz->bv_last[p] = BV(0, i, j, k);
        zz->bv_last[zz->L0.sz + p] = S1(vel, i, j, k);
        zz->bv_last[2 * zz->L0.sz + p] = BV(4, i, j, k);
        zz->bv_last[3 * zz->L0.sz + p] = BV(5, i, j, k);
      }
  for (int k = -g; k < zz->mz + g; ++k)
    for (int j = -g; j < zz->my + g; ++j)
      for (int i = -g; i < zz->mx + g; ++i) cv_from_bv_point(zz, i, j, k);
  zz->step = 0;
  zz->t_phys = 0;
}

/* ---------------------------------------------------------- inviscid */

/* gxl_lib/Math.hpp:13-15 sgn (sgn(0) = +1) and Limiter.cuh:13-16 minmod */
static double minmod(double a, double b) {
  const double sa = a < 0 ? -1 : 1, sb = b < 0 ? -1 : 1;
  return 0.5 * (sa + sb) * fmin(fabs(a), fabs(b));
}

/* Reconstruction.cuh:20-72 + Reconstruction.cu:5-46.  pv is a line buffer of
 * nv entries per point; ip is the left point of the face.  Appends E (and
 * gamma for mixtures) at pv_l[nv], pv_l[nv+1]. */
static void reconstruct_order(const orc_zone *zz, const double *pv, int ip, int recon, double *pl, double *pr) {
  const int nv = zz->nv;
  const double kappa = 1.0 / 3.0;
  for (int l = 0; l < nv; ++l) {
    const double q0 = pv[(ip - 1) * nv + l], q1 = pv[ip * nv + l], q2 = pv[(ip + 1) * nv + l],
                 q3 = pv[(ip + 2) * nv + l];
    switch (recon) {
      case 2: {
        const double di = q1 - q0, di1 = q2 - q1, di2 = q3 - q2;
        const double dnl = minmod(di, di1), dpl = minmod(di1, di), dnr = minmod(di1, di2), dpr = minmod(di2, di1);
        pl[l] = q1 + 0.25 * ((1 - kappa) * dnl + (1 + kappa) * dpl);
        pr[l] = q2 - 0.25 * ((1 - kappa) * dpr + (1 + kappa) * dnr);
        break;
      }
      case 3: {
        const double di = q1 - q0, di1 = q2 - q1, di2 = q3 - q2;
        pl[l] = q1 + 0.5 * minmod(di, di1);
        pr[l] = q2 - 0.5 * minmod(di1, di2);
        break;
      }
      default:
        pl[l] = q1;
        pr[l] = q2;
    }
  }
  if (zz->cfg.mix) {
    const orc_mech *m = &zz->mech;
    double el = 0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
    double er = 0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
    double mil = 0, mir = 0;
    for (int l = 0; l < zz->ns; ++l) {
      mil += pl[5 + l] / m->mw[l];
      mir += pr[5 + l] / m->mw[l];
    }
    const double tl = pl[4] / (pl[0] * R_U * mil), tr = pr[4] / (pr[0] * R_U * mir);
    double hl[ORC_MAXS], hr[ORC_MAXS], cpl_i[ORC_MAXS], cpr_i[ORC_MAXS];
    orc_enthalpy_and_cp(tl, m, hl, cpl_i);
    orc_enthalpy_and_cp(tr, m, hr, cpr_i);
    double cpl = 0, cpr = 0, cvl = 0, cvr = 0;
    for (int l = 0; l < zz->ns; ++l) {
      cpl += cpl_i[l] * pl[l + 5];
      cpr += cpr_i[l] * pr[l + 5];
      cvl += pl[l + 5] * (cpl_i[l] - R_U / m->mw[l]);
      cvr += pr[l + 5] * (cpr_i[l] - R_U / m->mw[l]);
      el += hl[l] * pl[l + 5];
      er += hr[l] * pr[l + 5];
    }
    pl[nv] = pl[0] * el - pl[4];
    pr[nv] = pr[0] * er - pr[4];
    pl[nv + 1] = cpl / cvl;
    pr[nv + 1] = cpr / cvr;
  } else {
    const double el = 0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]);
    const double er = 0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]);
    pl[nv] = el * pl[0] + pl[4] / (GAMMA_AIR - 1);
    pr[nv] = er * pr[0] + pr[4] / (GAMMA_AIR - 1);
    pl[nv + 1] = GAMMA_AIR;
    pr[nv + 1] = GAMMA_AIR;
  }
}

static void reconstruct(const orc_zone *zz, const double *pv, int ip, double *pl, double *pr) {
  reconstruct_order(zz, pv, ip, zz->cfg.reconstruction, pl, pr);
}

/* AusmP.cu:6-94 compute_ausmPlus_flux; met/jac are the line buffers, ip the left point. */
static void flux_ausm(const orc_zone *zz, const double *pl, const double *pr, const double *met, const double *jac,
                      int ip, double *f) {
  const int nv = zz->nv;
  const double *ml = &met[ip * 3], *mr = &met[(ip + 1) * 3];
  const double jl = jac[ip], jr = jac[ip + 1];
  const double k1 = 0.5 * (jl * ml[0] + jr * mr[0]);
  const double k2 = 0.5 * (jl * ml[1] + jr * mr[1]);
  const double k3 = 0.5 * (jl * ml[2] + jr * mr[2]);
  const double gk = sqrt(k1 * k1 + k2 * k2 + k3 * k3);
  const double ul = (k1 * pl[1] + k2 * pl[2] + k3 * pl[3]) / gk;
  const double ur = (k1 * pr[1] + k2 * pr[2] + k3 * pr[3]) / gk;
  const double pL = pl[4], pR = pr[4], rl = pl[0], rr = pr[0];
  double gl = GAMMA_AIR, gr = GAMMA_AIR;
  if (zz->cfg.mix) {
    gl = pl[nv + 1];
    gr = pr[nv + 1];
  }
  const double c = 0.5 * (sqrt(gl * pL / rl) + sqrt(gr * pR / rr));
  const double mlc = ul / c, mrc = ur / c;
  double mlp, mrn, plp, prn;
  const double alpha = 3 / 16.0;
  if (fabs(mlc) > 1) {
    mlp = 0.5 * (mlc + fabs(mlc));
    plp = mlp / mlc;
  } else {
    const double a = (mlc + 1) * (mlc + 1) * 0.25;
    const double b = (mlc * mlc - 1) * (mlc * mlc - 1);
    mlp = a + 0.125 * b;
    plp = a * (2 - mlc) + alpha * mlc * b;
  }
  if (fabs(mrc) > 1) {
    mrn = 0.5 * (mrc - fabs(mrc));
    prn = mrn / mrc;
  } else {
    const double a = (mrc - 1) * (mrc - 1) * 0.25;
    const double b = (mrc * mrc - 1) * (mrc * mrc - 1);
    mrn = -a - 0.125 * b;
    prn = a * (2 + mrc) - alpha * mrc * b;
  }
  const double p_coeff = plp * pL + prn * pR;
  const double m_half = mlp + mrn;
  const double mpos = 0.5 * (m_half + fabs(m_half)), mneg = 0.5 * (m_half - fabs(m_half));
  const double mass = c * (rl * mpos + rr * mneg);
  const double coeff = mass * gk;
  const double *s = mass >= 0 ? pl : pr;
  f[0] = coeff;
  f[1] = coeff * s[1] + p_coeff * k1;
  f[2] = coeff * s[2] + p_coeff * k2;
  f[3] = coeff * s[3] + p_coeff * k3;
  f[4] = coeff * (s[nv] + s[4]) / s[0];
  for (int l = 5; l < nv; ++l) f[l] = coeff * s[l];
}

/* Face metric and Jacobian of HLLC and LF at the face ip + 1/2: the
 * reference's (m0 + m1) / 2 and (J0 + J1) / 2 (HLLC.cu:64-71,
 * LaxFriedrichs.cu:22-35); with conservative metrics the face vector is
 * (J0 m0 + J1 m1) / 2 and the factor 1, as AUSM+ and Roe have it, so that
 * the face vectors of a cell close. */
static void face_vector(const orc_zone *zz, const double *met, const double *jac, int ip, double *kx, double *ky,
                        double *kz, double *ja) {
  if (zz->cfg.conservative_metric) {
    *kx = 0.5 * (jac[ip] * met[ip * 3] + jac[ip + 1] * met[(ip + 1) * 3]);
    *ky = 0.5 * (jac[ip] * met[ip * 3 + 1] + jac[ip + 1] * met[(ip + 1) * 3 + 1]);
    *kz = 0.5 * (jac[ip] * met[ip * 3 + 2] + jac[ip + 1] * met[(ip + 1) * 3 + 2]);
    *ja = 1;
    return;
  }
  *kx = 0.5 * (met[ip * 3] + met[(ip + 1) * 3]);
  *ky = 0.5 * (met[ip * 3 + 1] + met[(ip + 1) * 3 + 1]);
  *kz = 0.5 * (met[ip * 3 + 2] + met[(ip + 1) * 3 + 2]);
  *ja = 0.5 * (jac[ip] + jac[ip + 1]);
}

/* HLLC.cu:10-143 compute_hllc_flux */
static void flux_hllc(const orc_zone *zz, const double *pl, const double *pr, const double *met, const double *jac,
                      int ip, double *f) {
  const int nv = zz->nv;
  const double rlc = sqrt(pl[0]) / (sqrt(pl[0]) + sqrt(pr[0])), rrc = sqrt(pr[0]) / (sqrt(pl[0]) + sqrt(pr[0]));
  const double ut = pl[1] * rlc + pr[1] * rrc, vt = pl[2] * rlc + pr[2] * rrc, wt = pl[3] * rlc + pr[3] * rrc;
  const double hl = (pl[nv] + pl[4]) / pl[0], hr = (pr[nv] + pr[4]) / pr[0];
  const double ht = hl * rlc + hr * rrc;
  const double ekt = 0.5 * (ut * ut + vt * vt + wt * wt);
  double gamma = GAMMA_AIR, ct;
  double svm[ORC_MAXS + 4];
  for (int l = 0; l < nv - 5; ++l) svm[l] = rlc * pl[l + 5] + rrc * pr[l + 5];
  if (!zz->cfg.mix) {
    ct = sqrt((gamma - 1) * (ht - ekt));
  } else {
    const orc_mech *m = &zz->mech;
    double mw_inv = 0;
    for (int l = 0; l < zz->ns; ++l) mw_inv += svm[l] / m->mw[l];
    const double tl = pl[4] / pl[0], tr = pr[4] / pr[0];
    const double t = (rlc * tl + rrc * tr) / (R_U * mw_inv);
    double h_i[ORC_MAXS], cp_i[ORC_MAXS];
    orc_enthalpy_and_cp(t, m, h_i, cp_i);
    double cvv = 0, cpp = 0;
    for (int l = 0; l < zz->ns; ++l) {
      cpp += cp_i[l] * svm[l];
      cvv += svm[l] * (cp_i[l] - R_U / m->mw[l]);
    }
    gamma = cpp / cvv;
    ct = sqrt(gamma * R_U * mw_inv * t);
  }
  double kx, ky, kz, ja;
  face_vector(zz, met, jac, ip, &kx, &ky, &kz, &ja);
  const double gK = sqrt(kx * kx + ky * ky + kz * kz);
  const double Utb = (kx * ut + ky * vt + kz * wt) / gK;
  double gl = GAMMA_AIR, gr = GAMMA_AIR;
  if (zz->cfg.mix) {
    gl = pl[nv + 1];
    gr = pr[nv + 1];
  }
  const double Ul = kx * pl[1] + ky * pl[2] + kz * pl[3];
  const double cl = sqrt(gl * pl[4] / pl[0]);
  const double sl = fmin(Ul / gK - cl, Utb - ct);
  if (sl >= 0) {
    const double rU = pl[0] * Ul;
    f[0] = ja * rU;
    f[1] = ja * (rU * pl[1] + pl[4] * kx);
    f[2] = ja * (rU * pl[2] + pl[4] * ky);
    f[3] = ja * (rU * pl[3] + pl[4] * kz);
    f[4] = ja * ((pl[nv] + pl[4]) * Ul);
    for (int l = 5; l < nv; ++l) f[l] = ja * rU * pl[l];
    return;
  }
  const double Ur = kx * pr[1] + ky * pr[2] + kz * pr[3];
  const double cr = sqrt(gr * pr[4] / pr[0]);
  const double sr = fmax(Ur / gK + cr, Utb + ct);
  if (sr < 0) {
    const double rU = pr[0] * Ur;
    f[0] = ja * rU;
    f[1] = ja * (rU * pr[1] + pr[4] * kx);
    f[2] = ja * (rU * pr[2] + pr[4] * ky);
    f[3] = ja * (rU * pr[3] + pr[4] * kz);
    f[4] = ja * ((pr[nv] + pr[4]) * Ur);
    for (int l = 5; l < nv; ++l) f[l] = ja * rU * pr[l];
    return;
  }
  const double sm = ((pr[0] * Ur * (sr - Ur / gK) - pl[0] * Ul * (sl - Ul / gK)) / gK + pl[4] - pr[4]) /
                    (pr[0] * (sr - Ur / gK) - pl[0] * (sl - Ul / gK));
  const double pm = pl[0] * (sl - Ul / gK) * (sm - Ul / gK) + pl[4];
  if (sm >= 0) {
    const double pC = 1.0 / (sl - sm);
    const double Q = ja * pC * sm * (sl * gK - Ul) * pl[0];
    f[0] = Q;
    const double dP = (sl * pm - sm * pl[4]) * pC * ja;
    f[1] = Q * pl[1] + dP * kx;
    f[2] = Q * pl[2] + dP * ky;
    f[3] = Q * pl[3] + dP * kz;
    f[4] = Q * pl[nv] / pl[0] + pC * ja * (sl * pm * sm * gK - sm * pl[4] * Ul);
    for (int l = 5; l < nv; ++l) f[l] = Q * pl[l];
  } else {
    const double pC = 1.0 / (sr - sm);
    const double Q = ja * pC * sm * (sr * gK - Ur) * pr[0];
    f[0] = Q;
    const double dP = (sr * pm - sm * pr[4]) * pC * ja;
    f[1] = Q * pr[1] + dP * kx;
    f[2] = Q * pr[2] + dP * ky;
    f[3] = Q * pr[3] + dP * kz;
    f[4] = Q * pr[nv] / pr[0] + pC * ja * (sr * pm * sm * gK - sm * pr[4] * Ur);
    for (int l = 5; l < nv; ++l) f[l] = Q * pr[l];
  }
}

/* LaxFriedrichs.cu:14-49 compute_lf_flux<Air> */
static void flux_lf(const orc_zone *zz, const double *pl, const double *pr, const double *met, const double *jac, int ip,
                    double *f) {
  const int nv = zz->nv;
  double kx, ky, kz, ja;
  face_vector(zz, met, jac, ip, &kx, &ky, &kz, &ja);
  const double gK = sqrt(kx * kx + ky * ky + kz * kz);
  const double Ukl = pl[1] * kx + pl[2] * ky + pl[3] * kz, Ukr = pr[1] * kx + pr[2] * ky + pr[3] * kz;
  const double cl = sqrt(GAMMA_AIR * pl[4] / pl[0]), cr = sqrt(GAMMA_AIR * pr[4] / pr[0]);
  const double sr = fmax(fabs(Ukl) + cl * gK, fabs(Ukr) + cr * gK);
  const double hj = 0.5 * ja;
  const double rUl = pl[0] * Ukl, rUr = pr[0] * Ukr;
  f[0] = (rUl + rUr - sr * (pr[0] - pl[0])) * hj;
  f[1] = (rUl * pl[1] + rUr * pr[1] + kx * (pl[4] + pr[4]) - sr * (pr[1] * pr[0] - pl[1] * pl[0])) * hj;
  f[2] = (rUl * pl[2] + rUr * pr[2] + ky * (pl[4] + pr[4]) - sr * (pr[2] * pr[0] - pl[2] * pl[0])) * hj;
  f[3] = (rUl * pl[3] + rUr * pr[3] + kz * (pl[4] + pr[4]) - sr * (pr[3] * pr[0] - pl[3] * pl[0])) * hj;
  f[4] = ((pl[nv] + pl[4]) * Ukl + (pr[nv] + pr[4]) * Ukr - sr * (pr[nv] - pl[nv])) * hj;
}

/* Roe.cu:60-87 compute_entropy_fix_delta at one point: eps (|U| + |V| (+|W|)
 * + a * mean |grad xi|), 2-D without the third direction. */
static double efix_delta(const orc_zone *zz, int i, int j, int k) {
  const int64_t p = at(zz->Lg, i, j, k);
  const double *m = &zz->metric[p * 9];
  const double u = BV(1, i, j, k), v = BV(2, i, j, k), w = BV(3, i, j, k);
  const double U = fabs(u * m[0] + v * m[1] + w * m[2]);
  const double V = fabs(u * m[3] + v * m[4] + w * m[5]);
  const double W = fabs(u * m[6] + v * m[7] + w * m[8]);
  const double kx = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
  const double ky = sqrt(m[3] * m[3] + m[4] * m[4] + m[5] * m[5]);
  const double kz = sqrt(m[6] * m[6] + m[7] * m[7] + m[8] * m[8]);
  const double a = zz->acoustic[p];
  if (zz->mz == 1) return zz->cfg.entropy_fix_factor * (U + V + a * 0.5 * (kx + ky));
  return zz->cfg.entropy_fix_factor * (U + V + W + a * (kx + ky + kz) / 3.0);
}

/* Roe.cu:186-375: the half sum of the left and right fluxes, each with its
 * own J * metric, minus 1/2 |A| dQ from the characteristic decomposition at
 * the Roe average (species rows included; the mixture's gamma and c from the
 * Roe-averaged temperature), with Harten's entropy fix on the face average
 * of the per-point delta.  parity pinned by: consistency F(q, q), the Roe
 * property on a uniform grid (supersonic face flux = F(q_L)). */
static void flux_roe(const orc_zone *zz, const double *pl, const double *pr, const double *met, const double *jac,
                     const double *efd, int ip, double *f) {
  const int nv = zz->nv, ns = zz->ns;
  for (int l = 0; l < nv; ++l) f[l] = 0;
  for (int sd = 0; sd < 2; ++sd) {
    const double *q = sd ? pr : pl;
    const int pt = ip + sd;
    const double Jx = jac[pt] * met[pt * 3], Jy = jac[pt] * met[pt * 3 + 1], Jz = jac[pt] * met[pt * 3 + 2];
    const double Uk = q[1] * Jx + q[2] * Jy + q[3] * Jz;
    const double co = Uk * q[0];
    f[0] += 0.5 * co;
    f[1] += 0.5 * (co * q[1] + q[4] * Jx);
    f[2] += 0.5 * (co * q[2] + q[4] * Jy);
    f[3] += 0.5 * (co * q[3] + q[4] * Jz);
    f[4] += 0.5 * Uk * (q[4] + q[nv]);
    for (int l = 5; l < nv; ++l) f[l] += 0.5 * co * q[l];
  }
  const double dl = sqrt(pl[0]), dr = sqrt(pr[0]), id = 1.0 / (dl + dr);
  const double u = (dl * pl[1] + dr * pr[1]) * id, v = (dl * pl[2] + dr * pr[2]) * id,
               w = (dl * pl[3] + dr * pr[3]) * id;
  const double ek = 0.5 * (u * u + v * v + w * w);
  const double h = (dl * (pl[nv] + pl[4]) / pl[0] + dr * (pr[nv] + pr[4]) / pr[0]) * id;
  double gamma = GAMMA_AIR, c = sqrt((gamma - 1) * (h - ek)), mw = MW_AIR;
  double svm[ORC_MAXS + 4] = {0}, hi[ORC_MAXS] = {0};
  for (int l = 0; l < nv - 5; ++l) svm[l] = (dl * pl[l + 5] + dr * pr[l + 5]) * id;
  if (zz->cfg.mix) {
    const orc_mech *m = &zz->mech;
    double mwi = 0;
    for (int l = 0; l < ns; ++l) mwi += svm[l] / m->mw[l];
    const double t = (dl * (pl[4] / pl[0]) + dr * (pr[4] / pr[0])) * id / (R_U * mwi);
    double cpi[ORC_MAXS];
    orc_enthalpy_and_cp(t, m, hi, cpi);
    double cp = 0, cv = 0;
    for (int l = 0; l < ns; ++l) {
      cp += cpi[l] * svm[l];
      cv += svm[l] * (cpi[l] - R_U / m->mw[l]);
    }
    gamma = cp / cv;
    c = sqrt(gamma * R_U * mwi * t);
    mw = 1.0 / mwi;
  }
  double kx = 0.5 * (met[ip * 3] + met[(ip + 1) * 3]);
  double ky = 0.5 * (met[ip * 3 + 1] + met[(ip + 1) * 3 + 1]);
  double kz = 0.5 * (met[ip * 3 + 2] + met[(ip + 1) * 3 + 2]);
  const double gk = sqrt(kx * kx + ky * ky + kz * kz);
  double Uk = kx * u + ky * v + kz * w;
  double lam[3] = {Uk - gk * c, Uk, Uk + gk * c};
  const double d = 0.5 * (efd[ip] + efd[ip + 1]);
  for (int a = 0; a < 3; ++a) {
    lam[a] = fabs(lam[a]);
    if (lam[a] < d) lam[a] = 0.5 * (lam[a] * lam[a] / d + d);
  }
  kx /= gk;
  ky /= gk;
  kz /= gk;
  Uk /= gk;
  const double ja = 0.5 * (jac[ip] + jac[ip + 1]);
  double dq[5 + ORC_MAXS + 4] = {0};
  dq[0] = ja * (pr[0] - pl[0]);
  for (int l = 1; l < nv; ++l) dq[l] = ja * (pr[0] * pr[l] - pl[0] * pl[l]);
  dq[4] = ja * (pr[nv] - pl[nv]);
  double c1 = (gamma - 1) * (ek * dq[0] - u * dq[1] - v * dq[2] - w * dq[3] + dq[4]) / (c * c);
  double c2 = (kx * dq[1] + ky * dq[2] + kz * dq[3] - Uk * dq[0]) / c;
  for (int l = 0; l < ns; ++l) c1 += (mw / zz->mech.mw[l] - hi[l] * (gamma - 1) / (c * c)) * dq[l + 5];
  double c3 = dq[0] - c1;
  double L[5 + ORC_MAXS + 4] = {0};
  L[0] = 0.5 * (c1 - c2);
  L[1] = kx * c3 - ((kz * v - ky * w) * dq[0] - kz * dq[2] + ky * dq[3]) / c;
  L[2] = ky * c3 - ((kx * w - kz * u) * dq[0] - kx * dq[3] + kz * dq[1]) / c;
  L[3] = kz * c3 - ((ky * u - kx * v) * dq[0] - ky * dq[1] + kx * dq[2]) / c;
  L[4] = 0.5 * (c1 + c2);
  for (int l = 0; l < nv - 5; ++l) L[l + 5] = dq[l + 5] - svm[l] * dq[0];
  double b[5 + ORC_MAXS + 4];
  b[0] = -lam[0] * L[0];
  for (int l = 1; l < nv; ++l) b[l] = -lam[1] * L[l];
  b[4] = -lam[2] * L[4];
  const double c0 = kx * b[1] + ky * b[2] + kz * b[3];
  c1 = c0 + b[0] + b[4];
  c2 = c * (b[4] - b[0]);
  c3 = 0;
  for (int l = 0; l < ns; ++l) c3 += (hi[l] - mw / zz->mech.mw[l] * c * c / (gamma - 1)) * b[l + 5];
  f[0] += 0.5 * c1;
  f[1] += 0.5 * (u * c1 + kx * c2 - c * (kz * b[2] - ky * b[3]));
  f[2] += 0.5 * (v * c1 + ky * c2 - c * (kx * b[3] - kz * b[1]));
  f[3] += 0.5 * (w * c1 + kz * c2 - c * (ky * b[1] - kx * b[2]));
  f[4] += 0.5 * (h * c1 + Uk * c2 - c * c * c0 / (gamma - 1) +
                 c * ((kz * v - ky * w) * b[1] + (kx * w - kz * u) * b[2] + (ky * u - kx * v) * b[3]) + c3);
  for (int l = 0; l < nv - 5; ++l) f[5 + l] += 0.5 * (b[l + 5] + svm[l] * c1);
}

static void riemann(const orc_zone *zz, const double *pl, const double *pr, const double *met, const double *jac,
                    const double *efd, int ip, double *f) {
  switch (zz->cfg.inviscid_scheme) {
    case 1: flux_lf(zz, pl, pr, met, jac, ip, f); break;
    case 2: /* with AWENO (inviscid_type 1) the reference runs AUSM+ (InviscidScheme.cu:290-303) */
      if (zz->inviscid_type == 2) {
        flux_roe(zz, pl, pr, met, jac, efd, ip, f);
        break;
      }
      flux_ausm(zz, pl, pr, met, jac, ip, f);
      break;
    case 4: flux_hllc(zz, pl, pr, met, jac, ip, f); break;
    default: flux_ausm(zz, pl, pr, met, jac, ip, f); break;
  }
}

/* AWENO.cu:151-206 WENO5: characteristic projection of the 6-point stencil
 * and the alternative-WENO weights (linear 1/16, 5/8, 5/16; tau5 = (b0-b2)^2) */
static void weno5_values(const double *v, double *vm, double *vp);

static void weno5(const double *L, const double *cvb, int ip, int row, double *vm, double *vp) {
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (int s = -2; s < 4; ++s) {
    const double *U = &cvb[(ip + s) * 7];
    for (int c = 0; c < 5; ++c) v[s + 2] += L[row * 5 + c] * U[c];
  }
  weno5_values(v, vm, vp);
}

/* the WENO5 weights and candidate stencils on six characteristic values of
 * points i-2 .. i+3: v-(i+1/2) and v+(i+1/2) */
static void weno5_values(const double *v, double *vm, double *vp) {
  double v0 = 0.125 * (3 * v[0] - 10 * v[1] + 15 * v[2]);
  double v1 = 0.125 * (-v[1] + 6 * v[2] + 3 * v[3]);
  double v2 = 0.125 * (3 * v[2] + 6 * v[3] - v[4]);
  const double t12 = 13.0 / 12;
  double b0 = t12 * (v[0] + v[2] - 2 * v[1]) * (v[0] + v[2] - 2 * v[1]) +
              0.25 * (v[0] + 3 * v[2] - 4 * v[1]) * (v[0] + 3 * v[2] - 4 * v[1]);
  double b1 = t12 * (v[1] + v[3] - 2 * v[2]) * (v[1] + v[3] - 2 * v[2]) + 0.25 * (v[3] - v[1]) * (v[3] - v[1]);
  double b2 = t12 * (v[2] + v[4] - 2 * v[3]) * (v[2] + v[4] - 2 * v[3]) +
              0.25 * (3 * v[2] + v[4] - 4 * v[3]) * (3 * v[2] + v[4] - 4 * v[3]);
  double tau = (b0 - b2) * (b0 - b2);
  const double eps = 1e-6;
  const double o16 = 1.0 / 16, f8 = 5.0 / 8, f16 = 5.0 / 16;
  double a0 = o16 + o16 * tau / ((eps + b0) * (eps + b0));
  double a1 = f8 + f8 * tau / ((eps + b1) * (eps + b1));
  double a2 = f16 + f16 * tau / ((eps + b2) * (eps + b2));
  *vm = (a0 * v0 + a1 * v1 + a2 * v2) / (a0 + a1 + a2);
  v0 = v1;
  v1 = v2;
  v2 = 0.125 * (15 * v[3] - 10 * v[4] + 3 * v[5]);
  b0 = t12 * (v[1] + v[3] - 2 * v[2]) * (v[1] + v[3] - 2 * v[2]) +
       0.25 * (v[1] + 3 * v[3] - 4 * v[2]) * (v[1] + 3 * v[3] - 4 * v[2]);
  b1 = t12 * (v[2] + v[4] - 2 * v[3]) * (v[2] + v[4] - 2 * v[3]) + 0.25 * (v[4] - v[2]) * (v[4] - v[2]);
  b2 = t12 * (v[3] + v[5] - 2 * v[4]) * (v[3] + v[5] - 2 * v[4]) +
       0.25 * (3 * v[3] + v[5] - 4 * v[4]) * (3 * v[3] + v[5] - 4 * v[4]);
  tau = (b0 - b2) * (b0 - b2);
  a0 = f16 + f16 * tau / ((eps + b0) * (eps + b0));
  a1 = f8 + f8 * tau / ((eps + b1) * (eps + b1));
  a2 = o16 + o16 * tau / ((eps + b2) * (eps + b2));
  *vp = (a0 * v0 + a1 * v1 + a2 * v2) / (a0 + a1 + a2);
}

/* AWENO.cu:17-149 AWENO_interpolation<Air>: Roe average at the face, left
 * eigenvectors along the normalised averaged metric, WENO5 per
 * characteristic, right eigenvectors back to conservative, then primitives.
 * cvb holds (rho, rho u, rho v, rho w, rho E, p, T) per point. */
static void aweno_air_m(const double *cvb, const double *met, int ip, double *pl, double *pr, double *Lout,
                        double *Rout) {
  const double *cl = &cvb[ip * 7], *cr = &cvb[(ip + 1) * 7];
  const double rlc = sqrt(cl[0]) / (sqrt(cl[0]) + sqrt(cr[0]));
  const double rrc = sqrt(cr[0]) / (sqrt(cl[0]) + sqrt(cr[0]));
  const double um = rlc * cl[1] / cl[0] + rrc * cr[1] / cr[0];
  const double vm = rlc * cl[2] / cl[0] + rrc * cr[2] / cr[0];
  const double wm = rlc * cl[3] / cl[0] + rrc * cr[3] / cr[0];
  const double ekm = 0.5 * (um * um + vm * vm + wm * wm);
  const double gm1 = GAMMA_AIR - 1;
  const double hl = (cl[4] + cl[5]) / cl[0], hr = (cr[4] + cr[5]) / cr[0];
  const double hm = rlc * hl + rrc * hr;
  const double cm2 = gm1 * (hm - ekm);
  const double cm = sqrt(cm2);
  double kx = 0.5 * (met[ip * 3] + met[(ip + 1) * 3]);
  double ky = 0.5 * (met[ip * 3 + 1] + met[(ip + 1) * 3 + 1]);
  double kz = 0.5 * (met[ip * 3 + 2] + met[(ip + 1) * 3 + 2]);
  const double gK = sqrt(kx * kx + ky * ky + kz * kz);
  kx /= gK;
  ky /= gK;
  kz /= gK;
  const double Uk = kx * um + ky * vm + kz * wm;
  const double alpha = gm1 * ekm;
  double LR[25];
  LR[0] = (alpha + Uk * cm) / cm2 * 0.5;
  LR[1] = -(gm1 * um + kx * cm) / cm2 * 0.5;
  LR[2] = -(gm1 * vm + ky * cm) / cm2 * 0.5;
  LR[3] = -(gm1 * wm + kz * cm) / cm2 * 0.5;
  LR[4] = gm1 / cm2 * 0.5;
  LR[5] = kx * (1 - alpha / cm2) - (kz * vm - ky * wm) / cm;
  LR[6] = kx * gm1 * um / cm2;
  LR[7] = (kx * gm1 * vm + kz * cm) / cm2;
  LR[8] = (kx * gm1 * wm - ky * cm) / cm2;
  LR[9] = -kx * gm1 / cm2;
  LR[10] = ky * (1 - alpha / cm2) - (kx * wm - kz * um) / cm;
  LR[11] = (ky * gm1 * um - kz * cm) / cm2;
  LR[12] = ky * gm1 * vm / cm2;
  LR[13] = (ky * gm1 * wm + kx * cm) / cm2;
  LR[14] = -ky * gm1 / cm2;
  LR[15] = kz * (1 - alpha / cm2) - (ky * um - kx * vm) / cm;
  LR[16] = (kz * gm1 * um + ky * cm) / cm2;
  LR[17] = (kz * gm1 * vm - kx * cm) / cm2;
  LR[18] = kz * gm1 * wm / cm2;
  LR[19] = -kz * gm1 / cm2;
  LR[20] = (alpha - Uk * cm) / cm2 * 0.5;
  LR[21] = -(gm1 * um - kx * cm) / cm2 * 0.5;
  LR[22] = -(gm1 * vm - ky * cm) / cm2 * 0.5;
  LR[23] = -(gm1 * wm - kz * cm) / cm2 * 0.5;
  LR[24] = gm1 / cm2 * 0.5;
  double vp[5], vmn[5];
  if (Lout) memcpy(Lout, LR, sizeof(LR));
  for (int l = 0; l < 5; ++l) weno5(LR, cvb, ip, l, &vmn[l], &vp[l]);
  LR[0] = 1.0;
  LR[1] = kx;
  LR[2] = ky;
  LR[3] = kz;
  LR[4] = 1.0;
  LR[5] = um - kx * cm;
  LR[6] = kx * um;
  LR[7] = ky * um - kz * cm;
  LR[8] = kz * um + ky * cm;
  LR[9] = um + kx * cm;
  LR[10] = vm - ky * cm;
  LR[11] = kx * vm + kz * cm;
  LR[12] = ky * vm;
  LR[13] = kz * vm - kx * cm;
  LR[14] = vm + ky * cm;
  LR[15] = wm - kz * cm;
  LR[16] = kx * wm - ky * cm;
  LR[17] = ky * wm + kx * cm;
  LR[18] = kz * wm;
  LR[19] = wm + kz * cm;
  LR[20] = hm - Uk * cm;
  LR[21] = kx * alpha / gm1 + (kz * vm - ky * wm) * cm;
  LR[22] = ky * alpha / gm1 + (kx * wm - kz * um) * cm;
  LR[23] = kz * alpha / gm1 + (ky * um - kx * vm) * cm;
  LR[24] = hm + Uk * cm;
  if (Rout) memcpy(Rout, LR, sizeof(LR));
  double ql[5], qr[5];
  for (int a = 0; a < 5; ++a) {
    ql[a] = 0;
    qr[a] = 0;
    for (int b = 0; b < 5; ++b) {
      ql[a] += LR[a * 5 + b] * vmn[b];
      qr[a] += LR[a * 5 + b] * vp[b];
    }
  }
  pl[0] = ql[0];
  pl[1] = ql[1] / ql[0];
  pl[2] = ql[2] / ql[0];
  pl[3] = ql[3] / ql[0];
  pl[4] = gm1 * (ql[4] - 0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]) * ql[0]);
  pl[5] = ql[4];
  pr[0] = qr[0];
  pr[1] = qr[1] / qr[0];
  pr[2] = qr[2] / qr[0];
  pr[3] = qr[3] / qr[0];
  pr[4] = gm1 * (qr[4] - 0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]) * qr[0]);
  pr[5] = qr[4];
}

static void aweno_air(const double *cvb, const double *met, int ip, double *pl, double *pr) {
  aweno_air_m(cvb, met, ip, pl, pr, NULL, NULL);
}

/* ---- AWENO-5 for multi-species mixtures (and transported scalars).
 * The reference's mixture interpolation is an empty stub (AWENO.cu:10-15),
 * so this is the build's definition: the air algorithm (AWENO.cu:17-149) on
 * the 5 + n_scalar conservative variables, with the characteristic fields of
 * the frozen multi-species Euler Jacobian at the face's Roe average -- the
 * eigen-system the mixture Roe flux uses (Roe.cu:186-327): acoustic waves
 * from c1 = dp/c^2 (frozen thermodynamics: gamma and h_k at the Roe-averaged
 * temperature, and the total enthalpy ek~ + sum Y~ h_k(T~) that makes L and R
 * exact inverses), the entropy and shear combinations of air, and one wave
 * Q_{5+l} - Y~_l Q_0 per scalar.  Each reconstructed side is turned into
 * primitives by a Newton inversion of its internal energy to 1e-12, started
 * at the Roe-averaged temperature; gamma of each side at its temperature. */
typedef struct {
  double u, v, w, h, ek, c, gm1, mw, nx, ny, nz, U, t;
  double y[ORC_MAXS + 4], hi[ORC_MAXS];
} mix_eig;

/* Roe average of the two face points cl, cr = (Q[nv], p, T) and the unit
 * normal of the face-averaged metric row k = (k0 + k1)/2. */
static void mix_roe_state(const orc_zone *zz, const double *cl, const double *cr, const double *k0, const double *k1,
                          mix_eig *E) {
  const int nv = zz->nv, ns = zz->ns, nsc = nv - 5;
  const double dl = sqrt(cl[0]), dr = sqrt(cr[0]), id = 1.0 / (dl + dr);
  E->u = (dl * cl[1] / cl[0] + dr * cr[1] / cr[0]) * id;
  E->v = (dl * cl[2] / cl[0] + dr * cr[2] / cr[0]) * id;
  E->w = (dl * cl[3] / cl[0] + dr * cr[3] / cr[0]) * id;
  E->ek = 0.5 * (E->u * E->u + E->v * E->v + E->w * E->w);
  E->h = (dl * (cl[4] + cl[nv]) / cl[0] + dr * (cr[4] + cr[nv]) / cr[0]) * id;
  for (int l = 0; l < nsc; ++l) E->y[l] = (dl * cl[5 + l] / cl[0] + dr * cr[5 + l] / cr[0]) * id;
  double gamma = GAMMA_AIR;
  E->mw = MW_AIR;
  E->t = 0;
  if (zz->cfg.mix) {
    const orc_mech *m = &zz->mech;
    double mwi = 0;
    for (int l = 0; l < ns; ++l) mwi += E->y[l] / m->mw[l];
    E->t = (dl * (cl[nv] / cl[0]) + dr * (cr[nv] / cr[0])) * id / (R_U * mwi);
    double cpi[ORC_MAXS];
    orc_enthalpy_and_cp(E->t, m, E->hi, cpi);
    double cp = 0, cv = 0;
    for (int l = 0; l < ns; ++l) {
      cp += cpi[l] * E->y[l];
      cv += E->y[l] * (cpi[l] - R_U / m->mw[l]);
    }
    gamma = cp / cv;
    E->c = sqrt(gamma * R_U * mwi * E->t);
    E->mw = 1.0 / mwi;
    /* the total enthalpy consistent with T~: L R = I needs
     * H~ - ek~ = sum Y~ h(T~), which the sqrt(rho) average of H does not
     * satisfy when cp depends on T */
    double hs = 0;
    for (int l = 0; l < ns; ++l) hs += E->y[l] * E->hi[l];
    E->h = E->ek + hs;
  } else {
    E->c = sqrt((gamma - 1) * (E->h - E->ek));
  }
  E->gm1 = gamma - 1;
  double kx = 0.5 * (k0[0] + k1[0]), ky = 0.5 * (k0[1] + k1[1]), kz = 0.5 * (k0[2] + k1[2]);
  const double gk = sqrt(kx * kx + ky * ky + kz * kz);
  E->nx = kx / gk;
  E->ny = ky / gk;
  E->nz = kz / gk;
  E->U = E->nx * E->u + E->ny * E->v + E->nz * E->w;
}

/* characteristic variables w = L Q of a conservative vector Q */
static void mix_left(const orc_zone *zz, const mix_eig *E, const double *Q, double *wv) {
  const int nv = zz->nv, ns = zz->ns;
  const double c = E->c, c2 = c * c;
  double c1 = E->gm1 * (E->ek * Q[0] - E->u * Q[1] - E->v * Q[2] - E->w * Q[3] + Q[4]) / c2;
  for (int l = 0; l < ns; ++l) c1 += (E->mw / zz->mech.mw[l] - E->hi[l] * E->gm1 / c2) * Q[5 + l];
  const double c2v = (E->nx * Q[1] + E->ny * Q[2] + E->nz * Q[3] - E->U * Q[0]) / c;
  const double c3 = Q[0] - c1;
  wv[0] = 0.5 * (c1 - c2v);
  wv[1] = E->nx * c3 - ((E->nz * E->v - E->ny * E->w) * Q[0] - E->nz * Q[2] + E->ny * Q[3]) / c;
  wv[2] = E->ny * c3 - ((E->nx * E->w - E->nz * E->u) * Q[0] - E->nx * Q[3] + E->nz * Q[1]) / c;
  wv[3] = E->nz * c3 - ((E->ny * E->u - E->nx * E->v) * Q[0] - E->ny * Q[1] + E->nx * Q[2]) / c;
  wv[4] = 0.5 * (c1 + c2v);
  for (int l = 5; l < nv; ++l) wv[l] = Q[l] - E->y[l - 5] * Q[0];
}

/* conservative vector Q = R w of characteristic amplitudes w */
static void mix_right(const orc_zone *zz, const mix_eig *E, const double *wv, double *Q) {
  const int nv = zz->nv, ns = zz->ns;
  const double c = E->c;
  const double c0 = E->nx * wv[1] + E->ny * wv[2] + E->nz * wv[3];
  const double C1 = c0 + wv[0] + wv[4], C2 = c * (wv[4] - wv[0]);
  double C3 = 0;
  for (int l = 0; l < ns; ++l) C3 += (E->hi[l] - E->mw / zz->mech.mw[l] * c * c / E->gm1) * wv[5 + l];
  Q[0] = C1;
  Q[1] = E->u * C1 + E->nx * C2 - c * (E->nz * wv[2] - E->ny * wv[3]);
  Q[2] = E->v * C1 + E->ny * C2 - c * (E->nx * wv[3] - E->nz * wv[1]);
  Q[3] = E->w * C1 + E->nz * C2 - c * (E->ny * wv[1] - E->nx * wv[2]);
  Q[4] = E->h * C1 + E->U * C2 - c * c * c0 / E->gm1 +
         c * ((E->nz * E->v - E->ny * E->w) * wv[1] + (E->nx * E->w - E->nz * E->u) * wv[2] +
              (E->ny * E->u - E->nx * E->v) * wv[3]) +
         C3;
  for (int l = 5; l < nv; ++l) Q[l] = wv[l] + E->y[l - 5] * C1;
}

/* temperature of a mixture state from its internal energy per unit mass e:
 * Newton on sum Y h(T) - R T sum Y/W = e, from t0, until |dT| <= 1e-12 T
 * (at most 30 iterations) */
static double mix_temperature(const orc_mech *m, const double *Y, double e, double t0) {
  const int ns = m->n_spec;
  double mwi = 0;
  for (int l = 0; l < ns; ++l) mwi += Y[l] / m->mw[l];
  double t = t0;
  for (int it = 0; it < 30; ++it) {
    double h[ORC_MAXS], cp[ORC_MAXS];
    orc_enthalpy_and_cp(t, m, h, cp);
    double f = 0, fp = 0;
    for (int l = 0; l < ns; ++l) {
      f += Y[l] * h[l];
      fp += Y[l] * cp[l];
    }
    f -= R_U * t * mwi + e;
    fp -= R_U * mwi;
    const double dt = f / fp;
    t -= dt;
    if (fabs(dt) <= 1e-12 * t) break;
  }
  return t;
}

/* primitives (rho, u, v, w, p, scalars; rho E at [nv], gamma at [nv+1]) of
 * a reconstructed conservative side Q */
static void mix_side(const orc_zone *zz, const mix_eig *E, const double *Q, double *pv) {
  const int nv = zz->nv, ns = zz->ns;
  const double rho = Q[0], ri = 1.0 / rho;
  pv[0] = rho;
  pv[1] = Q[1] * ri;
  pv[2] = Q[2] * ri;
  pv[3] = Q[3] * ri;
  for (int l = 5; l < nv; ++l) pv[l] = Q[l] * ri;
  const double ek = 0.5 * (pv[1] * pv[1] + pv[2] * pv[2] + pv[3] * pv[3]);
  pv[nv] = Q[4];
  if (zz->cfg.mix) {
    const orc_mech *m = &zz->mech;
    double Y[ORC_MAXS], mwi = 0;
    for (int l = 0; l < ns; ++l) {
      Y[l] = pv[5 + l];
      mwi += Y[l] / m->mw[l];
    }
    const double t = mix_temperature(m, Y, Q[4] * ri - ek, E->t);
    pv[4] = rho * R_U * t * mwi;
    double cpi[ORC_MAXS];
    orc_cp(t, m, cpi);
    double cp = 0, cv = 0;
    for (int l = 0; l < ns; ++l) {
      cp += cpi[l] * Y[l];
      cv += Y[l] * (cpi[l] - R_U / m->mw[l]);
    }
    pv[nv + 1] = cp / cv;
  } else {
    pv[4] = (GAMMA_AIR - 1) * (Q[4] - ek * rho);
    pv[nv + 1] = GAMMA_AIR;
  }
}

/* A reconstructed side is admissible when its density and pressure are
 * positive, its gamma exceeds 1, everything is finite, and each mass
 * fraction lies in [-0.01, 1.01]. */
static int mix_side_ok(const orc_zone *zz, const double *pv) {
  const int nv = zz->nv;
  if (!(pv[0] > 0) || !(pv[4] > 0) || !(pv[nv + 1] > 1) || !isfinite(pv[4]) || !isfinite(pv[nv]) ||
      !isfinite(pv[nv + 1]))
    return 0;
  for (int l = 5; l < nv; ++l) {
    if (!isfinite(pv[l])) return 0;
    if (l < 5 + zz->ns && (pv[l] < -0.01 || pv[l] > 1.01)) return 0;
  }
  return 1;
}

/* AWENO interpolation of a mixture face: cm holds (Q[nv], p, T) per point
 * (stride nv + 2), met the metric rows of the line, ip the face's left point.
 * Where either reconstructed side is not admissible (WENO overshoot at a
 * jump no candidate stencil avoids, e.g. a jet port two cells wide), both
 * sides take the first-order values of the two adjacent points (pv holds
 * their primitives, stride nv). */
static void aweno_mix(const orc_zone *zz, const double *cm, const double *pv, const double *met, int ip, double *pl,
                      double *pr) {
  const int nv = zz->nv, st = nv + 2;
  mix_eig E;
  mix_roe_state(zz, &cm[ip * st], &cm[(ip + 1) * st], &met[ip * 3], &met[(ip + 1) * 3], &E);
  double wv[6][5 + ORC_MAXS + 4];
  for (int s = 0; s < 6; ++s) mix_left(zz, &E, &cm[(ip - 2 + s) * st], wv[s]);
  double wm[5 + ORC_MAXS + 4], wp[5 + ORC_MAXS + 4];
  for (int l = 0; l < nv; ++l) {
    const double v[6] = {wv[0][l], wv[1][l], wv[2][l], wv[3][l], wv[4][l], wv[5][l]};
    weno5_values(v, &wm[l], &wp[l]);
  }
  double ql[5 + ORC_MAXS + 4], qr[5 + ORC_MAXS + 4];
  mix_right(zz, &E, wm, ql);
  mix_right(zz, &E, wp, qr);
  mix_side(zz, &E, ql, pl);
  mix_side(zz, &E, qr, pr);
  if (!mix_side_ok(zz, pl) || !mix_side_ok(zz, pr)) reconstruct_order(zz, pv, ip, 1, pl, pr);
}

/* line gather helpers: position s on line (dir, a, b) -> (i,j,k) */
static void line_ijk(int dir, int a, int b, int s, int *i, int *j, int *k) {
  if (dir == 0) { *i = s; *j = a; *k = b; }
  else if (dir == 1) { *i = a; *j = s; *k = b; }
  else { *i = a; *j = b; *k = s; }
}

/* InviscidScheme.cu:10-158 (pv path), :160-311 (AWENO path) and
 * AWENO.cu:213-264 (CDSPart1D<Air>) for one direction. */
static void inviscid_dir(orc_zone *zz, int dir) {
  const int mext[3] = {zz->mx, zz->my, zz->mz};
  const int n = mext[dir], g = zz->g, nv = zz->nv;
  const int oa = dir == 0 ? 1 : 0, ob = dir == 2 ? 1 : 2;
  const int na = mext[oa], nb = mext[ob];
  const int npt = n + 2 * g;
  /* AWENO with species or transported scalars takes the mixture
   * interpolation, whose CDS node flux carries every component */
  const int mixaw = zz->inviscid_type == 1 && (zz->cfg.mix || zz->nt);
  const int nfn = mixaw ? nv : 5, stm = nv + 2;
  double *pv = malloc(sizeof(double) * npt * (nv + 2) + 8);
  double *cvb = malloc(sizeof(double) * npt * 7 + 8);
  double *cm = malloc(sizeof(double) * npt * stm + 8);
  double *met = malloc(sizeof(double) * npt * 3 + 8);
  double *jac = malloc(sizeof(double) * npt + 8);
  double *efd = malloc(sizeof(double) * npt + 8);
  double *fc = malloc(sizeof(double) * (n + 1) * nv + 8);
  double *fn = malloc(sizeof(double) * npt * nfn + 8);
  for (int b = 0; b < nb; ++b)
    for (int a = 0; a < na; ++a) {
      for (int s = -g; s < n + g; ++s) {
        int i, j, k;
        line_ijk(dir, a, b, s, &i, &j, &k);
        const int q = s + g;
        const int64_t p = at(zz->Lg, i, j, k);
        for (int l = 0; l < 5; ++l) pv[q * nv + l] = BV(l, i, j, k);
        for (int l = 0; l < zz->nsc; ++l) pv[q * nv + 5 + l] = SV(l, i, j, k);
        for (int l = 0; l < 5; ++l) cvb[q * 7 + l] = CV(l, i, j, k);
        cvb[q * 7 + 5] = BV(4, i, j, k);
        cvb[q * 7 + 6] = BV(5, i, j, k);
        if (mixaw) {
          for (int l = 0; l < nv; ++l) cm[q * stm + l] = CV(l, i, j, k);
          cm[q * stm + nv] = BV(4, i, j, k);
          cm[q * stm + nv + 1] = BV(5, i, j, k);
        }
        for (int c = 0; c < 3; ++c) met[q * 3 + c] = zz->metric[p * 9 + dir * 3 + c];
        jac[q] = zz->jac[p];
        efd[q] = (zz->inviscid_type == 2 && s >= -1 && s <= n) ? efix_delta(zz, i, j, k) : 0.0;
      }
      /* face fluxes for faces s+1/2, s = -1 .. n-1 */
      for (int s = -1; s < n; ++s) {
        const int q = s + g;
        double pl[ORC_MAXS + 16], pr[ORC_MAXS + 16];
        if (mixaw)
          aweno_mix(zz, cm, pv, met, q, pl, pr);
        else if (zz->inviscid_type == 1)
          aweno_air(cvb, met, q, pl, pr);
        else
          reconstruct(zz, pv, q, pl, pr);
        riemann(zz, pl, pr, met, jac, efd, q, &fc[(s + 1) * nv]);
      }
      for (int s = 0; s < n; ++s) {
        int i, j, k;
        line_ijk(dir, a, b, s, &i, &j, &k);
        for (int l = 0; l < nv; ++l) DQ(l, i, j, k) -= fc[(s + 1) * nv + l] - fc[s * nv + l];
      }
      if (zz->inviscid_type == 1) {
        /* CDSPart1D<Air>: node fluxes on [-g, n-g), correction on [0, n-2g) */
        for (int s = -g; s < n - g; ++s) {
          int i, j, k;
          line_ijk(dir, a, b, s, &i, &j, &k);
          const int q = s + g;
          const double *mm = &met[q * 3];
          const double J = jac[q];
          const double u = BV(1, i, j, k), v = BV(2, i, j, k), w = BV(3, i, j, k);
          const double Uk = mm[0] * u + mm[1] * v + mm[2] * w;
          const double rU = BV(0, i, j, k) * Uk;
          const double p = BV(4, i, j, k);
          fn[q * nfn + 0] = rU * J;
          fn[q * nfn + 1] = J * (rU * u + mm[0] * p);
          fn[q * nfn + 2] = J * (rU * v + mm[1] * p);
          fn[q * nfn + 3] = J * (rU * w + mm[2] * p);
          fn[q * nfn + 4] = J * Uk * (CV(4, i, j, k) + p);
          /* scalars (mixture build definition): J U rho Y */
          for (int l = 5; l < nfn; ++l) fn[q * nfn + l] = J * Uk * CV(l, i, j, k);
        }
        const double c1 = 19.0 / 3840, c2 = -13.0 / 320, c3 = 17.0 / 256;
        for (int s = 0; s < n - 2 * g; ++s) {
          int i, j, k;
          line_ijk(dir, a, b, s, &i, &j, &k);
          const int q = s + g;
          for (int l = 0; l < nfn; ++l)
            DQ(l, i, j, k) -= c1 * (fn[(q + 3) * nfn + l] - fn[(q - 3) * nfn + l]) +
                              c2 * (fn[(q + 2) * nfn + l] - fn[(q - 2) * nfn + l]) +
                              c3 * (fn[(q + 1) * nfn + l] - fn[(q - 1) * nfn + l]);
        }
      }
    }
  free(pv);
  free(cvb);
  free(cm);
  free(met);
  free(jac);
  free(efd);
  free(fc);
  free(fn);
}

/* ----------------------------------------------------------- viscous */

/* ViscousScheme.cuh:22-288 (fv), 290-557 (gv), 559-821 (hv): second-order
 * central face gradients, Stokes stress, Fourier heat flux and (mixtures)
 * mixture-averaged diffusion with the correction velocity.  Face between
 * p0 and p0 + e_dir. */
static void viscous_face(orc_zone *zz, int dir, int i, int j, int k, double *fv) {
  const int nv = zz->nv, ns = zz->ns;
  int e[3] = {0, 0, 0};
  e[dir] = 1;
  const int i1 = i + e[0], j1 = j + e[1], k1 = k + e[2];
  const double *m = &zz->metric[at(zz->Lg, i, j, k) * 9];
  const double *m1 = &zz->metric[at(zz->Lg, i1, j1, k1) * 9];
  double mm[9];
  for (int c = 0; c < 9; ++c) mm[c] = 0.5 * (m[c] + m1[c]);
  /* derivative along computational direction d of field f at the face */
#define FLD(l, a, b, c) (l < 6 ? BV(l, a, b, c) : SV(l - 6, a, b, c))
  double d_c[3][ORC_MAXS + 8];
  const int nf = 6 + zz->nsc;
  for (int l = 0; l < nf; ++l) {
    if (l == 0 || l == 4) {
      d_c[0][l] = d_c[1][l] = d_c[2][l] = 0;
      if (!(l == 4 && zz->cfg.gradPInDiffusionFlux)) continue;
    }
    for (int d = 0; d < 3; ++d) {
      if (d == dir) {
        d_c[d][l] = FLD(l, i1, j1, k1) - FLD(l, i, j, k);
      } else {
        int o[3] = {0, 0, 0};
        o[d] = 1;
        d_c[d][l] = 0.25 * (FLD(l, i + o[0], j + o[1], k + o[2]) - FLD(l, i - o[0], j - o[1], k - o[2]) +
                            FLD(l, i1 + o[0], j1 + o[1], k1 + o[2]) - FLD(l, i1 - o[0], j1 - o[1], k1 - o[2]));
      }
    }
  }
  /* chain rule: f_x = f_xi xi_x + f_eta eta_x + f_zeta zeta_x */
#define GX(l, c) (d_c[0][l] * mm[0 * 3 + (c)] + d_c[1][l] * mm[1 * 3 + (c)] + d_c[2][l] * mm[2 * 3 + (c)])
  const double u_x = GX(1, 0), u_y = GX(1, 1), u_z = GX(1, 2);
  const double v_x = GX(2, 0), v_y = GX(2, 1), v_z = GX(2, 2);
  const double w_x = GX(3, 0), w_y = GX(3, 1), w_z = GX(3, 2);
  const double mul = 0.5 * (S1(mul, i, j, k) + S1(mul, i1, j1, k1));
  const double mut = zz->nt ? 0.5 * (S1(mut, i, j, k) + S1(mut, i1, j1, k1)) : 0.0;
  const double vis = mul + mut;
  double txx = vis * (4 * u_x - 2 * v_y - 2 * w_z) / 3.0;
  double tyy = vis * (4 * v_y - 2 * u_x - 2 * w_z) / 3.0;
  double tzz = vis * (4 * w_z - 2 * u_x - 2 * v_y) / 3.0;
  const double txy = vis * (u_y + v_x), txz = vis * (u_z + w_x), tyz = vis * (v_z + w_y);
  if (zz->nt) {
    /* SST: -2/3 rho k on the normal stresses, faded out within 1e-3 of a wall */
    const double dyd = fmin(0.5 * (S1(wdist, i, j, k) + S1(wdist, i1, j1, k1)) / 1e-3, 1.0);
    const double rk = -2.0 / 3 * 0.5 * (BV(0, i, j, k) * SV(ns, i, j, k) + BV(0, i1, j1, k1) * SV(ns, i1, j1, k1)) * dyd;
    txx += rk;
    tyy += rk;
    tzz += rk;
  }
  const double J0 = zz->jac[at(zz->Lg, i, j, k)], J1 = zz->jac[at(zz->Lg, i1, j1, k1)];
  const double kx = 0.5 * (m[dir * 3 + 0] * J0 + m1[dir * 3 + 0] * J1);
  const double ky = 0.5 * (m[dir * 3 + 1] * J0 + m1[dir * 3 + 1] * J1);
  const double kz = 0.5 * (m[dir * 3 + 2] * J0 + m1[dir * 3 + 2] * J1);
  fv[0] = 0;
  fv[1] = kx * txx + ky * txy + kz * txz;
  fv[2] = kx * txy + ky * tyy + kz * tyz;
  fv[3] = kx * txz + ky * tyz + kz * tzz;
  const double um = 0.5 * (BV(1, i, j, k) + BV(1, i1, j1, k1));
  const double vm = 0.5 * (BV(2, i, j, k) + BV(2, i1, j1, k1));
  const double wm = 0.5 * (BV(3, i, j, k) + BV(3, i1, j1, k1));
  const double t_x = GX(5, 0), t_y = GX(5, 1), t_z = GX(5, 2);
  double cond = 0.5 * (S1(cond, i, j, k) + S1(cond, i1, j1, k1));
  if (zz->nt) cond += 0.5 * (S1(tcond, i, j, k) + S1(tcond, i1, j1, k1));
  fv[4] = um * fv[1] + vm * fv[2] + wm * fv[3] + cond * (kx * t_x + ky * t_y + kz * t_z);
  if (zz->cfg.mix) {
    const orc_mech *me = &zz->mech;
    double diffusivity[ORC_MAXS], yk[ORC_MAXS], ddf[ORC_MAXS];
    double sum_gy_w = 0, sum_rhoDY = 0, corr = 0, mw_tot = 0;
    for (int l = 0; l < ns; ++l) {
      yk[l] = 0.5 * (SV(l, i, j, k) + SV(l, i1, j1, k1));
      diffusivity[l] = 0.5 * (zz->rhoD[l * zz->Lg.sz + at(zz->Lg, i, j, k)] +
                              zz->rhoD[l * zz->Lg.sz + at(zz->Lg, i1, j1, k1)]) +
                       mut / zz->cfg.Sct;
      const double y_x = GX(6 + l, 0), y_y = GX(6 + l, 1), y_z = GX(6 + l, 2);
      const double gy = kx * y_x + ky * y_y + kz * y_z;
      ddf[l] = gy;
      corr += diffusivity[l] * gy;
      sum_gy_w += gy / me->mw[l];
      mw_tot += yk[l] / me->mw[l];
      sum_rhoDY += diffusivity[l] * yk[l];
    }
    mw_tot = 1.0 / mw_tot;
    corr -= mw_tot * sum_rhoDY * sum_gy_w;
    if (zz->cfg.gradPInDiffusionFlux) {
      const double p_x = GX(4, 0), p_y = GX(4, 1), p_z = GX(4, 2);
      const double gp = (kx * p_x + ky * p_y + kz * p_z) / (0.5 * (BV(4, i1, j1, k1) + BV(4, i, j, k)));
      for (int l = 0; l < ns; ++l) {
        ddf[l] += (mw_tot / me->mw[l] - 1) * yk[l] * gp;
        corr += (mw_tot / me->mw[l] - 1) * yk[l] * gp * diffusivity[l];
      }
    }
    double h[ORC_MAXS];
    const double tm = 0.5 * (BV(5, i, j, k) + BV(5, i1, j1, k1));
    orc_enthalpy(tm, me, h);
    for (int l = 0; l < ns; ++l) {
      const double dflux = diffusivity[l] * (ddf[l] - mw_tot * yk[l] * sum_gy_w) - yk[l] * corr;
      fv[5 + l] = dflux;
      fv[4] += h[l] * dflux;
    }
  }
  if (zz->nt) {
    /* SST diffusion of k and omega, sigma blended with F1 at the face */
    const double k_x = GX(6 + ns, 0), k_y = GX(6 + ns, 1), k_z = GX(6 + ns, 2);
    const double o_x = GX(7 + ns, 0), o_y = GX(7 + ns, 1), o_z = GX(7 + ns, 2);
    const double wd = 0.5 * (S1(wdist, i, j, k) + S1(wdist, i1, j1, k1));
    double f1 = 1;
    if (wd > 1e-25) {
      const double km = 0.5 * (SV(ns, i, j, k) + SV(ns, i1, j1, k1));
      const double om = 0.5 * (SV(ns + 1, i, j, k) + SV(ns + 1, i1, j1, k1));
      const double p1 = sqrt(km) / (0.09 * om * wd);
      const double rhom = 0.5 * (BV(0, i, j, k) + BV(0, i1, j1, k1));
      const double d2 = wd * wd;
      const double p2 = 500 * mul / (rhom * d2 * om);
      const double cdkw = fmax(1e-20, 2 * rhom * SST_SIGMA_OMEGA2 / om * (k_x * o_x + k_y * o_y + k_z * o_z));
      const double p3 = 4 * rhom * SST_SIGMA_OMEGA2 * km / (cdkw * d2);
      const double a1 = fmin(fmax(p1, p2), p3);
      f1 = tanh(a1 * a1 * a1 * a1);
    }
    const double sk = SST_SIGMA_K2 + (SST_SIGMA_K1 - SST_SIGMA_K2) * f1;
    const double sw = SST_SIGMA_OMEGA2 + (SST_SIGMA_OMEGA1 - SST_SIGMA_OMEGA2) * f1;
    fv[5 + ns] = (mul + mut * sk) * (kx * k_x + ky * k_y + kz * k_z);
    fv[6 + ns] = (mul + mut * sw) * (kx * o_x + ky * o_y + kz * o_z);
  }
  (void)nv;
#undef GX
#undef FLD
}

/* SchemeSelector.cuh:49-157 compute_viscous_flux: dq += fv(i+1/2) - fv(i-1/2) */
static void viscous_dir(orc_zone *zz, int dir) {
  const int nv = zz->nv;
  const int mext[3] = {zz->mx, zz->my, zz->mz};
  const int n = mext[dir];
  const int oa = dir == 0 ? 1 : 0, ob = dir == 2 ? 1 : 2;
  double *fv = malloc(sizeof(double) * (n + 1) * nv);
  for (int b = 0; b < mext[ob]; ++b)
    for (int a = 0; a < mext[oa]; ++a) {
      for (int s = -1; s < n; ++s) {
        int i, j, k;
        line_ijk(dir, a, b, s, &i, &j, &k);
        viscous_face(zz, dir, i, j, k, &fv[(s + 1) * nv]);
      }
      for (int s = 0; s < n; ++s) {
        int i, j, k;
        line_ijk(dir, a, b, s, &i, &j, &k);
        for (int l = 0; l < nv; ++l) DQ(l, i, j, k) += fv[(s + 1) * nv + l] - fv[s * nv + l];
      }
    }
  free(fv);
}

/* ------------------------------------------------------------- SST */

/* central point gradient 0.5 (xi_c dF/di + eta_c dF/dj + zeta_c dF/dk) of
 * bv component l (l < 6) or sv component l - 6, as written in SST.cu:52-157 */
static void point_grad(const orc_zone *zz, int l, int i, int j, int k, double *g) {
  const double *m = &zz->metric[at(zz->Lg, i, j, k) * 9];
#define F_(a, b, c) (l < 6 ? BV(l, a, b, c) : SV(l - 6, a, b, c))
  const double di = F_(i + 1, j, k) - F_(i - 1, j, k);
  const double dj = F_(i, j + 1, k) - F_(i, j - 1, k);
  const double dk = F_(i, j, k + 1) - F_(i, j, k - 1);
#undef F_
  for (int c = 0; c < 3; ++c) g[c] = 0.5 * (m[c] * di + m[3 + c] * dj + m[6 + c] * dk);
}

/* SST.cu:44-97 compute_mut: mu_t = a1 rho k / max(a1 omega, |vorticity| F2) */
static double sst_mut(const orc_zone *zz, int i, int j, int k, double mul) {
  double gu[3], gv[3], gw[3];
  point_grad(zz, 1, i, j, k, gu);
  point_grad(zz, 2, i, j, k, gv);
  point_grad(zz, 3, i, j, k, gw);
  const double rho = BV(0, i, j, k), tke = SV(zz->ns, i, j, k), omega = SV(zz->ns + 1, i, j, k);
  const double vort = sqrt((gv[0] - gu[1]) * (gv[0] - gu[1]) + (gw[0] - gu[2]) * (gw[0] - gu[2]) +
                           (gw[1] - gv[2]) * (gw[1] - gv[2]));
  double f2 = 1;
  const double dy = S1(wdist, i, j, k);
  if (dy > 1e-25) {
    const double p1 = 2 * sqrt(tke) / (0.09 * omega * dy);
    const double p2 = 500 * mul / (rho * dy * dy * omega);
    const double a2 = fmax(p1, p2);
    f2 = tanh(a2 * a2);
  }
  const double den = fmax(SST_A1 * omega, vort * f2);
  return den > 1e-25 ? SST_A1 * rho * tke / den : 0.0;
}

/* FieldOperation.cuh:175-195 initialize_mut on [-1, m], with the laminar
 * viscosity from T (Sutherland, or the mixture rule at the point's Y) */
static void initialize_mut(orc_zone *zz) {
  for (int k = -1; k <= zz->mz; ++k)
    for (int j = -1; j <= zz->my; ++j)
      for (int i = -1; i <= zz->mx; ++i) {
        const double t = BV(5, i, j, k);
        double mul = orc_sutherland(t);
        if (zz->cfg.mix) {
          double Y[ORC_MAXS], mw = 0;
          for (int l = 0; l < zz->ns; ++l) {
            Y[l] = SV(l, i, j, k);
            mw += Y[l] / zz->mech.mw[l];
          }
          mul = wilke_viscosity(&zz->mech, t, 1 / mw, Y);
        }
        S1(mut, i, j, k) = sst_mut(zz, i, j, k, mul);
      }
}

/* SST.cu:24-42 compressibility corrections beta* xi* F(Mt) */
double orc_sst_wilcox(double mt) {
  const double mt0 = 0.25;
  return mt - mt0 > 0 ? SST_BETA_STAR * 2 * (mt * mt - mt0 * mt0) : 0.0;
}

double orc_sst_zeman(double mt, double gamma_p1) {
  const double mt0 = 0.25 * sqrt(2.0 / gamma_p1);
  const double dmt = mt - mt0;
  if (dmt <= 0) return 0.0;
  const double lambda2 = 0.66 * 0.66;
  return SST_BETA_STAR * 0.75 * (1 - exp(-0.5 * gamma_p1 * dmt * dmt / lambda2));
}

void orc_sst_constants(double *gamma1, double *gamma2) {
  *gamma1 = SST_GAMMA1;
  *gamma2 = SST_GAMMA2;
}

/* SST.cu:99-247 compute_source_and_mut + SourceTerm.cuh:19-28: F1/F2 blend,
 * mu_t, the k and omega sources (production limited to 10x, replaced by
 * mu_t |vorticity|^2), compressibility correction, the source Jacobian
 * diagonal and lambda_t = mu_t cp / Prt */
static void sst_source_point(orc_zone *zz, int i, int j, int k) {
  const int ns = zz->ns;
  double gu[3], gv[3], gw[3], gk[3], go[3];
  point_grad(zz, 1, i, j, k, gu);
  point_grad(zz, 2, i, j, k, gv);
  point_grad(zz, 3, i, j, k, gw);
  point_grad(zz, 6 + ns, i, j, k, gk);
  point_grad(zz, 7 + ns, i, j, k, go);
  const double u_x = gu[0], u_y = gu[1], u_z = gu[2], v_x = gv[0], v_y = gv[1], v_z = gv[2];
  const double w_x = gw[0], w_y = gw[1], w_z = gw[2];
  const double rho = BV(0, i, j, k);
  const double omega = SV(ns + 1, i, j, k), tke = SV(ns, i, j, k);
  const double inter = 2 * rho * SST_SIGMA_OMEGA2 / omega * (gk[0] * go[0] + gk[1] * go[1] + gk[2] * go[2]);
  const double rhoK = rho * tke;
  const double vort = sqrt((v_x - u_y) * (v_x - u_y) + (w_x - u_z) * (w_x - u_z) + (w_y - v_z) * (w_y - v_z));
  double f1 = 1, f2 = 1;
  const double dy = S1(wdist, i, j, k);
  if (dy > 1e-25) {
    const double p1 = sqrt(tke) / (0.09 * omega * dy);
    const double d2 = dy * dy;
    const double p2 = 500 * S1(mul, i, j, k) / (rho * d2 * omega);
    const double a2 = fmax(2 * p1, p2);
    f2 = tanh(a2 * a2);
    const double cdkw = fmax(1e-20, inter);
    const double p3 = 4 * rho * SST_SIGMA_OMEGA2 * tke / (cdkw * d2);
    const double a1 = fmin(fmax(p1, p2), p3);
    f1 = tanh(a1 * a1 * a1 * a1);
  }
  double mut = 0;
  const double den = fmax(SST_A1 * omega, vort * f2);
  if (den > 1e-25) mut = SST_A1 * rhoK / den;
  S1(mut, i, j, k) = mut;
  double beta = SST_BETA2 + (SST_BETA1 - SST_BETA2) * f1;
  double beta_star = SST_BETA_STAR;
  const int corr = zz->cfg.compressibility_correction;
  if (corr) {
    const double gam = zz->cfg.mix ? S1(gam, i, j, k) : GAMMA_AIR;
    const double mt = sqrt(2 * tke / (gam * BV(4, i, j, k) / rho));
    double c;
    if (corr == 1)
      c = orc_sst_wilcox(mt);
    else if (corr == 2)
      c = SST_BETA_STAR * mt * mt;
    else
      c = orc_sst_zeman(mt, gam + 1.0);
    beta_star += c;
    beta -= c;
  }
  if (mut > 1e-25) {
    const double divU = u_x + v_y + w_z;
    double prod_k = mut * (2 * (u_x * u_x + v_y * v_y + w_z * w_z) - 2.0 / 3 * divU * divU + (u_y + v_x) * (u_y + v_x) +
                           (u_z + w_x) * (u_z + w_x) + (v_z + w_y) * (v_z + w_y)) -
                    2.0 / 3 * rhoK * divU;
    const double prod_k2 = mut * vort * vort;
    if (prod_k > 10 * prod_k2) prod_k = prod_k2;
    const double diss_k = beta_star * rhoK * omega;
    const double J = zz->jac[at(zz->Lg, i, j, k)];
    DQ(5 + ns, i, j, k) += J * (prod_k - diss_k);
    const double gamma = SST_GAMMA2 + (SST_GAMMA1 - SST_GAMMA2) * f1;
    const double prod_w = gamma * rho / mut * prod_k + (1 - f1) * inter;
    const double diss_w = beta * rho * omega * omega;
    DQ(6 + ns, i, j, k) += J * (prod_w - diss_w);
  }
  if (zz->cfg.turb_implicit == 1) {
    const int64_t p = at(zz->L0, i, j, k);
    zz->turb_jac[p] = -2 * beta_star * omega;
    zz->turb_jac[zz->L0.sz + p] = -2 * beta * omega;
  }
  const double cp = zz->cfg.mix ? S1(cp, i, j, k) : GAMMA_AIR * R_U / MW_AIR / (GAMMA_AIR - 1);
  S1(tcond, i, j, k) = mut * cp / zz->cfg.Prt;
}

/* SourceTerm.cuh:11-37 compute_source: SST (mu_t and the k, omega sources),
 * then the finite-rate chemistry */
static void source(orc_zone *zz) {
  if (zz->nt)
    for (int k = 0; k < zz->mz; ++k)
      for (int j = 0; j < zz->my; ++j)
        for (int i = 0; i < zz->mx; ++i) sst_source_point(zz, i, j, k);
  if (!(zz->cfg.mix && zz->cfg.reaction == 1) || zz->cfg.split_chem) return;
  const int ns = zz->ns;
  for (int k = 0; k < zz->mz; ++k)
    for (int j = 0; j < zz->my; ++j)
      for (int i = 0; i < zz->mx; ++i) {
        double Y[ORC_MAXS], om[ORC_MAXS], jc[ORC_MAXS * ORC_MAXS], jd[ORC_MAXS];
        for (int l = 0; l < ns; ++l) Y[l] = SV(l, i, j, k);
        const int meth = zz->cfg.chemSrcMethod;
        orc_chem_source_point(&zz->mech, BV(0, i, j, k), BV(5, i, j, k), Y, om, meth == 1 ? jc : NULL,
                              meth == 2 ? jd : NULL);
        const double J = zz->jac[at(zz->Lg, i, j, k)];
        for (int l = 0; l < ns; ++l) DQ(5 + l, i, j, k) += J * om[l];
        const int64_t p = at(zz->L0, i, j, k);
        if (meth == 1)
          for (int c = 0; c < ns * ns; ++c) zz->chem_jac[c * zz->L0.sz + p] = jc[c];
        if (meth == 2)
          for (int c = 0; c < ns; ++c) zz->chem_jac[c * zz->L0.sz + p] = jd[c];
      }
}

void orc_zone_rhs(orc_zone *zz) {
  memset(zz->dq, 0, sizeof(double) * zz->L0.sz * zz->nv);
  source(zz);
  for (int d = 0; d < zz->dim; ++d) inviscid_dir(zz, d);
  if (zz->cfg.viscous_order != 0)
    for (int d = 0; d < zz->dim; ++d) viscous_dir(zz, d);
}

/* ----------------------------------------------------------- time */

/* TimeAdvanceFunc.cuh:50-100 local_time_step + TimeAdvanceFunc.cu:155-201 global min */
double orc_zone_global_dt(orc_zone *zz) {
  double dt = 1e+6;
  for (int k = 0; k < zz->mz; ++k)
    for (int j = 0; j < zz->my; ++j)
      for (int i = 0; i < zz->mx; ++i) {
        const double *m = &zz->metric[at(zz->Lg, i, j, k) * 9];
        const double gxi = sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        const double geta = sqrt(m[3] * m[3] + m[4] * m[4] + m[5] * m[5]);
        const double gzeta = sqrt(m[6] * m[6] + m[7] * m[7] + m[8] * m[8]);
        const double u = BV(1, i, j, k), v = BV(2, i, j, k), w = BV(3, i, j, k);
        const double U = u * m[0] + v * m[1] + w * m[2];
        const double V = u * m[3] + v * m[4] + w * m[5];
        const double W = u * m[6] + v * m[7] + w * m[8];
        const double a = S1(acoustic, i, j, k);
        const int64_t p = at(zz->L0, i, j, k);
        double *isr = &zz->inv_sr[p * 3];
        isr[0] = fabs(U) + a * gxi;
        isr[1] = fabs(V) + a * geta;
        isr[2] = 0;
        if (zz->dim == 3) isr[2] = fabs(W) + a * gzeta;
        const double sri = isr[0] + isr[1] + isr[2];
        double gamma = GAMMA_AIR;
        if (zz->cfg.mix) gamma = S1(gam, i, j, k);
        const double c1 = fmax(gamma, 4.0 / 3.0) / BV(0, i, j, k);
        double c2 = S1(mul, i, j, k) / zz->cfg.Pr;
        if (zz->nt) c2 += S1(mut, i, j, k) / zz->cfg.Prt;
        double srv = gxi * gxi + geta * geta;
        if (zz->dim == 3) srv += gzeta * gzeta;
        srv *= c1 * c2;
        zz->visc_sr[p] = srv;
        zz->dt_local[p] = zz->cfg.cfl / (sri + srv);
        if (zz->dt_local[p] < dt) dt = zz->dt_local[p];
      }
  return dt;
}

/* RK.cuh:10-14 SSPRK3 coefficients and RK.cuh:197-253 update_cv_and_bv_rk */
static const double RK_A[3] = {1.0, 0.75, 1.0 / 3.0}, RK_B[3] = {0.0, 0.25, 2.0 / 3.0},
                    RK_C[3] = {1.0, 0.25, 2.0 / 3.0};

static void rk_update(orc_zone *zz, double dt, int rk) {
  const double *A = RK_A, *B = RK_B, *C = RK_C;
  for (int k = 0; k < zz->mz; ++k)
    for (int j = 0; j < zz->my; ++j)
      for (int i = 0; i < zz->mx; ++i) {
        const double dtj = dt / zz->jac[at(zz->Lg, i, j, k)];
        for (int l = 0; l < zz->nv; ++l)
          CV(l, i, j, k) = A[rk] * QN(l, i, j, k) + B[rk] * CV(l, i, j, k) + C[rk] * dtj * DQ(l, i, j, k);
        if (zz->mz == 1) CV(3, i, j, k) = 0;
        BV(0, i, j, k) = CV(0, i, j, k);
        const double ri = 1.0 / CV(0, i, j, k);
        BV(1, i, j, k) = CV(1, i, j, k) * ri;
        BV(2, i, j, k) = CV(2, i, j, k) * ri;
        BV(3, i, j, k) = CV(3, i, j, k) * ri;
        const double V2 = BV(1, i, j, k) * BV(1, i, j, k) + BV(2, i, j, k) * BV(2, i, j, k) +
                          BV(3, i, j, k) * BV(3, i, j, k);
        for (int l = 0; l < zz->nsc; ++l) SV(l, i, j, k) = CV(5 + l, i, j, k) * ri;
        if (zz->cfg.mix) {
          double Y[ORC_MAXS], p;
          for (int l = 0; l < zz->ns; ++l) Y[l] = SV(l, i, j, k);
          const double v3[3] = {BV(1, i, j, k), BV(2, i, j, k), BV(3, i, j, k)};
          const double t = orc_temperature_newton(&zz->mech, BV(0, i, j, k), CV(4, i, j, k), v3, Y, BV(5, i, j, k),
                                                  &p, NULL);
          BV(5, i, j, k) = t;
          BV(4, i, j, k) = p;
        } else {
          BV(4, i, j, k) = (GAMMA_AIR - 1) * (CV(4, i, j, k) - 0.5 * BV(0, i, j, k) * V2);
          BV(5, i, j, k) = BV(4, i, j, k) * MW_AIR * ri / R_U;
        }
        S1(vel, i, j, k) = sqrt(V2);
      }
}

static void store_last_step(orc_zone *zz) {
  for (int k = 0; k < zz->mz; ++k)
    for (int j = 0; j < zz->my; ++j)
      for (int i = 0; i < zz->mx; ++i) {
        const int64_t p = at(zz->L0, i, j, k);
        zz->bv_last[p] = BV(0, i, j, k);
        zz->bv_last[zz->L0.sz + p] = S1(vel, i, j, k);
        zz->bv_last[2 * zz->L0.sz + p] = BV(4, i, j, k);
        zz->bv_last[3 * zz->L0.sz + p] = BV(5, i, j, k);
      }
}

/* Residual.cuh:57-161 + TimeAdvanceFunc.cu:19-33: RMS of the change of
 * (rho, |V|, p, T), scaled by the first step's value (floor 1e-20). */
static void residual_sums(orc_zone *zz, double *s) {
  for (int k = 0; k < zz->mz; ++k)
    for (int j = 0; j < zz->my; ++j)
      for (int i = 0; i < zz->mx; ++i) {
        const int64_t p = at(zz->L0, i, j, k);
        const double cur[4] = {BV(0, i, j, k), S1(vel, i, j, k), BV(4, i, j, k), BV(5, i, j, k)};
        for (int l = 0; l < 4; ++l) {
          const double d = cur[l] - zz->bv_last[l * zz->L0.sz + p];
          zz->bv_last[l * zz->L0.sz + p] = d * d;
          s[l] += d * d;
        }
      }
}

/* Residual.cuh:113-139: RMS over all n points, scaled by the first step's */
static double residual_finish(const double *s, int64_t n, int step, double *res, double *res_scale) {
  for (int l = 0; l < 4; ++l) res[l] = sqrt(s[l] / (double)n);
  if (step == 1)
    for (int l = 0; l < 4; ++l) {
      res_scale[l] = res[l];
      if (res_scale[l] < 1e-20) res_scale[l] = 1e-20;
    }
  for (int l = 0; l < 4; ++l) res[l] /= res_scale[l];
  double err = res[0];
  for (int l = 1; l < 4; ++l)
    if (fabs(res[l]) > err) err = res[l];
  return err;
}

static double residual(orc_zone *zz, int step) {
  double s[4] = {0, 0, 0, 0};
  residual_sums(zz, s);
  return residual_finish(s, (int64_t)zz->mx * zz->my * zz->mz, step, zz->res, zz->res_scale);
}

/* RK.cuh:89-193 RK3_bv loop (single block, no IO/monitors) */
void orc_zone_run_rk3(orc_zone *zz, int nsteps, double *dt_hist, double *res_hist) {
  const int64_t ncv = zz->Lg.sz * zz->nv;
  const int os = zz->cfg.output_screen > 0 ? zz->cfg.output_screen : 1;
  for (int it = 0; it < nsteps; ++it) {
    const int step = ++zz->step;
    double dt = zz->cfg.fixed_time_step ? zz->cfg.dt : 1e+6;
    if (step % os == 0) store_last_step(zz);
    memcpy(zz->qn, zz->cv, sizeof(double) * ncv);
    for (int rk = 0; rk < 3; ++rk) {
      orc_zone_rhs(zz);
      if (!zz->cfg.fixed_time_step) dt = orc_zone_global_dt(zz);
      rk_update(zz, dt, rk);
      apply_bcs(zz, 1);
      if (zz->dim == 2) eliminate_k_gradient(zz, 1);
      update_properties(zz);
      if (dt_hist) dt_hist[it * 3 + rk] = dt;
    }
    zz->t_phys += dt;
    if (step % os == 0 || step == 1) {
      const double err = residual(zz, step);
      if (res_hist) {
        for (int l = 0; l < 4; ++l) res_hist[it * 5 + l] = zz->res[l];
        res_hist[it * 5 + 4] = err;
      }
    } else if (res_hist) {
      for (int l = 0; l < 5; ++l) res_hist[it * 5 + l] = NAN;
    }
  }
}

/* ------------------------------------------------------- test hooks */
/* Point-level access to the oracle's own face routines so the tests can pin
 * them against closed forms (consistency, supersonic upwinding, exactness of
 * WENO5 on low-order data, L R = I, the RK3 stability polynomial). */

/* ----------------------------------------------------------- steady path */

/* FiniteRateChem.cu:306-362 solve_chem_system / 234-256 EPI, EPI_for_dq0:
 * (diag I - dt J) x = dq_species in place on the cell's dq */
static void epi_in_place(orc_zone *zz, int i, int j, int k, double diag, double dt) {
  const int ns = zz->ns;
  const int64_t p = at(zz->L0, i, j, k);
  double lhs[ORC_MAXS * ORC_MAXS], rhs[ORC_MAXS];
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b < ns; ++b)
      lhs[a * ns + b] = (a == b ? diag : 0.0) - dt * zz->chem_jac[(a * ns + b) * zz->L0.sz + p];
  for (int l = 0; l < ns; ++l) rhs[l] = DQ(5 + l, i, j, k);
  orc_lu_solve(lhs, rhs, ns);
  for (int l = 0; l < ns; ++l) DQ(5 + l, i, j, k) = rhs[l];
}

/* finite-rate chemistry inside the implicit treatment: off when the
 * chemistry is operator-split (the right-hand side then has no source) */
static int is_fr(const orc_zone *zz) { return zz->cfg.mix && zz->cfg.reaction == 1 && !zz->cfg.split_chem; }

/* DPLUR.cuh:14-78 compute_DQ_0 */
static void dplur_dq0(orc_zone *zz, double *dq0) {
  const int nv = zz->nv, ns = zz->ns;
  for (int k = 0; k < zz->mz; ++k)
    for (int j = 0; j < zz->my; ++j)
      for (int i = 0; i < zz->mx; ++i) {
        const int64_t p = at(zz->L0, i, j, k);
        const double dt = zz->dt_local[p];
        for (int l = 0; l < nv; ++l) DQ(l, i, j, k) *= dt;
        const double *sr = &zz->inv_sr[p * 3];
        const double diag = 1 + dt * (sr[0] + sr[1] + sr[2]);
        for (int l = 0; l < 5; ++l) DQ(l, i, j, k) /= diag;
        if (is_fr(zz) && zz->cfg.chemSrcMethod == 1) {
          epi_in_place(zz, i, j, k, diag, dt);
        } else if (is_fr(zz) && zz->cfg.chemSrcMethod == 2) {
          for (int l = 0; l < ns; ++l) DQ(5 + l, i, j, k) /= diag - dt * zz->chem_jac[l * zz->L0.sz + p];
        } else {
          for (int l = 0; l < ns; ++l) DQ(5 + l, i, j, k) /= diag;
        }
        /* SST.cu:249-258 implicit_treat_for_dq0 */
        for (int l = 0; l < zz->nt; ++l)
          DQ(5 + ns + l, i, j, k) /= zz->cfg.turb_implicit == 1 ? diag - dt * zz->turb_jac[l * zz->L0.sz + p] : diag;
      }
  memcpy(dq0, zz->dq, sizeof(double) * zz->L0.sz * nv);
}

/* DPLUR.cuh:80-137 compute_jacobian_times_dq: (A +- lam I) dq at one point
 * with primitive state (rho, u, v, w, p, T), |V|, the scalars sv (species
 * then k, omega), the mixture gamma (mix only), the metric row m of
 * direction dir and the signed spectral radius lam */
static void jac_times_dq_point(int mix, int ns, int nsc, const orc_mech *mech, const double *b6, double vel,
                               const double *sv, double gam, const double *m, double lam, const double *dqin,
                               double *out) {
  const int nv = 5 + nsc;
  const double u = b6[1], v = b6[2], w = b6[3];
  const double U = m[0] * u + m[1] * v + m[2] * w;
  const double l1 = U + lam;
  const double e = 0.5 * vel * vel;
  double gamma = GAMMA_AIR, b3 = 0, b4 = 0, h = 0;
  double dq[5 + ORC_MAXS + 4];
  for (int l = 0; l < nv; ++l) dq[l] = dqin[l];
  if (!mix) {
    h = gamma / (gamma - 1) * b6[4] / b6[0] + e;
  } else {
    double hs[ORC_MAXS];
    const double t = b6[5];
    orc_enthalpy(t, mech, hs);
    gamma = gam;
    for (int l = 0; l < ns; ++l) {
      b3 += R_U * t / mech->mw[l] * dq[5 + l];
      b4 += hs[l] * dq[5 + l];
      h += sv[l] * hs[l];
    }
    b3 *= gamma;
    b4 *= gamma - 1;
    h += e;
  }
  const double b1 = m[0] * dq[1] + m[1] * dq[2] + m[2] * dq[3] - U * dq[0];
  const double b2 = (gamma - 1) * (e * dq[0] - u * dq[1] - v * dq[2] - w * dq[3] + dq[4]);
  out[0] = b1 + l1 * dq[0];
  out[1] = u * b1 + m[0] * b2 + l1 * dq[1] + m[0] * (b3 - b4);
  out[2] = v * b1 + m[1] * b2 + l1 * dq[2] + m[1] * (b3 - b4);
  out[3] = w * b1 + m[2] * b2 + l1 * dq[3] + m[2] * (b3 - b4);
  out[4] = h * b1 + U * b2 + l1 * dq[4] + U * (b3 - b4);
  for (int l = 0; l < nsc; ++l) out[5 + l] = l1 * dq[5 + l] + sv[l] * b1;
}

/* the same at zone point (i,j,k) with its current dq */
static void jac_times_dq(const orc_zone *zz, int i, int j, int k, int dir, double lam, double *out) {
  const int64_t pg = at(zz->Lg, i, j, k);
  double b6[6], sv[ORC_MAXS + 4], dq[5 + ORC_MAXS + 4];
  for (int l = 0; l < 6; ++l) b6[l] = BV(l, i, j, k);
  for (int l = 0; l < zz->nsc; ++l) sv[l] = SV(l, i, j, k);
  for (int l = 0; l < zz->nv; ++l) dq[l] = DQ(l, i, j, k);
  jac_times_dq_point(zz->cfg.mix, zz->ns, zz->nsc, &zz->mech, b6, zz->vel[pg], sv, zz->cfg.mix ? zz->gam[pg] : 0.0,
                     &zz->metric[pg * 9 + dir * 3], lam, dq, out);
}

/* DPLUR.cu:20-52 set_wall_dq_to_0 */
static void wall_dq_zero(orc_zone *zz) {
  for (int n = 0; n < zz->nbc; ++n) {
    const int b = zz->order[n];
    if (zz->cfg.bc[b].type != 2) continue;
    for (int k = zz->rs[b][2]; k <= zz->re[b][2]; ++k)
      for (int j = zz->rs[b][1]; j <= zz->re[b][1]; ++j)
        for (int i = zz->rs[b][0]; i <= zz->re[b][0]; ++i)
          for (int l = 0; l < zz->nv; ++l) DQ(l, i, j, k) = 0;
  }
}

/* DPLUR.cuh:139-257 DPLUR_inner_iteration: dqk from the neighbours' current
 * dq (interior neighbours only, no halo exchange), then dq = dqk */
static void dplur_inner(orc_zone *zz, const double *dq0, double *dqk) {
  const int nv = zz->nv, ns = zz->ns;
  const int ext[3] = {zz->mx, zz->my, zz->mz};
  for (int k = 0; k < zz->mz; ++k)
    for (int j = 0; j < zz->my; ++j)
      for (int i = 0; i < zz->mx; ++i) {
        double tot[5 + ORC_MAXS + 4] = {0}, jd[5 + ORC_MAXS + 4];
        const int c[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) {
          if (c[d] > 0) {
            int n3[3] = {i, j, k};
            n3[d] -= 1;
            jac_times_dq(zz, n3[0], n3[1], n3[2], d, zz->inv_sr[at(zz->L0, n3[0], n3[1], n3[2]) * 3 + d], jd);
            for (int l = 0; l < nv; ++l) tot[l] += 0.5 * jd[l];
          }
        }
        for (int d = 0; d < 3; ++d) {
          if (c[d] != ext[d] - 1) {
            int n3[3] = {i, j, k};
            n3[d] += 1;
            jac_times_dq(zz, n3[0], n3[1], n3[2], d, -zz->inv_sr[at(zz->L0, n3[0], n3[1], n3[2]) * 3 + d], jd);
            for (int l = 0; l < nv; ++l) tot[l] -= 0.5 * jd[l];
          }
        }
        const int64_t p = at(zz->L0, i, j, k);
        const double dt = zz->dt_local[p];
        const double *sr = &zz->inv_sr[p * 3];
        const double diag = 1 + dt * (sr[0] + sr[1] + sr[2]);
        for (int l = 0; l < 5; ++l) dqk[l * zz->L0.sz + p] = dq0[l * zz->L0.sz + p] + dt * tot[l] / diag;
        if (is_fr(zz) && zz->cfg.chemSrcMethod == 1) {
          double lhs[ORC_MAXS * ORC_MAXS], rhs[ORC_MAXS];
          for (int a = 0; a < ns; ++a)
            for (int b = 0; b < ns; ++b)
              lhs[a * ns + b] = (a == b ? diag : 0.0) - dt * zz->chem_jac[(a * ns + b) * zz->L0.sz + p];
          for (int l = 0; l < ns; ++l) rhs[l] = tot[5 + l] * dt;
          orc_lu_solve(lhs, rhs, ns);
          for (int l = 0; l < ns; ++l) dqk[(5 + l) * zz->L0.sz + p] = dq0[(5 + l) * zz->L0.sz + p] + rhs[l];
        } else if (is_fr(zz) && zz->cfg.chemSrcMethod == 2) {
          for (int l = 0; l < ns; ++l)
            dqk[(5 + l) * zz->L0.sz + p] =
                dq0[(5 + l) * zz->L0.sz + p] + dt * tot[5 + l] / (diag - dt * zz->chem_jac[l * zz->L0.sz + p]);
        } else {
          for (int l = 0; l < ns; ++l) dqk[(5 + l) * zz->L0.sz + p] = dq0[(5 + l) * zz->L0.sz + p] + dt * tot[5 + l] / diag;
        }
        /* SST.cu:260-273 implicit_treat_for_dqk */
        for (int l = 0; l < zz->nt; ++l) {
          const int c = 5 + ns + l;
          const double dg = zz->cfg.turb_implicit == 1 ? diag - dt * zz->turb_jac[l * zz->L0.sz + p] : diag;
          dqk[c * zz->L0.sz + p] = dq0[c * zz->L0.sz + p] + dt * tot[c] / dg;
        }
      }
  memcpy(zz->dq, dqk, sizeof(double) * zz->L0.sz * nv);
}

/* ImplicitTreatmentHPP.cuh:7-52 */
void orc_zone_implicit(orc_zone *zz) {
  const int ns = zz->ns;
  if (zz->cfg.implicit_method == 1) {
    const int64_t n = zz->L0.sz * zz->nv;
    double *dq0 = malloc(sizeof(double) * n + 8), *dqk = malloc(sizeof(double) * n + 8);
    dplur_dq0(zz, dq0);
    for (int it = 0; it < zz->cfg.dplur_inner_step; ++it) {
      wall_dq_zero(zz);
      dplur_inner(zz, dq0, dqk);
    }
    /* DPLUR.cu:5-17 convert_dq_back_to_dqDt */
    for (int k = 0; k < zz->mz; ++k)
      for (int j = 0; j < zz->my; ++j)
        for (int i = 0; i < zz->mx; ++i)
          for (int l = 0; l < zz->nv; ++l) DQ(l, i, j, k) /= zz->dt_local[at(zz->L0, i, j, k)];
    free(dq0);
    free(dqk);
    return;
  }
  for (int k = 0; k < zz->mz; ++k)
    for (int j = 0; j < zz->my; ++j)
      for (int i = 0; i < zz->mx; ++i) {
        const int64_t p = at(zz->L0, i, j, k);
        const double dt = zz->dt_local[p];
        if (is_fr(zz) && zz->cfg.chemSrcMethod == 1) epi_in_place(zz, i, j, k, 1.0, dt);
        if (is_fr(zz) && zz->cfg.chemSrcMethod == 2) /* FiniteRateChem.cu:420-432 DA */
          for (int l = 0; l < ns; ++l) DQ(5 + l, i, j, k) /= 1 - dt * zz->chem_jac[l * zz->L0.sz + p];
        /* SST.cu:8-22 implicit_treat_for_SST */
        if (zz->nt && zz->cfg.turb_implicit == 1)
          for (int l = 0; l < 2; ++l) DQ(5 + ns + l, i, j, k) /= 1 - dt * zz->turb_jac[l * zz->L0.sz + p];
      }
}

/* FieldOperation.cuh:221-296 update_bv: local-dt update of the primitives */
static void update_bv(orc_zone *zz) {
  for (int k = 0; k < zz->mz; ++k)
    for (int j = 0; j < zz->my; ++j)
      for (int i = 0; i < zz->mx; ++i) {
        const int64_t pg = at(zz->Lg, i, j, k);
        const double rho_n = BV(0, i, j, k);
        /* compute_total_energy_1_point from the old primitives */
        const double u0 = BV(1, i, j, k), v0 = BV(2, i, j, k), w0 = BV(3, i, j, k);
        double E_n = 0.5 * (u0 * u0 + v0 * v0 + w0 * w0);
        if (zz->cfg.mix) {
          double h[ORC_MAXS];
          orc_enthalpy(BV(5, i, j, k), &zz->mech, h);
          for (int l = 0; l < zz->ns; ++l) E_n += h[l] * SV(l, i, j, k);
          E_n = E_n * rho_n - BV(4, i, j, k);
        } else {
          E_n = E_n * rho_n + BV(4, i, j, k) / (GAMMA_AIR - 1);
        }
        const double dtj = zz->dt_local[at(zz->L0, i, j, k)] / zz->jac[pg];
        BV(0, i, j, k) += DQ(0, i, j, k) * dtj;
        const double ri = 1.0 / BV(0, i, j, k);
        BV(1, i, j, k) = ri * (rho_n * BV(1, i, j, k) + DQ(1, i, j, k) * dtj);
        BV(2, i, j, k) = ri * (rho_n * BV(2, i, j, k) + DQ(2, i, j, k) * dtj);
        if (zz->mz == 1)
          BV(3, i, j, k) = 0;
        else
          BV(3, i, j, k) = ri * (rho_n * BV(3, i, j, k) + DQ(3, i, j, k) * dtj);
        const double E = E_n + DQ(4, i, j, k) * dtj;
        for (int l = 0; l < zz->nsc; ++l) SV(l, i, j, k) = ri * (rho_n * SV(l, i, j, k) + DQ(5 + l, i, j, k) * dtj);
        const double V2 = BV(1, i, j, k) * BV(1, i, j, k) + BV(2, i, j, k) * BV(2, i, j, k) +
                          BV(3, i, j, k) * BV(3, i, j, k);
        if (zz->inviscid_type == 1) {
          /* build definition: the AWENO path reconstructs cv, so the steady
           * update keeps cv the updated conservative state (the reference
           * leaves cv at its initial values, FieldOperation.cuh:221-296) */
          const double rho = BV(0, i, j, k);
          CV(0, i, j, k) = rho;
          CV(1, i, j, k) = rho * BV(1, i, j, k);
          CV(2, i, j, k) = rho * BV(2, i, j, k);
          CV(3, i, j, k) = rho * BV(3, i, j, k);
          CV(4, i, j, k) = E;
          for (int l = 0; l < zz->nsc; ++l) CV(5 + l, i, j, k) = rho * SV(l, i, j, k);
        }
        if (zz->cfg.mix) {
          double Y[ORC_MAXS], p;
          for (int l = 0; l < zz->ns; ++l) Y[l] = SV(l, i, j, k);
          const double v3[3] = {BV(1, i, j, k), BV(2, i, j, k), BV(3, i, j, k)};
          const double t = orc_temperature_newton(&zz->mech, BV(0, i, j, k), E, v3, Y, BV(5, i, j, k), &p, NULL);
          BV(5, i, j, k) = t;
          BV(4, i, j, k) = p;
        } else {
          BV(4, i, j, k) = (GAMMA_AIR - 1) * (E - 0.5 * BV(0, i, j, k) * V2);
          BV(5, i, j, k) = BV(4, i, j, k) * MW_AIR * ri / R_U;
        }
        zz->vel[pg] = sqrt(V2);
      }
}

/* one point-implicit step dt of every interior cell at its density and
 * total energy: gather, orc_chem_batch, scatter (Y, T, p and the rho Y rows
 * of cv) */
static void chem_split_interior(orc_zone *zz, double dt) {
  const int ns = zz->ns;
  const int64_t n = (int64_t)zz->mx * zz->my * zz->mz;
  double *rho = malloc(sizeof(double) * n * 21 + 8);
  double *T = rho + n, *Y = T + n, *dY = Y + ns * n, *dT = dY + ns * n;
  int64_t c = 0;
  for (int k = 0; k < zz->mz; ++k)
    for (int j = 0; j < zz->my; ++j)
      for (int i = 0; i < zz->mx; ++i, ++c) {
        rho[c] = BV(0, i, j, k);
        T[c] = BV(5, i, j, k);
        for (int l = 0; l < ns; ++l) Y[l * n + c] = SV(l, i, j, k);
      }
  orc_chem_batch(&zz->mech, n, rho, T, Y, dt, dY, dT);
  c = 0;
  for (int k = 0; k < zz->mz; ++k)
    for (int j = 0; j < zz->my; ++j)
      for (int i = 0; i < zz->mx; ++i, ++c) {
        double mwi = 0;
        for (int l = 0; l < ns; ++l) {
          const double y = SV(l, i, j, k) + dY[l * n + c];
          SV(l, i, j, k) = y;
          CV(5 + l, i, j, k) = BV(0, i, j, k) * y;
          mwi += y / zz->mech.mw[l];
        }
        BV(5, i, j, k) += dT[c];
        BV(4, i, j, k) = BV(0, i, j, k) * BV(5, i, j, k) * R_U * mwi;
      }
  free(rho);
}

void orc_zone_chem_step(orc_zone *zz, double dt) {
  chem_split_interior(zz, dt);
  apply_bcs(zz, 1);
  if (zz->dim == 2) eliminate_k_gradient(zz, 1);
  update_properties(zz);
}

/* SteadySim.cuh:52-124 (single block, no data communication) */
void orc_zone_run_steady(orc_zone *zz, int nsteps, double *res_hist) {
  const int os = zz->cfg.output_screen > 0 ? zz->cfg.output_screen : 1;
  if (zz->step == 0) store_last_step(zz);
  for (int it = 0; it < nsteps; ++it) {
    const int step = ++zz->step;
    if (step % os == 0) store_last_step(zz);
    orc_zone_rhs(zz);
    (void)orc_zone_global_dt(zz);
    orc_zone_implicit(zz);
    update_bv(zz);
    if (zz->cfg.split_chem && zz->cfg.steady_chem_dt > 0) chem_split_interior(zz, zz->cfg.steady_chem_dt);
    /* ghost cv follows for the AWENO path, which reads it */
    apply_bcs(zz, zz->inviscid_type == 1);
    if (zz->dim == 2) eliminate_k_gradient(zz, zz->inviscid_type == 1);
    update_properties(zz);
    if (step % os == 0 || step == 1) {
      const double err = residual(zz, step);
      if (res_hist) {
        for (int l = 0; l < 4; ++l) res_hist[it * 5 + l] = zz->res[l];
        res_hist[it * 5 + 4] = err;
      }
    } else if (res_hist) {
      for (int l = 0; l < 5; ++l) res_hist[it * 5 + l] = NAN;
    }
  }
}

/* air face flux: pl/pr = (rho,u,v,w,p) on each side, met6 = metric row of
 * the sweep direction at the left and right points, jac2 their Jacobians */
void orc_face_flux_air(int scheme, const double *pl5, const double *pr5, const double *met6, const double *jac2,
                       double *f5) {
  orc_zone zz;
  memset(&zz, 0, sizeof(zz));
  zz.nv = 5;
  zz.cfg.inviscid_scheme = scheme;
  double pl[8], pr[8];
  for (int l = 0; l < 5; ++l) {
    pl[l] = pl5[l];
    pr[l] = pr5[l];
  }
  pl[5] = 0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]) * pl[0] + pl[4] / (GAMMA_AIR - 1);
  pr[5] = 0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]) * pr[0] + pr[4] / (GAMMA_AIR - 1);
  riemann(&zz, pl, pr, met6, jac2, NULL, 0, f5);
}

/* Roe face flux of air (scheme 2) with the entropy-fix deltas of the two points */
void orc_face_flux_roe_air(const double *pl5, const double *pr5, const double *met6, const double *jac2,
                           const double *delta2, double *f5) {
  orc_zone zz;
  memset(&zz, 0, sizeof(zz));
  zz.nv = 5;
  zz.cfg.inviscid_scheme = 2;
  double pl[8], pr[8];
  for (int l = 0; l < 5; ++l) {
    pl[l] = pl5[l];
    pr[l] = pr5[l];
  }
  pl[5] = 0.5 * (pl[1] * pl[1] + pl[2] * pl[2] + pl[3] * pl[3]) * pl[0] + pl[4] / (GAMMA_AIR - 1);
  pr[5] = 0.5 * (pr[1] * pr[1] + pr[2] * pr[2] + pr[3] * pr[3]) * pr[0] + pr[4] / (GAMMA_AIR - 1);
  flux_roe(&zz, pl, pr, met6, jac2, delta2, 0, f5);
}

/* WENO5 of scalar data v[0..5] at points i-2..i+3: v-(i+1/2), v+(i+1/2) */
void orc_weno5_scalar(const double *v6, double *vm, double *vp) {
  double L[25] = {0};
  double cvb[6 * 7] = {0};
  L[0] = 1.0;
  for (int s = 0; s < 6; ++s) cvb[s * 7] = v6[s];
  weno5(L, cvb, 2, 0, vm, vp);
}

/* AWENO characteristic matrices at the face between two states
 * c7 = (rho, rho u, rho v, rho w, rho E, p, T) and the face metric rows */
void orc_aweno_matrices(const double *cl7, const double *cr7, const double *met6, double *L25, double *R25) {
  double cvb[6 * 7];
  for (int s = 0; s < 6; ++s)
    for (int c = 0; c < 7; ++c) cvb[s * 7 + c] = s < 3 ? cl7[c] : cr7[c];
  double met[18];
  for (int s = 0; s < 6; ++s)
    for (int c = 0; c < 3; ++c) met[s * 3 + c] = s < 3 ? met6[c] : met6[3 + c];
  double pl[8], pr[8];
  aweno_air_m(cvb, met, 2, pl, pr, L25, R25);
}

/* Face flux of Riemann solver `scheme` (1 LF, 2 Roe, 3 AUSM+, 4 HLLC) for one
 * state pair of air (mech NULL) or a mixture: pl/pr = rho, u, v, w, p, then
 * the n_spec mass fractions (mixtures); met6 the metric rows of the two
 * points, jac2 their Jacobians, delta2 their entropy-fix deltas (Roe; NULL
 * = 0).  The states pass through the first-order reconstruction, which
 * appends rho E (and gamma for mixtures) as the sweeps do. */
void orc_face_flux(int scheme, const orc_mech *mech, const double *plx, const double *prx, const double *met6,
                   const double *jac2, const double *delta2, double *f) {
  orc_zone *zz = calloc(1, sizeof(orc_zone));
  zz->ns = mech ? mech->n_spec : 0;
  zz->nsc = zz->ns;
  zz->nv = 5 + zz->ns;
  zz->cfg.mix = mech != NULL;
  if (mech) zz->mech = *mech;
  zz->cfg.inviscid_scheme = scheme;
  zz->cfg.reconstruction = 1;
  zz->inviscid_type = scheme == 2 ? 2 : 0;
  const int nv = zz->nv;
  double pv[4 * (5 + ORC_MAXS)], met[12], jac[4], efd[4], pl[ORC_MAXS + 16], pr[ORC_MAXS + 16];
  for (int q = 0; q < 4; ++q) {
    const double *src = q < 2 ? plx : prx;
    for (int l = 0; l < nv; ++l) pv[q * nv + l] = src[l];
    for (int c = 0; c < 3; ++c) met[q * 3 + c] = met6[(q < 2 ? 0 : 3) + c];
    jac[q] = jac2[q < 2 ? 0 : 1];
    efd[q] = delta2 ? delta2[q < 2 ? 0 : 1] : 0.0;
  }
  reconstruct(zz, pv, 1, pl, pr);
  riemann(zz, pl, pr, met, jac, efd, 1, f);
  free(zz);
}

/* Characteristic matrices of the mixture AWENO at the face between two
 * points c = (Q[5 + nsc], p, T) with metric rows k0, k1: L and R
 * (row-major (5 + nsc)^2) from the projections of unit vectors.  mech NULL:
 * air with nsc transported scalars. */
void orc_aweno_mix_matrices(const orc_mech *mech, int nsc, const double *cl, const double *cr, const double *k0,
                            const double *k1, double *L, double *R) {
  orc_zone *zz = calloc(1, sizeof(orc_zone));
  zz->ns = mech ? mech->n_spec : 0;
  zz->nsc = nsc;
  zz->nv = 5 + nsc;
  zz->cfg.mix = mech != NULL;
  if (mech) zz->mech = *mech;
  const int nv = zz->nv;
  mix_eig E;
  mix_roe_state(zz, cl, cr, k0, k1, &E);
  double e[5 + ORC_MAXS + 4], col[5 + ORC_MAXS + 4];
  for (int c = 0; c < nv; ++c) {
    for (int l = 0; l < nv; ++l) e[l] = l == c ? 1.0 : 0.0;
    mix_left(zz, &E, e, col);
    for (int r = 0; r < nv; ++r) L[r * nv + c] = col[r];
    mix_right(zz, &E, e, col);
    for (int r = 0; r < nv; ++r) R[r * nv + c] = col[r];
  }
  free(zz);
}

/* Mixture AWENO face states from six stencil points: c6 holds
 * (Q[5 + nsc], p, T) of points i-2 .. i+3 (stride 7 + nsc), k0/k1 the metric
 * rows of points i and i+1; pl/pr receive (rho, u, v, w, p, scalars, rho E,
 * gamma).  Returns 1 when the sides failed the admissibility test and took
 * the first-order values. */
int orc_aweno_mix_sides(const orc_mech *mech, int nsc, const double *c6, const double *k0, const double *k1,
                        double *pl, double *pr) {
  orc_zone *zz = calloc(1, sizeof(orc_zone));
  zz->ns = mech ? mech->n_spec : 0;
  zz->nsc = nsc;
  zz->nv = 5 + nsc;
  zz->cfg.mix = mech != NULL;
  if (mech) zz->mech = *mech;
  const int nv = zz->nv, st = nv + 2;
  double met[6 * 3], pv[6 * (5 + ORC_MAXS + 4)];
  for (int q = 0; q < 6; ++q) {
    const double *k = q < 3 ? k0 : k1;
    for (int c = 0; c < 3; ++c) met[q * 3 + c] = k[c];
    const double *Q = &c6[q * st];
    pv[q * nv] = Q[0];
    for (int l = 1; l < 4; ++l) pv[q * nv + l] = Q[l] / Q[0];
    pv[q * nv + 4] = Q[nv];
    for (int l = 5; l < nv; ++l) pv[q * nv + l] = Q[l] / Q[0];
  }
  aweno_mix(zz, c6, pv, met, 2, pl, pr);
  double fl[5 + ORC_MAXS + 6], fr[5 + ORC_MAXS + 6];
  reconstruct_order(zz, pv, 2, 1, fl, fr);
  int same = 1;
  for (int l = 0; l < nv + 2; ++l) same = same && fl[l] == pl[l] && fr[l] == pr[l];
  free(zz);
  return same;
}

/* Reconstruction.cuh:20-72 of air at one face: pv4 holds the primitive
 * states (rho, u, v, w, p) of points i-1, i, i+1, i+2; pl/pr receive the
 * face states (and rho E at index 5). */
void orc_reconstruct_air(int reconstruction, const double *pv4, double *pl, double *pr) {
  orc_zone *zz = calloc(1, sizeof(orc_zone));
  zz->nv = 5;
  zz->cfg.reconstruction = reconstruction;
  reconstruct(zz, pv4, 1, pl, pr);
  free(zz);
}

/* DPLUR Jacobian product (A + lam I) dq at one state: b6 = rho, u, v, w, p,
 * T; sv the n_scalar scalars (species first); gam the mixture gamma (mixtures);
 * m the metric row of the direction. */
void orc_jac_times_dq(const orc_mech *mech, int n_scalar, const double *b6, const double *sv, double gam,
                      const double *m3, double lam, const double *dq, double *out) {
  const double vel = sqrt(b6[1] * b6[1] + b6[2] * b6[2] + b6[3] * b6[3]);
  jac_times_dq_point(mech != NULL, mech ? mech->n_spec : 0, n_scalar, mech, b6, vel, sv, gam, m3, lam, dq, out);
}

/* SourceTerm.cuh:11-37 alone: dq = the SST and chemistry source terms */
void orc_zone_source(orc_zone *zz) {
  memset(zz->dq, 0, sizeof(double) * zz->L0.sz * zz->nv);
  source(zz);
}

/* TimeAdvanceFunc.cu:6-17 store_last_step on the zone's current state */
void orc_zone_store_last(orc_zone *zz) { store_last_step(zz); }

/* Residual.cuh:57-161 against the last store: res[0..3] scaled as at `step`
 * (step 1 sets the scale), res[4] = their maximum */
void orc_zone_residual(orc_zone *zz, int step, double *res5) {
  res5[4] = residual(zz, step);
  for (int l = 0; l < 4; ++l) res5[l] = zz->res[l];
}

/* the SSP-RK3 coefficients used by the stage update */
void orc_ssprk3_coeffs(double *a, double *b, double *c) {
  for (int r = 0; r < 3; ++r) {
    a[r] = RK_A[r];
    b[r] = RK_B[r];
    c[r] = RK_C[r];
  }
}

/* ======================================================= multi-block mesh */

/* gxl_lib/Math.hpp:13-15 sgn of an integer (sgn(0) = +1) */
static int isgn(int a) { return a < 0 ? -1 : 1; }

/* Mesh.cpp:34-187 InnerFace constructor + register_boundary: s7 = the
 * source line (i1 i2 j1 j2 k1 k2 ...), t7 = the target line (... block id,
 * 1-based), 1-based in the file.  In 3-D the negative pair marks the
 * non-normal direction whose correspondence the line states.  Returns 0, or
 * -1 for a 2-D pair of faces with different normal directions, whose target
 * ranges the reference shifts by the source's direction (Mesh.cpp:69-72) and
 * so does not describe. */
int orc_inner_face_register(const int *s7, const int *t7, int ngg, int dim, orc_inner_face *f) {
  memset(f, 0, sizeof(*f));
  for (int i = 0; i < 3; ++i) {
    f->rs[i] = s7[2 * i];
    f->re[i] = s7[2 * i + 1];
    f->ts[i] = t7[2 * i];
    f->te[i] = t7[2 * i + 1];
    f->loop[i] = f->tloop[i] = 1;
  }
  f->tblk = t7[6] - 1;
  f->dir = f->tdir = 1;
  for (int i = 0; i < dim; ++i)
    if (f->rs[i] == f->re[i]) {
      f->face = i;
      if (f->rs[i] == 1) f->dir = -1;
      f->rs[i] -= 1;
      f->re[i] -= 1;
      break;
    }
  for (int i = 0; i < dim; ++i)
    if (f->ts[i] == f->te[i]) {
      f->tface = i;
      if (f->ts[i] == 1) f->tdir = -1;
      f->ts[i] -= 1;
      f->te[i] -= 1;
      break;
    }
  f->src_tar[f->tface] = -(f->dir * f->tdir) * (f->face + 1);
  if (dim == 2) {
    if (f->face != f->tface) return -1;
    const int nf = 1 - f->face, tnf = 1 - f->tface;
    f->rs[nf] = abs(f->rs[nf]) - 1;
    f->re[nf] = abs(f->re[nf]) - 1;
    f->ts[nf] = abs(f->ts[nf]) - 1;
    f->te[nf] = abs(f->te[nf]) - 1;
    f->src_tar[tnf] = isgn((f->te[tnf] - f->ts[tnf]) * (f->re[nf] - f->rs[nf])) * (nf + 1);
    f->src_tar[2] = 3;
    f->rs[2] = f->re[2] = f->ts[2] = f->te[2] = 0;
  } else {
    int nf = 0, tnf = 0;
    for (int i = 0; i < dim; ++i)
      if (f->rs[i] < 0) {
        nf = i;
        f->rs[i] = -f->rs[i] - 1;
        f->re[i] = -f->re[i] - 1;
        break;
      }
    for (int i = 0; i < dim; ++i)
      if (f->ts[i] < 0) {
        tnf = i;
        f->ts[i] = -f->ts[i] - 1;
        f->te[i] = -f->te[i] - 1;
        break;
      }
    f->src_tar[tnf] = isgn((f->te[tnf] - f->ts[tnf]) * (f->re[nf] - f->rs[nf])) * (nf + 1);
    nf = 3 - f->face - nf;
    tnf = 3 - f->tface - tnf;
    f->rs[nf] -= 1;
    f->re[nf] -= 1;
    f->ts[tnf] -= 1;
    f->te[tnf] -= 1;
    f->src_tar[tnf] = isgn((f->te[tnf] - f->ts[tnf]) * (f->re[nf] - f->rs[nf])) * (nf + 1);
  }
  /* the face's own range becomes the face layer plus its ngg ghost layers */
  f->rs[f->face] = f->dir < 0 ? -ngg : f->re[f->face];
  f->re[f->face] = f->dir < 0 ? 0 : f->re[f->face] + ngg;
  const int sgn_tar = isgn(f->src_tar[f->tface]);
  if (f->tdir > 0)
    f->ts[f->tface] = sgn_tar > 0 ? f->te[f->tface] - ngg : f->te[f->tface];
  else
    f->ts[f->tface] = sgn_tar > 0 ? 0 : ngg;
  for (int i = 0; i < 3; ++i) {
    f->loop[i] = f->rs[i] <= f->re[i] ? 1 : -1;
    f->tloop[i] = isgn(f->src_tar[i]);
    f->src_tar[i] = abs(f->src_tar[i]) - 1;
  }
  for (int i = 0; i < 3; ++i) f->n[i] = abs(f->rs[i] - f->re[i]) + 1;
  return 0;
}

struct orc_mesh {
  int nb, dim, step;
  orc_zone **z;
  int *nface, *face0;        /* faces of block b: face0[b] .. face0[b] + nface[b] - 1 */
  orc_inner_face *faces;
  double res[4], res_scale[4];
  double t_phys;
};

/* Mesh.cpp:946-994 init_inner_ghost_grid: the ngg + 1 ghost node layers
 * beyond an inner face are the neighbour block's nodes */
static void inner_ghost_grid(orc_mesh *M) {
  for (int b = 0; b < M->nb; ++b) {
    orc_zone *B = M->z[b];
    for (int q = 0; q < M->nface[b]; ++q) {
      const orc_inner_face *f = &M->faces[M->face0[b] + q];
      const orc_zone *N = M->z[f->tblk];
      int mn[3] = {f->rs[0], f->rs[1], f->rs[2]}, mxx[3] = {f->re[0], f->re[1], f->re[2]},
          mt[3] = {f->ts[0], f->ts[1], f->ts[2]};
      if (f->dir < 0) {
        mn[f->face] -= 1;
        mxx[f->face] -= 1;
      } else {
        mxx[f->face] += 1;
        mn[f->face] += 1;
      }
      mt[f->tface] += f->tdir > 0 ? -1 : 1;
      int d[3];
      for (int i1 = mn[0]; f->loop[0] * (i1 - mxx[0]) != 1; i1 += f->loop[0]) {
        d[0] = i1 - mn[0];
        for (int j1 = mn[1]; f->loop[1] * (j1 - mxx[1]) != 1; j1 += f->loop[1]) {
          d[1] = j1 - mn[1];
          for (int k1 = mn[2]; f->loop[2] * (k1 - mxx[2]) != 1; k1 += f->loop[2]) {
            d[2] = k1 - mn[2];
            const int i2 = mt[0] + f->tloop[0] * d[f->src_tar[0]];
            const int j2 = mt[1] + f->tloop[1] * d[f->src_tar[1]];
            const int k2 = mt[2] + f->tloop[2] * d[f->src_tar[2]];
            const int64_t p = at(B->Lc, i1, j1, k1), pn = at(N->Lc, i2, j2, k2);
            B->x[p] = N->x[pn];
            B->y[p] = N->y[pn];
            B->z[p] = N->z[pn];
          }
        }
      }
    }
  }
}

/* DataCommunication.cuh:27-56, 58-110: for each block's inner faces in
 * order, the face layer is averaged with the neighbour's (both written, |V|
 * of both set) and the ghost layers take the neighbour's interior values;
 * with_cv recomputes cv (and |V|) of this block's points. */
static void inner_exchange(orc_mesh *M, int with_cv) {
  for (int b = 0; b < M->nb; ++b) {
    orc_zone *zz = M->z[b];
    for (int q = 0; q < M->nface[b]; ++q) {
      const orc_inner_face *f = &M->faces[M->face0[b] + q];
      orc_zone *tz = M->z[f->tblk];
      const int face_at = f->dir > 0 ? f->rs[f->face] : f->re[f->face];
      for (int n0 = 0; n0 < f->n[0]; ++n0)
        for (int n1 = 0; n1 < f->n[1]; ++n1)
          for (int n2 = 0; n2 < f->n[2]; ++n2) {
            const int nn[3] = {n0, n1, n2};
            int idx[3], it[3], d[3];
            for (int i = 0; i < 3; ++i) {
              d[i] = f->loop[i] * nn[i];
              idx[i] = f->rs[i] + d[i];
            }
            for (int i = 0; i < 3; ++i) it[i] = f->ts[i] + f->tloop[i] * d[f->src_tar[i]];
            const int64_t p = at(zz->Lg, idx[0], idx[1], idx[2]), pt = at(tz->Lg, it[0], it[1], it[2]);
            const int64_t sz = zz->Lg.sz, tsz = tz->Lg.sz;
            if (idx[f->face] == face_at) {
              for (int l = 0; l < 6; ++l) {
                const double ave = 0.5 * (tz->bv[l * tsz + pt] + zz->bv[l * sz + p]);
                zz->bv[l * sz + p] = ave;
                tz->bv[l * tsz + pt] = ave;
              }
              for (int l = 0; l < zz->nsc; ++l) {
                const double ave = 0.5 * (tz->sv[l * tsz + pt] + zz->sv[l * sz + p]);
                zz->sv[l * sz + p] = ave;
                tz->sv[l * tsz + pt] = ave;
              }
              const double u = zz->bv[sz + p], v = zz->bv[2 * sz + p], w = zz->bv[3 * sz + p];
              const double vel = sqrt(u * u + v * v + w * w);
              zz->vel[p] = vel;
              tz->vel[pt] = vel;
            } else {
              for (int l = 0; l < 6; ++l) zz->bv[l * sz + p] = tz->bv[l * tsz + pt];
              for (int l = 0; l < zz->nsc; ++l) zz->sv[l * sz + p] = tz->sv[l * tsz + pt];
            }
            if (with_cv) cv_from_bv_point(zz, idx[0], idx[1], idx[2]);
          }
    }
  }
}

orc_mesh *orc_mesh_create(int nb, const int *dims, int ngg, const orc_cfg *cfgs, const orc_mech *mech,
                          const double *const *xyz, const int *nface, const int *lines) {
  orc_mesh *M = (orc_mesh *)calloc(1, sizeof(orc_mesh));
  M->nb = nb;
  M->z = (orc_zone **)calloc((size_t)nb, sizeof(orc_zone *));
  M->nface = (int *)calloc((size_t)nb, sizeof(int));
  M->face0 = (int *)calloc((size_t)nb, sizeof(int));
  int nf = 0;
  for (int b = 0; b < nb; ++b) {
    M->nface[b] = nface[b];
    M->face0[b] = nf;
    nf += nface[b];
  }
  M->faces = (orc_inner_face *)calloc((size_t)(nf > 0 ? nf : 1), sizeof(orc_inner_face));
  M->dim = dims[2] == 1 ? 2 : 3;
  for (int q = 0; q < nf; ++q)
    if (orc_inner_face_register(&lines[q * 14], &lines[q * 14 + 7], ngg, M->dim, &M->faces[q])) {
      orc_mesh_destroy(M);
      return NULL;
    }
  for (int b = 0; b < nb; ++b) {
    const int mx = dims[3 * b], my = dims[3 * b + 1], mz = dims[3 * b + 2];
    orc_zone *zz = orc_zone_create(mx, my, mz, ngg, &cfgs[b], mech, NULL, NULL, NULL);
    M->z[b] = zz;
    const double *x = xyz[3 * b], *y = xyz[3 * b + 1], *z = xyz[3 * b + 2];
    for (int k = 0; k < mz; ++k)
      for (int j = 0; j < my; ++j)
        for (int i = 0; i < mx; ++i) {
          const int64_t s = ((int64_t)k * my + j) * mx + i, p = at(zz->Lc, i, j, k);
          zz->x[p] = x[s];
          zz->y[p] = y[s];
          zz->z[p] = z[s];
        }
    ghost_faces(mx, my, mz, ngg, M->dim, zz->Lc, zz->x, zz->y, zz->z);
  }
  inner_ghost_grid(M);
  for (int b = 0; b < nb; ++b) {
    orc_zone *zz = M->z[b];
    ghost_edges(zz->mx, zz->my, zz->mz, ngg, zz->Lc, zz->x, zz->y, zz->z);
    metrics(zz->mx, zz->my, zz->mz, ngg, zz->Lc, zz->Lg, zz->x, zz->y, zz->z, zz->jac, zz->metric);
    zone_register(zz);
  }
  return M;
}

void orc_mesh_destroy(orc_mesh *M) {
  if (!M) return;
  for (int b = 0; b < M->nb; ++b) orc_zone_destroy(M->z[b]);
  free(M->z);
  free(M->nface);
  free(M->face0);
  free(M->faces);
  free(M);
}

orc_zone *orc_mesh_zone(orc_mesh *M, int b) { return (b >= 0 && b < M->nb) ? M->z[b] : NULL; }

const orc_inner_face *orc_mesh_face(orc_mesh *M, int q) { return &M->faces[q]; }

/* Driver.cu:40-98 over all blocks (data communication without cv after the
 * boundary conditions) followed by RK.cuh:38-48 */
void orc_mesh_initialize(orc_mesh *M) {
  for (int b = 0; b < M->nb; ++b) {
    orc_zone *zz = M->z[b];
    if (zz->dim == 2) eliminate_k_gradient(zz, 0);
  }
  for (int b = 0; b < M->nb; ++b) {
    orc_zone *zz = M->z[b];
    if (zz->cfg.mix) update_properties(zz);
    apply_bcs(zz, 0);
  }
  for (int b = 0; b < M->nb; ++b) {
    orc_zone *zz = M->z[b];
    const int g = zz->g;
    for (int k = -g; k < zz->mz + g; ++k)
      for (int j = -g; j < zz->my + g; ++j)
        for (int i = -g; i < zz->mx + g; ++i)
          S1(vel, i, j, k) = sqrt(BV(1, i, j, k) * BV(1, i, j, k) + BV(2, i, j, k) * BV(2, i, j, k) +
                                  BV(3, i, j, k) * BV(3, i, j, k));
    if (zz->nt) initialize_mut(zz);
  }
  inner_exchange(M, 0);
  for (int b = 0; b < M->nb; ++b) {
    orc_zone *zz = M->z[b];
    const int g = zz->g;
    update_properties(zz);
    store_last_step(zz);
    for (int k = -g; k < zz->mz + g; ++k)
      for (int j = -g; j < zz->my + g; ++j)
        for (int i = -g; i < zz->mx + g; ++i) cv_from_bv_point(zz, i, j, k);
    zz->step = 0;
    zz->t_phys = 0;
  }
  M->step = 0;
  M->t_phys = 0;
}

static double mesh_residual(orc_mesh *M, int step) {
  double s[4] = {0, 0, 0, 0};
  int64_t n = 0;
  for (int b = 0; b < M->nb; ++b) {
    double sb[4] = {0, 0, 0, 0};
    residual_sums(M->z[b], sb);
    for (int l = 0; l < 4; ++l) s[l] += sb[l];
    n += (int64_t)M->z[b]->mx * M->z[b]->my * M->z[b]->mz;
  }
  return residual_finish(s, n, step, M->res, M->res_scale);
}

static void mesh_hist(orc_mesh *M, int step, int os, int it, double *res_hist) {
  if (step % os == 0 || step == 1) {
    const double err = mesh_residual(M, step);
    if (res_hist) {
      for (int l = 0; l < 4; ++l) res_hist[it * 5 + l] = M->res[l];
      res_hist[it * 5 + 4] = err;
    }
  } else if (res_hist) {
    for (int l = 0; l < 5; ++l) res_hist[it * 5 + l] = NAN;
  }
}

/* RK.cuh:89-193 over all blocks: dt is the minimum over every block */
void orc_mesh_run_rk3(orc_mesh *M, int nsteps, double *dt_hist, double *res_hist) {
  orc_zone *z0 = M->z[0];
  const int os = z0->cfg.output_screen > 0 ? z0->cfg.output_screen : 1;
  for (int it = 0; it < nsteps; ++it) {
    const int step = ++M->step;
    double dt = z0->cfg.fixed_time_step ? z0->cfg.dt : 1e+6;
    for (int b = 0; b < M->nb; ++b) {
      orc_zone *zz = M->z[b];
      if (step % os == 0) store_last_step(zz);
      memcpy(zz->qn, zz->cv, sizeof(double) * zz->Lg.sz * zz->nv);
    }
    for (int rk = 0; rk < 3; ++rk) {
      for (int b = 0; b < M->nb; ++b) orc_zone_rhs(M->z[b]);
      if (!z0->cfg.fixed_time_step) {
        dt = 1e+6;
        for (int b = 0; b < M->nb; ++b) {
          const double d = orc_zone_global_dt(M->z[b]);
          if (d < dt) dt = d;
        }
      }
      for (int b = 0; b < M->nb; ++b) {
        rk_update(M->z[b], dt, rk);
        apply_bcs(M->z[b], 1);
      }
      inner_exchange(M, 1);
      for (int b = 0; b < M->nb; ++b) {
        if (M->z[b]->dim == 2) eliminate_k_gradient(M->z[b], 1);
      }
      for (int b = 0; b < M->nb; ++b) update_properties(M->z[b]);
      if (dt_hist) dt_hist[it * 3 + rk] = dt;
    }
    M->t_phys += dt;
    for (int b = 0; b < M->nb; ++b) {
      M->z[b]->t_phys = M->t_phys;
      M->z[b]->step = step;
    }
    mesh_hist(M, step, os, it, res_hist);
  }
}

/* SteadySim.cuh:52-124 over all blocks (cv follows at the boundaries and
 * inner faces for the AWENO path, which reads it) */
void orc_mesh_run_steady(orc_mesh *M, int nsteps, double *res_hist) {
  orc_zone *z0 = M->z[0];
  const int os = z0->cfg.output_screen > 0 ? z0->cfg.output_screen : 1;
  const int with_cv = z0->inviscid_type == 1;
  if (M->step == 0)
    for (int b = 0; b < M->nb; ++b) store_last_step(M->z[b]);
  for (int it = 0; it < nsteps; ++it) {
    const int step = ++M->step;
    for (int b = 0; b < M->nb; ++b) {
      orc_zone *zz = M->z[b];
      if (step % os == 0) store_last_step(zz);
      orc_zone_rhs(zz);
      (void)orc_zone_global_dt(zz);
      orc_zone_implicit(zz);
      update_bv(zz);
      if (zz->cfg.split_chem && zz->cfg.steady_chem_dt > 0) chem_split_interior(zz, zz->cfg.steady_chem_dt);
      apply_bcs(zz, with_cv);
    }
    inner_exchange(M, with_cv);
    for (int b = 0; b < M->nb; ++b) {
      if (M->z[b]->dim == 2) eliminate_k_gradient(M->z[b], with_cv);
    }
    for (int b = 0; b < M->nb; ++b) {
      update_properties(M->z[b]);
      M->z[b]->step = step;
    }
    mesh_hist(M, step, os, it, res_hist);
  }
}

void orc_mesh_exchange(orc_mesh *M, int with_cv) { inner_exchange(M, with_cv); }
