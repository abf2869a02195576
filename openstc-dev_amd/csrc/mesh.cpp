// Host-side grid set-up for one structured block: ghost nodes by point
// reflection with parallelogram edges/corners (Mesh.cpp:790-944; 2-D
// k-ghosts at z0+k, Mesh.cpp:848) and the Jacobian / metrics by second-order
// central differences (Mesh.cpp:259-521).  One-time work; the result is
// uploaded as component-major arrays for coalesced device reads.
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "ostc_internal.h"

namespace ostc {

namespace {

}  // namespace

// Mesh.cpp:790-858: reflected ghost nodes on the six faces (ngg + 1 layers),
// 2-D k-ghosts at z0 + k.
void ghost_node_faces(int mx, int my, int mz, int ngg, const Layout &L, double *A, bool is_z) {
  auto a = [&](int i, int j, int k) -> double & { return A[L.at(i, j, k)]; };
  const int G = ngg + 1;
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j) {
      for (int i = 1; i <= G; ++i) a(-i, j, k) = 2 * a(0, j, k) - a(i, j, k);
      for (int i = mx; i < mx + G; ++i) a(i, j, k) = 2 * a(mx - 1, j, k) - a(2 * (mx - 1) - i, j, k);
    }
  for (int k = 0; k < mz; ++k)
    for (int i = 0; i < mx; ++i) {
      for (int j = 1; j <= G; ++j) a(i, -j, k) = 2 * a(i, 0, k) - a(i, j, k);
      for (int j = my; j < my + G; ++j) a(i, j, k) = 2 * a(i, my - 1, k) - a(i, 2 * (my - 1) - j, k);
    }
  if (mz > 1) {
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        for (int k = 1; k <= G; ++k) a(i, j, -k) = 2 * a(i, j, 0) - a(i, j, k);
        for (int k = mz; k < mz + G; ++k) a(i, j, k) = 2 * a(i, j, mz - 1) - a(i, j, 2 * (mz - 1) - k);
      }
  } else {
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        for (int k = 1; k <= G; ++k) a(i, j, -k) = a(i, j, 0) - (is_z ? k : 0);
        for (int k = mz; k < mz + G; ++k) a(i, j, k) = a(i, j, mz - 1) + (is_z ? k - mz + 1 : 0);
      }
  }
}

// Mesh.cpp:861-937: parallelogram edges and corners from the face ghosts
// (run after an inner face's ghosts were taken from the neighbour block).
void ghost_node_edges(int mx, int my, int mz, int ngg, const Layout &L, double *A) {
  auto a = [&](int i, int j, int k) -> double & { return A[L.at(i, j, k)]; };
  const int G = ngg + 1;
  // edges and corners: x-y edges over the interior k range
  for (int i = 1; i <= G; ++i)
    for (int j = 1; j <= G; ++j)
      for (int k = 0; k < mz; ++k) {
        const int i1 = -i, j1 = -j, i2 = mx - 1 + i, j2 = my - 1 + j;
        a(i1, j1, k) = a(i1, 0, k) + a(0, j1, k) - a(0, 0, k);
        a(i2, j2, k) = a(i2, my - 1, k) + a(mx - 1, j2, k) - a(mx - 1, my - 1, k);
        a(i1, j2, k) = a(i1, my - 1, k) + a(0, j2, k) - a(0, my - 1, k);
        a(i2, j1, k) = a(i2, 0, k) + a(mx - 1, j1, k) - a(mx - 1, 0, k);
      }
  // x-z edges over the interior j range
  for (int i = 1; i <= G; ++i)
    for (int j = 0; j < my; ++j)
      for (int k = 1; k <= G; ++k) {
        const int i1 = -i, k1 = -k, i2 = mx - 1 + i, k2 = mz - 1 + k;
        a(i1, j, k1) = a(i1, j, 0) + a(0, j, k1) - a(0, j, 0);
        a(i2, j, k2) = a(i2, j, mz - 1) + a(mx - 1, j, k2) - a(mx - 1, j, mz - 1);
        a(i1, j, k2) = a(i1, j, mz - 1) + a(0, j, k2) - a(0, j, mz - 1);
        a(i2, j, k1) = a(i2, j, 0) + a(mx - 1, j, k1) - a(mx - 1, j, 0);
      }
  // y-z edges and corners over the full i range
  for (int i = -G; i < mx + G; ++i)
    for (int j = 1; j <= G; ++j)
      for (int k = 1; k <= G; ++k) {
        const int j1 = -j, k1 = -k, j2 = my - 1 + j, k2 = mz - 1 + k;
        a(i, j1, k1) = a(i, j1, 0) + a(i, 0, k1) - a(i, 0, 0);
        a(i, j2, k2) = a(i, j2, mz - 1) + a(i, my - 1, k2) - a(i, my - 1, mz - 1);
        a(i, j1, k2) = a(i, j1, mz - 1) + a(i, 0, k2) - a(i, 0, mz - 1);
        a(i, j2, k1) = a(i, j2, 0) + a(i, my - 1, k1) - a(i, my - 1, 0);
      }
}

namespace {

void ghost_nodes(int mx, int my, int mz, int ngg, const Layout &L, double *A, bool is_z) {
  ghost_node_faces(mx, my, mz, ngg, L, A, is_z);
  ghost_node_edges(mx, my, mz, ngg, L, A);
}

struct Metric9 {
  double J;
  double m[9];
};

bool point_metric(const Layout &Lc, const double *X, const double *Y, const double *Z, int i, int j, int k,
                  Metric9 &out) {
  auto d = [&](const double *A, int di, int dj, int dk) {
    return (A[Lc.at(i + di, j + dj, k + dk)] - A[Lc.at(i - di, j - dj, k - dk)]) * 0.5;
  };
  const double x1 = d(X, 1, 0, 0), y1 = d(Y, 1, 0, 0), z1 = d(Z, 1, 0, 0);
  const double x2 = d(X, 0, 1, 0), y2 = d(Y, 0, 1, 0), z2 = d(Z, 0, 1, 0);
  const double x3 = d(X, 0, 0, 1), y3 = d(Y, 0, 0, 1), z3 = d(Z, 0, 0, 1);
  const double J = x1 * y2 * z3 + x2 * y3 * z1 + x3 * y1 * z2 - x1 * y3 * z2 - x2 * y1 * z3 - x3 * y2 * z1;
  if (J <= 0) return false;
  out.J = J;
  out.m[0] = (y2 * z3 - y3 * z2) / J;
  out.m[1] = (x3 * z2 - x2 * z3) / J;
  out.m[2] = (x2 * y3 - x3 * y2) / J;
  out.m[3] = (y3 * z1 - y1 * z3) / J;
  out.m[4] = (x1 * z3 - x3 * z1) / J;
  out.m[5] = (x3 * y1 - x1 * y3) / J;
  out.m[6] = (y1 * z2 - y2 * z1) / J;
  out.m[7] = (x2 * z1 - x1 * z2) / J;
  out.m[8] = (x1 * y2 - x2 * y1) / J;
  return true;
}

// Conservative symmetric metrics (the option conservative_metric; the
// reference's metrics above are non-conservative and a uniform flow on a
// curved grid does not stay uniform with them).  Row r of J d(xi_r)/d(x_c):
//   J xi_c   = 1/2 [D_zeta(A^c_eta) - D_eta(A^c_zeta)]
//   J eta_c  = 1/2 [D_xi(A^c_zeta)  - D_zeta(A^c_xi)]
//   J zeta_c = 1/2 [D_eta(A^c_xi)   - D_xi(A^c_eta)]
// with A^x_d = y_d z - y z_d, A^y_d = z_d x - z x_d, A^z_d = x_d y - x y_d
// (Thomas & Lombard 1979; Vinokur & Yee 2002).  D_d, also for the inner
// derivatives, is the difference the flux scheme applies along d at that
// index: (f[+1] - f[-1]) / 2 (the face average of J * metric differenced),
// plus for AWENO-5 on cells [0, n - 2 ngg), where the CDS correction acts,
// c3 (f[+1] - f[-1]) + c2 (f[+2] - f[-2]) + c1 (f[+3] - f[-3]).  Operators
// along different directions commute, so the discrete metric identities
// hold to rounding and a uniform flow has a zero right-hand side.  J stays
// the determinant; the stored row is divided by it.  A row whose stencil
// leaves the node array keeps the non-conservative row (the fluxes never
// read such nodes).
constexpr double kFspW1 = 145.0 / 256, kFspW2 = -13.0 / 320, kFspW3 = 19.0 / 3840;

// D_d of F at (i, j, k) on the (ngg + 1)-ghosted node layout; NaN when the
// stencil leaves it
double fsp_diff(const Layout &Lc, const double *F, int d, int i, int j, int k, int n, int ngg, bool aweno) {
  const int c = d == 0 ? i : (d == 1 ? j : k);
  const bool wide = aweno && c >= 0 && c < n - 2 * ngg;
  const int w = wide ? 3 : 1;
  if (c - w < -(ngg + 1) || c + w > n + ngg) return std::nan("");
  const int64_t st = d == 0 ? 1 : (d == 1 ? Lc.d2 : Lc.d1);
  const int64_t p = Lc.at(i, j, k);
  if (!wide) return (F[p + st] - F[p - st]) * 0.5;
  return (F[p + st] - F[p - st]) * kFspW1 + (F[p + 2 * st] - F[p - 2 * st]) * kFspW2 +
         (F[p + 3 * st] - F[p - 3 * st]) * kFspW3;
}

void fsp_rows(int mx, int my, int mz, int ngg, bool aweno, const Layout &Lc, const Layout &L, const double *X,
              const double *Y, const double *Z, std::vector<Metric9> &M) {
  const int n[3] = {mx, my, mz};
  const double *C[3] = {X, Y, Z};
  // A[c * 3 + d] on the node layout
  std::vector<double> A((size_t)9 * Lc.sz);
  for (int k = -ngg - 1; k <= mz + ngg; ++k)
    for (int j = -ngg - 1; j <= my + ngg; ++j)
      for (int i = -ngg - 1; i <= mx + ngg; ++i) {
        const int64_t p = Lc.at(i, j, k);
        for (int d = 0; d < 3; ++d) {
          double der[3];
          for (int c = 0; c < 3; ++c) der[c] = fsp_diff(Lc, C[c], d, i, j, k, n[d], ngg, aweno);
          for (int c = 0; c < 3; ++c) {
            const int a = (c + 1) % 3, b = (c + 2) % 3;
            A[(size_t)(c * 3 + d) * Lc.sz + p] = der[a] * C[b][p] - C[a][p] * der[b];
          }
        }
      }
  for (int k = -ngg; k < mz + ngg; ++k)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int i = -ngg; i < mx + ngg; ++i) {
        Metric9 &v = M[L.at(i, j, k)];
        for (int r = 0; r < 3; ++r) {
          const int d1 = (r + 1) % 3, d2 = (r + 2) % 3;
          double row[3];
          bool ok = true;
          for (int c = 0; c < 3; ++c) {
            row[c] = (fsp_diff(Lc, &A[(size_t)(c * 3 + d1) * Lc.sz], d2, i, j, k, n[d2], ngg, aweno) -
                      fsp_diff(Lc, &A[(size_t)(c * 3 + d2) * Lc.sz], d1, i, j, k, n[d1], ngg, aweno)) *
                     0.5;
            if (std::isnan(row[c])) ok = false;
          }
          if (!ok) continue;
          for (int c = 0; c < 3; ++c) v.m[r * 3 + c] = row[c] / v.J;
        }
      }
}

}  // namespace

bool compute_metrics(int mx, int my, int mz, int ngg, const double *x, const double *y, const double *z,
                     std::vector<double> &jac, std::vector<double> &metric, std::string &err,
                     std::vector<double> *xyz, bool conservative, bool aweno) {
  const Layout Lc(mx, my, mz, ngg + 1), L(mx, my, mz, ngg);
  std::vector<double> X(Lc.sz, 0), Y(Lc.sz, 0), Z(Lc.sz, 0);
  for (int k = 0; k < mz; ++k)
    for (int j = 0; j < my; ++j)
      for (int i = 0; i < mx; ++i) {
        const size_t s = ((size_t)k * my + j) * mx + i;
        X[Lc.at(i, j, k)] = x[s];
        Y[Lc.at(i, j, k)] = y[s];
        Z[Lc.at(i, j, k)] = z[s];
      }
  ghost_nodes(mx, my, mz, ngg, Lc, X.data(), false);
  ghost_nodes(mx, my, mz, ngg, Lc, Y.data(), false);
  ghost_nodes(mx, my, mz, ngg, Lc, Z.data(), true);
  // interim AoS store, repacked component-major at the end
  std::vector<Metric9> M(L.sz);
  auto put = [&](int i, int j, int k, int di, int dj, int dk) {
    Metric9 v;
    if (point_metric(Lc, X.data(), Y.data(), Z.data(), i, j, k, v))
      M[L.at(i, j, k)] = v;
    else
      M[L.at(i, j, k)] = M[L.at(i + di, j + dj, k + dk)];
  };
  // interior (a non-positive Jacobian here is an error, Mesh.cpp:282-288)
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j) {
      const int k0 = mz == 1 ? 0 : 1, k1 = mz == 1 ? 1 : mz - 1;
      for (int k = k0; k < k1; ++k) {
        Metric9 v;
        if (!point_metric(Lc, X.data(), Y.data(), Z.data(), i, j, k, v)) {
          err = "negative Jacobian at interior point (" + std::to_string(i) + "," + std::to_string(j) + "," +
                std::to_string(k) + ")";
          return false;
        }
        M[L.at(i, j, k)] = v;
      }
    }
  // the six ghost slabs in the reference's order, each falling back to its inner neighbour
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j)
      for (int k = 0; k >= -ngg; --k) put(i, j, k, 0, 0, 1);
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 1; j < my - 1; ++j)
      for (int k = mz - 1; k < mz + ngg; ++k) put(i, j, k, 0, 0, -1);
  for (int i = 1; i < mx - 1; ++i)
    for (int j = 0; j >= -ngg; --j)
      for (int k = -ngg; k < mz + ngg; ++k) put(i, j, k, 0, 1, 0);
  for (int i = 1; i < mx - 1; ++i)
    for (int j = my - 1; j < my + ngg; ++j)
      for (int k = -ngg; k < mz + ngg; ++k) put(i, j, k, 0, -1, 0);
  for (int i = 0; i >= -ngg; --i)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int k = -ngg; k < mz + ngg; ++k) put(i, j, k, 1, 0, 0);
  for (int i = mx - 1; i < mx + ngg; ++i)
    for (int j = -ngg; j < my + ngg; ++j)
      for (int k = -ngg; k < mz + ngg; ++k) put(i, j, k, -1, 0, 0);
  if (conservative) fsp_rows(mx, my, mz, ngg, aweno, Lc, L, X.data(), Y.data(), Z.data(), M);
  if (xyz) {
    // node coordinates on the ngg-ghosted layout, component-major
    xyz->assign((size_t)L.sz * 3, 0);
    for (int k = -ngg; k < mz + ngg; ++k)
      for (int j = -ngg; j < my + ngg; ++j)
        for (int i = -ngg; i < mx + ngg; ++i) {
          const int64_t p = L.at(i, j, k), q = Lc.at(i, j, k);
          (*xyz)[p] = X[q];
          (*xyz)[(size_t)L.sz + p] = Y[q];
          (*xyz)[(size_t)2 * L.sz + p] = Z[q];
        }
  }
  jac.assign(L.sz, 0);
  metric.assign((size_t)L.sz * 9, 0);
  for (int64_t p = 0; p < L.sz; ++p) {
    jac[p] = M[p].J;
    for (int c = 0; c < 9; ++c) metric[(size_t)c * L.sz + p] = M[p].m[c];
  }
  return true;
}

// gxl_lib/Math.hpp:13-15 sgn of an integer (sgn(0) = +1)
static int isgn(int a) { return a < 0 ? -1 : 1; }

// Mesh.cpp:34-187 InnerFace constructor + register_boundary.
bool register_inner_face(const int *s7, const int *t7, int ngg, int dim, InnerFace &f, std::string &err) {
  f = InnerFace{};
  for (int i = 0; i < 3; ++i) {
    f.rs[i] = s7[2 * i];
    f.re[i] = s7[2 * i + 1];
    f.ts[i] = t7[2 * i];
    f.te[i] = t7[2 * i + 1];
  }
  f.tblk = t7[6] - 1;
  for (int i = 0; i < dim; ++i)
    if (f.rs[i] == f.re[i]) {
      f.face = i;
      if (f.rs[i] == 1) f.dir = -1;
      f.rs[i] -= 1;
      f.re[i] -= 1;
      break;
    }
  for (int i = 0; i < dim; ++i)
    if (f.ts[i] == f.te[i]) {
      f.tface = i;
      if (f.ts[i] == 1) f.tdir = -1;
      f.ts[i] -= 1;
      f.te[i] -= 1;
      break;
    }
  f.src_tar[f.tface] = -(f.dir * f.tdir) * (f.face + 1);
  if (dim == 2) {
    // Mesh.cpp:69-72 shifts the target's ranges by the source's tangential
    // direction, which only describes faces with the same normal direction
    if (f.face != f.tface) {
      err = "2-D inner faces must join faces of the same index direction (i with i, j with j)";
      return false;
    }
    const int nf = 1 - f.face, tnf = 1 - f.tface;
    f.rs[nf] = std::abs(f.rs[nf]) - 1;
    f.re[nf] = std::abs(f.re[nf]) - 1;
    f.ts[nf] = std::abs(f.ts[nf]) - 1;
    f.te[nf] = std::abs(f.te[nf]) - 1;
    f.src_tar[tnf] = isgn((f.te[tnf] - f.ts[tnf]) * (f.re[nf] - f.rs[nf])) * (nf + 1);
    f.src_tar[2] = 3;
    f.rs[2] = f.re[2] = f.ts[2] = f.te[2] = 0;
  } else {
    // the negative pair marks the tangential direction whose correspondence
    // the line states; the third direction follows from the other two
    int nf = -1, tnf = -1;
    for (int i = 0; i < dim; ++i)
      if (f.rs[i] < 0) {
        nf = i;
        f.rs[i] = -f.rs[i] - 1;
        f.re[i] = -f.re[i] - 1;
        break;
      }
    for (int i = 0; i < dim; ++i)
      if (f.ts[i] < 0) {
        tnf = i;
        f.ts[i] = -f.ts[i] - 1;
        f.te[i] = -f.te[i] - 1;
        break;
      }
    if (nf < 0 || tnf < 0 || nf == f.face || tnf == f.tface) {
      err = "3-D inner face lines need one negative index pair (not the normal direction) on each side";
      return false;
    }
    f.src_tar[tnf] = isgn((f.te[tnf] - f.ts[tnf]) * (f.re[nf] - f.rs[nf])) * (nf + 1);
    nf = 3 - f.face - nf;
    tnf = 3 - f.tface - tnf;
    f.rs[nf] -= 1;
    f.re[nf] -= 1;
    f.ts[tnf] -= 1;
    f.te[tnf] -= 1;
    f.src_tar[tnf] = isgn((f.te[tnf] - f.ts[tnf]) * (f.re[nf] - f.rs[nf])) * (nf + 1);
  }
  // the face's own range becomes the face layer plus its ngg ghost layers
  f.rs[f.face] = f.dir < 0 ? -ngg : f.re[f.face];
  f.re[f.face] = f.dir < 0 ? 0 : f.re[f.face] + ngg;
  const int sgn_tar = isgn(f.src_tar[f.tface]);
  if (f.tdir > 0)
    f.ts[f.tface] = sgn_tar > 0 ? f.te[f.tface] - ngg : f.te[f.tface];
  else
    f.ts[f.tface] = sgn_tar > 0 ? 0 : ngg;
  for (int i = 0; i < 3; ++i) {
    f.loop[i] = f.rs[i] <= f.re[i] ? 1 : -1;
    f.tloop[i] = isgn(f.src_tar[i]);
    f.src_tar[i] = std::abs(f.src_tar[i]) - 1;
    f.n[i] = std::abs(f.rs[i] - f.re[i]) + 1;
  }
  return true;
}

// Mesh.cpp:946-994 init_inner_ghost_grid for one face of block B with
// neighbour N (both on their (ngg + 1)-ghosted node layouts).
void inner_ghost_nodes(const InnerFace &f, const Layout &LB, double *const *B, const Layout &LN,
                       const double *const *N) {
  int mn[3] = {f.rs[0], f.rs[1], f.rs[2]}, mxx[3] = {f.re[0], f.re[1], f.re[2]}, mt[3] = {f.ts[0], f.ts[1], f.ts[2]};
  if (f.dir < 0) {
    mn[f.face] -= 1;
    mxx[f.face] -= 1;
  } else {
    mxx[f.face] += 1;
    mn[f.face] += 1;
  }
  mt[f.tface] += f.tdir > 0 ? -1 : 1;
  int d[3];
  for (int i1 = mn[0]; f.loop[0] * (i1 - mxx[0]) != 1; i1 += f.loop[0]) {
    d[0] = i1 - mn[0];
    for (int j1 = mn[1]; f.loop[1] * (j1 - mxx[1]) != 1; j1 += f.loop[1]) {
      d[1] = j1 - mn[1];
      for (int k1 = mn[2]; f.loop[2] * (k1 - mxx[2]) != 1; k1 += f.loop[2]) {
        d[2] = k1 - mn[2];
        const int i2 = mt[0] + f.tloop[0] * d[f.src_tar[0]];
        const int j2 = mt[1] + f.tloop[1] * d[f.src_tar[1]];
        const int k2 = mt[2] + f.tloop[2] * d[f.src_tar[2]];
        const int64_t p = LB.at(i1, j1, k1), q = LN.at(i2, j2, k2);
        for (int c = 0; c < 3; ++c) B[c][p] = N[c][q];
      }
    }
  }
}

}  // namespace ostc
