// Grid metrics on the device (Mesh.cpp:259-521 compute_jac_metric): the
// Jacobian and d(xi_r)/d(x_c) of every point of the ngg-ghosted layout by
// second-order central differences of the (ngg + 1)-ghosted node
// coordinates, then the reference's fall-back for ghost points with a
// non-positive Jacobian (copy of the inner neighbour), slab by slab in the
// reference's order; optionally the conservative metric rows on top.  Contraction into FMAs is off in this file so the
// results equal the host computation bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "flux_kernels.h"

#pragma clang fp contract(off)

namespace ostc {

namespace {

constexpr int kGTB = 256;

struct GridDims {
  int mx, my, mz, g;
  int64_t cd2, cd1;  // node layout (g + 1 ghosts)
  int64_t d2, d1, sz;  // metric layout (g ghosts)
  __device__ int64_t cat(int i, int j, int k) const {
    return (int64_t)(k + g + 1) * cd1 + (int64_t)(j + g + 1) * cd2 + (i + g + 1);
  }
  __device__ int64_t at(int i, int j, int k) const { return (int64_t)(k + g) * d1 + (int64_t)(j + g) * d2 + (i + g); }
};

// one point: J and the nine metrics, written component-major; bad[p] = 1
// when J <= 0 (the point then takes a neighbour's values in the slab pass),
// *interior_bad set when that happens inside the interior range
__global__ void __launch_bounds__(kGTB) k_metrics(GridDims G, const double *__restrict__ X,
                                                  const double *__restrict__ Y, const double *__restrict__ Z,
                                                  double *__restrict__ jac, double *__restrict__ met,
                                                  unsigned char *__restrict__ bad, int *__restrict__ interior_bad) {
  const int64_t t = (int64_t)blockIdx.x * kGTB + threadIdx.x;
  if (t >= G.sz) return;
  const int nx = G.mx + 2 * G.g, ny = G.my + 2 * G.g;
  const int i = (int)(t % nx) - G.g;
  const int j = (int)((t / nx) % ny) - G.g;
  const int k = (int)(t / ((int64_t)nx * ny)) - G.g;
  auto d = [&](const double *__restrict__ A, int di, int dj, int dk) {
    return (A[G.cat(i + di, j + dj, k + dk)] - A[G.cat(i - di, j - dj, k - dk)]) * 0.5;
  };
  const double x1 = d(X, 1, 0, 0), y1 = d(Y, 1, 0, 0), z1 = d(Z, 1, 0, 0);
  const double x2 = d(X, 0, 1, 0), y2 = d(Y, 0, 1, 0), z2 = d(Z, 0, 1, 0);
  const double x3 = d(X, 0, 0, 1), y3 = d(Y, 0, 0, 1), z3 = d(Z, 0, 0, 1);
  const double J = x1 * y2 * z3 + x2 * y3 * z1 + x3 * y1 * z2 - x1 * y3 * z2 - x2 * y1 * z3 - x3 * y2 * z1;
  const int64_t p = t;
  const bool ok = J > 0;
  bad[p] = ok ? 0 : 1;
  if (!ok) {
    const bool k_in = G.mz == 1 ? k == 0 : (k >= 1 && k < G.mz - 1);
    if (i >= 1 && i < G.mx - 1 && j >= 1 && j < G.my - 1 && k_in) atomicOr(interior_bad, 1);
    return;
  }
  jac[p] = J;
  met[p] = (y2 * z3 - y3 * z2) / J;
  met[G.sz + p] = (x3 * z2 - x2 * z3) / J;
  met[2 * G.sz + p] = (x2 * y3 - x3 * y2) / J;
  met[3 * G.sz + p] = (y3 * z1 - y1 * z3) / J;
  met[4 * G.sz + p] = (x1 * z3 - x3 * z1) / J;
  met[5 * G.sz + p] = (x3 * y1 - x1 * y3) / J;
  met[6 * G.sz + p] = (y1 * z2 - y2 * z1) / J;
  met[7 * G.sz + p] = (x2 * z1 - x1 * z2) / J;
  met[8 * G.sz + p] = (x1 * y2 - x2 * y1) / J;
}

// One ghost slab of Mesh.cpp:259-521: lines along `axis` (0 i, 1 j, 2 k),
// one thread per line over the two other index ranges [a0, a1) x [b0, b1);
// each line runs from layer `from` outward by `step` for g + 1 layers and a
// point with a non-positive Jacobian copies its inner neighbour.
__global__ void __launch_bounds__(kGTB) k_metric_slab(GridDims G, int axis, int from, int step, int a0, int a1,
                                                      int b0, int b1, double *__restrict__ jac,
                                                      double *__restrict__ met,
                                                      const unsigned char *__restrict__ bad) {
  const int64_t t = (int64_t)blockIdx.x * kGTB + threadIdx.x;
  const int na = a1 - a0, nb = b1 - b0;
  if (na <= 0 || nb <= 0 || t >= (int64_t)na * nb) return;
  const int a = a0 + (int)(t % na), b = b0 + (int)(t / na);
  for (int l = 0; l <= G.g; ++l) {
    const int c = from + step * l;
    int ijk[3], nbr[3];
    // the two free directions in increasing order, the line along axis
    const int u = axis == 0 ? 1 : 0, v = axis == 2 ? 1 : 2;
    ijk[axis] = c;
    ijk[u] = a;
    ijk[v] = b;
    nbr[axis] = c - step;
    nbr[u] = a;
    nbr[v] = b;
    const int64_t p = G.at(ijk[0], ijk[1], ijk[2]);
    if (!bad[p]) continue;
    const int64_t q = G.at(nbr[0], nbr[1], nbr[2]);
    jac[p] = jac[q];
#pragma unroll
    for (int r = 0; r < 9; ++r) met[r * G.sz + p] = met[r * G.sz + q];
  }
}

// Conservative symmetric metrics (the option conservative_metric), the same
// arithmetic as the host pass: row r of J d(xi_r)/d(x_c) is
// 1/2 [D_{r+2}(A^c_{r+1}) - D_{r+1}(A^c_{r+2})] (directions mod 3) with
// A^c_d = a_d b - a b_d for the coordinates (a, b) following c, and D_d the
// scheme's difference along d at that index: (f[+1] - f[-1]) / 2, or for
// AWENO-5 on cells [0, n - 2 g) the face difference plus the CDS
// correction.  The discrete metric identities then hold to rounding.
constexpr double kFspW1 = 145.0 / 256, kFspW2 = -13.0 / 320, kFspW3 = 19.0 / 3840;

struct FspDims {
  GridDims G;
  int aweno;
  __device__ int n(int d) const { return d == 0 ? G.mx : (d == 1 ? G.my : G.mz); }
  __device__ int64_t cstride(int d) const { return d == 0 ? 1 : (d == 1 ? G.cd2 : G.cd1); }
};

// D_d of F at node (i, j, k) of the (g + 1)-ghosted layout; NaN when the
// stencil leaves it
__device__ inline double fsp_diff(const FspDims &S, const double *__restrict__ F, int d, int i, int j, int k) {
  const int c = d == 0 ? i : (d == 1 ? j : k);
  const int n = S.n(d), g = S.G.g;
  const bool wide = S.aweno && c >= 0 && c < n - 2 * g;
  const int w = wide ? 3 : 1;
  if (c - w < -(g + 1) || c + w > n + g) return __builtin_nan("");
  const int64_t st = S.cstride(d);
  const int64_t p = S.G.cat(i, j, k);
  if (!wide) return (F[p + st] - F[p - st]) * 0.5;
  return (F[p + st] - F[p - st]) * kFspW1 + (F[p + 2 * st] - F[p - 2 * st]) * kFspW2 +
         (F[p + 3 * st] - F[p - 3 * st]) * kFspW3;
}

// A[(c * 3 + d) * csz + p] at every node of the (g + 1)-ghosted layout
__global__ void __launch_bounds__(kGTB) k_fsp_products(FspDims S, int64_t csz, const double *__restrict__ X,
                                                       const double *__restrict__ Y, const double *__restrict__ Z,
                                                       double *__restrict__ A) {
  const int64_t t = (int64_t)blockIdx.x * kGTB + threadIdx.x;
  if (t >= csz) return;
  const int g1 = S.G.g + 1;
  const int nx = S.G.mx + 2 * g1, ny = S.G.my + 2 * g1;
  const int i = (int)(t % nx) - g1;
  const int j = (int)((t / nx) % ny) - g1;
  const int k = (int)(t / ((int64_t)nx * ny)) - g1;
  const double *C[3] = {X, Y, Z};
  for (int d = 0; d < 3; ++d) {
    double der[3];
    for (int c = 0; c < 3; ++c) der[c] = fsp_diff(S, C[c], d, i, j, k);
    for (int c = 0; c < 3; ++c) {
      const int a = (c + 1) % 3, b = (c + 2) % 3;
      A[(int64_t)(c * 3 + d) * csz + t] = der[a] * C[b][t] - C[a][t] * der[b];
    }
  }
}

// the conservative rows of every point of the g-ghosted layout whose
// stencils stay on the nodes, divided by the point's J
__global__ void __launch_bounds__(kGTB) k_fsp_rows(FspDims S, int64_t csz, const double *__restrict__ A,
                                                   const double *__restrict__ jac, double *__restrict__ met) {
  const int64_t t = (int64_t)blockIdx.x * kGTB + threadIdx.x;
  const GridDims &G = S.G;
  if (t >= G.sz) return;
  const int nx = G.mx + 2 * G.g, ny = G.my + 2 * G.g;
  const int i = (int)(t % nx) - G.g;
  const int j = (int)((t / nx) % ny) - G.g;
  const int k = (int)(t / ((int64_t)nx * ny)) - G.g;
  const double J = jac[t];
  for (int r = 0; r < 3; ++r) {
    const int d1 = (r + 1) % 3, d2 = (r + 2) % 3;
    double row[3];
    bool ok = true;
    for (int c = 0; c < 3; ++c) {
      row[c] = (fsp_diff(S, A + (int64_t)(c * 3 + d1) * csz, d2, i, j, k) -
                fsp_diff(S, A + (int64_t)(c * 3 + d2) * csz, d1, i, j, k)) *
               0.5;
      if (__builtin_isnan(row[c])) ok = false;
    }
    if (!ok) continue;
    for (int c = 0; c < 3; ++c) met[(int64_t)(r * 3 + c) * G.sz + t] = row[c] / J;
  }
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kGTB - 1) / kGTB); }

}  // namespace

hipError_t launch_metrics(int mx, int my, int mz, int g, const double *X, const double *Y, const double *Z,
                          double *jac, double *met, unsigned char *bad, int *interior_bad, hipStream_t s,
                          int conservative, int aweno, double *fsp_scratch) {
  GridDims G;
  G.mx = mx;
  G.my = my;
  G.mz = mz;
  G.g = g;
  G.cd2 = mx + 2 * (g + 1);
  G.cd1 = (int64_t)(my + 2 * (g + 1)) * G.cd2;
  G.d2 = mx + 2 * g;
  G.d1 = (int64_t)(my + 2 * g) * G.d2;
  G.sz = G.d1 * (mz + 2 * g);
  hipLaunchKernelGGL(k_metrics, dim3(nblk(G.sz)), dim3(kGTB), 0, s, G, X, Y, Z, jac, met, bad, interior_bad);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the six slabs in the order of Mesh.cpp (k low, k high, j low, j high, i low, i high)
  struct Slab {
    int axis, from, step, a0, a1, b0, b1;
  };
  const Slab slabs[6] = {
      {2, 0, -1, 1, mx - 1, 1, my - 1},           {2, mz - 1, 1, 1, mx - 1, 1, my - 1},
      {1, 0, -1, 1, mx - 1, -g, mz + g},          {1, my - 1, 1, 1, mx - 1, -g, mz + g},
      {0, 0, -1, -g, my + g, -g, mz + g},         {0, mx - 1, 1, -g, my + g, -g, mz + g},
  };
  for (const Slab &sl : slabs) {
    const int64_t n = (int64_t)std::max(0, sl.a1 - sl.a0) * std::max(0, sl.b1 - sl.b0);
    if (n == 0) continue;
    hipLaunchKernelGGL(k_metric_slab, dim3(nblk(n)), dim3(kGTB), 0, s, G, sl.axis, sl.from, sl.step, sl.a0, sl.a1,
                       sl.b0, sl.b1, jac, met, (const unsigned char *)bad);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (conservative) {
    if (!fsp_scratch) return hipErrorInvalidValue;
    FspDims S;
    S.G = G;
    S.aweno = aweno;
    const int64_t csz = G.cd1 * (mz + 2 * (g + 1));
    hipLaunchKernelGGL(k_fsp_products, dim3(nblk(csz)), dim3(kGTB), 0, s, S, csz, X, Y, Z, fsp_scratch);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fsp_rows, dim3(nblk(G.sz)), dim3(kGTB), 0, s, S, csz, (const double *)fsp_scratch,
                       (const double *)jac, met);
    e = hipGetLastError();
  }
  return e;
}

}  // namespace ostc
