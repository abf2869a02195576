// MLP chemistry surrogate: training-set generation with the FP64
// point-implicit integrator, BF16-MFMA training (forward, backward, Adam),
// gated inference and the per-cell FP64 fallback.
#include "surrogate.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <mutex>
#include <numeric>
#include <vector>

#include "device_chem.h"
#include "flux_kernels.h"
#include "mlp.h"

namespace ostc {

namespace {

constexpr int kDin[kSurLayers] = {11, 256, 256, 256};
constexpr int kDout[kSurLayers] = {256, 256, 256, kSurOut};
constexpr int kKpad[kSurLayers] = {16, 256, 256, 256};
constexpr int kLd[kSurLayers] = {24, 264, 264, 264};
constexpr int kOutPad[kSurLayers] = {256, 256, 256, 16};
constexpr int kXld = 24, kHld = 264, kYld = 16, kRawLd = 16;
constexpr int kNS = 9;
constexpr double kYFloor = 1e-12;
constexpr int64_t kInferChunk = 1 << 20;

// Standardisation, conservation rows and gate thresholds (kernel argument).
struct SurNorm {
  double x_mean[kSurIn], x_istd[kSurIn], x_lo[kSurIn], x_hi[kSurIn];
  double c_mean[kSurOut], c_std[kSurOut];  // of u = asinh(c / c_scale)
  double c_scale[kSurOut];
  double basis[kNS * kSurNull];  // dY_k = sum_j basis[k * 6 + j] c_j
  double cons[4 * kNS];          // element mass-fraction rows, then the mass row
  int n_cons;
  double cons_tol, dy_max, dT_max, sig_min, dt;
  // the input bounds on log(Y + 1e-12) and log p exponentiated, so the gate
  // tests Y + 1e-12 and p directly (log is monotone: the same test)
  double y_lo[kNS], y_hi[kNS], p_lo, p_hi;
};

// (log(Y + 1e-12), T, log p) of one cell; p from the equation of state as
// in the point-implicit batch kernel
__device__ inline void raw_features(const DMech *__restrict__ m, double rho, double T, const double *Y, double *x) {
  double mwi = 0;
#pragma unroll
  for (int l = 0; l < kNS; ++l) mwi += Y[l] / m->mw[l];
  const double mw = 1 / mwi;
  const double p = rho * T * (kRuD / mw);
#pragma unroll
  for (int l = 0; l < kNS; ++l) x[l] = log(fmax(Y[l], 0.0) + kYFloor);
  x[9] = T;
  x[10] = log(p);
}

__device__ inline void load_cell(int64_t n, int64_t c, const double *__restrict__ Y_a, double *Y) {
#pragma unroll
  for (int l = 0; l < kNS; ++l) Y[l] = Y_a[(int64_t)l * n + c];
}

// raw features F [n x 16] and raw targets Tt [n x 16]: null-space coordinates
// of dY, dT and sigma_min(I - dt J)
__global__ void k_sur_raw(const DMech *__restrict__ m, int64_t n, const double *__restrict__ rho,
                          const double *__restrict__ T, const double *__restrict__ Y_a, const double *__restrict__ dY,
                          const double *__restrict__ dT, const double *__restrict__ sig, SurNorm nm,
                          float *__restrict__ F, float *__restrict__ Tt) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  double Y[kNS], x[kSurIn], d[kNS];
  load_cell(n, c, Y_a, Y);
  raw_features(m, rho[c], T[c], Y, x);
#pragma unroll
  for (int i = 0; i < kRawLd; ++i) F[c * kRawLd + i] = i < kSurIn ? (float)x[i] : 0.f;
  load_cell(n, c, dY, d);
#pragma unroll
  for (int j = 0; j < kSurNull; ++j) {
    double s = 0;
#pragma unroll
    for (int k = 0; k < kNS; ++k) s += nm.basis[k * kSurNull + j] * d[k];
    Tt[c * kRawLd + j] = (float)s;
  }
  Tt[c * kRawLd + 6] = (float)dT[c];
  Tt[c * kRawLd + 7] = (float)sig[c];
#pragma unroll
  for (int i = kSurOut; i < kRawLd; ++i) Tt[c * kRawLd + i] = 0.f;
}

// per-block column sums, sums of squares, minima and maxima of A [n x 16]
__global__ void __launch_bounds__(256) k_colstats(const float *__restrict__ A, int64_t n, double *__restrict__ part) {
  const int col = threadIdx.x & 15, rl = threadIdx.x >> 4;
  double s = 0, q = 0, lo = 1e300, hi = -1e300;
  for (int64_t r = (int64_t)blockIdx.x * 16 + rl; r < n; r += (int64_t)gridDim.x * 16) {
    const double v = A[r * kRawLd + col];
    s += v;
    q += v * v;
    lo = fmin(lo, v);
    hi = fmax(hi, v);
  }
  __shared__ double sh[4][16][16];
  sh[0][rl][col] = s;
  sh[1][rl][col] = q;
  sh[2][rl][col] = lo;
  sh[3][rl][col] = hi;
  __syncthreads();
  if (rl == 0) {
    for (int r = 1; r < 16; ++r) {
      s += sh[0][r][col];
      q += sh[1][r][col];
      lo = fmin(lo, sh[2][r][col]);
      hi = fmax(hi, sh[3][r][col]);
    }
    double *o = part + ((int64_t)blockIdx.x * 16 + col) * 4;
    o[0] = s;
    o[1] = q;
    o[2] = lo;
    o[3] = hi;
  }
}

// u = asinh(c / scale) of the raw targets, in place
__global__ void k_sur_transform(int64_t n, float *__restrict__ Tt, SurNorm nm) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
#pragma unroll
  for (int j = 0; j < kSurOut; ++j) Tt[c * kRawLd + j] = (float)asinh((double)Tt[c * kRawLd + j] / nm.c_scale[j]);
}

__device__ inline void write_x_row(__bf16 *__restrict__ X, int64_t r, const double *xs) {
  __bf16 *o = X + r * kXld;
#pragma unroll
  for (int i = 0; i < kXld; ++i) o[i] = (__bf16)(i < kSurIn ? (float)xs[i] : (i == 16 ? 1.f : 0.f));
}

// standardise: F -> X (BF16, augmented), Tt in place
__global__ void k_sur_standardize(int64_t n, const float *__restrict__ F, float *__restrict__ Tt, SurNorm nm,
                                  __bf16 *__restrict__ X) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  double xs[kSurIn];
#pragma unroll
  for (int i = 0; i < kSurIn; ++i) xs[i] = ((double)F[c * kRawLd + i] - nm.x_mean[i]) * nm.x_istd[i];
  write_x_row(X, c, xs);
#pragma unroll
  for (int j = 0; j < kSurOut; ++j) Tt[c * kRawLd + j] = (float)(((double)Tt[c * kRawLd + j] - nm.c_mean[j]) / nm.c_std[j]);
}

// host-supplied standardised inputs [B x 11] -> X (augmented BF16)
__global__ void k_pack_x(int64_t B, const float *__restrict__ x, __bf16 *__restrict__ X) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= B) return;
  double xs[kSurIn];
#pragma unroll
  for (int i = 0; i < kSurIn; ++i) xs[i] = x[r * kSurIn + i];
  write_x_row(X, r, xs);
}

// standardised network inputs of cells c0 .. c0+count
__global__ void k_sur_features(const DMech *__restrict__ m, int64_t n, int64_t c0, int64_t count,
                               const double *__restrict__ rho, const double *__restrict__ T,
                               const double *__restrict__ Y_a, SurNorm nm, __bf16 *__restrict__ X) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= count) return;
  const int64_t c = c0 + t;
  double Y[kNS], x[kSurIn];
  load_cell(n, c, Y_a, Y);
  raw_features(m, rho[c], T[c], Y, x);
#pragma unroll
  for (int i = 0; i < kSurIn; ++i) x[i] = (x[i] - nm.x_mean[i]) * nm.x_istd[i];
  write_x_row(X, t, x);
}

// Decode the network output of cells c0 .. c0+count, apply the gate and
// either write (dY, dT) or append the cell to the fallback list.
//   gate: inputs inside the training range, finite outputs, |dY| and |dT|
//   within the accepted activity, predicted sigma_min(I - dt J) >= sig_min
//   (near-singular cells, where the point-implicit step is close to a pole,
//   stay with the integrator), element / mass conservation error of the
//   written state (conservation rows of dY plus the mass that mass fractions
//   outside [0, 1] would need clipped) <= cons_tol, converged temperature.
// The outputs are u_j = asinh(c_j / scale_j) standardised; c_j are the
// null-space coordinates of dY (j < 6), dT (j = 6) and sigma_min (j = 7).
// T is the Newton temperature at the unchanged internal energy started
// from T + dT_net, as in the point-implicit batch kernel.
// Decode one cell's network output o[0..6], apply the gate and, if the cell
// is accepted, write (dY, dT).  Returns 0 for an accepted cell, else the
// first gate condition that failed: 1 input outside the training range,
// 2 non-finite output, 3 |dY_k| above dy_max, 4 |dT| above dT_max,
// 5 sigma_min below sig_min, 6 conservation error above cons_tol,
// 7 temperature inversion failed.
__device__ inline int finish_cell(const DMech *__restrict__ m, const SurNorm &nm, int64_t n, int64_t c,
                                   const double *__restrict__ rho_a, const double *__restrict__ T_a,
                                   const double *__restrict__ Y_a, const float *o, double *__restrict__ dY,
                                   double *__restrict__ dT) {
  bool ok = true;
  int code = 0;
  const double rho = rho_a[c], T = T_a[c];
  double Y[kNS];
  load_cell(n, c, Y_a, Y);
  {
    // the training-range test of the raw features (raw_features), on
    // Y + 1e-12, T and p against the exponentiated bounds
    double mwi = 0;
#pragma unroll
    for (int l = 0; l < kNS; ++l) mwi += Y[l] / m->mw[l];
    const double p = rho * T * (kRuD / (1 / mwi));
#pragma unroll
    for (int l = 0; l < kNS; ++l) {
      const double y = fmax(Y[l], 0.0) + kYFloor;
      ok = ok && y >= nm.y_lo[l] && y <= nm.y_hi[l];
    }
    ok = ok && T >= nm.x_lo[9] && T <= nm.x_hi[9] && p >= nm.p_lo && p <= nm.p_hi;
    if (!ok) code = 1;
  }
  double cj[kSurOut];
#pragma unroll
  for (int j = 0; j < kSurOut; ++j) {
    // sinh u = (e^u - e^-u) / 2 from one exp: near u = 0 the difference
    // loses relative digits but its absolute error stays ~1e-16 c_scale,
    // far below the network's own error
    const double e = exp(nm.c_mean[j] + nm.c_std[j] * (double)o[j]);
    cj[j] = nm.c_scale[j] * (0.5 * (e - 1.0 / e));
  }
  double dy[kNS];
  double amax = 0;
#pragma unroll
  for (int k = 0; k < kNS; ++k) {
    double s = 0;
#pragma unroll
    for (int j = 0; j < kSurNull; ++j) s += nm.basis[k * kSurNull + j] * cj[j];
    dy[k] = s;
    amax = fmax(amax, fabs(s));
  }
  const double dTn = cj[6];
  if (code == 0 && !(isfinite(amax) && isfinite(dTn))) code = 2;
  if (code == 0 && !(amax <= nm.dy_max)) code = 3;
  if (code == 0 && !(fabs(dTn) <= nm.dT_max)) code = 4;
  if (code == 0 && !(cj[7] >= nm.sig_min)) code = 5;  // NaN fails as well
  ok = code == 0;
  double err = 0, neg = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    double s = 0;
#pragma unroll
    for (int k = 0; k < kNS; ++k) s += nm.cons[r * kNS + k] * dy[k];
    if (r < nm.n_cons) err = fmax(err, fabs(s));
  }
#pragma unroll
  for (int k = 0; k < kNS; ++k) neg += fmax(0.0, -(Y[k] + dy[k])) + fmax(0.0, Y[k] + dy[k] - 1.0);
  if (code == 0 && !(fmax(err, neg) <= nm.cons_tol)) code = 6;
  ok = code == 0;
  if (ok) {
    double mwi = 0, h[kNS];
#pragma unroll
    for (int l = 0; l < kNS; ++l) mwi += Y[l] / m->mw[l];
    const double mw = 1 / mwi;
    const double p_old = rho * T * (kRuD / mw);
    d_enthalpy<kNS>(T, m, h);
    double rhoE = 0;
#pragma unroll
    for (int l = 0; l < kNS; ++l) rhoE += h[l] * Y[l];
    rhoE = rhoE * rho - p_old;
    double Yn[kNS];
#pragma unroll
    for (int l = 0; l < kNS; ++l) Yn[l] = Y[l] + dy[l];
    const double t0 = fmin(fmax(T + dTn, 0.5 * T), 2.0 * T + 1000.0);
    double pn;
    const double Tn = d_temperature<kNS>(m, rho, rhoE, 0.0, 0.0, 0.0, Yn, t0, pn);
    ok = isfinite(Tn) && Tn > 0;
    if (ok) {
#pragma unroll
      for (int l = 0; l < kNS; ++l) dY[(int64_t)l * n + c] = dy[l];
      dT[c] = Tn - T;
    } else {
      code = 7;
    }
  }
  return code;
}

// Gate codes of one workgroup (index 8: forced fallback, a network trained
// for another dt) summed in LDS, then added to one of kReasonCopies striped
// copies of the counters (the host sums the copies), so thousands of
// workgroups do not serialise on one cache line; reasons may be null.
constexpr int kReasonCopies = 16;

// block-aggregated append of the cells with fb set to the fallback list: one
// global atomic per workgroup of 4 wavefronts
__device__ inline void append_fallback(bool fb, int64_t c, int *__restrict__ fb_count, int *__restrict__ fb_idx,
                                       int *s_wave, int &s_base) {
  const unsigned long long mask = __ballot(fb);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_wave[wave] = (int)__popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    const int total = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    s_base = total ? atomicAdd(fb_count, total) : 0;
  }
  __syncthreads();
  int base = s_base;
  for (int w = 0; w < wave; ++w) base += s_wave[w];
  if (fb) fb_idx[base + (int)__popcll(mask & ((1ull << lane) - 1ull))] = (int)c;
}

__global__ void __launch_bounds__(256) k_sur_finish(const DMech *__restrict__ m, SurNorm nm, int64_t n, int64_t c0,
                                                    int64_t count, const double *__restrict__ rho_a,
                                                    const double *__restrict__ T_a, const double *__restrict__ Y_a,
                                                    const float *__restrict__ out, int ld, int force_fallback,
                                                    double *__restrict__ dY, double *__restrict__ dT,
                                                    int *__restrict__ fb_count, int *__restrict__ fb_idx,
                                                    int *__restrict__ reasons) {
  __shared__ int s_codes[9];
  __shared__ int s_wave[4];
  __shared__ int s_base;
  if (threadIdx.x < 9) s_codes[threadIdx.x] = 0;
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool active = t < count;
  const int64_t c = c0 + t;
  int code = 8;
  if (active && !force_fallback) code = finish_cell(m, nm, n, c, rho_a, T_a, Y_a, out + t * ld, dY, dT);
  __syncthreads();  // s_codes cleared
  if (active && code > 0) atomicAdd(&s_codes[code], 1);
  append_fallback(active && code != 0, c, fb_count, fb_idx, s_wave, s_base);  // synchronises the block
  if (reasons && threadIdx.x >= 1 && threadIdx.x <= 8 && s_codes[threadIdx.x])
    atomicAdd(reasons + (blockIdx.x % kReasonCopies) * 16 + threadIdx.x, s_codes[threadIdx.x]);
}

// ------------------------------------------------------------ fused network
// The whole network of R cells in one workgroup of 4 wavefronts: features ->
// 3 hidden layers -> output layer, activations resident in LDS, weights read
// as MFMA B fragments straight from L2 (every workgroup reads the same 280 kB
// of BF16 weights) with a two-K-step register prefetch, so the K loop has no
// barrier.  Hidden layers: wave w computes output columns [64 w, 64 w + 64)
// for all R rows (R/16 x 4 tiles of v_mfma_f32_16x16x32_bf16, FP32
// accumulation); the output layer (8 columns, one 16-wide tile) splits the
// rows over the waves.  Same roundings as the layer-by-layer path: BF16
// operands, FP32 accumulation, the bias from the BF16 weight copy added in
// FP32.  Writes the standardised outputs out[c * 8 + j]; the gate runs in
// k_sur_finish.  LDS: R x 264 BF16 (66 kB for R = 128), two workgroups per CU.
template <int R, int PF>
__global__ void __launch_bounds__(256, 2) k_sur_net(const DMech *__restrict__ m, SurNorm nm, int64_t n,
                                                    int64_t count, const double *__restrict__ rho_a,
                                                    const double *__restrict__ T_a, const double *__restrict__ Y_a,
                                                    const __bf16 *__restrict__ wb, int64_t off1, int64_t off2,
                                                    int64_t off3, float *__restrict__ out) {
  static_assert(R % 64 == 0 && R <= 256, "rows per workgroup");
  constexpr int RT = R / 16;
  __shared__ __bf16 act[R][kHld];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = 8 * (lane >> 4);
  const int64_t r0 = (int64_t)blockIdx.x * R;
  const bf16x8 zero8 = {};

  // layer-0 B fragments (K = 24: 11 weights, zeros, the bias column, zeros)
  bf16x8 b0[4];
#pragma unroll
  for (int b = 0; b < 4; ++b)
    b0[b] = lk < kLd[0] ? *(const bf16x8 *)(wb + (int64_t)(wave * 64 + b * 16 + lr) * kLd[0] + lk) : zero8;
  for (int t = tid; t < R; t += 256) {
    const int64_t c = r0 + t;
    double x[kSurIn];
    if (c < count) {
      double Y[kNS];
      load_cell(n, c, Y_a, Y);
      raw_features(m, rho_a[c], T_a[c], Y, x);
#pragma unroll
      for (int i = 0; i < kSurIn; ++i) x[i] = (x[i] - nm.x_mean[i]) * nm.x_istd[i];
    } else {
#pragma unroll
      for (int i = 0; i < kSurIn; ++i) x[i] = 0;
    }
    // zeros beyond the 11 inputs: the bias column of W0 meets a zero here
#pragma unroll
    for (int i = 0; i < 32; ++i) act[t][i] = (__bf16)(i < kSurIn ? (float)x[i] : 0.f);
  }
  __syncthreads();

  // Hidden-layer tiles are computed transposed, C[neuron][cell] = W x act^T
  // (the weight fragment as the A operand, the activation fragment as B:
  // the same registers as act x W^T), so that a lane holds four consecutive
  // neurons of one cell and the epilogue stores them with one 8-byte write.
  f32x4 acc[RT][4];
  typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
  auto epilogue = [&](const __bf16 *W, int ld, int kp) {
    float bias[4][4];
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) bias[b][e] = (float)W[(int64_t)(wave * 64 + b * 16 + 4 * (lane >> 4) + e) * ld + kp];
    __syncthreads();  // every wave has read its last fragments of act
#pragma unroll
    for (int a = 0; a < RT; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        bf16x4 v;
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = (__bf16)gelu(acc[a][b][e] + bias[b][e]);
        *(bf16x4 *)&act[a * 16 + lr][wave * 64 + b * 16 + 4 * (lane >> 4)] = v;
      }
    __syncthreads();
  };

  // layer 0: one K step
#pragma unroll
  for (int a = 0; a < RT; ++a) {
    const bf16x8 af = *(const bf16x8 *)&act[a * 16 + lr][lk];
#pragma unroll
    for (int b = 0; b < 4; ++b)
      acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b0[b], af, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
  }
  epilogue(wb, kLd[0], kKpad[0]);

  // hidden layers 1, 2: 8 K steps of 32, B fragments prefetched PF steps ahead
  for (int l = 1; l < 3; ++l) {
    const __bf16 *W = l == 1 ? wb + off1 : wb + off2;
    const __bf16 *wrow[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) wrow[b] = W + (int64_t)(wave * 64 + b * 16 + lr) * kHld + lk;
    bf16x8 bq[PF + 1][4];
#pragma unroll
    for (int q = 0; q < PF; ++q)
#pragma unroll
      for (int b = 0; b < 4; ++b) bq[q][b] = *(const bf16x8 *)(wrow[b] + 32 * q);
#pragma unroll
    for (int a = 0; a < RT; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      if (ks + PF < 8)
#pragma unroll
        for (int b = 0; b < 4; ++b) bq[(ks + PF) % (PF + 1)][b] = *(const bf16x8 *)(wrow[b] + 32 * (ks + PF));
#pragma unroll
      for (int a = 0; a < RT; ++a) {
        const bf16x8 af = *(const bf16x8 *)&act[a * 16 + lr][32 * ks + lk];
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[ks % (PF + 1)][b], af, acc[a][b], 0, 0, 0);
      }
    }
    epilogue(W, kHld, 256);
  }

  // output layer: 8 columns (rows 8..15 of W3 are zero)
  {
    const __bf16 *W3 = wb + off3;
    constexpr int RW = RT / 4;
    f32x4 o[RW];
#pragma unroll
    for (int a = 0; a < RW; ++a) o[a] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const bf16x8 bfr = *(const bf16x8 *)&W3[(int64_t)lr * kLd[3] + 32 * ks + lk];
#pragma unroll
      for (int a = 0; a < RW; ++a) {
        const bf16x8 af = *(const bf16x8 *)&act[(wave * RW + a) * 16 + lr][32 * ks + lk];
        o[a] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfr, o[a], 0, 0, 0);
      }
    }
    if (lr < kSurOut) {
      const float b3 = (float)W3[(int64_t)lr * kLd[3] + kKpad[3]];
#pragma unroll
      for (int a = 0; a < RW; ++a)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int64_t c = r0 + (wave * RW + a) * 16 + 4 * (lane >> 4) + e;
          if (c < count) out[c * kSurOut + lr] = o[a][e] + b3;
        }
    }
  }
}

// k_sur_net with R = 128 and the hidden and output layers pipelined over
// the two halves (64 cells) of the workgroup: while the MFMAs of one half
// (or of the next layer's first half) run, the same wave applies the
// bias + GELU epilogue to the other half's tiles, two tiles per K step, so
// the matrix cores and the VALU (the GELU's exp and reciprocal) overlap.
// Cells are independent, so a half needs only its own rows of the previous
// layer; the halves alternate in the single activation buffer, with one
// barrier between phases:
//   M1(0) | M1(1)+E1(0) | M2(0)+E1(1) | M2(1)+E2(0) | O(0)+E2(1) | O(1)
// (Ml(h): layer-l products of half h, El(h): its epilogue, O: output layer).
// Every product and epilogue is the one of k_sur_net, in the same order.
template <int PF>
__global__ void __launch_bounds__(256, 2) k_sur_net_pipe(const DMech *__restrict__ m, SurNorm nm, int64_t n,
                                                         int64_t count, const double *__restrict__ rho_a,
                                                         const double *__restrict__ T_a,
                                                         const double *__restrict__ Y_a,
                                                         const __bf16 *__restrict__ wb, int64_t off1, int64_t off2,
                                                         int64_t off3, float *__restrict__ out) {
  constexpr int R = 128, RT = 8;
  __shared__ __bf16 act[R][kHld];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = 8 * (lane >> 4);
  const int64_t r0 = (int64_t)blockIdx.x * R;
  const bf16x8 zero8 = {};
  typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;

  bf16x8 b0[4];
#pragma unroll
  for (int b = 0; b < 4; ++b)
    b0[b] = lk < kLd[0] ? *(const bf16x8 *)(wb + (int64_t)(wave * 64 + b * 16 + lr) * kLd[0] + lk) : zero8;
  for (int t = tid; t < R; t += 256) {
    const int64_t c = r0 + t;
    double x[kSurIn];
    if (c < count) {
      double Y[kNS];
      load_cell(n, c, Y_a, Y);
      raw_features(m, rho_a[c], T_a[c], Y, x);
#pragma unroll
      for (int i = 0; i < kSurIn; ++i) x[i] = (x[i] - nm.x_mean[i]) * nm.x_istd[i];
    } else {
#pragma unroll
      for (int i = 0; i < kSurIn; ++i) x[i] = 0;
    }
#pragma unroll
    for (int i = 0; i < 32; ++i) act[t][i] = (__bf16)(i < kSurIn ? (float)x[i] : 0.f);
  }
  __syncthreads();

  f32x4 acc[RT][4];
  auto load_bias = [&](const __bf16 *W, int ld, int kp, float (&bias)[4][4]) {
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) bias[b][e] = (float)W[(int64_t)(wave * 64 + b * 16 + 4 * (lane >> 4) + e) * ld + kp];
  };
  // epilogue of tile (a, b): bias, GELU, BF16, one 8-byte LDS store
  auto epi_tile = [&](int a, int b, const float (&bias)[4][4]) {
    bf16x4 v;
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = (__bf16)gelu(acc[a][b][e] + bias[b][e]);
    *(bf16x4 *)&act[a * 16 + lr][wave * 64 + b * 16 + 4 * (lane >> 4)] = v;
  };

  // layer 0 (one K step) and its epilogue, not pipelined
#pragma unroll
  for (int a = 0; a < RT; ++a) {
    const bf16x8 af = *(const bf16x8 *)&act[a * 16 + lr][lk];
#pragma unroll
    for (int b = 0; b < 4; ++b)
      acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b0[b], af, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
  }
  {
    float bias0[4][4];
    load_bias(wb, kLd[0], kKpad[0], bias0);
    __syncthreads();
#pragma unroll
    for (int a = 0; a < RT; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) epi_tile(a, b, bias0);
    __syncthreads();
  }

  // layer products of half h (row tiles 4h .. 4h+3), 8 K steps with the
  // weight fragments prefetched PF steps ahead; with eh >= 0 the 16
  // epilogue tiles of half eh ride along, two per K step
  auto mma_half = [&](const __bf16 *W, int h, int eh, const float (&ebias)[4][4]) {
    const __bf16 *wrow[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) wrow[b] = W + (int64_t)(wave * 64 + b * 16 + lr) * kHld + lk;
    bf16x8 bq[PF + 1][4];
#pragma unroll
    for (int q = 0; q < PF; ++q)
#pragma unroll
      for (int b = 0; b < 4; ++b) bq[q][b] = *(const bf16x8 *)(wrow[b] + 32 * q);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[4 * h + a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      if (ks + PF < 8)
#pragma unroll
        for (int b = 0; b < 4; ++b) bq[(ks + PF) % (PF + 1)][b] = *(const bf16x8 *)(wrow[b] + 32 * (ks + PF));
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const bf16x8 af = *(const bf16x8 *)&act[(4 * h + a) * 16 + lr][32 * ks + lk];
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[4 * h + a][b] =
              __builtin_amdgcn_mfma_f32_16x16x32_bf16(bq[ks % (PF + 1)][b], af, acc[4 * h + a][b], 0, 0, 0);
      }
      if (eh >= 0) {
        epi_tile(4 * eh + (2 * ks) / 4, (2 * ks) % 4, ebias);
        epi_tile(4 * eh + (2 * ks + 1) / 4, (2 * ks + 1) % 4, ebias);
      }
    }
  };
  // output layer of half h: wave w takes row tile 4h + w (8 outputs, one
  // 16-wide tile; rows 8..15 of W3 are zero), with the epilogue tiles of
  // half eh riding along
  const __bf16 *W3 = wb + off3;
  auto out_half = [&](int h, int eh, const float (&ebias)[4][4]) {
    const int a = 4 * h + wave;
    f32x4 o = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
      const bf16x8 bfr = *(const bf16x8 *)&W3[(int64_t)lr * kLd[3] + 32 * ks + lk];
      const bf16x8 af = *(const bf16x8 *)&act[a * 16 + lr][32 * ks + lk];
      o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bfr, o, 0, 0, 0);
      if (eh >= 0) {
        epi_tile(4 * eh + (2 * ks) / 4, (2 * ks) % 4, ebias);
        epi_tile(4 * eh + (2 * ks + 1) / 4, (2 * ks + 1) % 4, ebias);
      }
    }
    if (lr < kSurOut) {
      const float b3 = (float)W3[(int64_t)lr * kLd[3] + kKpad[3]];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t c = r0 + a * 16 + 4 * (lane >> 4) + e;
        if (c < count) out[c * kSurOut + lr] = o[e] + b3;
      }
    }
  };

  const __bf16 *W1 = wb + off1, *W2 = wb + off2;
  float b1[4][4], b2[4][4];
  load_bias(W1, kHld, 256, b1);
  load_bias(W2, kHld, 256, b2);
  mma_half(W1, 0, -1, b1);
  __syncthreads();  // half 0 of the layer-1 input is read
  mma_half(W1, 1, 0, b1);
  __syncthreads();  // layer-1 output of half 0 written, half-1 input read
  mma_half(W2, 0, 1, b1);
  __syncthreads();
  mma_half(W2, 1, 0, b2);
  __syncthreads();
  out_half(0, 1, b2);
  __syncthreads();
  out_half(1, -1, b2);
}

inline unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double unit(uint64_t &s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

struct SurrogateState {
  int device = 0;
  SurNorm norm{};
  SurrogateGate gate{};
  double x_min[kSurIn]{}, x_max[kSurIn]{};
  bool ready = false;
  int n_reac = 0;  // reactions of the mechanism (sizes the integrator's LDS park)
  double dt_train = 0;
  int64_t woff[kSurLayers]{}, nw = 0;
  float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  __bf16 *wb = nullptr;
  float *aug = nullptr;  // [1, 0 x 7]: the augmented columns of an activation row
  int adam_t = 0;
  // training workspace for batches up to tcap rows
  int64_t tcap = 0;
  __bf16 *tX = nullptr, *tH[3] = {};
  float *tT = nullptr, *tZ[3] = {}, *tY = nullptr, *tG = nullptr, *tD[2] = {}, *loss = nullptr;
  // inference workspace for chunks up to icap rows, fallback list up to fcap cells
  int64_t icap = 0, fcap = 0;
  __bf16 *iX = nullptr, *iH[2] = {};
  float *iOut = nullptr;
  int *fb_idx = nullptr, *fb_count = nullptr;
  int *reasons = nullptr;  // gate codes 1..8 of the last inference (kReasonCopies x 16 ints)
  float *fb_out = nullptr;  // network outputs [fcap x 8] of the fused path
  int kernel = 3;  // inference: 3 fused network, 128 cells per workgroup, halves pipelined (default);
                   // 1 the same unpipelined; 2 fused, 64 cells; 0 layer-by-layer GEMMs
  hipEvent_t ev[4] = {};
  double last_ms[3] = {0, 0, 0};  // network, gate, fallback
  int64_t last_fallback = 0;

  void free_train() {
    for (void *p : {(void *)tX, (void *)tH[0], (void *)tH[1], (void *)tH[2], (void *)tT, (void *)tZ[0],
                    (void *)tZ[1], (void *)tZ[2], (void *)tY, (void *)tG, (void *)tD[0], (void *)tD[1]})
      (void)hipFree(p);
    tX = nullptr;
    tH[0] = tH[1] = tH[2] = nullptr;
    tT = tY = tG = nullptr;
    tZ[0] = tZ[1] = tZ[2] = nullptr;
    tD[0] = tD[1] = nullptr;
    tcap = 0;
  }
  void free_infer() {
    for (void *p : {(void *)iX, (void *)iH[0], (void *)iH[1], (void *)iOut}) (void)hipFree(p);
    iX = nullptr;
    iH[0] = iH[1] = nullptr;
    iOut = nullptr;
    icap = 0;
  }
  ~SurrogateState() {
    (void)hipSetDevice(device);
    free_train();
    free_infer();
    for (void *p : {(void *)w, (void *)g, (void *)m1, (void *)m2, (void *)wb, (void *)aug, (void *)loss,
                    (void *)fb_idx, (void *)fb_count, (void *)fb_out, (void *)reasons})
      (void)hipFree(p);
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

namespace {

#define SCHK(expr)                                                                  \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) {                                                         \
      err = std::string(#expr) + ": " + hipGetErrorString(e_);                      \
      return 1;                                                                     \
    }                                                                               \
  } while (0)

template <typename T>
hipError_t dalloc(T **p, int64_t count) {
  return hipMalloc((void **)p, (size_t)std::max<int64_t>(count, 1) * sizeof(T));
}

int ensure_train_ws(SurrogateState *s, int64_t B, hipStream_t st, std::string &err) {
  if (B <= s->tcap) return 0;
  s->free_train();
  SCHK(dalloc(&s->tX, B * kXld));
  for (int l = 0; l < 3; ++l) {
    SCHK(dalloc(&s->tH[l], B * kHld));
    SCHK(launch_fill_cols_bf16(s->tH[l], B, kHld, 256, s->aug, st));
    SCHK(dalloc(&s->tZ[l], B * kHld));
  }
  SCHK(dalloc(&s->tT, B * kYld));
  SCHK(dalloc(&s->tY, B * kYld));
  SCHK(hipMemsetAsync(s->tY, 0, (size_t)B * kYld * sizeof(float), st));
  SCHK(dalloc(&s->tG, B * kYld));
  SCHK(dalloc(&s->tD[0], B * kHld));
  SCHK(dalloc(&s->tD[1], B * kHld));
  s->tcap = B;
  return 0;
}

int ensure_infer_ws(SurrogateState *s, int64_t chunk, int64_t n, hipStream_t st, std::string &err) {
  if (chunk > s->icap) {
    s->free_infer();
    SCHK(dalloc(&s->iX, chunk * kXld));
    for (int l = 0; l < 2; ++l) {
      SCHK(dalloc(&s->iH[l], chunk * kHld));
      SCHK(launch_fill_cols_bf16(s->iH[l], chunk, kHld, 256, s->aug, st));
    }
    SCHK(dalloc(&s->iOut, chunk * kYld));
    s->icap = chunk;
  }
  if (n > s->fcap) {
    (void)hipFree(s->fb_idx);
    (void)hipFree(s->fb_out);
    s->fb_idx = nullptr;
    s->fb_out = nullptr;
    SCHK(dalloc(&s->fb_idx, n));
    SCHK(dalloc(&s->fb_out, n * kSurOut));
    s->fcap = n;
  }
  return 0;
}

int split_for(int M, int N, int K) {
  const int tiles = ((M + 63) / 64) * ((N + 63) / 64);
  int sp = std::max(1, 512 / tiles);
  return std::max(1, std::min(sp, K / 256));
}

// forward pass; Z (saved pre-activations) may be null (inference)
int forward(SurrogateState *s, int64_t M, const __bf16 *X, __bf16 *const *H, float *const *Z, float *Y, bool bf16w,
            hipStream_t st, std::string &err) {
  const void *A = X;
  for (int l = 0; l < kSurLayers; ++l) {
    GemmArgs g;
    g.M = (int)M;
    g.N = kDout[l];
    g.K = kLd[l];
    g.lda = kLd[l];
    g.ldb = kLd[l];
    g.transB = 1;
    const void *Bw = bf16w ? (const void *)(s->wb + s->woff[l]) : (const void *)(s->w + s->woff[l]);
    if (l < kSurLayers - 1) {
      g.mode = EPI_GELU;
      g.Z = Z ? Z[l] : nullptr;
      g.Cb = H[l];
      g.ldc = kHld;
    } else {
      g.mode = EPI_STORE;
      g.Cf = Y;
      g.ldc = kYld;
    }
    SCHK(launch_gemm(g, A, true, Bw, bf16w, 1, st));
    A = H[l];
  }
  return 0;
}

// loss + backward into s->g (zeroed here); forward must have filled tZ, tH, tY
int backward(SurrogateState *s, int64_t B, const __bf16 *X, const float *T, hipStream_t st, std::string &err) {
  SCHK(hipMemsetAsync(s->g, 0, (size_t)s->nw * sizeof(float), st));
  SCHK(launch_mse(s->tY, T, (int)B, kSurOut, kYld, s->tG, s->loss, st));
  // dZ of the current layer: G for the output layer, then tD ping-pong
  const float *dZ = s->tG;
  int ldz = kYld;
  for (int l = kSurLayers - 1; l >= 0; --l) {
    // weight (and bias) gradient: dW_l = dZ^T [h, 1]
    GemmArgs gw;
    gw.M = kDout[l];
    gw.N = kKpad[l] + 1;
    gw.K = (int)B;
    gw.transA = 1;
    gw.lda = ldz;
    gw.ldb = l == 0 ? kXld : kHld;
    gw.mode = EPI_ATOMIC;
    gw.Cf = s->g + s->woff[l];
    gw.ldc = kLd[l];
    const void *Hin = l == 0 ? (const void *)X : (const void *)s->tH[l - 1];
    SCHK(launch_gemm(gw, dZ, false, Hin, true, split_for(gw.M, gw.N, gw.K), st));
    if (l == 0) break;
    // dZ_{l-1} = (dZ_l W_l) * GELU'(Z_{l-1})
    GemmArgs gd;
    gd.M = (int)B;
    gd.N = kDin[l];
    gd.K = kOutPad[l];
    gd.lda = ldz;
    gd.ldb = kLd[l];
    gd.mode = EPI_GELU_GRAD;
    gd.Zin = s->tZ[l - 1];
    float *dst = s->tD[(l + 1) & 1];
    gd.Cf = dst;
    gd.ldc = kHld;
    SCHK(launch_gemm(gd, dZ, false, s->w + s->woff[l], false, 1, st));
    dZ = dst;
    ldz = kHld;
  }
  return 0;
}

void finish_norm(SurrogateState *s) {
  for (int i = 0; i < kSurIn; ++i) {
    const double r = s->x_max[i] - s->x_min[i];
    s->norm.x_lo[i] = s->x_min[i] - s->gate.margin * r - 1e-6;
    s->norm.x_hi[i] = s->x_max[i] + s->gate.margin * r + 1e-6;
  }
  for (int l = 0; l < kNS; ++l) {
    s->norm.y_lo[l] = std::exp(s->norm.x_lo[l]);
    s->norm.y_hi[l] = std::exp(s->norm.x_hi[l]);
  }
  s->norm.p_lo = std::exp(s->norm.x_lo[10]);
  s->norm.p_hi = std::exp(s->norm.x_hi[10]);
  s->norm.cons_tol = s->gate.cons_tol;
  s->norm.dy_max = s->gate.dy_max;
  s->norm.dT_max = s->gate.dT_max;
  s->norm.sig_min = s->gate.sig_min;
  s->norm.dt = s->dt_train;
}

}  // namespace

int surrogate_create(const Mechanism &mech, int device, SurrogateState **out, std::string &err) {
  *out = nullptr;
  if (mech.n_spec != kNS) {
    err = "the surrogate is built for 9-species mechanisms";
    return 1;
  }
  auto s = std::make_unique<SurrogateState>();
  s->device = device;
  s->n_reac = mech.n_reac;
  SCHK(hipSetDevice(device));
  // conservation rows: mass fraction of each element per unit dY_k, then mass
  const int ne = (int)mech.elements.size();
  std::vector<std::vector<double>> rows;
  for (int j = 0; j < ne; ++j) {
    const double aw = atomic_weight(mech.elements[j]);
    if (aw <= 0) {
      err = "unknown element " + mech.elements[j];
      return 1;
    }
    std::vector<double> r(kNS);
    for (int k = 0; k < kNS; ++k) r[k] = mech.elem_comp[(size_t)k * ne + j] * aw / mech.mw[k];
    rows.push_back(r);
  }
  rows.push_back(std::vector<double>(kNS, 1.0));
  if ((int)rows.size() > 4) {
    err = "too many elements for the surrogate's conservation gate";
    return 1;
  }
  s->norm.n_cons = (int)rows.size();
  for (size_t r = 0; r < rows.size(); ++r)
    for (int k = 0; k < kNS; ++k) s->norm.cons[r * kNS + k] = rows[r][k];
  // orthonormal basis of the row space (modified Gram-Schmidt), then of its complement
  std::vector<std::vector<double>> q;
  auto orth = [&](std::vector<double> v, double tol) {
    const double n0 = std::sqrt(std::inner_product(v.begin(), v.end(), v.begin(), 0.0));
    for (int pass = 0; pass < 2; ++pass)
      for (auto &u : q) {
        const double d = std::inner_product(u.begin(), u.end(), v.begin(), 0.0);
        for (int k = 0; k < kNS; ++k) v[k] -= d * u[k];
      }
    const double nn = std::sqrt(std::inner_product(v.begin(), v.end(), v.begin(), 0.0));
    if (nn <= tol * n0) return false;
    for (auto &e : v) e /= nn;
    q.push_back(v);
    return true;
  };
  for (auto &r : rows) orth(r, 1e-10);
  const size_t rank = q.size();
  int nnull = 0;
  for (int k = 0; k < kNS && nnull < kSurNull + 1; ++k) {
    std::vector<double> e(kNS, 0.0);
    e[k] = 1;
    if (orth(e, 1e-8)) ++nnull;
  }
  if (nnull > kSurNull) {
    err = "conservation null space larger than the network's 6 outputs";
    return 1;
  }
  for (int j = 0; j < kSurNull; ++j)
    for (int k = 0; k < kNS; ++k) s->norm.basis[k * kSurNull + j] = j < nnull ? q[rank + j][k] : 0.0;
  for (int i = 0; i < kSurIn; ++i) {
    s->norm.x_mean[i] = 0;
    s->norm.x_istd[i] = 1;
    s->x_min[i] = -1e300;
    s->x_max[i] = 1e300;
  }
  for (int j = 0; j < kSurOut; ++j) {
    s->norm.c_mean[j] = 0;
    s->norm.c_std[j] = 1;
    s->norm.c_scale[j] = 1;
  }
  finish_norm(s.get());
  // parameters
  int64_t off = 0;
  for (int l = 0; l < kSurLayers; ++l) {
    s->woff[l] = off;
    off += (int64_t)kOutPad[l] * kLd[l];
  }
  s->nw = off;
  SCHK(dalloc(&s->w, off));
  SCHK(dalloc(&s->g, off));
  SCHK(dalloc(&s->m1, off));
  SCHK(dalloc(&s->m2, off));
  SCHK(dalloc(&s->wb, off));
  SCHK(hipMemset(s->w, 0, off * sizeof(float)));
  SCHK(dalloc(&s->aug, 8));
  const float aug[8] = {1, 0, 0, 0, 0, 0, 0, 0};
  SCHK(hipMemcpy(s->aug, aug, sizeof(aug), hipMemcpyHostToDevice));
  SCHK(dalloc(&s->loss, 1));
  SCHK(dalloc(&s->fb_count, 1));
  SCHK(dalloc(&s->reasons, kReasonCopies * 16));
  for (auto &e : s->ev) SCHK(hipEventCreate(&e));
  *out = s.release();
  return 0;
}

void surrogate_destroy(SurrogateState *s) { delete s; }

bool surrogate_ready(const SurrogateState *s) { return s && s->ready; }

int64_t surrogate_param_count(const SurrogateState *s) { return s ? s->nw : 0; }

void surrogate_set_gate(SurrogateState *s, const SurrogateGate &g) {
  s->gate = g;
  finish_norm(s);
}

SurrogateGate surrogate_get_gate(const SurrogateState *s) { return s->gate; }

int surrogate_get_norm(const SurrogateState *s, double *o) {
  for (int i = 0; i < kSurIn; ++i) {
    o[i] = s->norm.x_mean[i];
    o[11 + i] = 1.0 / s->norm.x_istd[i];
    o[22 + i] = s->x_min[i];
    o[33 + i] = s->x_max[i];
  }
  for (int j = 0; j < kSurOut; ++j) {
    o[44 + j] = s->norm.c_mean[j];
    o[52 + j] = s->norm.c_std[j];
  }
  o[60] = s->dt_train;
  for (int k = 0; k < kNS * kSurNull; ++k) o[61 + k] = s->norm.basis[k];
  for (int j = 0; j < kSurOut; ++j) o[115 + j] = s->norm.c_scale[j];
  return 0;
}

static int upload_params(SurrogateState *s, const std::vector<float> &h, hipStream_t st, std::string &err) {
  SCHK(hipMemcpyAsync(s->w, h.data(), (size_t)s->nw * sizeof(float), hipMemcpyHostToDevice, st));
  SCHK(hipMemsetAsync(s->m1, 0, (size_t)s->nw * sizeof(float), st));
  SCHK(hipMemsetAsync(s->m2, 0, (size_t)s->nw * sizeof(float), st));
  SCHK(launch_to_bf16(s->w, s->wb, s->nw, st));
  SCHK(hipStreamSynchronize(st));
  s->adam_t = 0;
  return 0;
}

int surrogate_init_params(SurrogateState *s, uint64_t seed, hipStream_t st, std::string &err) {
  SCHK(hipSetDevice(s->device));
  std::vector<float> h((size_t)s->nw, 0.f);
  uint64_t rng = seed;
  for (int l = 0; l < kSurLayers; ++l) {
    // uniform with variance 2 / fan_in (GELU hidden layers), 1 / fan_in (output layer)
    const double a = std::sqrt((l < kSurLayers - 1 ? 6.0 : 3.0) / kDin[l]);
    for (int r = 0; r < kDout[l]; ++r)
      for (int c = 0; c < kDin[l]; ++c) h[(size_t)(s->woff[l] + (int64_t)r * kLd[l] + c)] = (float)(a * (2 * unit(rng) - 1));
  }
  return upload_params(s, h, st, err);
}

int surrogate_get_params(SurrogateState *s, float *host, hipStream_t st, std::string &err) {
  SCHK(hipSetDevice(s->device));
  SCHK(hipMemcpyAsync(host, s->w, (size_t)s->nw * sizeof(float), hipMemcpyDeviceToHost, st));
  SCHK(hipStreamSynchronize(st));
  return 0;
}

int surrogate_set_params(SurrogateState *s, const float *host, hipStream_t st, std::string &err) {
  SCHK(hipSetDevice(s->device));
  std::vector<float> h(host, host + s->nw);
  // keep the padding exactly zero
  for (int l = 0; l < kSurLayers; ++l)
    for (int r = 0; r < kOutPad[l]; ++r)
      for (int c = 0; c < kLd[l]; ++c)
        if (r >= kDout[l] || (c >= kDin[l] && c != kKpad[l])) h[(size_t)(s->woff[l] + (int64_t)r * kLd[l] + c)] = 0.f;
  return upload_params(s, h, st, err);
}

static int stage_batch(SurrogateState *s, int64_t B, const float *x, const float *t, hipStream_t st,
                       std::string &err) {
  if (B <= 0 || B % 8) {
    err = "batch size must be a positive multiple of 8";
    return 1;
  }
  if (ensure_train_ws(s, B, st, err)) return 1;
  float *xin = nullptr;
  SCHK(dalloc(&xin, B * kSurIn));
  SCHK(hipMemcpyAsync(xin, x, (size_t)B * kSurIn * sizeof(float), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_pack_x, dim3(nb256(B)), dim3(256), 0, st, B, (const float *)xin, s->tX);
  SCHK(hipGetLastError());
  if (t) {
    std::vector<float> tp((size_t)B * kYld, 0.f);
    for (int64_t r = 0; r < B; ++r)
      for (int j = 0; j < kSurOut; ++j) tp[(size_t)(r * kYld + j)] = t[r * kSurOut + j];
    SCHK(hipMemcpyAsync(s->tT, tp.data(), tp.size() * sizeof(float), hipMemcpyHostToDevice, st));
    SCHK(hipStreamSynchronize(st));
  } else {
    SCHK(hipStreamSynchronize(st));
  }
  SCHK(hipFree(xin));
  return 0;
}

int surrogate_forward(SurrogateState *s, int64_t B, const float *x, float *y, hipStream_t st, std::string &err) {
  SCHK(hipSetDevice(s->device));
  if (stage_batch(s, B, x, nullptr, st, err)) return 1;
  if (forward(s, B, s->tX, s->tH, s->tZ, s->tY, false, st, err)) return 1;
  std::vector<float> yp((size_t)B * kYld);
  SCHK(hipMemcpyAsync(yp.data(), s->tY, yp.size() * sizeof(float), hipMemcpyDeviceToHost, st));
  SCHK(hipStreamSynchronize(st));
  for (int64_t r = 0; r < B; ++r)
    for (int j = 0; j < kSurOut; ++j) y[r * kSurOut + j] = yp[(size_t)(r * kYld + j)];
  return 0;
}

int surrogate_gradient(SurrogateState *s, int64_t B, const float *x, const float *t, float *grad, double *loss,
                       hipStream_t st, std::string &err) {
  SCHK(hipSetDevice(s->device));
  if (stage_batch(s, B, x, t, st, err)) return 1;
  SCHK(hipMemsetAsync(s->loss, 0, sizeof(float), st));
  if (forward(s, B, s->tX, s->tH, s->tZ, s->tY, false, st, err)) return 1;
  if (backward(s, B, s->tX, s->tT, st, err)) return 1;
  float l = 0;
  SCHK(hipMemcpyAsync(&l, s->loss, sizeof(float), hipMemcpyDeviceToHost, st));
  SCHK(hipMemcpyAsync(grad, s->g, (size_t)s->nw * sizeof(float), hipMemcpyDeviceToHost, st));
  SCHK(hipStreamSynchronize(st));
  if (loss) *loss = l;
  return 0;
}

int surrogate_train(SurrogateState *s, const DMech *m, int64_t n, const double *rho, const double *T,
                    const double *Y, double dt, const SurrogateTrainCfg &cfg, double *loss_hist, hipStream_t st,
                    std::string &err) {
  SCHK(hipSetDevice(s->device));
  const int64_t B = cfg.batch;
  if (B <= 0 || B % 8 || n < B || n >= (int64_t)1 << 31) {
    err = "training needs a batch that is a positive multiple of 8, and batch <= n < 2^31";
    return 1;
  }
  if (!(dt > 0)) {
    err = "dt must be positive";
    return 1;
  }
  // 1. targets from the FP64 point-implicit integrator
  double *dY = nullptr, *dTt = nullptr;
  SCHK(dalloc(&dY, n * kNS));
  SCHK(dalloc(&dTt, n));
  SCHK(launch_chem_batch(m, s->n_reac, n, rho, T, Y, dt, dY, dTt, st));
  double *sig = nullptr;
  SCHK(dalloc(&sig, n));
  SCHK(launch_epi_sigma(m, n, rho, T, Y, dt, sig, st));
  // 2. raw features / targets and their statistics
  float *F = nullptr, *Tt = nullptr;
  SCHK(dalloc(&F, n * kRawLd));
  SCHK(dalloc(&Tt, n * kRawLd));
  hipLaunchKernelGGL(k_sur_raw, dim3(nb256(n)), dim3(256), 0, st, m, n, rho, T, Y, (const double *)dY,
                     (const double *)dTt, (const double *)sig, s->norm, F, Tt);
  SCHK(hipGetLastError());
  SCHK(hipFree(dY));
  SCHK(hipFree(dTt));
  SCHK(hipFree(sig));
  // asinh scale of each target: the median |c_j| of an evenly strided sample
  {
    const int64_t ns = std::min<int64_t>(n, 65536), stride = n / ns;
    std::vector<float> smp((size_t)ns * kRawLd);
    SCHK(hipMemcpy2DAsync(smp.data(), kRawLd * sizeof(float), Tt, (size_t)stride * kRawLd * sizeof(float),
                          kRawLd * sizeof(float), (size_t)ns, hipMemcpyDeviceToHost, st));
    SCHK(hipStreamSynchronize(st));
    std::vector<double> col((size_t)ns);
    for (int j = 0; j < kSurOut; ++j) {
      double amax = 0;
      for (int64_t r = 0; r < ns; ++r) {
        col[(size_t)r] = std::fabs((double)smp[(size_t)(r * kRawLd + j)]);
        amax = std::max(amax, col[(size_t)r]);
      }
      std::nth_element(col.begin(), col.begin() + ns / 2, col.end());
      double med = col[(size_t)(ns / 2)];
      if (!(med > 1e-9 * amax) || !(med > 0)) med = amax > 0 ? amax : 1.0;
      s->norm.c_scale[j] = med;
    }
  }
  hipLaunchKernelGGL(k_sur_transform, dim3(nb256(n)), dim3(256), 0, st, n, Tt, s->norm);
  SCHK(hipGetLastError());
  const int nblk = 1024;
  double *part = nullptr;
  SCHK(dalloc(&part, (int64_t)nblk * 16 * 4 * 2));
  hipLaunchKernelGGL(k_colstats, dim3(nblk), dim3(256), 0, st, (const float *)F, n, part);
  hipLaunchKernelGGL(k_colstats, dim3(nblk), dim3(256), 0, st, (const float *)Tt, n, part + nblk * 16 * 4);
  SCHK(hipGetLastError());
  std::vector<double> hp((size_t)nblk * 16 * 4 * 2);
  SCHK(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  SCHK(hipStreamSynchronize(st));
  SCHK(hipFree(part));
  auto stats = [&](int which, int col, double &mean, double &sd, double &lo, double &hi) {
    double sm = 0, sq = 0;
    lo = 1e300;
    hi = -1e300;
    for (int b = 0; b < nblk; ++b) {
      const double *o = &hp[(size_t)(((which * nblk + b) * 16 + col) * 4)];
      sm += o[0];
      sq += o[1];
      lo = std::min(lo, o[2]);
      hi = std::max(hi, o[3]);
    }
    mean = sm / (double)n;
    sd = std::sqrt(std::max(0.0, sq / (double)n - mean * mean));
    if (!(sd > 1e-9 * (1.0 + std::fabs(mean)))) sd = 1.0;  // constant column
  };
  for (int i = 0; i < kSurIn; ++i) {
    double sd;
    stats(0, i, s->norm.x_mean[i], sd, s->x_min[i], s->x_max[i]);
    s->norm.x_istd[i] = 1.0 / sd;
  }
  for (int j = 0; j < kSurOut; ++j) {
    double lo, hi;
    stats(1, j, s->norm.c_mean[j], s->norm.c_std[j], lo, hi);
  }
  finish_norm(s);
  // 3. standardised training set
  __bf16 *Xall = nullptr;
  SCHK(dalloc(&Xall, n * kXld));
  hipLaunchKernelGGL(k_sur_standardize, dim3(nb256(n)), dim3(256), 0, st, n, (const float *)F, Tt, s->norm, Xall);
  SCHK(hipGetLastError());
  SCHK(hipFree(F));
  if (ensure_train_ws(s, B, st, err)) return 1;
  int *perm = nullptr;
  SCHK(dalloc(&perm, n));
  std::vector<int> hperm((size_t)n);
  const int64_t steps = n / B;
  const int64_t total = steps * cfg.epochs;
  int64_t gstep = 0;
  SCHK(hipMemsetAsync(s->m1, 0, (size_t)s->nw * sizeof(float), st));
  SCHK(hipMemsetAsync(s->m2, 0, (size_t)s->nw * sizeof(float), st));
  s->adam_t = 0;
  s->ready = false;
  uint64_t rng = cfg.seed ^ 0x5DEECE66Dull;
  for (int64_t e = 0; e < cfg.epochs; ++e) {
    for (int64_t i = 0; i < n; ++i) hperm[(size_t)i] = (int)i;
    for (int64_t i = n - 1; i > 0; --i) std::swap(hperm[(size_t)i], hperm[(size_t)(splitmix(rng) % (uint64_t)(i + 1))]);
    SCHK(hipMemcpyAsync(perm, hperm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, st));
    SCHK(hipMemsetAsync(s->loss, 0, sizeof(float), st));
    for (int64_t k = 0; k < steps; ++k, ++gstep) {
      const int *idx = perm + k * B;
      SCHK(launch_gather_rows_bf16(Xall, idx, (int)B, kXld, s->tX, st));
      SCHK(launch_gather_rows_f32(Tt, idx, (int)B, kYld, s->tT, st));
      if (forward(s, B, s->tX, s->tH, s->tZ, s->tY, false, st, err)) return 1;
      if (backward(s, B, s->tX, s->tT, st, err)) return 1;
      const double frac = total > 1 ? (double)gstep / (double)(total - 1) : 1.0;
      const double lr = cfg.lr_min + 0.5 * (cfg.lr - cfg.lr_min) * (1 + std::cos(M_PI * frac));
      SCHK(launch_adam(s->w, s->g, s->m1, s->m2, s->nw, (float)lr, (float)cfg.beta1, (float)cfg.beta2,
                       (float)cfg.eps, ++s->adam_t, st));
    }
    float l = 0;
    SCHK(hipMemcpyAsync(&l, s->loss, sizeof(float), hipMemcpyDeviceToHost, st));
    SCHK(hipStreamSynchronize(st));
    if (loss_hist) loss_hist[e] = (double)l / (double)steps;
  }
  SCHK(launch_to_bf16(s->w, s->wb, s->nw, st));
  SCHK(hipStreamSynchronize(st));
  SCHK(hipFree(perm));
  SCHK(hipFree(Xall));
  SCHK(hipFree(Tt));
  s->dt_train = dt;
  finish_norm(s);
  s->ready = true;
  return 0;
}

// fused network of n cells into out [n x 8]
static int launch_net(SurrogateState *s, const DMech *m, int64_t n, const double *rho, const double *T,
                      const double *Y, float *out, hipStream_t st, std::string &err) {
  if (s->kernel == 2)
    hipLaunchKernelGGL((k_sur_net<64, 2>), dim3((unsigned)((n + 63) / 64)), dim3(256), 0, st, m, s->norm, n, n, rho, T, Y,
                       (const __bf16 *)s->wb, s->woff[1], s->woff[2], s->woff[3], out);
  else if (s->kernel == 3)
    hipLaunchKernelGGL((k_sur_net_pipe<1>), dim3((unsigned)((n + 127) / 128)), dim3(256), 0, st, m, s->norm, n, n,
                       rho, T, Y, (const __bf16 *)s->wb, s->woff[1], s->woff[2], s->woff[3], out);
  else
    hipLaunchKernelGGL((k_sur_net<128, 1>), dim3((unsigned)((n + 127) / 128)), dim3(256), 0, st, m, s->norm, n, n, rho,
                       T, Y, (const __bf16 *)s->wb, s->woff[1], s->woff[2], s->woff[3], out);
  SCHK(hipGetLastError());
  return 0;
}

void surrogate_set_kernel(SurrogateState *s, int kernel) { s->kernel = kernel; }

int surrogate_network_outputs(SurrogateState *s, const DMech *m, int64_t n, const double *rho, const double *T,
                              const double *Y, float *host_out, hipStream_t st, std::string &err) {
  SCHK(hipSetDevice(s->device));
  if (n <= 0) return 0;
  if (n >= (int64_t)1 << 31) {
    err = "batches are limited to 2^31 - 1 cells";
    return 1;
  }
  float *raw = nullptr;
  SCHK(dalloc(&raw, n * kSurOut));
  if (s->kernel != 0) {
    if (launch_net(s, m, n, rho, T, Y, raw, st, err)) return 1;
  } else {
    const int64_t chunk = std::min<int64_t>(n, kInferChunk);
    if (ensure_infer_ws(s, (chunk + 7) / 8 * 8, 0, st, err)) return 1;
    std::vector<float> tmp;
    for (int64_t c0 = 0; c0 < n; c0 += chunk) {
      const int64_t cnt = std::min(chunk, n - c0);
      hipLaunchKernelGGL(k_sur_features, dim3(nb256(cnt)), dim3(256), 0, st, m, n, c0, cnt, rho, T, Y, s->norm,
                         s->iX);
      SCHK(hipGetLastError());
      __bf16 *H[3] = {s->iH[0], s->iH[1], s->iH[0]};
      if (forward(s, cnt, s->iX, H, nullptr, s->iOut, true, st, err)) return 1;
      SCHK(hipMemcpy2DAsync(raw + c0 * kSurOut, kSurOut * sizeof(float), s->iOut, kYld * sizeof(float),
                            kSurOut * sizeof(float), (size_t)cnt, hipMemcpyDeviceToDevice, st));
    }
  }
  SCHK(hipMemcpyAsync(host_out, raw, (size_t)n * kSurOut * sizeof(float), hipMemcpyDeviceToHost, st));
  SCHK(hipStreamSynchronize(st));
  SCHK(hipFree(raw));
  return 0;
}

int surrogate_infer(SurrogateState *s, const DMech *m, int64_t n, const double *rho, const double *T,
                    const double *Y, double dt, double *dY, double *dT, int64_t *n_fallback, hipStream_t st,
                    std::string &err) {
  if (!surrogate_ready(s)) {
    err = "surrogate not trained or loaded";
    return 1;
  }
  if (n >= (int64_t)1 << 31) {
    err = "surrogate batches are limited to 2^31 - 1 cells";
    return 1;
  }
  SCHK(hipSetDevice(s->device));
  if (n <= 0) {
    if (n_fallback) *n_fallback = 0;
    return 0;
  }
  // a network trained for another dt is not used: every cell falls back
  const int force = std::fabs(dt - s->dt_train) > 1e-12 * s->dt_train ? 1 : 0;
  const int64_t chunk = std::min<int64_t>(n, kInferChunk);
  if (ensure_infer_ws(s, s->kernel == 0 ? (chunk + 7) / 8 * 8 : 0, n, st, err)) return 1;
  SCHK(hipEventRecord(s->ev[0], st));
  SCHK(hipMemsetAsync(s->fb_count, 0, sizeof(int), st));
  SCHK(hipMemsetAsync(s->reasons, 0, kReasonCopies * 16 * sizeof(int), st));
  if (s->kernel != 0) {
    if (!force && launch_net(s, m, n, rho, T, Y, s->fb_out, st, err)) return 1;
    SCHK(hipEventRecord(s->ev[3], st));
    hipLaunchKernelGGL(k_sur_finish, dim3(nb256(n)), dim3(256), 0, st, m, s->norm, n, (int64_t)0, n, rho, T, Y,
                       (const float *)s->fb_out, kSurOut, force, dY, dT, s->fb_count, s->fb_idx, s->reasons);
    SCHK(hipGetLastError());
  }
  for (int64_t c0 = 0; c0 < n && s->kernel == 0; c0 += chunk) {
    const int64_t cnt = std::min(chunk, n - c0);
    if (!force) {
      hipLaunchKernelGGL(k_sur_features, dim3(nb256(cnt)), dim3(256), 0, st, m, n, c0, cnt, rho, T, Y, s->norm,
                         s->iX);
      SCHK(hipGetLastError());
      __bf16 *H[3] = {s->iH[0], s->iH[1], s->iH[0]};
      if (forward(s, cnt, s->iX, H, nullptr, s->iOut, true, st, err)) return 1;
    }
    hipLaunchKernelGGL(k_sur_finish, dim3(nb256(cnt)), dim3(256), 0, st, m, s->norm, n, c0, cnt, rho, T, Y,
                       (const float *)s->iOut, kYld, force, dY, dT, s->fb_count, s->fb_idx, s->reasons);
    SCHK(hipGetLastError());
  }
  SCHK(hipEventRecord(s->ev[1], st));
  int nf = 0;
  SCHK(hipMemcpyAsync(&nf, s->fb_count, sizeof(int), hipMemcpyDeviceToHost, st));
  SCHK(hipStreamSynchronize(st));
  if (nf < 0 || nf > n) {
    err = "fallback count out of range";
    return 1;
  }
  SCHK(launch_chem_batch_subset(m, s->n_reac, n, nf, s->fb_idx, rho, T, Y, dt, dY, dT, st));
  SCHK(hipEventRecord(s->ev[2], st));
  SCHK(hipEventSynchronize(s->ev[2]));
  float a = 0, b = 0, c = 0;
  if (s->kernel != 0) {
    SCHK(hipEventElapsedTime(&a, s->ev[0], s->ev[3]));
    SCHK(hipEventElapsedTime(&c, s->ev[3], s->ev[1]));
  } else {
    SCHK(hipEventElapsedTime(&a, s->ev[0], s->ev[1]));
  }
  SCHK(hipEventElapsedTime(&b, s->ev[1], s->ev[2]));
  s->last_ms[0] = a;
  s->last_ms[1] = c;
  s->last_ms[2] = b;
  s->last_fallback = nf;
  if (n_fallback) *n_fallback = nf;
  return 0;
}

// ------------------------------------------------- allocation-free inference
// Fixed device scratch of surrogate_infer_noalloc: the network outputs
// (kSurOut FP32 per cell), the fallback index list and the fallback counter
// for batches of up to kNoAllocCells cells.  Each GPU the library runs on
// holds its own copy until the process ends (36 MiB + 256 B).  Host code
// never passes `scratch` itself to a kernel: its device address is taken
// once per device with hipGetSymbolAddress (noalloc_scratch_init).
namespace {
constexpr int64_t kNoAllocOutBytes = kNoAllocCells * kSurOut * (int64_t)sizeof(float);
constexpr int64_t kNoAllocIdxBytes = kNoAllocCells * (int64_t)sizeof(int);
constexpr int64_t kNoAllocBytes = kNoAllocOutBytes + kNoAllocIdxBytes + 256;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kNoAllocMaxDevices] = {};

__global__ void k_zero_int(int *__restrict__ p) {
  if (threadIdx.x == 0) p[0] = 0;
}
}  // namespace

int noalloc_scratch_init(int device, std::string &err) {
  if (device < 0 || device >= kNoAllocMaxDevices) {
    err = "device index out of range for the fixed scratch";
    return 1;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  int cur = -1;
  SCHK(hipGetDevice(&cur));
  if (cur != device) {
    err = "fixed scratch set-up needs its device current";
    return 1;
  }
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return 1;
  }
  g_scratch_ptr[device] = (unsigned char *)ptr;
  return 0;
}

int surrogate_infer_noalloc(SurrogateState *s, const DMech *m, int64_t n, const double *rho, const double *T,
                            const double *Y, double dt, double *dY, double *dT, int64_t *n_fallback, hipStream_t st,
                            std::string &err) {
  if (!surrogate_ready(s)) {
    err = "surrogate not trained or loaded";
    return 1;
  }
  if (n > kNoAllocCells) {
    err = "the allocation-free surrogate call takes at most " + std::to_string(kNoAllocCells) + " cells, got " +
          std::to_string(n);
    return 1;
  }
  if (s->kernel == 0) {
    err = "the allocation-free surrogate call needs a fused inference kernel (1, 2 or 3), not 0";
    return 1;
  }
  unsigned char *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (s->device >= 0 && s->device < kNoAllocMaxDevices) base = g_scratch_ptr[s->device];
  }
  if (!base) {
    err = "fixed scratch not set up on this device";
    return 1;
  }
  SCHK(hipSetDevice(s->device));
  if (n <= 0) {
    if (n_fallback) *n_fallback = 0;
    return 0;
  }
  float *out = (float *)base;
  int *idx = (int *)(base + kNoAllocOutBytes);
  int *count = (int *)(base + kNoAllocOutBytes + kNoAllocIdxBytes);
  const int force = std::fabs(dt - s->dt_train) > 1e-12 * s->dt_train ? 1 : 0;
  SCHK(hipEventRecord(s->ev[0], st));
  // the counter is cleared by a kernel, not a memset node
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, count);
  SCHK(hipGetLastError());
  if (!force && launch_net(s, m, n, rho, T, Y, out, st, err)) return 1;
  SCHK(hipEventRecord(s->ev[3], st));
  hipLaunchKernelGGL(k_sur_finish, dim3(nb256(n)), dim3(256), 0, st, m, s->norm, n, (int64_t)0, n, rho, T, Y,
                     (const float *)out, kSurOut, force, dY, dT, count, idx, (int *)nullptr);
  SCHK(hipGetLastError());
  SCHK(hipEventRecord(s->ev[1], st));
  int nf = 0;
  SCHK(hipMemcpyAsync(&nf, count, sizeof(int), hipMemcpyDeviceToHost, st));
  SCHK(hipStreamSynchronize(st));
  if (nf < 0 || nf > n) {
    err = "fallback count out of range";
    return 1;
  }
  SCHK(launch_chem_batch_subset(m, s->n_reac, n, nf, idx, rho, T, Y, dt, dY, dT, st));
  SCHK(hipEventRecord(s->ev[2], st));
  SCHK(hipEventSynchronize(s->ev[2]));
  float a = 0, b = 0, c = 0;
  SCHK(hipEventElapsedTime(&a, s->ev[0], s->ev[3]));
  SCHK(hipEventElapsedTime(&c, s->ev[3], s->ev[1]));
  SCHK(hipEventElapsedTime(&b, s->ev[1], s->ev[2]));
  s->last_ms[0] = a;
  s->last_ms[1] = c;
  s->last_ms[2] = b;
  s->last_fallback = nf;
  if (n_fallback) *n_fallback = nf;
  return 0;
}

// ---------------------------------------------------------------- file IO
namespace {
const char kMagic[8] = {'O', 'S', 'T', 'C', 'S', 'U', 'R', '3'};
}

int surrogate_save(SurrogateState *s, const std::string &path, hipStream_t st, std::string &err) {
  if (!surrogate_ready(s)) {
    err = "nothing to save: surrogate not trained or loaded";
    return 1;
  }
  std::vector<float> w((size_t)s->nw);
  if (surrogate_get_params(s, w.data(), st, err)) return 1;
  std::ofstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open " + path;
    return 1;
  }
  const int32_t dims[4] = {kSurIn, kSurOut, kSurLayers, 256};
  f.write(kMagic, 8);
  f.write((const char *)dims, sizeof(dims));
  f.write((const char *)&s->dt_train, sizeof(double));
  f.write((const char *)s->norm.x_mean, sizeof(s->norm.x_mean));
  f.write((const char *)s->norm.x_istd, sizeof(s->norm.x_istd));
  f.write((const char *)s->x_min, sizeof(s->x_min));
  f.write((const char *)s->x_max, sizeof(s->x_max));
  f.write((const char *)s->norm.c_mean, sizeof(s->norm.c_mean));
  f.write((const char *)s->norm.c_std, sizeof(s->norm.c_std));
  f.write((const char *)s->norm.c_scale, sizeof(s->norm.c_scale));
  f.write((const char *)s->norm.basis, sizeof(s->norm.basis));
  const int64_t nw = s->nw;
  f.write((const char *)&nw, sizeof(nw));
  f.write((const char *)w.data(), (std::streamsize)(w.size() * sizeof(float)));
  if (!f) {
    err = "write failed: " + path;
    return 1;
  }
  return 0;
}

int surrogate_load(SurrogateState *s, const std::string &path, hipStream_t st, std::string &err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open " + path;
    return 1;
  }
  char magic[8];
  int32_t dims[4];
  f.read(magic, 8);
  f.read((char *)dims, sizeof(dims));
  if (!f || std::memcmp(magic, kMagic, 8) != 0 || dims[0] != kSurIn || dims[1] != kSurOut ||
      dims[2] != kSurLayers || dims[3] != 256) {
    err = path + ": not a surrogate file of this network shape";
    return 1;
  }
  SurNorm nm = s->norm;
  double dt = 0, xmin[kSurIn], xmax[kSurIn], basis[kNS * kSurNull];
  f.read((char *)&dt, sizeof(double));
  f.read((char *)nm.x_mean, sizeof(nm.x_mean));
  f.read((char *)nm.x_istd, sizeof(nm.x_istd));
  f.read((char *)xmin, sizeof(xmin));
  f.read((char *)xmax, sizeof(xmax));
  f.read((char *)nm.c_mean, sizeof(nm.c_mean));
  f.read((char *)nm.c_std, sizeof(nm.c_std));
  f.read((char *)nm.c_scale, sizeof(nm.c_scale));
  f.read((char *)basis, sizeof(basis));
  int64_t nw = 0;
  f.read((char *)&nw, sizeof(nw));
  if (!f || nw != s->nw) {
    err = path + ": truncated or wrong parameter count";
    return 1;
  }
  for (int k = 0; k < kNS * kSurNull; ++k)
    if (std::fabs(basis[k] - s->norm.basis[k]) > 1e-12) {
      err = path + ": trained for a mechanism with other species or elements";
      return 1;
    }
  std::vector<float> w((size_t)nw);
  f.read((char *)w.data(), (std::streamsize)(w.size() * sizeof(float)));
  if (!f) {
    err = path + ": truncated parameters";
    return 1;
  }
  if (surrogate_set_params(s, w.data(), st, err)) return 1;
  s->norm = nm;
  std::copy(xmin, xmin + kSurIn, s->x_min);
  std::copy(xmax, xmax + kSurIn, s->x_max);
  s->dt_train = dt;
  finish_norm(s);
  s->ready = true;
  return 0;
}

int surrogate_gate_reasons(SurrogateState *s, int64_t *out9, hipStream_t st, std::string &err) {
  int h[kReasonCopies * 16] = {};
  SCHK(hipMemcpyAsync(h, s->reasons, sizeof(h), hipMemcpyDeviceToHost, st));
  SCHK(hipStreamSynchronize(st));
  out9[0] = s->last_fallback;
  for (int r = 1; r <= 8; ++r) {
    out9[r] = 0;
    for (int k = 0; k < kReasonCopies; ++k) out9[r] += h[k * 16 + r];
  }
  return 0;
}

int surrogate_last_stats(const SurrogateState *s, double *out4) {
  out4[0] = (double)s->last_fallback;
  out4[1] = s->last_ms[0];
  out4[2] = s->last_ms[1];
  out4[3] = s->last_ms[2];
  return 0;
}

}  // namespace ostc
