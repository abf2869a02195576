// BF16 MFMA kernels of the chemistry surrogate MLP (gfx950).
//
// One tiled GEMM covers every product of training and inference:
//   C[M x N] = op(A)[M x K] * op(B)[K x N]      (f32 accumulation)
// with op = identity or transpose of row-major operands held in FP32 or
// BF16, staged through LDS as BF16, multiplied with
// v_mfma_f32_16x16x32_bf16 and finished by a fused epilogue (bias, GELU,
// GELU' of a saved pre-activation, FP32 / BF16 stores, or FP32 atomic
// accumulation for split-K weight gradients).
//
// Tile: 64 x 64 outputs per workgroup of 4 wavefronts (each 32 x 32 = 2 x 2
// MFMA tiles), K step 32.  LDS rows are padded to 40 BF16 (80 B) so the 16-B
// fragment reads of a wavefront spread over the banks.
//
// MFMA 16x16x32 BF16 operand layout (lane l of a wave64):
//   A fragment: row l%16, k = 8*(l/16) + 0..7
//   B fragment: col l%16, k = 8*(l/16) + 0..7
//   C/D f32x4 : col l%16, row = 4*(l/16) + 0..3
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp.h"

namespace ostc {

namespace {

constexpr int BM = 64, BN = 64, BK = 32, LDK = 40;

__device__ inline __bf16 to_bf(float x) { return (__bf16)x; }
__device__ inline float from_bf(__bf16 x) { return (float)x; }


template <typename T>
__device__ inline void load8(const T *p, float *v) {
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = (float)p[e];
}

template <typename TA, typename TB>
__global__ void __launch_bounds__(256) k_gemm(GemmArgs g, const TA *__restrict__ A, const TB *__restrict__ B) {
  __shared__ __bf16 As[BM][LDK];
  __shared__ __bf16 Bs[BN][LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  // split-K: blockIdx.z owns K range [kz0, kz1)
  // (chunks are whole K steps so every 8-element load stays 16-byte aligned)
  const int kchunk = (((g.K + (int)gridDim.z - 1) / (int)gridDim.z) + BK - 1) / BK * BK;
  const int kz0 = blockIdx.z * kchunk, kz1 = min(g.K, kz0 + kchunk);
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  for (int k0 = kz0; k0 < kz1; k0 += BK) {
    // ---- stage A tile (64 x 32) ----
    {
      float v[8];
      if (!g.transA) {  // A stored M x K (lda), k contiguous
        const int m = tid >> 2, kk = (tid & 3) * 8;
        const int gm = m0 + m, gk = k0 + kk;
        if (gm < g.M && gk < kz1)
          load8(A + (int64_t)gm * g.lda + gk, v);
        else
#pragma unroll
          for (int e = 0; e < 8; ++e) v[e] = 0.f;
#pragma unroll
        for (int e = 0; e < 8; ++e) As[m][kk + e] = (gk + e < kz1) ? to_bf(v[e]) : to_bf(0.f);
      } else {  // A stored K x M (lda), m contiguous
        const int k = tid >> 3, mm = (tid & 7) * 8;
        const int gk = k0 + k, gm = m0 + mm;
        if (gk < kz1 && gm < g.M)
          load8(A + (int64_t)gk * g.lda + gm, v);
        else
#pragma unroll
          for (int e = 0; e < 8; ++e) v[e] = 0.f;
#pragma unroll
        for (int e = 0; e < 8; ++e) As[mm + e][k] = (gm + e < g.M) ? to_bf(v[e]) : to_bf(0.f);
      }
    }
    // ---- stage B tile (32 x 64), held as Bs[n][k] ----
    {
      float v[8];
      if (!g.transB) {  // B stored K x N (ldb), n contiguous
        const int k = tid >> 3, nn = (tid & 7) * 8;
        const int gk = k0 + k, gn = n0 + nn;
        if (gk < kz1 && gn < g.N)
          load8(B + (int64_t)gk * g.ldb + gn, v);
        else
#pragma unroll
          for (int e = 0; e < 8; ++e) v[e] = 0.f;
#pragma unroll
        for (int e = 0; e < 8; ++e) Bs[nn + e][k] = (gn + e < g.N) ? to_bf(v[e]) : to_bf(0.f);
      } else {  // B stored N x K (ldb), k contiguous
        const int n = tid >> 2, kk = (tid & 3) * 8;
        const int gn = n0 + n, gk = k0 + kk;
        if (gn < g.N && gk < kz1)
          load8(B + (int64_t)gn * g.ldb + gk, v);
        else
#pragma unroll
          for (int e = 0; e < 8; ++e) v[e] = 0.f;
#pragma unroll
        for (int e = 0; e < 8; ++e) Bs[n][kk + e] = (gk + e < kz1) ? to_bf(v[e]) : to_bf(0.f);
      }
    }
    __syncthreads();
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int r = wm * 32 + t * 16 + (lane & 15);
      const int c = wn * 32 + t * 16 + (lane & 15);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        af[t][e] = As[r][8 * (lane >> 4) + e];
        bfr[t][e] = Bs[c][8 * (lane >> 4) + e];
      }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }

  // ---- epilogue ----
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = m0 + wm * 32 + a * 16 + 4 * (lane >> 4) + e;
        const int col = n0 + wn * 32 + b * 16 + (lane & 15);
        if (row >= g.M || col >= g.N) continue;
        float v = acc[a][b][e];
        const int64_t o = (int64_t)row * g.ldc + col;
        if (g.mode == EPI_ATOMIC) {
          atomicAdd(&g.Cf[o], v);
          continue;
        }
        if (g.bias) v += g.bias[col];
        if (g.mode == EPI_GELU) {
          if (g.Z) g.Z[o] = v;
          v = gelu(v);
        } else if (g.mode == EPI_GELU_GRAD) {
          v *= gelu_grad(g.Zin[o]);
        }
        if (g.Cf) g.Cf[o] = v;
        if (g.Cb) g.Cb[o] = to_bf(v);
      }
}

}  // namespace

hipError_t launch_gemm(const GemmArgs &g, const void *A, bool a_bf16, const void *B, bool b_bf16, int split_k,
                       hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.K <= 0) return hipSuccess;
  auto r8 = [](int v) { return (v + 7) / 8 * 8; };
  // every 8-wide operand load must stay inside its row
  if (g.K % 8 || g.lda % 8 || g.ldb % 8) return hipErrorInvalidValue;
  if (g.transA ? g.lda < r8(g.M) : g.lda < g.K) return hipErrorInvalidValue;
  if (g.transB ? g.ldb < g.K : g.ldb < r8(g.N)) return hipErrorInvalidValue;
  if (g.ldc < g.N || (!g.Cf && !g.Cb)) return hipErrorInvalidValue;
  if (g.mode == EPI_ATOMIC && !g.Cf) return hipErrorInvalidValue;
  if (g.mode == EPI_GELU_GRAD && !g.Zin) return hipErrorInvalidValue;
  if (split_k < 1) split_k = 1;
  if (split_k > 1 && g.mode != EPI_ATOMIC) return hipErrorInvalidValue;
  const dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM, split_k), blk(256);
  if (a_bf16 && b_bf16)
    hipLaunchKernelGGL((k_gemm<__bf16, __bf16>), grid, blk, 0, s, g, (const __bf16 *)A, (const __bf16 *)B);
  else if (a_bf16 && !b_bf16)
    hipLaunchKernelGGL((k_gemm<__bf16, float>), grid, blk, 0, s, g, (const __bf16 *)A, (const float *)B);
  else if (!a_bf16 && b_bf16)
    hipLaunchKernelGGL((k_gemm<float, __bf16>), grid, blk, 0, s, g, (const float *)A, (const __bf16 *)B);
  else
    hipLaunchKernelGGL((k_gemm<float, float>), grid, blk, 0, s, g, (const float *)A, (const float *)B);
  return hipGetLastError();
}

namespace {

__global__ void k_to_bf16(const float *__restrict__ src, __bf16 *__restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = to_bf(src[i]);
}

__global__ void k_fill_cols_bf16(__bf16 *__restrict__ dst, int64_t rows, int cols, int c0,
                                 const float *__restrict__ vals) {
  const int w = cols - c0;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= rows * w) return;
  const int64_t r = i / w;
  const int c = (int)(i - r * w);
  dst[r * cols + c0 + c] = to_bf(vals[c]);
}

// MSE over the first nout of ld columns: loss = sum (y - t)^2 / (M nout),
// G = d loss / d y (zero in the padding columns)
__global__ void k_mse(const float *__restrict__ Y, const float *__restrict__ T, int M, int nout, int ld,
                      float *__restrict__ G, float *__restrict__ loss_acc) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  float l = 0.f;
  if (i < (int64_t)M * ld) {
    const int c = (int)(i % ld);
    float g = 0.f;
    if (c < nout) {
      const float d = Y[i] - T[i];
      l = d * d;
      g = 2.f * d / ((float)M * (float)nout);
    }
    G[i] = g;
  }
  for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
  __shared__ float part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0) {
    const float b = part[0] + part[1] + part[2] + part[3];
    if (b != 0.f) atomicAdd(loss_acc, b / ((float)M * (float)nout));
  }
}

// Adam with bias correction on FP32 master weights
__global__ void k_adam(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1, float *__restrict__ m2,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// gather rows idx[0..m) of a row-major matrix with `cols` columns
template <typename T>
__global__ void k_gather_rows(const T *__restrict__ src, const int *__restrict__ idx, int m, int cols,
                              T *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)m * cols) return;
  const int r = (int)(i / cols), c = (int)(i % cols);
  dst[i] = src[(int64_t)idx[r] * cols + c];
}

inline unsigned nb(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

hipError_t launch_to_bf16(const float *src, __bf16 *dst, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_to_bf16, dim3(nb(n)), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

hipError_t launch_fill_cols_bf16(__bf16 *dst, int64_t rows, int cols, int c0, const float *vals, hipStream_t s) {
  if (rows <= 0 || c0 >= cols) return hipSuccess;
  hipLaunchKernelGGL(k_fill_cols_bf16, dim3(nb(rows * (cols - c0))), dim3(256), 0, s, dst, rows, cols, c0, vals);
  return hipGetLastError();
}

hipError_t launch_mse(const float *Y, const float *T, int M, int nout, int ld, float *G, float *loss_acc,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_mse, dim3(nb((int64_t)M * ld)), dim3(256), 0, s, Y, T, M, nout, ld, G, loss_acc);
  return hipGetLastError();
}

hipError_t launch_adam(float *w, const float *g, float *m1, float *m2, int64_t n, float lr, float b1, float b2,
                       float eps, int step, hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(nb(n)), dim3(256), 0, s, w, g, m1, m2, n, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t launch_gather_rows_bf16(const __bf16 *src, const int *idx, int m, int cols, __bf16 *dst, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_rows<__bf16>, dim3(nb((int64_t)m * cols)), dim3(256), 0, s, src, idx, m, cols, dst);
  return hipGetLastError();
}

hipError_t launch_gather_rows_f32(const float *src, const int *idx, int m, int cols, float *dst, hipStream_t s) {
  if (m <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather_rows<float>, dim3(nb((int64_t)m * cols)), dim3(256), 0, s, src, idx, m, cols, dst);
  return hipGetLastError();
}

}  // namespace ostc
