// BF16 MFMA GEMM and the small kernels of MLP training (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ostc {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

// GELU, tanh form: x (1 + tanh u) / 2 = x / (1 + exp(-2u)),
// u = sqrt(2/pi) (x + 0.044715 x^3); and its derivative.  The logistic
// factor uses the hardware reciprocal (1 ulp; exp(-2u) = inf gives 0), as
// the value is rounded to BF16 next; training and inference share it.
// exp(-2u) = 2^(-x (a + b x^2)), a = 2 sqrt(2/pi) / ln 2, b = 0.044715 a: one
// multiply-add and the hardware base-2 exponential
constexpr float kGeluA = 2.3022082f, kGeluB = 0.10294324f;
__device__ inline float gelu_logistic(float x) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-x * __builtin_fmaf(kGeluB, x * x, kGeluA)));
}
__device__ inline float gelu(float x) { return x * gelu_logistic(x); }
__device__ inline float gelu_grad(float x) {
  const float s = gelu_logistic(x);
  return s + x * s * (1.0f - s) * 1.5957691216057308f * (1.0f + 0.134145f * x * x);
}

enum GemmEpilogue {
  EPI_STORE = 0,      // C = acc (+ bias)
  EPI_GELU = 1,       // Z = acc (+ bias) saved if Z != null, C = GELU(Z)
  EPI_GELU_GRAD = 2,  // C = acc * GELU'(Zin)
  EPI_ATOMIC = 3      // Cf += acc (FP32 atomics; split-K accumulation)
};

// C[M x N] = op(A)[M x K] op(B)[K x N].  Row-major operands:
//   transA = 0: A is M x K (lda >= K);      transA = 1: A is K x M (lda >= M rounded up to 8)
//   transB = 0: B is K x N (ldb >= N rounded up to 8);  transB = 1: B is N x K (ldb >= K)
// K, lda and ldb are multiples of 8 and the base pointers 16-byte aligned.
// Outputs: Cf (FP32) and/or Cb (BF16), both with leading dimension ldc;
// Z / Zin (FP32) share ldc.
struct GemmArgs {
  int M = 0, N = 0, K = 0;
  int lda = 0, ldb = 0, ldc = 0;
  int transA = 0, transB = 0;
  int mode = EPI_STORE;
  const float *bias = nullptr;
  float *Cf = nullptr;
  __bf16 *Cb = nullptr;
  float *Z = nullptr;
  const float *Zin = nullptr;
};

hipError_t launch_gemm(const GemmArgs &g, const void *A, bool a_bf16, const void *B, bool b_bf16, int split_k,
                       hipStream_t s);
hipError_t launch_to_bf16(const float *src, __bf16 *dst, int64_t n, hipStream_t s);
// dst[r * cols + c] = v for all r < rows and c0 <= c < cols (fills the augmented columns)
hipError_t launch_fill_cols_bf16(__bf16 *dst, int64_t rows, int cols, int c0, const float *vals, hipStream_t s);
hipError_t launch_mse(const float *Y, const float *T, int M, int nout, int ld, float *G, float *loss_acc,
                      hipStream_t s);
hipError_t launch_adam(float *w, const float *g, float *m1, float *m2, int64_t n, float lr, float b1, float b2,
                       float eps, int step, hipStream_t s);
// dst row r = src row idx[r] (cols values per row)
hipError_t launch_gather_rows_bf16(const __bf16 *src, const int *idx, int m, int cols, __bf16 *dst, hipStream_t s);
hipError_t launch_gather_rows_f32(const float *src, const int *idx, int m, int cols, float *dst, hipStream_t s);

}  // namespace ostc
