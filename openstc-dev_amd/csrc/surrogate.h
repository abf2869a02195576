// MLP chemistry surrogate (no counterpart in the reference): a 3x256 GELU
// (tanh form) network mapping (log Y, T, log p) of a cell to its point-implicit
// increments (dY, dT) over one dt, trained and run with BF16 MFMA and FP32
// accumulation, with a per-cell fallback to the FP64 integrator.
//
// Network: 11 -> 256 -> 256 -> 256 -> 8.  The 8 outputs encode the 6
// coordinates c_j of dY in an orthonormal basis of the null space of the
// element/mass conservation rows (so dY conserves elements and mass to FP64
// rounding by construction), dT, and sigma, the smallest singular value of
// the point-implicit matrix I - dt J (the step's conditioning, which the gate
// uses), each as u_j = asinh(c_j / s_j) with s_j the median |c_j| of the
// training set (the point-implicit increments span many decades).  Inputs and
// u are standardised with the mean and standard deviation of the training set.
//
// Parameter layout (FP32, one contiguous array): layer l is a row-major
// [out_pad_l x ld_l] matrix W_l with the bias in column kpad_l
// (augmented-input form  z = [h, 1, 0...] W_l^T):
//   l  in   kpad  ld   out  out_pad
//   0  11   16    24   256  256
//   1  256  256   264  256  256
//   2  256  256   264  256  256
//   3  256  256   264  8    16
// Entries outside (out x in) and the bias column are zero.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "device_common.h"
#include "ostc_internal.h"

namespace ostc {

struct SurrogateState;

struct SurrogateTrainCfg {
  int64_t epochs = 8, batch = 8192;
  double lr = 2e-3, lr_min = 2e-5, beta1 = 0.9, beta2 = 0.999, eps = 1e-8;
  uint64_t seed = 2004;
};

struct SurrogateGate {
  double cons_tol = 1e-6;  // element / mass conservation error bound (fallback above)
  // largest |dY_k| and |dT| [K] accepted from the network: larger increments
  // (the stiff, fast-reacting cells, about 7% of the 900-2400 K H2/air batch
  // at dt = 1e-7 s) go to the FP64 integrator
  double dy_max = 3e-3;
  double dT_max = 50.0;
  double margin = 0.01;    // input range widening, as a fraction of the training range
  // smallest predicted sigma_min(I - dt J) accepted: near-singular cells, where
  // the point-implicit step is close to a pole and not smooth, stay with the
  // integrator
  double sig_min = 0.08;
};

constexpr int kSurIn = 11, kSurOut = 8, kSurNull = 6, kSurLayers = 4;

// create for a 9-species mechanism (element tables give the conservation rows)
int surrogate_create(const Mechanism &mech, int device, SurrogateState **out, std::string &err);
void surrogate_destroy(SurrogateState *s);
bool surrogate_ready(const SurrogateState *s);
int64_t surrogate_param_count(const SurrogateState *s);
int surrogate_init_params(SurrogateState *s, uint64_t seed, hipStream_t st, std::string &err);
int surrogate_get_params(SurrogateState *s, float *host, hipStream_t st, std::string &err);
int surrogate_set_params(SurrogateState *s, const float *host, hipStream_t st, std::string &err);
// one forward+backward pass on standardised inputs x [B x 11] and targets t [B x 8]
// (host arrays); grad gets d loss / d params in the parameter layout
int surrogate_gradient(SurrogateState *s, int64_t B, const float *x, const float *t, float *grad, double *loss,
                       hipStream_t st, std::string &err);
// standardised forward pass of the training path (FP32 master weights): y [B x 8]
int surrogate_forward(SurrogateState *s, int64_t B, const float *x, float *y, hipStream_t st, std::string &err);
// generate targets with the FP64 integrator on the device states, fit the
// standardisation and train; loss_hist[e] = mean training loss of epoch e
int surrogate_train(SurrogateState *s, const DMech *m, int64_t n, const double *rho, const double *T,
                    const double *Y, double dt, const SurrogateTrainCfg &cfg, double *loss_hist, hipStream_t st,
                    std::string &err);
void surrogate_set_gate(SurrogateState *s, const SurrogateGate &g);
SurrogateGate surrogate_get_gate(const SurrogateState *s);
int surrogate_save(SurrogateState *s, const std::string &path, hipStream_t st, std::string &err);
int surrogate_load(SurrogateState *s, const std::string &path, hipStream_t st, std::string &err);
// standardisation: x mean/std/min/max (4 x 11), u mean/std (2 x 8), dt, basis (54), asinh scales (8)
int surrogate_get_norm(const SurrogateState *s, double *out123);
// inference kernel: 1 fused (128 cells per workgroup, default), 2 fused (64), 0 layer-by-layer GEMMs
void surrogate_set_kernel(SurrogateState *s, int kernel);
// raw (standardised) network outputs [n x 8] of device states, copied to the host
int surrogate_network_outputs(SurrogateState *s, const DMech *m, int64_t n, const double *rho, const double *T,
                              const double *Y, float *host_out, hipStream_t st, std::string &err);
// fallback count, network ms (features + MLP), gate ms and fallback ms of the last inference
int surrogate_last_stats(const SurrogateState *s, double *out4);
// fallback count and the cells each gate condition sent to the fallback in
// the last inference: 1 input range, 2 non-finite output, 3 |dY| > dy_max,
// 4 |dT| > dT_max, 5 sigma_min < sig_min, 6 conservation, 7 temperature
// inversion, 8 dt other than the training dt (out9[0] the total)
int surrogate_gate_reasons(SurrogateState *s, int64_t *out9, hipStream_t st, std::string &err);
int surrogate_infer(SurrogateState *s, const DMech *m, int64_t n, const double *rho, const double *T,
                    const double *Y, double dt, double *dY, double *dT, int64_t *n_fallback, hipStream_t stream,
                    std::string &err);

// Allocation-free inference: the same kernels as surrogate_infer with a fused
// network kernel, on a fixed per-device scratch array (network outputs,
// fallback list and counter) of at most kNoAllocCells cells.
// noalloc_scratch_init takes that array's device address once per device
// (the device must be current); calls that overlap on one GPU share it and race.
constexpr int64_t kNoAllocCells = 1 << 20;
int noalloc_scratch_init(int device, std::string &err);
int surrogate_infer_noalloc(SurrogateState *s, const DMech *m, int64_t n, const double *rho, const double *T,
                            const double *Y, double dt, double *dY, double *dT, int64_t *n_fallback,
                            hipStream_t stream, std::string &err);

}  // namespace ostc
