// MLP chemistry surrogate (no counterpart in the reference): a 3x256 GELU
// network mapping (log Y, T, log p) of a cell to its point-implicit
// increments (dY, dT) over one dt, run with BF16 MFMA and FP32 accumulation,
// with a per-cell fallback to the FP64 integrator.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "device_common.h"

namespace ostc {

struct SurrogateState;

void surrogate_destroy(SurrogateState *s);
bool surrogate_ready(const SurrogateState *s);
int surrogate_infer(SurrogateState *s, const DMech *m, int64_t n, const double *rho, const double *T,
                    const double *Y, double dt, double *dY, double *dT, int64_t *n_fallback, hipStream_t stream,
                    std::string &err);

}  // namespace ostc
