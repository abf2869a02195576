// MLP chemistry surrogate.  Until a network is trained or loaded the
// surrogate reports itself as not ready and ostc_chem_batch(mode 1) fails.
#include "surrogate.h"

namespace ostc {

struct SurrogateState {
  bool ready = false;
};

void surrogate_destroy(SurrogateState *s) { delete s; }

bool surrogate_ready(const SurrogateState *s) { return s && s->ready; }

int surrogate_infer(SurrogateState *s, const DMech *, int64_t, const double *, const double *, const double *, double,
                    double *, double *, int64_t *, hipStream_t, std::string &err) {
  if (!surrogate_ready(s)) {
    err = "surrogate not ready";
    return 1;
  }
  return 0;
}

}  // namespace ostc
