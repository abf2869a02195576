// Fused flux sweeps with the k-omega SST scalars (device code in
// flux_device.h).  This translation unit is built at -O1: at -O2 and -O3
// the FP64 k-direction sweep of the 9-species + SST case
// (sweep_jk<2, 9, 2, 0, double>) has given results that differ from run to
// run in the k row and from the CPU oracle, while the same source built at
// -O1 is bit-reproducible and matches the oracle to 1e-15 (4-8 % slower).
#include "flux_device.h"

namespace ostc {

hipError_t launch_flux_sweep_sst(const ZoneView &z, const DParams &P, const DMech *m, int dir, bool init, bool fp32,
                                 hipStream_t s) {
  const bool aweno = P.inviscid_type == 1;
  if (P.nt != 2) return hipErrorNotSupported;
  return with_species(P.ns, [&](auto c) -> hipError_t {
    constexpr int NS = decltype(c)::value;
    if (aweno)
      return fp32 ? launch_sweep_t<NS, 2, 1, float>(z, P, m, dir, init, s)
                  : launch_sweep_t<NS, 2, 1, double>(z, P, m, dir, init, s);
    return fp32 ? launch_sweep_t<NS, 2, 0, float>(z, P, m, dir, init, s)
                : launch_sweep_t<NS, 2, 0, double>(z, P, m, dir, init, s);
  });
}


}  // namespace ostc
