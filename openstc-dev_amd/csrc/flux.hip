// Fused per-direction flux sweeps for gfx950 (device code in
// flux_device.h): the laminar instantiations (air and the 4- and 9-species
// mixtures) and the dispatch; the k-omega SST instantiations are in
// flux_sst.hip.
#include "flux_device.h"

namespace ostc {

hipError_t launch_flux_sweep(const ZoneView &z, const DParams &P, const DMech *m, int dir, bool init, bool fp32,
                             hipStream_t s, const StepScratch *dt_w, bool *dt_done) {
  const bool aweno = P.inviscid_type == 1;
  double *dtp = dt_w ? dt_w->partial : nullptr, *dto = dt_w ? dt_w->dt : nullptr;
  const int dtn = dt_w ? dt_w->max_blocks : 0;
  if (P.nt == 0)
    return with_species(P.ns, [&](auto c) -> hipError_t {
      constexpr int NS = decltype(c)::value;
      if (aweno)
        return fp32 ? launch_sweep_t<NS, 0, 1, float>(z, P, m, dir, init, s, dtp, dto, dtn, dt_done)
                    : launch_sweep_t<NS, 0, 1, double>(z, P, m, dir, init, s, dtp, dto, dtn, dt_done);
      return fp32 ? launch_sweep_t<NS, 0, 0, float>(z, P, m, dir, init, s)
                  : launch_sweep_t<NS, 0, 0, double>(z, P, m, dir, init, s);
    });
  // SST: k and omega are transported after the species (the AWENO path
  // projects them on their own characteristic waves); flux_sst.hip
  return launch_flux_sweep_sst(z, P, m, dir, init, fp32, s);
}

}  // namespace ostc
