// Metropolis sweep kernels for FP32 two-row links.
#include "metro_kernels.hpp"

namespace klft_internal {
int launch_passes_f32rows(klft_field_t f, void *links, int variant, int p0, int p1, const klft_metropolis_params *p,
                          uint64_t sweep, int nsweeps, uint32_t chain, unsigned long long *counter,
                          int counter_stride, int *visit) {
  return klft_metro::launch_passes_typed<float, true>(f, links, variant, p0, p1, p, sweep, nsweeps, chain, counter,
                                                     counter_stride, visit);
}
}  // namespace klft_internal
