// Stage kernels of one closure (ICE_M), Riemann solver (ICE_S) and shock-sensor
// setting (ICE_SD, 0 or 1).  The Makefile compiles this file once for each
// combination, so that the kernel instantiations compile in parallel.
#include "launch.hpp"

namespace ice {
template <int M, int S, bool SD> void launch_stage_unit(const StageLaunch& L, const FamArgs& a) {
  if (L.coupled == 2) launch_stage_kernels<M, S, SD, 2>(L, a);
  else if (L.coupled) launch_stage_kernels<M, S, SD, 1>(L, a);
  else launch_stage_kernels<M, S, SD, 0>(L, a);
}
template <int M, int S, bool SD> void launch_batched_unit(const BatchLaunch& B) {
  if (B.coupled == 2) launch_batched_stage<M, S, SD, 2>(B);
  else if (B.coupled) launch_batched_stage<M, S, SD, 1>(B);
  else launch_batched_stage<M, S, SD, 0>(B);
}
template <int M, int S, bool SD> int persistent_wgs_unit(int cus, int cp) {
  return cp == 2 ? persistent_wgs_of<M, S, SD, 2>(cus)
                 : (cp ? persistent_wgs_of<M, S, SD, 1>(cus) : persistent_wgs_of<M, S, SD, 0>(cus));
}

template void launch_stage_unit<ICE_M, ICE_S, (ICE_SD != 0)>(const StageLaunch&, const FamArgs&);
template void launch_batched_unit<ICE_M, ICE_S, (ICE_SD != 0)>(const BatchLaunch&);
template int persistent_wgs_unit<ICE_M, ICE_S, (ICE_SD != 0)>(int, int);
}  // namespace ice
