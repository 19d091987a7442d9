// Stage kernels of the anisotropic Gaussian (AG) closure: Rusanov or HLLE.
#include "launch.hpp"

namespace ice {
void launch_stage_ag(const StageLaunch& L, const FamArgs& a) {
  if (L.riemann == 2) launch_sd_cp<2, 2>(L, a);
  else launch_sd_cp<2, 1>(L, a);
}
void launch_dt_ag(hipStream_t s, const int* wg, int nwg, int nb, const FamArgs& a) {
  launch_dt_kernel<2>(s, wg, nwg, nb, a);
}
void launch_ghosts_ag(hipStream_t s, const int* grp, int ngrp, const FamArgs& a) {
  launch_ghost_kernels<2>(s, grp, ngrp, a);
}
void launch_births_ag(hipStream_t s, const FamArgs& a, const long long* cell, const int* first, const double* pk,
                      int n) {
  launch_births_kernel<2>(s, a, cell, first, pk, n);
}
void launch_batched_ag(const BatchLaunch& B) {
  if (B.riemann == 2) launch_batched_sd_cp<2, 2>(B);
  else launch_batched_sd_cp<2, 1>(B);
}
void launch_batched_dt_ag(const BatchLaunch& B) { launch_batched_dt_kernel<2>(B); }
int persistent_wgs_ag(int cus, int cp) { return persistent_wgs_unit<2, 1, true>(cus, cp); }
}  // namespace ice
