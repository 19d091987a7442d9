// Stage kernels of the monokinetic (MK) closure: Saurel, Rusanov or HLLE.
#include "launch.hpp"

namespace ice {
void launch_stage_mk(const StageLaunch& L, const FamArgs& a) {
  if (L.riemann == 0) launch_sd_cp<0, 0>(L, a);
  else if (L.riemann == 1) launch_sd_cp<0, 1>(L, a);
  else launch_sd_cp<0, 2>(L, a);
}
void launch_dt_mk(hipStream_t s, const int* wg, int nwg, int nb, const FamArgs& a) {
  launch_dt_kernel<0>(s, wg, nwg, nb, a);
}
void launch_ghosts_mk(hipStream_t s, const int* grp, int ngrp, const FamArgs& a) {
  launch_ghost_kernels<0>(s, grp, ngrp, a);
}
void launch_births_mk(hipStream_t s, const FamArgs& a, const long long* cell, const int* first, const double* pk,
                      int n) {
  launch_births_kernel<0>(s, a, cell, first, pk, n);
}
void launch_batched_mk(const BatchLaunch& B) {
  if (B.riemann == 0) launch_batched_sd_cp<0, 0>(B);
  else if (B.riemann == 1) launch_batched_sd_cp<0, 1>(B);
  else launch_batched_sd_cp<0, 2>(B);
}
void launch_batched_dt_mk(const BatchLaunch& B) { launch_batched_dt_kernel<0>(B); }
int persistent_wgs_mk(int cus, int cp) { return persistent_wgs_unit<0, 1, true>(cus, cp); }
}  // namespace ice
