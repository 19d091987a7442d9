// Launch of one family's stage kernels, one translation unit per closure so the
// kernel instantiations compile in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "stage_kernels.hpp"

namespace ice {

struct StageLaunch {
  hipStream_t stream;
  const BlkDev* blk;        // host copies of the blocks
  const Tile* tiles;        // stage tile of each block (tx = 0: untiled kernel)
  int nb;
  const int* grp;           // family's ghost records
  int ngrp;
  const int* own;           // owner records of its connection faces
  int nown;
  int nbcell;
  int riemann, sd, coupled, irs, limiter, ndir;
  int timing;                        // bracket each stage launch with events
  std::vector<hipEvent_t>* kev;
};

inline unsigned grid_cells(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

template <int M, int S, bool SD, bool CP>
void launch_tiled(const StageLaunch& L, FamArgs& a, int b) {
  const Tile T = L.tiles[b];
  dim3 grid((unsigned)(T.ntx * T.nty * T.ntz)), blk(kThreads);
  if (L.irs) {
    hipLaunchKernelGGL((k_stage_tiled<M, S, SD, CP, true, -1>), grid, blk, 0, L.stream, a, b, T);
  } else if (L.limiter == 1) {
    hipLaunchKernelGGL((k_stage_tiled<M, S, SD, CP, false, 1>), grid, blk, 0, L.stream, a, b, T);
  } else if (L.limiter == 3) {
    hipLaunchKernelGGL((k_stage_tiled<M, S, SD, CP, false, 3>), grid, blk, 0, L.stream, a, b, T);
  } else {
    hipLaunchKernelGGL((k_stage_tiled<M, S, SD, CP, false, -1>), grid, blk, 0, L.stream, a, b, T);
  }
}

template <int M, int S, bool SD, bool CP>
void launch_stage_kernels(const StageLaunch& L, FamArgs a) {
  constexpr int NC = Closure<M>::NC;
  const int nb = L.nb;
  a.list = L.grp;
  a.nlist = L.ngrp;
  if (L.ngrp) {
    hipLaunchKernelGGL(k_ghost1<M>, dim3(grid_cells(L.ngrp)), dim3(kThreads), 0, L.stream, a);
    hipLaunchKernelGGL(k_ghost2<M>, dim3(grid_cells(L.ngrp)), dim3(kThreads), 0, L.stream, a);
  }
  if (SD)
    for (int b = 0; b < nb; ++b)
      hipLaunchKernelGGL(k_beta<M>, dim3(grid_cells(L.blk[b].cells)), dim3(kThreads), 0, L.stream, a, b);
  if (L.nown) {
    a.list = L.own;
    a.nlist = L.nown;
    hipLaunchKernelGGL((k_conn<M, S, SD>), dim3(grid_cells(L.nown)), dim3(kThreads), 0, L.stream, a);
  }
  if (L.nbcell) hipLaunchKernelGGL((k_bcell<M, S, SD>), dim3(grid_cells(L.nbcell)), dim3(kThreads), 0, L.stream, a);
  for (int b = 0; b < nb; ++b) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool timed = L.timing && !L.irs;
    if (timed) {
      (void)hipEventCreate(&e0);
      (void)hipEventCreate(&e1);
      (void)hipEventRecord(e0, L.stream);
    }
    if (L.tiles[b].tx == 0) {
      if (L.irs)
        hipLaunchKernelGGL((k_stage<M, S, SD, CP, true>), dim3(grid_cells(L.blk[b].cells)), dim3(kThreads), 0, L.stream,
                           a, b);
      else
        hipLaunchKernelGGL((k_stage<M, S, SD, CP, false>), dim3(grid_cells(L.blk[b].cells)), dim3(kThreads), 0,
                           L.stream, a, b);
    } else {
      launch_tiled<M, S, SD, CP>(L, a, b);
    }
    if (timed) {
      (void)hipEventRecord(e1, L.stream);
      L.kev->push_back(e0);
      L.kev->push_back(e1);
    }
  }
  if (L.irs) {
    for (int d = 0; d < L.ndir; ++d)
      for (int sw = 1; sw <= 2; ++sw)
        for (int b = 0; b < nb; ++b)
          hipLaunchKernelGGL(k_irs<NC>, dim3(grid_cells(L.blk[b].cells)), dim3(kThreads), 0, L.stream, a, b, d, sw);
    for (int b = 0; b < nb; ++b)
      hipLaunchKernelGGL(k_update<M>, dim3(grid_cells(L.blk[b].cells)), dim3(kThreads), 0, L.stream, a, b);
  }
}

template <int M, int S> void launch_sd_cp(const StageLaunch& L, const FamArgs& a) {
  if (L.sd && L.coupled) launch_stage_kernels<M, S, true, true>(L, a);
  else if (L.sd) launch_stage_kernels<M, S, true, false>(L, a);
  else if (L.coupled) launch_stage_kernels<M, S, false, true>(L, a);
  else launch_stage_kernels<M, S, false, false>(L, a);
}

template <int M> void launch_dt_kernel(hipStream_t stream, const BlkDev* blk, int nb, const FamArgs& a) {
  for (int b = 0; b < nb; ++b)
    hipLaunchKernelGGL(k_dt<M>, dim3(grid_cells(blk[b].cells)), dim3(kThreads), 0, stream, a, b);
}

// defined in stage_mk.hip, stage_ig.hip, stage_ag.hip
void launch_stage_mk(const StageLaunch& L, const FamArgs& a);
void launch_stage_ig(const StageLaunch& L, const FamArgs& a);
void launch_stage_ag(const StageLaunch& L, const FamArgs& a);
void launch_dt_mk(hipStream_t s, const BlkDev* blk, int nb, const FamArgs& a);
void launch_dt_ig(hipStream_t s, const BlkDev* blk, int nb, const FamArgs& a);
void launch_dt_ag(hipStream_t s, const BlkDev* blk, int nb, const FamArgs& a);

}  // namespace ice
