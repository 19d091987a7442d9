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
  int riemann, sd, irs, limiter, ndir;
  int coupled;              // 0 uncoupled, 1 coupled, 2 coupled with an evaporating family
  const int* irs_wg;        // first workgroup of each block in the all-blocks cell launches (IRS along i,
  int irs_nwg;              // the update, the dt pass), and the workgroups of such a launch
  const int* irs_wl[2];     // the same for the lines along j and k of the marching smoothing
  int irs_nwl[2];
  int timing;                        // bracket each stage launch with events
  std::vector<hipEvent_t>* kev;
  std::vector<hipEvent_t>* lkev;     // and each deferred interface-solve launch
  int lk_grid;                       // workgroups of k_lk_finish (its grid-stride loop)
};

// The queue of deferred interface solves: cleared before a stage's kernels, then
// drained by k_lk_finish after them
inline void clear_lk_queue(const FamArgs& a, hipStream_t stream) {
  if (a.lkq_g) hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(64), 0, stream, a.lkq_n);
}
template <int M, bool WRITE_R> void launch_lk_finish(const FamArgs& a, int grid, hipStream_t stream) {
  if (a.lkq_g) hipLaunchKernelGGL((k_lk_finish<M, WRITE_R>), dim3((unsigned)grid), dim3(kThreads), 0, stream, a);
}

inline unsigned grid_cells(long long n) { return (unsigned)((n + kThreads - 1) / kThreads); }

// Implicit residual smoothing of a family over every block: one launch per direction,
// both Jacobi sweeps fused, from R into T, then T into R, and so on: along i one
// thread per cell, along j and k one per line and slot, marching; with a direction
// j the march along j also forms the smoothing along i (k_irs_march FOLD_I), so
// direction i has no launch of its own.  Returns the buffer that holds the
// smoothed residual.
template <int NC>
const double* launch_irs(const FamArgs& a, int ndir, const int* wg, int nwg, const int* const* wl, const int* nwl,
                         int nb, hipStream_t stream) {
  const double* src = a.R;
  double* dst = a.T;
  for (int d = 0; d < ndir; ++d) {
    if (d == 0 && ndir > 1) continue;  // folded into the march along j
    if (d == 0)
      hipLaunchKernelGGL(k_irs<NC>, dim3((unsigned)nwg), dim3(kThreads), 0, stream, a, d, src, dst, wg, nb);
    else if (d == 1)
      hipLaunchKernelGGL((k_irs_march<NC, true>), dim3((unsigned)nwl[0], NC), dim3(kThreads), 0, stream, a, d, src,
                         dst, wl[0], nb);
    else
      hipLaunchKernelGGL((k_irs_march<NC, false>), dim3((unsigned)nwl[d - 1], NC), dim3(kThreads), 0, stream, a, d,
                         src, dst, wl[d - 1], nb);
    double* t = const_cast<double*>(src);
    src = dst;
    dst = t;
  }
  return src;
}

// The smoothing followed by the RK update (k_update): with two or three
// directions the last direction's march does the update itself
// (k_irs_march_update), and the smoothing along i is formed inside the march along
// j.  R holds the residual on entry.
template <int M>
void launch_irs_update(const FamArgs& a, int ndir, const int* wg, int nwg, const int* const* wl, const int* nwl, int nb,
                       hipStream_t stream) {
  constexpr int NC = Closure<M>::NC;
  if (ndir < 2) {
    const double* res = launch_irs<NC>(a, ndir, wg, nwg, wl, nwl, nb, stream);
    hipLaunchKernelGGL(k_update<M>, dim3((unsigned)nwg), dim3(kThreads), 0, stream, a, res, wg, nb);
  } else if (ndir == 2) {
    hipLaunchKernelGGL((k_irs_march_update<M, true>), dim3((unsigned)nwl[0]), dim3(kThreads), 0, stream, a, 1, a.R,
                       wl[0], nb);
  } else {
    const double* res = launch_irs<NC>(a, 2, wg, nwg, wl, nwl, nb, stream);
    hipLaunchKernelGGL((k_irs_march_update<M, false>), dim3((unsigned)nwl[1]), dim3(kThreads), 0, stream, a, 2, res,
                       wl[1], nb);
  }
}

template <int M, int S, bool SD, int CP>
void launch_tiled(const StageLaunch& L, FamArgs& a, int b) {
  const Tile T = L.tiles[b];
  dim3 grid((unsigned)(T.ntx * T.nty * T.ntz)), blk(kThreads);
  if (L.irs) {
    if (L.limiter == 3) hipLaunchKernelGGL((k_stage_tiled<M, S, SD, CP, true, 3>), grid, blk, 0, L.stream, a, b, T);
    else hipLaunchKernelGGL((k_stage_tiled<M, S, SD, CP, true, -1>), grid, blk, 0, L.stream, a, b, T);
  } else if (L.limiter == 1) {
    hipLaunchKernelGGL((k_stage_tiled<M, S, SD, CP, false, 1>), grid, blk, 0, L.stream, a, b, T);
  } else if (L.limiter == 3) {
    hipLaunchKernelGGL((k_stage_tiled<M, S, SD, CP, false, 3>), grid, blk, 0, L.stream, a, b, T);
  } else {
    hipLaunchKernelGGL((k_stage_tiled<M, S, SD, CP, false, -1>), grid, blk, 0, L.stream, a, b, T);
  }
}

template <int M, int S, bool SD, int CP>
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
  if (CP == 2) clear_lk_queue(a, L.stream);
  for (int b = 0; b < nb; ++b) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool timed = L.timing;
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
  if (CP == 2 && a.lkq_g) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (L.timing) {
      (void)hipEventCreate(&e0);
      (void)hipEventCreate(&e1);
      (void)hipEventRecord(e0, L.stream);
    }
    if (L.irs) launch_lk_finish<M, true>(a, L.lk_grid, L.stream);
    else launch_lk_finish<M, false>(a, L.lk_grid, L.stream);
    if (L.timing) {
      (void)hipEventRecord(e1, L.stream);
      L.lkev->push_back(e0);
      L.lkev->push_back(e1);
    }
  }
  if (L.irs) launch_irs_update<M>(a, L.ndir, L.irs_wg, L.irs_nwg, L.irs_wl, L.irs_nwl, nb, L.stream);
}

// One closure, Riemann solver and sensor setting's stage kernels: instantiated in
// stage_unit.hip, one object per combination
template <int M, int S, bool SD> void launch_stage_unit(const StageLaunch& L, const FamArgs& a);

template <int M, int S> void launch_sd_cp(const StageLaunch& L, const FamArgs& a) {
  if (L.sd) launch_stage_unit<M, S, true>(L, a);
  else launch_stage_unit<M, S, false>(L, a);
}

// the two ghost layers of a family's current state (compute_ghost, fill_second_ghost)
template <int M> void launch_ghost_kernels(hipStream_t stream, const int* grp, int ngrp, FamArgs a) {
  a.list = grp;
  a.nlist = ngrp;
  if (!ngrp) return;
  hipLaunchKernelGGL(k_ghost1<M>, dim3(grid_cells(ngrp)), dim3(kThreads), 0, stream, a);
  hipLaunchKernelGGL(k_ghost2<M>, dim3(grid_cells(ngrp)), dim3(kThreads), 0, stream, a);
}

template <int M>
void launch_births_kernel(hipStream_t stream, const FamArgs& a, const long long* cell, const int* first,
                          const double* pk, int ncell) {
  if (ncell > 0)
    hipLaunchKernelGGL(k_add_births<M>, dim3(grid_cells(ncell)), dim3(kThreads), 0, stream, a, cell, first, pk, ncell);
}

// the dt pass of a family over every block, one launch (wg: the all-blocks cell table)
template <int M> void launch_dt_kernel(hipStream_t stream, const int* wg, int nwg, int nb, const FamArgs& a) {
  hipLaunchKernelGGL(k_dt_blocks<M>, dim3((unsigned)nwg), dim3(kThreads), 0, stream, a, wg, nb);
}

// ------------------------------------------------------------------ batched path
// One closure group (families of one closure and Riemann solver) of the batched
// step: its families' record kernels, then the persistent kernels over the
// group's work queues.
struct BatchLaunch {
  hipStream_t stream;
  const FamArgs* d_fams;       // device copies, indexed by WorkItem::fam (family number - 1)
  const FamArgs* h_fams;       // host copies of the same
  const int* fam_ids;          // the group's families
  int nfam;
  const StageLaunch* per_fam;  // record lists of each family of the group
  WorkQueues Q;
  int grid;                    // persistent workgroups
  int dispatch;                // 0: persistent workgroups over the queues, 1, 2: one workgroup per item
  int nitems;
  int irs, ndir, nb;           // implicit residual smoothing: the stage writes R, then the sweeps
  const int* irs_wg;           // first workgroup of each block in the all-blocks cell launches
  int irs_nwg;
  const int* irs_wl[2];        // the same for the lines along j and k
  int irs_nwl[2];
  const BlkDev* blk;           // host copies of the blocks
  int riemann, sd, limiter;
  int coupled;                 // 0 uncoupled, 1 coupled, 2 coupled with an evaporating family in the group
  int lk_grid;                 // workgroups of k_lk_finish
  // recorded before the stage kernel, after it and after the deferred interface
  // solves when set (kernel timing)
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
};

// the deferred interface solves of the group's families, after the stage kernel
template <int M, int CP> void finish_lk_group(const BatchLaunch& B, bool write_r) {
  if (CP != 2) return;
  for (int f = 0; f < B.nfam; ++f) {
    const FamArgs& a = B.h_fams[B.fam_ids[f]];
    if (write_r) launch_lk_finish<M, true>(a, B.lk_grid, B.stream);
    else launch_lk_finish<M, false>(a, B.lk_grid, B.stream);
  }
}

inline void reset_tickets(const BatchLaunch& B) {
  hipLaunchKernelGGL(k_zero_tickets, dim3(1), dim3(64), 0, B.stream, B.Q.ticket, B.Q.nq);
}

template <int M, int S, bool SD, int CP> void launch_batched_stage(const BatchLaunch& B) {
  for (int f = 0; f < B.nfam; ++f) {
    const StageLaunch& L = B.per_fam[f];
    FamArgs a = B.h_fams[B.fam_ids[f]];
    a.list = L.grp;
    a.nlist = L.ngrp;
    if (L.ngrp) {
      hipLaunchKernelGGL(k_ghost1<M>, dim3(grid_cells(L.ngrp)), dim3(kThreads), 0, B.stream, a);
      hipLaunchKernelGGL(k_ghost2<M>, dim3(grid_cells(L.ngrp)), dim3(kThreads), 0, B.stream, a);
    }
  }
  if (SD) {
    reset_tickets(B);
    hipLaunchKernelGGL(k_beta_persist<M>, dim3(B.grid), dim3(kThreads), 0, B.stream, B.d_fams, B.Q);
  }
  for (int f = 0; f < B.nfam; ++f) {
    const StageLaunch& L = B.per_fam[f];
    FamArgs a = B.h_fams[B.fam_ids[f]];
    if (L.nown) {
      a.list = L.own;
      a.nlist = L.nown;
      hipLaunchKernelGGL((k_conn<M, S, SD>), dim3(grid_cells(L.nown)), dim3(kThreads), 0, B.stream, a);
    }
    if (L.nbcell) hipLaunchKernelGGL((k_bcell<M, S, SD>), dim3(grid_cells(L.nbcell)), dim3(kThreads), 0, B.stream, a);
  }
  reset_tickets(B);
  if (CP == 2)
    for (int f = 0; f < B.nfam; ++f) clear_lk_queue(B.h_fams[B.fam_ids[f]], B.stream);
  if (B.ev0) (void)hipEventRecord(B.ev0, B.stream);
  dim3 blk(kThreads);
  if (B.irs) {
    // residual into R, smoothed per block and direction, then the update
    const dim3 items((unsigned)B.nitems), grid(B.grid);
    if (B.dispatch != 0 && B.limiter == 3)
      hipLaunchKernelGGL((k_stage_items<M, S, SD, CP, true, 3>), items, blk, 0, B.stream, B.d_fams, B.Q.items);
    else if (B.dispatch != 0)
      hipLaunchKernelGGL((k_stage_items<M, S, SD, CP, true, -1>), items, blk, 0, B.stream, B.d_fams, B.Q.items);
    else if (B.limiter == 3)
      hipLaunchKernelGGL((k_stage_persist<M, S, SD, CP, true, 3>), grid, blk, 0, B.stream, B.d_fams, B.Q);
    else
      hipLaunchKernelGGL((k_stage_persist<M, S, SD, CP, true, -1>), grid, blk, 0, B.stream, B.d_fams, B.Q);
    if (B.ev1) (void)hipEventRecord(B.ev1, B.stream);
    finish_lk_group<M, CP>(B, true);
    if (B.ev2) (void)hipEventRecord(B.ev2, B.stream);
    for (int f = 0; f < B.nfam; ++f)
      launch_irs_update<M>(B.h_fams[B.fam_ids[f]], B.ndir, B.irs_wg, B.irs_nwg, B.irs_wl, B.irs_nwl, B.nb, B.stream);
    return;
  }
  if (B.dispatch != 0) {
    dim3 grid((unsigned)B.nitems);
    if (B.limiter == 1) hipLaunchKernelGGL((k_stage_items<M, S, SD, CP, false, 1>), grid, blk, 0, B.stream, B.d_fams, B.Q.items);
    else if (B.limiter == 3) hipLaunchKernelGGL((k_stage_items<M, S, SD, CP, false, 3>), grid, blk, 0, B.stream, B.d_fams, B.Q.items);
    else hipLaunchKernelGGL((k_stage_items<M, S, SD, CP, false, -1>), grid, blk, 0, B.stream, B.d_fams, B.Q.items);
  } else {
    dim3 grid(B.grid);
    if (B.limiter == 1) hipLaunchKernelGGL((k_stage_persist<M, S, SD, CP, false, 1>), grid, blk, 0, B.stream, B.d_fams, B.Q);
    else if (B.limiter == 3) hipLaunchKernelGGL((k_stage_persist<M, S, SD, CP, false, 3>), grid, blk, 0, B.stream, B.d_fams, B.Q);
    else hipLaunchKernelGGL((k_stage_persist<M, S, SD, CP, false, -1>), grid, blk, 0, B.stream, B.d_fams, B.Q);
  }
  if (B.ev1) (void)hipEventRecord(B.ev1, B.stream);
  finish_lk_group<M, CP>(B, false);
  if (B.ev2) (void)hipEventRecord(B.ev2, B.stream);
}

template <int M, int S, bool SD> void launch_batched_unit(const BatchLaunch& B);
template <int M, int S, bool SD> int persistent_wgs_unit(int cus, int cp);

template <int M, int S> void launch_batched_sd_cp(const BatchLaunch& B) {
  if (B.sd) launch_batched_unit<M, S, true>(B);
  else launch_batched_unit<M, S, false>(B);
}

// the dt pass of the group's families, one launch each over every block
template <int M> void launch_batched_dt_kernel(const BatchLaunch& B) {
  for (int f = 0; f < B.nfam; ++f)
    launch_dt_kernel<M>(B.stream, B.irs_wg, B.irs_nwg, B.nb, B.h_fams[B.fam_ids[f]]);
}

// persistent workgroups of the stage kernel on this device
template <int M, int S, bool SD, int CP> int persistent_wgs_of(int cus) {
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_stage_persist<M, S, SD, CP, false, -1>,
                                                   kThreads, 0) != hipSuccess || per < 1)
    per = 1;
  return cus * per;
}

// defined in stage_mk.hip, stage_ig.hip, stage_ag.hip
void launch_batched_mk(const BatchLaunch& B);
void launch_batched_ig(const BatchLaunch& B);
void launch_batched_ag(const BatchLaunch& B);
void launch_batched_dt_mk(const BatchLaunch& B);
void launch_batched_dt_ig(const BatchLaunch& B);
void launch_batched_dt_ag(const BatchLaunch& B);
int persistent_wgs_mk(int cus, int cp);
int persistent_wgs_ig(int cus, int cp);
int persistent_wgs_ag(int cus, int cp);

// defined in stage_mk.hip, stage_ig.hip, stage_ag.hip
void launch_ghosts_mk(hipStream_t s, const int* grp, int ngrp, const FamArgs& a);
void launch_ghosts_ig(hipStream_t s, const int* grp, int ngrp, const FamArgs& a);
void launch_ghosts_ag(hipStream_t s, const int* grp, int ngrp, const FamArgs& a);
void launch_births_mk(hipStream_t s, const FamArgs& a, const long long* cell, const int* first, const double* pk, int n);
void launch_births_ig(hipStream_t s, const FamArgs& a, const long long* cell, const int* first, const double* pk, int n);
void launch_births_ag(hipStream_t s, const FamArgs& a, const long long* cell, const int* first, const double* pk, int n);
void launch_stage_mk(const StageLaunch& L, const FamArgs& a);
void launch_stage_ig(const StageLaunch& L, const FamArgs& a);
void launch_stage_ag(const StageLaunch& L, const FamArgs& a);
void launch_dt_mk(hipStream_t s, const int* wg, int nwg, int nb, const FamArgs& a);
void launch_dt_ig(hipStream_t s, const int* wg, int nwg, int nb, const FamArgs& a);
void launch_dt_ag(hipStream_t s, const int* wg, int nwg, int nb, const FamArgs& a);

}  // namespace ice
