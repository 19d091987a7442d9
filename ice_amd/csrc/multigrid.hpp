// Grid sequencing between two levels of a run (Mod_Multigrid.f90:93-131,
// Lib_Multigrid.f90:40-191): a coarse level's block has half the cells of the
// fine level's block in each direction (a flat k stays one cell).  One thread per
// coarse cell; each level's state is the current buffer of its own context.
//   k_restrict  fine -> coarse: the 8 (flat: 4) fine cells' conservative states
//               times their volumes, summed in the reference's order, over the
//               coarse volume, back to primitives (fine2coarse_prim);
//   k_prolong   coarse -> fine: each of the 8 (4) fine cells of a coarse cell
//               from the coarse cell and its neighbours, trilinear weights
//               27/64 9/64 3/64 1/64 (flat: 9/16 3/16 1/16), neighbours clamped
//               to the block (coarse2fine_prim).
#pragma once
#include <hip/hip_runtime.h>

#include "stage_kernels.hpp"

namespace ice {

// the coarse cell (1-based) of thread t of block C
ICE_DEV bool coarse_cell(const BlkDev& C, long long t, int& i, int& j, int& k) {
  if (t >= C.cells) return false;
  i = (int)(t % C.ni) + 1;
  long long r = t / C.ni;
  j = (int)(r % C.nj) + 1;
  k = (int)(r / C.nj) + 1;
  return true;
}

template <int M>
__global__ __launch_bounds__(kThreads) void k_restrict(FamArgs fa, FamArgs ca, int b) {
  constexpr int NC = Closure<M>::NC;
  const BlkDev F = fa.blk[b], Cb = ca.blk[b];
  int i, j, k;
  if (!coarse_cell(Cb, (long long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const bool flat = F.nk == 1;
  const int i2 = 2 * i, j2 = 2 * j, k2 = flat ? 1 : 2 * k;
  const int i2d = i2 - 1, j2d = j2 - 1, k2d = flat ? 1 : k2 - 1;
  const int fi[8] = {i2d, i2, i2d, i2, i2d, i2, i2d, i2};
  const int fj[8] = {j2d, j2d, j2, j2, j2d, j2d, j2, j2};
  const int fk[8] = {k2d, k2d, k2d, k2d, k2, k2, k2, k2};
  const int nf = flat ? 4 : 8;
  double acc[NC], q[NC], c[NC];
  for (int m = 0; m < nf; ++m) {
    long long g = gidx(F, fi[m], fj[m], fk[m]);
    load_cell<NC>(fa.cur, fa.N, g, q);
    prim2cons<M>(q, c, fa.mat);
    double v = fa.vol[g];
#pragma unroll
    for (int s = 0; s < NC; ++s) acc[s] = m == 0 ? c[s] * v : acc[s] + c[s] * v;
  }
  const long long gc = gidx(Cb, i, j, k);
  const double cv = ca.vol[gc];
#pragma unroll
  for (int s = 0; s < NC; ++s) acc[s] = acc[s] / cv;
  cons2prim<M>(acc, q, ca.mat);
  store_cell<NC>(ca.prim_w, ca.N, gc, q);
}

template <int NC>
__global__ __launch_bounds__(kThreads) void k_prolong(FamArgs ca, FamArgs fa, int b) {
  const BlkDev F = fa.blk[b], Cb = ca.blk[b];
  int i, j, k;
  if (!coarse_cell(Cb, (long long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const bool flat = F.nk == 1;
  double w[8];
  if (flat) {
    const double a1 = 9.0 / 16.0, a2 = 3.0 / 16.0, a3 = 1.0 / 16.0, a4 = 0.0;
    w[0] = a1; w[1] = a2; w[2] = a2; w[3] = a4; w[4] = a3; w[5] = a4; w[6] = a4; w[7] = a4;
  } else {
    const double a1 = 27.0 / 64.0, a2 = 9.0 / 64.0, a3 = 3.0 / 64.0, a4 = 1.0 / 64.0;
    w[0] = a1; w[1] = a2; w[2] = a2; w[3] = a2; w[4] = a3; w[5] = a3; w[6] = a3; w[7] = a4;
  }
  const int im = i - 1 > 1 ? i - 1 : 1, ip = i + 1 < Cb.ni ? i + 1 : Cb.ni;
  const int jm = j - 1 > 1 ? j - 1 : 1, jp = j + 1 < Cb.nj ? j + 1 : Cb.nj;
  const int km = k - 1 > 1 ? k - 1 : 1, kp = k + 1 < Cb.nk ? k + 1 : Cb.nk;
  const int i2 = 2 * i, j2 = 2 * j, k2 = flat ? 1 : 2 * k;
  const int i2d = i2 - 1, j2d = j2 - 1, k2d = flat ? 1 : k2 - 1;
  const double* P = ca.cur;
  int counter = 0;
  for (int kk = k2d; kk <= k2; ++kk)
    for (int jj = j2d; jj <= j2; ++jj)
      for (int ii = i2d; ii <= i2; ++ii, ++counter) {
        // the neighbour in each direction: the lower one for the first fine cell of
        // the pair, the upper one for the second (the masks of :144-162)
        const int ni = (counter & 1) ? ip : im;
        const int nj = (counter & 2) ? jp : jm;
        const int nk = (counter & 4) ? kp : km;
        const long long g0 = gidx(Cb, i, j, k), g1 = gidx(Cb, ni, j, k), g2 = gidx(Cb, i, nj, k);
        const long long g3 = gidx(Cb, i, j, nk), g4 = gidx(Cb, ni, nj, k), g5 = gidx(Cb, ni, j, nk);
        const long long g6 = gidx(Cb, i, nj, nk), g7 = gidx(Cb, ni, nj, nk);
        const long long gf = gidx(F, ii, jj, kk);
        for (int s = 0; s < NC; ++s) {
          const double* p = P + (long long)s * ca.N;
          fa.prim_w[(long long)s * fa.N + gf] = w[0] * p[g0] + w[1] * p[g1] + w[2] * p[g2] + w[3] * p[g3] +
                                                w[4] * p[g4] + w[5] * p[g5] + w[6] * p[g6] + w[7] * p[g7];
        }
      }
}

}  // namespace ice
