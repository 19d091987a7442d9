// Particles entering through extrapolation faces (400): count_400_inflow
// (src/lib/diagnostic/Mod_Inflow400.f90:52-128) as two small kernels per stage
// of a family, both capturable in a graph (no memset, no host read):
//   k_i400_count  one workgroup per chunk of one boundary's 400 records: the
//                 records' measures reduced in the workgroup, then folded into the
//                 boundary's accumulators with four atomics (counts, and maxima of
//                 positive doubles as the maxima of their bits);
//   k_i400_fold   one workgroup: every boundary in (block, face) order folds its
//                 accumulators into its totals, logs a line where the reference
//                 prints one (slots by a workgroup prefix sum, so the log keeps
//                 the reference's order), and clears its accumulators.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/ice_gpu.h"
#include "stage_kernels.hpp"

namespace ice {

// One stage's accumulators of a boundary; zero between stages
struct I400Acc {
  unsigned int nface, nabove;
  unsigned long long xs, xa;  // bits of the largest measure, and of the largest above 10 x line_x
};

// A boundary's counts over the run (nstages ... line_x of Mod_Inflow400.f90:42-43)
struct I400Tot {
  int nstages, first_iter, first_faces, pad;
  double first_x, largest_x, line_x;
};

// The velocity of the closures' density floor, and the round-off bar on the measure
// (Mod_Inflow400.f90:28,31)
constexpr double kI400Floor = 1e-25;
constexpr double kI400RoundOff = 1e-12;

// The measure of a 400 record of family a in its stage input, or 0 when it does
// not count (Mod_Inflow400.f90:69-98)
ICE_DEV double i400_measure(const FamArgs& a, int closure, const DRec& r) {
  const BlkDev B = a.blk[r.b - 1];
  const long long N = a.N, g = gidx(B, r.i, r.j, r.k);
  const double* q = a.cur;
  double rho = q[g];
  if (!(rho > kRhoEmpty)) return 0.0;
  double nrm[3];
  bnd_normal(a, B, r, nrm);
  const double sg = (double)(1 - 2 * (r.f % 2));  // the normal out of the domain
  nrm[0] = sg * nrm[0];
  nrm[1] = sg * nrm[1];
  nrm[2] = sg * nrm[2];
  const double u = q[N + g], v = q[2 * N + g], w3 = q[3 * N + g];
  double w = u * nrm[0] + v * nrm[1] + w3 * nrm[2];
  if (!(w < -kI400Floor)) return 0.0;
  double x;
  if (closure == ICE_IG) {
    x = -w * sqrt(rho / q[4 * N + g]);
    if (!(x > kI400RoundOff)) return 0.0;
  } else if (closure == ICE_AG) {
    double p11 = q[4 * N + g], p12 = q[5 * N + g], p13 = q[6 * N + g];
    double p22 = q[7 * N + g], p23 = q[8 * N + g], p33 = q[9 * N + g];
    double pnn = p11 * (nrm[0] * nrm[0]) + p22 * (nrm[1] * nrm[1]) + p33 * (nrm[2] * nrm[2]) +
                 2.0 * (p12 * nrm[0] * nrm[1] + p13 * nrm[0] * nrm[2] + p23 * nrm[1] * nrm[2]);
    x = -w * sqrt(rho / dmax2(pnn, 2.2250738585072014e-308));
    if (!(x > kI400RoundOff)) return 0.0;
  } else {
    x = -w;
    if (!(x > kI400RoundOff * norm3(u, v, w3))) return 0.0;
  }
  return x;
}

// A chunk of one boundary's 400 records (the host sorts a family's records by
// boundary and cuts each boundary's run into chunks of at most kI400Chunk)
struct I400Chunk {
  int m, start, end, pad;
};
constexpr int kI400Chunk = 256;

// one workgroup of 256 per chunk: its records' counts and maxima reduced in the
// workgroup, then four atomics into the boundary's accumulators
__global__ __launch_bounds__(256) void k_i400_count(FamArgs a, int closure, const int* __restrict__ rec400,
                                                    const I400Chunk* __restrict__ chunks, I400Acc* __restrict__ acc,
                                                    const I400Tot* __restrict__ tot) {
  __shared__ unsigned long long red[4][4];
  const I400Chunk C = chunks[blockIdx.x];
  const double line = 10.0 * tot[C.m].line_x;
  unsigned long long nface = 0, nabove = 0, xs = 0, xa = 0;
  for (int t = C.start + (int)threadIdx.x; t < C.end; t += blockDim.x) {
    const DRec r = a.rec[rec400[t]];
    double x = i400_measure(a, closure, r);
    if (!(x > 0.0)) continue;
    const unsigned long long xb = (unsigned long long)__double_as_longlong(x);
    nface += 1;
    xs = xb > xs ? xb : xs;
    if (x > line) {
      nabove += 1;
      xa = xb > xa ? xb : xa;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    nface += __shfl_xor(nface, off);
    nabove += __shfl_xor(nabove, off);
    unsigned long long o = __shfl_xor(xs, off);
    xs = o > xs ? o : xs;
    o = __shfl_xor(xa, off);
    xa = o > xa ? o : xa;
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = nface;
    red[1][w] = nabove;
    red[2][w] = xs;
    red[3][w] = xa;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) {
      nface += red[0][k];
      nabove += red[1][k];
      xs = red[2][k] > xs ? red[2][k] : xs;
      xa = red[3][k] > xa ? red[3][k] : xa;
    }
    if (nface) {
      atomicAdd(&acc[C.m].nface, (unsigned)nface);
      atomicMax(&acc[C.m].xs, xs);
    }
    if (nabove) {
      atomicAdd(&acc[C.m].nabove, (unsigned)nabove);
      atomicMax(&acc[C.m].xa, xa);
    }
  }
}

// one workgroup of kThreads over the nbnd = 6 nb boundaries of family p (1-based)
__global__ __launch_bounds__(256) void k_i400_fold(int nbnd, int p, I400Acc* __restrict__ acc, I400Tot* __restrict__ tot,
                                                   ice_inflow400_line_t* __restrict__ lines, int* __restrict__ nlines,
                                                   const DevStatus* __restrict__ st) {
  __shared__ int scan[256];
  const int tid = threadIdx.x;
  const int iter = st->iter;
  int logged = *nlines;
  for (int base = 0; base < nbnd; base += 256) {
    const int m = base + tid;
    int flag = 0;
    ice_inflow400_line_t L;
    if (m < nbnd) {
      I400Acc A = acc[m];
      if (A.nface > 0) {
        I400Tot T = tot[m];
        double xs = __longlong_as_double((long long)A.xs);
        if (T.nstages == 0) {
          T.first_iter = iter;
          T.first_faces = (int)A.nface;
          T.first_x = xs;
        }
        T.nstages += 1;
        T.largest_x = dmax2(T.largest_x, xs);
        if (A.nabove > 0) {
          double xa = __longlong_as_double((long long)A.xa);
          flag = 1;
          L.iter = iter;
          L.block = m / 6 + 1;
          L.face = m % 6 + 1;
          L.family = p;
          L.faces = (int)A.nabove;
          L.relative = T.line_x > 0.0 ? 1 : 0;
          L.largest = xa;
          T.line_x = xa;
        }
        tot[m] = T;
        I400Acc z = {0u, 0u, 0ull, 0ull};
        acc[m] = z;
      }
    }
    // exclusive prefix sum of the flags over the workgroup (Hillis-Steele)
    scan[tid] = flag;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      int v = tid >= off ? scan[tid - off] : 0;
      __syncthreads();
      scan[tid] += v;
      __syncthreads();
    }
    const int total = scan[255];
    const int slot = logged + scan[tid] - flag;
    if (flag && slot < ICE_INFLOW400_LOG) lines[slot] = L;
    logged += total;
    __syncthreads();
  }
  if (tid == 0) *nlines = logged;
}

}  // namespace ice
