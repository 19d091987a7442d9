// ice_gpu_residual_norm_noalloc: residual_norm (Mod_Diagnostic.f90:22-69) for
// callers that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  It runs the kernels of
// ice_gpu_residual_norm on a file-scope device array in place of the partial-sum
// buffers that call allocates, and sums the totals on the device in the same order.
#pragma once

namespace {

// The residual norm's totals from the partial sums of every (block, family), summed
// in the order of residual_norm's host loop (blocks, then families: total over both,
// each family's over the blocks), then their square roots: one thread, nb*ng*5 adds.
__global__ void k_resnorm_totals(const double* __restrict__ sums, int nb, int ng, double* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double* total = out;
  double* fam = out + 5;
  for (int m = 0; m < 5; ++m) total[m] = 0.0;
  for (int m = 0; m < 5 * ng; ++m) fam[m] = 0.0;
  for (int b = 0; b < nb; ++b)
    for (int p = 0; p < ng; ++p)
      for (int m = 0; m < 5; ++m) {
        double x = sums[((long long)b * ng + p) * 5 + m];
        total[m] = total[m] + x;
        fam[p * 5 + m] = fam[p * 5 + m] + x;
      }
  for (int m = 0; m < 5; ++m) total[m] = sqrt(total[m]);
  for (int m = 0; m < 5 * ng; ++m) fam[m] = sqrt(fam[m]);
}

// Device memory of ice_gpu_residual_norm_noalloc, a file-scope array: no call
// allocates it.  It holds the partial sums of every (block, family) pair
// (kResWG * 5 doubles each), their sums (5 each) and the totals (5 + 5 ngroups).
// Each GPU holds its own copy until the process ends; the host reaches it only
// through the address hipGetSymbolAddress gives on that device (the array's name
// is a host-side stand-in).  Every call on one GPU shares it: calls that overlap
// on one GPU race.
constexpr long long kScratchDoubles = 196608;  // 1.5 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxScratchDevices = 64;
double* g_scratch[kMaxScratchDevices] = {nullptr};

// the device's scratch address, looked up once per device with the device current
int scratch_of(ice_gpu_ctx* c, double** out) {
  if (c->device < 0 || c->device >= kMaxScratchDevices)
    return fail(c, "[ICE::scratch] device %d beyond the %d this library serves", c->device, kMaxScratchDevices);
  if (!g_scratch[c->device]) {
    double* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return fail(c, "[ICE::scratch] hipGetSymbolAddress: %s", hipGetErrorString(e));
    g_scratch[c->device] = ptr;
  }
  *out = g_scratch[c->device];
  return 0;
}

// at ice_gpu_create, with the context's device current
int noalloc_setup(ice_gpu_ctx* c) {
  double* ptr = nullptr;
  return scratch_of(c, &ptr);
}

}  // namespace

extern "C" {

int ice_gpu_residual_norm_noalloc(ice_gpu_ctx* c, double* total, double* fam) {
  if (!c) return 1;
  if (!total || !fam) return fail(c, "[ICE::residual_norm_noalloc] null output");
  const int nb = (int)c->blk.size(), ng = c->cfg.ngroups;
  const long long pairs = (long long)nb * ng;
  const long long need = pairs * kResWG * 5 + pairs * 5 + 5 + 5ll * ng;
  if (need > kScratchDoubles)
    return fail(c, "[ICE::residual_norm_noalloc] %d blocks x %d families need %lld doubles of scratch, more than the %lld "
                   "it holds (%lld block-family pairs at %d families): use ice_gpu_residual_norm",
                nb, ng, need, kScratchDoubles, (kScratchDoubles - 5 - 5ll * ng) / (5ll * kResWG + 5), ng);
  double* sp = nullptr;
  if (scratch_of(c, &sp)) return 1;
  double* part = sp;
  double* sums = sp + pairs * kResWG * 5;
  double* out = sums + pairs * 5;
  for (int p = 0; p < ng; ++p) {
    const FamDev& F = c->fam[p];
    hipLaunchKernelGGL(k_resnorm_part, dim3(kResWG, nb), dim3(256), 0, c->stream, c->d_blk, F.buf[F.cur],
                       F.buf[F.old], c->N, F.nc - 2, part, ng, p);
  }
  hipLaunchKernelGGL(k_resnorm_sum, dim3((unsigned)((pairs * 5 + 255) / 256)), dim3(256), 0, c->stream, part, sums,
                     (int)pairs);
  hipLaunchKernelGGL(k_resnorm_totals, dim3(1), dim3(64), 0, c->stream, sums, nb, ng, out);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(total, out, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(fam, out + 5, 5 * (size_t)ng * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, "[ICE::residual_norm_noalloc] %s", hipGetErrorString(e));
  return 0;
}

}  // extern "C"
