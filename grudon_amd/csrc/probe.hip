// probe.hip -- random-gather ceiling of the device.
//
// The pull kernels spend their time on 4-byte gathers at random positions of
// a per-partition value array.  This measures the rate at which the device
// serves such gathers with the pull's own access structure (8 coalesced index
// loads, then 8 independent gathers per thread, 256-thread workgroups of 2048
// gathers), for indices drawn uniformly from a table of a given size.  The
// result is the ceiling against which a pull kernel's gather rate is read.
#include <algorithm>

#include "grd_device.hpp"
#include "grd_internal.hpp"

namespace grd {

__device__ __forceinline__ uint64_t probe_mix(uint64_t z)
{
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(kBlock) k_probe_fill(uint32_t *idx, uint64_t n, uint64_t table, uint64_t seed)
{
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
    idx[i] = (uint32_t)(probe_mix(seed + i * 0x9E3779B97F4A7C15ull) % table);
}

__global__ void __launch_bounds__(kBlock) k_probe_gather(const uint32_t *__restrict__ idx,
                                                         const float *__restrict__ tab, float *__restrict__ out,
                                                         uint64_t n)
{
  const uint64_t base = (uint64_t)blockIdx.x * 2048u;
  uint32_t s[8];
  float v[8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint64_t i = base + threadIdx.x + k * kBlock;
    s[k] = i < n ? idx[i] : 0u;
  }
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint64_t i = base + threadIdx.x + k * kBlock;
    v[k] = i < n ? tab[s[k]] : 0.0f;
  }
  float a = 0.0f;
#pragma unroll
  for (int k = 0; k < 8; k++) a += v[k];
  out[(uint64_t)blockIdx.x * kBlock + threadIdx.x] = a;
}

}  // namespace grd

using namespace grd;

extern "C" grd_status grd_probe_gather(grd_ctx *ctx, uint64_t table_entries, uint64_t n_gathers, double *ms_out,
                                       double *gathers_per_s)
{
  if (!ctx || !ms_out || !gathers_per_s) return fail(GRD_ERR_ARG, "grd_probe_gather: NULL argument");
  if (table_entries == 0 || table_entries > (1ull << 32) || n_gathers == 0 || n_gathers > (1ull << 31))
    return fail(GRD_ERR_ARG, "grd_probe_gather: table must hold 1..2^32 entries and gathers be 1..2^31");
  GRD_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  DevMem mem;
  uint32_t *idx = nullptr;
  float *tab = nullptr, *out = nullptr;
  const uint64_t nblk = (n_gathers + 2047) / 2048;
  grd_status rc = mem.alloc_n(&idx, n_gathers);
  if (rc == GRD_OK) rc = mem.alloc_n(&tab, table_entries);
  if (rc == GRD_OK) rc = mem.alloc_n(&out, nblk * kBlock);
  if (rc != GRD_OK) {
    mem.release();
    return rc;
  }
  hipLaunchKernelGGL(k_probe_fill, dim3(4096), dim3(kBlock), 0, st, idx, n_gathers, table_entries, 12345ull);
  hipLaunchKernelGGL(k_probe_fill, dim3(4096), dim3(kBlock), 0, st, (uint32_t *)tab, table_entries, 1ull << 20,
                     777ull);
  for (int r = 0; r < 2; r++)
    hipLaunchKernelGGL(k_probe_gather, dim3((uint32_t)nblk), dim3(kBlock), 0, st, idx, tab, out, n_gathers);
  hipEvent_t a = nullptr, b = nullptr;
  hipError_t e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  const int reps = 5;
  if (e == hipSuccess) e = hipEventRecord(a, st);
  for (int r = 0; r < reps && e == hipSuccess; r++)
    hipLaunchKernelGGL(k_probe_gather, dim3((uint32_t)nblk), dim3(kBlock), 0, st, idx, tab, out, n_gathers);
  if (e == hipSuccess) e = hipEventRecord(b, st);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  (void)hipStreamSynchronize(st);
  mem.release();
  if (e != hipSuccess) return fail(GRD_ERR_HIP, std::string("grd_probe_gather: ") + hipGetErrorString(e));
  *ms_out = ms / reps;
  *gathers_per_s = (double)n_gathers / (*ms_out * 1e-3);
  return GRD_OK;
}
