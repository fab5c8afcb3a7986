// probe.hip -- random-gather ceiling of the device (measurement library
// libgrudon_probe.so, loaded by bench.py only; not part of libgrudon_hip.so).
//
// The pull kernels spend their time on 4-byte gathers at random positions of
// a per-partition value array.  This measures the rate at which the device
// serves such gathers with the pull's own access structure (8 coalesced index
// loads, then 8 independent gathers per thread, 256-thread workgroups of 2048
// gathers), for indices drawn uniformly from a table of a given size.  The
// result is the ceiling against which a pull kernel's gather rate is read.
//
// It also holds a streaming-read kernel of known size, the yardstick by which
// the FETCH_SIZE counter of a pull dispatch is scaled to bytes.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace {

constexpr int kBlock = 256;
thread_local std::string g_err;

int fail(const std::string &m)
{
  g_err = m;
  return 1;
}

}  // namespace

namespace grdm {

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

// Streaming read of n floats, 4 per thread per step (16-byte loads), summed
// into one value per thread so that the loads cannot be dropped.
__global__ void __launch_bounds__(kBlock) k_probe_stream(const float4 *__restrict__ in, uint64_t n4,
                                                         float *__restrict__ out)
{
  float a = 0.0f;
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * kBlock) {
    float4 v = in[i];
    a += v.x + v.y + v.z + v.w;
  }
  out[blockIdx.x * kBlock + threadIdx.x] = a;
}

// Streaming write of n4 float4 values (16-byte stores), the value depending
// on the index so that the stores cannot be merged away.
__global__ void __launch_bounds__(kBlock) k_probe_write(float4 *__restrict__ out, uint64_t n4)
{
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * kBlock) {
    const float f = (float)(i & 1023);
    out[i] = make_float4(f, f + 1.0f, f + 2.0f, f + 3.0f);
  }
}

}  // namespace grdm

using namespace grdm;

#define PROBE_HIP(expr)                                                         \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      rc = fail(std::string(#expr) + ": " + hipGetErrorString(_e));             \
      goto done;                                                                \
    }                                                                           \
  } while (0)

extern "C" {

const char *grdm_last_error(void) { return g_err.c_str(); }

// Times n_gathers 4-byte gathers at uniform random positions of a table of
// table_entries floats on device `device`; returns 0 on success.  Allocates
// and frees its own buffers and stream.
int grdm_probe_gather(int device, uint64_t table_entries, uint64_t n_gathers, double *ms_out, double *gathers_per_s)
{
  if (!ms_out || !gathers_per_s) return fail("grdm_probe_gather: NULL argument");
  if (table_entries == 0 || table_entries > (1ull << 32) || n_gathers == 0 || n_gathers > (1ull << 31))
    return fail("grdm_probe_gather: table must hold 1..2^32 entries and gathers be 1..2^31");
  int rc = 0;
  const int reps = 5;
  const uint64_t nblk = (n_gathers + 2047) / 2048;
  uint32_t *idx = nullptr;
  float *tab = nullptr, *out = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  float ms = 0.0f;
  PROBE_HIP(hipSetDevice(device));
  PROBE_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  PROBE_HIP(hipMalloc(&idx, n_gathers * 4));
  PROBE_HIP(hipMalloc(&tab, table_entries * 4));
  PROBE_HIP(hipMalloc(&out, nblk * kBlock * 4));
  PROBE_HIP(hipEventCreate(&a));
  PROBE_HIP(hipEventCreate(&b));
  hipLaunchKernelGGL(k_probe_fill, dim3(4096), dim3(kBlock), 0, st, idx, n_gathers, table_entries, 12345ull);
  hipLaunchKernelGGL(k_probe_fill, dim3(4096), dim3(kBlock), 0, st, (uint32_t *)tab, table_entries, 1ull << 20,
                     777ull);
  for (int r = 0; r < 2; r++)
    hipLaunchKernelGGL(k_probe_gather, dim3((uint32_t)nblk), dim3(kBlock), 0, st, idx, tab, out, n_gathers);
  PROBE_HIP(hipGetLastError());
  PROBE_HIP(hipEventRecord(a, st));
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(k_probe_gather, dim3((uint32_t)nblk), dim3(kBlock), 0, st, idx, tab, out, n_gathers);
  PROBE_HIP(hipEventRecord(b, st));
  PROBE_HIP(hipEventSynchronize(b));
  PROBE_HIP(hipEventElapsedTime(&ms, a, b));
  *ms_out = ms / reps;
  *gathers_per_s = (double)n_gathers / (*ms_out * 1e-3);
done:
  if (st) (void)hipStreamSynchronize(st);
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (idx) (void)hipFree(idx);
  if (tab) (void)hipFree(tab);
  if (out) (void)hipFree(out);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}

// Reads (write == 0, k_probe_stream) or writes (write != 0, k_probe_write)
// `bytes` (a multiple of 16) once per launch, `launches` times; the
// FETCH_SIZE / WRITE_SIZE of one dispatch against `bytes` gives the
// counter's scale.  The buffer is larger than the 256 MB last-level cache
// when bytes is.  *ms_out = mean time per launch.
int grdm_probe_stream(int device, uint64_t bytes, int launches, int write, double *ms_out)
{
  if (!ms_out || bytes == 0 || bytes % 16 || launches < 1) return fail("grdm_probe_stream: bad argument");
  int rc = 0;
  const uint32_t grid = 4096;
  float *in = nullptr, *out = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t a = nullptr, b = nullptr;
  float ms = 0.0f;
  PROBE_HIP(hipSetDevice(device));
  PROBE_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  PROBE_HIP(hipMalloc(&in, bytes));
  PROBE_HIP(hipMalloc(&out, (uint64_t)grid * kBlock * 4));
  PROBE_HIP(hipMemsetAsync(in, 0, bytes, st));
  PROBE_HIP(hipEventCreate(&a));
  PROBE_HIP(hipEventCreate(&b));
  PROBE_HIP(hipEventRecord(a, st));
  for (int r = 0; r < launches; r++) {
    if (write) hipLaunchKernelGGL(k_probe_write, dim3(grid), dim3(kBlock), 0, st, (float4 *)in, bytes / 16);
    else hipLaunchKernelGGL(k_probe_stream, dim3(grid), dim3(kBlock), 0, st, (const float4 *)in, bytes / 16, out);
  }
  PROBE_HIP(hipGetLastError());
  PROBE_HIP(hipEventRecord(b, st));
  PROBE_HIP(hipEventSynchronize(b));
  PROBE_HIP(hipEventElapsedTime(&ms, a, b));
  *ms_out = ms / launches;
done:
  if (st) (void)hipStreamSynchronize(st);
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (in) (void)hipFree(in);
  if (out) (void)hipFree(out);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}

}  // extern "C"
