// gen.hip -- seeded graph generators on the device.
//
// Inputs only: the reference reads Galois .gr files produced offline
// (scripts/convert_and_partition.sh); the benchmark configurations name
// seeded RMAT and Erdos-Renyi graphs instead, generated here on the GPU and
// reproduced bit for bit by the CPU oracle's generator.
//
// Random numbers: hash(seed, index) = mix(mix(seed) ^ index) with the
// SplitMix64 finaliser `mix`.  RMAT edge e, level l (0 = most significant
// bit) uses 32 bits of hash(seed, e*16 + l/2) (high word for even l);
// quadrant thresholds are floor(p * 2^32).  ER pair k uses hash(seed, k):
// u = hi32 * n >> 32, v = lo32 * n >> 32, and both arcs are added.
// Duplicates and self-loops are removed after a radix sort (rocprim; setup
// only, not on the traversal path).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cmath>

#include "grd_device.hpp"
#include "grd_internal.hpp"

namespace grd {

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(kBlock) k_rmat(uint64_t E, int scale, uint64_t tA, uint64_t tAB, uint64_t tABC,
                                                 uint64_t hseed, int sym, uint64_t n, uint64_t *keys)
{
  for (uint64_t e = blockIdx.x * (uint64_t)kBlock + threadIdx.x; e < E; e += (uint64_t)gridDim.x * kBlock) {
    uint64_t s = 0, d = 0, h = 0;
    for (int l = 0; l < scale; l++) {
      if ((l & 1) == 0) h = mix64(hseed ^ ((e << 4) | (uint64_t)(l >> 1)));
      uint64_t r = (l & 1) ? (h & 0xFFFFFFFFull) : (h >> 32);
      uint64_t sb = r >= tAB ? 1 : 0;
      uint64_t db = ((r >= tA && r < tAB) || r >= tABC) ? 1 : 0;
      s = (s << 1) | sb;
      d = (d << 1) | db;
    }
    keys[e] = s * n + d;
    if (sym) keys[E + e] = d * n + s;
  }
}

__global__ void __launch_bounds__(kBlock) k_er(uint64_t M, uint64_t n, uint64_t hseed, uint64_t *keys)
{
  for (uint64_t k = blockIdx.x * (uint64_t)kBlock + threadIdx.x; k < M; k += (uint64_t)gridDim.x * kBlock) {
    uint64_t h = mix64(hseed ^ k);
    uint64_t u = ((h >> 32) * n) >> 32;
    uint64_t v = ((h & 0xFFFFFFFFull) * n) >> 32;
    keys[2 * k] = u * n + v;
    keys[2 * k + 1] = v * n + u;
  }
}

__global__ void __launch_bounds__(kBlock) k_keep(const uint64_t *__restrict__ keys, uint64_t cnt, uint64_t n,
                                                 uint32_t *__restrict__ keep)
{
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * kBlock) {
    uint64_t k = keys[i];
    bool ok = (i == 0 || keys[i - 1] != k) && (k / n != k % n);
    keep[i] = ok ? 1u : 0u;
  }
}

__global__ void __launch_bounds__(kBlock) k_compact(const uint64_t *__restrict__ keys, uint64_t cnt, uint64_t n,
                                                    const uint32_t *__restrict__ keep, const uint32_t *__restrict__ pos,
                                                    uint32_t *__restrict__ col, uint32_t *__restrict__ src)
{
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < cnt; i += (uint64_t)gridDim.x * kBlock) {
    if (!keep[i]) continue;
    uint64_t k = keys[i];
    col[pos[i]] = (uint32_t)(k % n);
    src[pos[i]] = (uint32_t)(k / n);
  }
}

// row pointers of the kept arcs, whose sources ascend: the first arc of each
// source writes its index into the row heads of that source and of the empty
// rows before it, so every entry of rp[0..n] is written exactly once
__global__ void __launch_bounds__(kBlock) k_gen_rowptr(const uint32_t *__restrict__ src, uint64_t m, uint64_t n,
                                                       uint32_t *__restrict__ rp)
{
  for (uint64_t j = blockIdx.x * (uint64_t)kBlock + threadIdx.x; j < m; j += (uint64_t)gridDim.x * kBlock) {
    const uint32_t s = src[j];
    const uint64_t lo = j ? (uint64_t)src[j - 1] + 1 : 0;
    for (uint64_t v = lo; v <= s; v++) rp[v] = (uint32_t)j;
    if (j + 1 == m)
      for (uint64_t v = (uint64_t)s + 1; v <= n; v++) rp[v] = (uint32_t)m;
  }
}

static int bits_for(uint64_t x)
{
  int b = 0;
  while (b < 64 && (x >> b) != 0) b++;
  return b;
}

// sort + dedup + CSR of `cnt` raw keys (s*n+d) already in d_keys
static grd_status keys_to_csr(grd_ctx *ctx, uint64_t *d_keys, uint64_t cnt, uint64_t n, std::vector<uint32_t> &rp,
                              std::vector<uint32_t> &col)
{
  hipStream_t st = ctx->stream;
  DevMem mem;
  uint64_t *d_alt = nullptr;
  GRD_TRY(mem.alloc_n(&d_alt, std::max<uint64_t>(cnt, 1)));
  const int end_bit = std::max(1, bits_for(n * n - 1));
  rocprim::double_buffer<uint64_t> db(d_keys, d_alt);
  size_t tmp_bytes = 0;
  GRD_HIP(rocprim::radix_sort_keys(nullptr, tmp_bytes, db, (size_t)cnt, 0, end_bit, st));
  void *tmp = nullptr;
  GRD_TRY(mem.alloc(&tmp, tmp_bytes));
  GRD_HIP(rocprim::radix_sort_keys(tmp, tmp_bytes, db, (size_t)cnt, 0, end_bit, st));
  uint64_t *keys = db.current();
  uint32_t *keep = nullptr, *pos = nullptr, *d_col = nullptr, *d_src = nullptr, *d_rp = nullptr;
  GRD_TRY(mem.alloc_n(&keep, cnt + 1));
  GRD_TRY(mem.alloc_n(&pos, cnt + 1));
  uint32_t grid = (uint32_t)std::min<uint64_t>(65535, (cnt + kBlock - 1) / kBlock + 1);
  hipLaunchKernelGGL(k_keep, dim3(grid), dim3(kBlock), 0, st, keys, cnt, n, keep);
  GRD_HIP(hipGetLastError());
  size_t sb = 0;
  GRD_HIP(rocprim::exclusive_scan(nullptr, sb, keep, pos, 0u, (size_t)cnt + 1, rocprim::plus<uint32_t>(), st));
  void *stmp = nullptr;
  GRD_TRY(mem.alloc(&stmp, sb));
  GRD_HIP(hipMemsetAsync(keep + cnt, 0, 4, st));
  GRD_HIP(rocprim::exclusive_scan(stmp, sb, keep, pos, 0u, (size_t)cnt + 1, rocprim::plus<uint32_t>(), st));
  uint32_t m = 0;
  GRD_HIP(hipMemcpyAsync(&m, pos + cnt, 4, hipMemcpyDeviceToHost, st));
  GRD_HIP(hipStreamSynchronize(st));
  GRD_TRY(mem.alloc_n(&d_col, std::max<uint32_t>(m, 1)));
  GRD_TRY(mem.alloc_n(&d_src, std::max<uint32_t>(m, 1)));
  GRD_TRY(mem.alloc_n(&d_rp, n + 1));
  hipLaunchKernelGGL(k_compact, dim3(grid), dim3(kBlock), 0, st, keys, cnt, n, keep, pos, d_col, d_src);
  GRD_HIP(hipGetLastError());
  if (m) {
    uint32_t g2 = (uint32_t)std::min<uint64_t>(65535, ((uint64_t)m + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_gen_rowptr, dim3(g2), dim3(kBlock), 0, st, d_src, (uint64_t)m, n, d_rp);
    GRD_HIP(hipGetLastError());
  } else {
    GRD_HIP(hipMemsetAsync(d_rp, 0, (n + 1) * 4, st));
  }
  rp.resize(n + 1);
  col.resize(m);
  GRD_HIP(hipMemcpyAsync(rp.data(), d_rp, (n + 1) * 4, hipMemcpyDeviceToHost, st));
  if (m) GRD_HIP(hipMemcpyAsync(col.data(), d_col, (uint64_t)m * 4, hipMemcpyDeviceToHost, st));
  GRD_HIP(hipStreamSynchronize(st));
  return GRD_OK;
}

grd_status generate_rmat(grd_ctx *ctx, int scale, int ef, double a, double b, double c, uint64_t seed, int sym,
                         std::vector<uint32_t> &rp, std::vector<uint32_t> &col)
{
  if (scale < 1 || scale > 30 || ef < 1) return fail(GRD_ERR_ARG, "rmat: need 1 <= scale <= 30 and edge_factor >= 1");
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return fail(GRD_ERR_ARG, "rmat: invalid quadrant probabilities");
  const uint64_t n = 1ull << scale;
  const uint64_t E = (uint64_t)ef << scale;
  const uint64_t raw = sym ? 2 * E : E;
  if (raw >= 0xFFFFFFF0ull) return fail(GRD_ERR_ARG, "rmat: raw edge count beyond the u32 limit");
  DevMem mem;
  uint64_t *keys = nullptr;
  GRD_TRY(mem.alloc_n(&keys, raw));
  const uint64_t tA = (uint64_t)(a * 4294967296.0);
  const uint64_t tAB = (uint64_t)((a + b) * 4294967296.0);
  const uint64_t tABC = (uint64_t)((a + b + c) * 4294967296.0);
  uint32_t grid = (uint32_t)std::min<uint64_t>(65535, (E + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(k_rmat, dim3(grid), dim3(kBlock), 0, ctx->stream, E, scale, tA, tAB, tABC, mix64(seed), sym, n,
                     keys);
  GRD_HIP(hipGetLastError());
  return keys_to_csr(ctx, keys, raw, n, rp, col);
}

grd_status generate_er(grd_ctx *ctx, uint32_t n, double deg, uint64_t seed, std::vector<uint32_t> &rp,
                       std::vector<uint32_t> &col)
{
  if (n < 2 || deg <= 0) return fail(GRD_ERR_ARG, "er: need n >= 2 and mean degree > 0");
  const uint64_t M = (uint64_t)((double)n * deg / 2.0);
  if (2 * M >= 0xFFFFFFF0ull) return fail(GRD_ERR_ARG, "er: edge count beyond the u32 limit");
  DevMem mem;
  uint64_t *keys = nullptr;
  GRD_TRY(mem.alloc_n(&keys, std::max<uint64_t>(2 * M, 1)));
  uint32_t grid = (uint32_t)std::min<uint64_t>(65535, (M + kBlock - 1) / kBlock + 1);
  hipLaunchKernelGGL(k_er, dim3(grid), dim3(kBlock), 0, ctx->stream, M, (uint64_t)n, mix64(seed), keys);
  GRD_HIP(hipGetLastError());
  return keys_to_csr(ctx, keys, 2 * M, n, rp, col);
}

void host_weights(uint32_t n, const uint32_t *rp, const uint32_t *col, uint64_t wseed, uint32_t *w, int nthreads)
{
  const uint64_t hs = mix64(wseed);
  parallel_for(nthreads, n, 8192, [&](uint64_t b, uint64_t e) {
    for (uint64_t u = b; u < e; u++)
      for (uint32_t x = rp[u]; x < rp[u + 1]; x++) {
        uint64_t v = col[x];
        uint64_t lo = u < v ? u : v, hi = u < v ? v : u;
        uint64_t h = mix64(hs ^ ((lo << 32) | hi));
        w[x] = 1u + (uint32_t)(((h >> 32) * 255ull) >> 32);
      }
  });
}

}  // namespace grd
