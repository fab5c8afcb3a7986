// layout_dev.hip -- the partitioned layout built on the device.
//
// Same rules and the same result, array for array, as the host construction
// in layout.cpp (the restatement of src/DistributedGraph.cpp:13-785):
//   master(g) = g % C, mirror(g) = g % M or mirror[g]
//   cross mirrors of m = {v : u -> v, mirror(u) = m, mirror(v) != m}
//   own vertices of a (master, mirror) pair by out-degree, highest first,
//   ties by gid; cross mirrors of a (mirror, master) region by compute slot
//   push CSR over own proxies, pull CSC over all proxies with every row
//   sorted by source proxy, cross proxies of each slot in proxy order.
// Every order above is produced by a stable radix sort (rocPRIM) of keys
// whose ties are already in the wanted order, so the device and host builds
// agree bit for bit (grd_graph_digest compares them).  Only the row-block
// cut, a sequential greedy pass over the CSC row pointers, runs on the host.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "grd_device.hpp"
#include "grd_internal.hpp"

namespace grd {

namespace {

constexpr uint32_t kNoneU = 0xFFFFFFFFu;

inline uint32_t grid_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(65535, (n + kBlock - 1) / kBlock)); }

#define GRID_LOOP(i, n) for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * kBlock)

__global__ void k_fill32(uint32_t *p, uint64_t n, uint32_t v) { GRID_LOOP(i, n) p[i] = v; }
__global__ void k_fill64(uint64_t *p, uint64_t n, uint64_t v) { GRID_LOOP(i, n) p[i] = v; }

// degrees, mirror ids, validity (err bit 0: row_ptr not monotone, bit 1:
// mirror id >= M), sinks and the largest degree
__global__ void k_vertex_init(uint32_t n, const uint32_t *__restrict__ rp, const uint32_t *__restrict__ mirror_in,
                              uint32_t M, uint32_t *__restrict__ deg, uint8_t *__restrict__ mir, uint32_t *err,
                              unsigned long long *sinks, uint32_t *maxdeg)
{
  uint64_t ns = 0;
  uint32_t mx = 0;
  GRID_LOOP(v, n)
  {
    const uint32_t a = rp[v], b = rp[v + 1];
    if (b < a) atomicOr(err, 1u);
    const uint32_t d = b - a;
    deg[v] = d;
    ns += (d == 0);
    mx = d > mx ? d : mx;
    const uint32_t mm = mirror_in ? mirror_in[v] : (uint32_t)(v % M);
    if (mm >= M) atomicOr(err, 2u);
    mir[v] = (uint8_t)(mm < M ? mm : 0);
  }
  ns = wave_sum(ns);
  mx = wave_max(mx);
  if (lane_id() == 0) {
    if (ns) atomicAdd(sinks, (unsigned long long)ns);
    if (mx) atomicMax(maxdeg, mx);
  }
}

// head[rp[u]] = u for rows with arcs (max-scanned into the source of every arc)
__global__ void k_heads(uint32_t n, const uint32_t *__restrict__ rp, uint32_t *__restrict__ head)
{
  GRID_LOOP(v, n)
  if (rp[v + 1] > rp[v]) head[rp[v]] = (uint32_t)v;
}

struct MaxOp {
  __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};

// column check (err bit 2) and the cross-mirror bitmaps
__global__ void k_cross(uint64_t m, uint32_t n, const uint32_t *__restrict__ src, const uint32_t *__restrict__ col,
                        const uint8_t *__restrict__ mir, uint32_t NW, uint64_t *crossw, uint32_t *err)
{
  GRID_LOOP(e, m)
  {
    const uint32_t v = col[e];
    if (v >= n) {
      atomicOr(err, 4u);
      continue;
    }
    const uint32_t mu = mir[src[e]];
    if (mir[v] != mu) {
      unsigned long long *wd = (unsigned long long *)(crossw + (uint64_t)mu * NW + (v >> 6));
      const unsigned long long bit = 1ull << (v & 63);
      if (!(__atomic_load_n(wd, __ATOMIC_RELAXED) & bit)) atomicOr(wd, bit);
    }
  }
}

// The per-pair counts below are few (C * M <= 4096): every workgroup counts
// into LDS and adds its nonzero counts to the global array once.
constexpr uint32_t kMaxPairs = 64 * 64;

// own count per pair (c, m) -> ccnt[c*M + m]
__global__ void k_own_count(uint32_t n, const uint8_t *__restrict__ mir, uint32_t C, uint32_t M, uint32_t *ccnt)
{
  __shared__ uint32_t s_c[kMaxPairs];
  const uint32_t np = C * M;
  for (uint32_t i = threadIdx.x; i < np; i += kBlock) s_c[i] = 0;
  __syncthreads();
  GRID_LOOP(v, n) atomicAdd(&s_c[(uint32_t)(v % C) * M + mir[v]], 1u);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < np; i += kBlock)
    if (s_c[i]) atomicAdd(&ccnt[i], s_c[i]);
}

// cross count per region (m, c) -> xcnt[m*C + c] and set bits per word
__global__ void k_cross_count(uint64_t nwords, uint32_t NW, const uint64_t *__restrict__ crossw, uint32_t C,
                              uint32_t M, uint32_t *xcnt, uint32_t *wcount)
{
  __shared__ uint32_t s_c[kMaxPairs];
  const uint32_t np = C * M;
  for (uint32_t i = threadIdx.x; i < np; i += kBlock) s_c[i] = 0;
  __syncthreads();
  GRID_LOOP(i, nwords)
  {
    uint64_t x = crossw[i];
    wcount[i] = (uint32_t)__popcll(x);
    if (!x) continue;
    const uint32_t mm = (uint32_t)(i / NW);
    if (C == 1) {
      atomicAdd(&s_c[mm], (uint32_t)__popcll(x));
    } else {
      while (x) {
        const uint32_t v = (uint32_t)(i % NW) * 64 + (uint32_t)__ffsll((unsigned long long)x) - 1;
        x &= x - 1;
        atomicAdd(&s_c[mm * C + v % C], 1u);
      }
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < np; i += kBlock)
    if (s_c[i]) atomicAdd(&xcnt[i], s_c[i]);
}

// sort keys of the own slots: (pair << 32) | ~deg, value gid
__global__ void k_own_keys(uint32_t n, const uint32_t *__restrict__ deg, const uint8_t *__restrict__ mir, uint32_t C,
                           uint32_t M, uint64_t *__restrict__ key, uint32_t *__restrict__ val)
{
  GRID_LOOP(v, n)
  {
    const uint64_t pair = (uint64_t)(v % C) * M + mir[v];
    key[v] = (pair << 32) | (uint32_t)~deg[v];
    val[v] = (uint32_t)v;
  }
}

// rank within the pair -> compute slot and own proxy
__global__ void k_own_assign(uint32_t n, const uint64_t *__restrict__ key, const uint32_t *__restrict__ val,
                             const uint32_t *__restrict__ gstart, const uint32_t *__restrict__ cbase,
                             const uint32_t *__restrict__ obase, uint32_t C, uint32_t M, uint32_t *__restrict__ gid_slot,
                             uint32_t *__restrict__ own_proxy)
{
  GRID_LOOP(i, n)
  {
    const uint32_t pair = (uint32_t)(key[i] >> 32), v = val[i];
    const uint32_t r = (uint32_t)i - gstart[pair];
    const uint32_t c = pair / M, mm = pair % M;
    gid_slot[v] = cbase[pair] + r;
    own_proxy[v] = obase[mm * C + c] + r;
  }
}

// cross keys: (region << 32) | slot, value gid; wpos = exclusive prefix of
// set bits per bitmap word
__global__ void k_cross_keys(uint64_t nwords, uint32_t NW, const uint64_t *__restrict__ crossw,
                             const uint32_t *__restrict__ wpos, const uint32_t *__restrict__ gid_slot, uint32_t C,
                             uint64_t *__restrict__ key, uint32_t *__restrict__ val)
{
  GRID_LOOP(i, nwords)
  {
    uint64_t x = crossw[i];
    uint32_t o = wpos[i];
    const uint32_t mm = (uint32_t)(i / NW);
    while (x) {
      const uint32_t v = (uint32_t)(i % NW) * 64 + (uint32_t)__ffsll((unsigned long long)x) - 1;
      x &= x - 1;
      key[o] = ((uint64_t)(mm * C + v % C) << 32) | gid_slot[v];
      val[o] = v;
      o++;
    }
  }
}

__global__ void k_cross_assign(uint64_t X, uint32_t n, const uint64_t *__restrict__ key, const uint32_t *__restrict__ val,
                               const uint32_t *__restrict__ xstart, const uint32_t *__restrict__ xbase, uint32_t C,
                               uint32_t *__restrict__ xp, uint32_t *__restrict__ proxy_gid)
{
  GRID_LOOP(i, X)
  {
    const uint32_t region = (uint32_t)(key[i] >> 32), v = val[i];
    const uint32_t p = xbase[region] + ((uint32_t)i - xstart[region]);
    xp[(uint64_t)(region / C) * n + v] = p;
    proxy_gid[p] = v;
  }
}

__global__ void k_slots(uint32_t n, const uint32_t *__restrict__ gid_slot, const uint32_t *__restrict__ own_proxy,
                        const uint32_t *__restrict__ deg, uint32_t *__restrict__ slot_gid, uint32_t *__restrict__ slot_deg,
                        uint32_t *__restrict__ proxy_gid, uint32_t *__restrict__ prp_cnt)
{
  GRID_LOOP(v, n)
  {
    const uint32_t s = gid_slot[v], p = own_proxy[v];
    slot_gid[s] = (uint32_t)v;
    slot_deg[s] = deg[v];
    proxy_gid[p] = (uint32_t)v;
    prp_cnt[p] = deg[v];
  }
}

__global__ void k_proxy_slot(uint32_t P, const uint32_t *__restrict__ proxy_gid, const uint32_t *__restrict__ gid_slot,
                             uint32_t *__restrict__ proxy_slot)
{
  GRID_LOOP(p, P)
  {
    const uint32_t v = proxy_gid[p];
    proxy_slot[p] = v != kNoneU ? gid_slot[v] : kNoneU;
  }
}

// push CSR columns over proxies; psrc = source proxy of each CSR position
__global__ void k_push_csr(uint64_t m, uint32_t n, const uint32_t *__restrict__ src, const uint32_t *__restrict__ rp,
                           const uint32_t *__restrict__ col, const uint32_t *__restrict__ w,
                           const uint8_t *__restrict__ mir, const uint32_t *__restrict__ own_proxy,
                           const uint32_t *__restrict__ xp, const uint32_t *__restrict__ prp, uint32_t *__restrict__ pcol,
                           uint32_t *__restrict__ pw, uint64_t *__restrict__ pval)
{
  GRID_LOOP(e, m)
  {
    const uint32_t u = src[e], v = col[e];
    const uint32_t pu = own_proxy[u];
    const uint32_t o = prp[pu] + (uint32_t)(e - rp[u]);
    const uint32_t mu = mir[u];
    pcol[o] = mir[v] == mu ? own_proxy[v] : xp[(uint64_t)mu * n + v];
    const uint32_t wt = w ? w[e] : 0u;
    if (w) pw[o] = wt;
    pval[o] = ((uint64_t)pu << 32) | wt;  // CSC value: source proxy, weight
  }
}

// Row pointers from sorted row keys, ptr[r] = number of keys below r: the
// first key of each run records its start, the last one the run's length,
// and the lengths (zero for rows without keys) are scanned.  Every pass is a
// flat grid over the keys or the rows, so long runs of empty rows (the
// zero-degree tail of a degree-ordered region) cost nothing extra.
__global__ void k_run_start(uint64_t count, const uint32_t *__restrict__ key, uint32_t *__restrict__ start)
{
  GRID_LOOP(e, count)
  {
    const uint32_t k = key[e];
    if (e == 0 || key[e - 1] != k) start[k] = (uint32_t)e;
  }
}

__global__ void k_run_len(uint64_t count, const uint32_t *__restrict__ key, const uint32_t *__restrict__ start,
                          uint32_t *__restrict__ len)
{
  GRID_LOOP(e, count)
  {
    const uint32_t k = key[e];
    if (e + 1 == count || key[e + 1] != k) len[k] = (uint32_t)(e + 1) - start[k];
  }
}

__global__ void k_csc_unpack(uint64_t m, const uint64_t *__restrict__ val, uint32_t *__restrict__ ccol,
                             uint32_t *__restrict__ cw)
{
  GRID_LOOP(e, m)
  {
    const uint64_t x = val[e];
    ccol[e] = (uint32_t)(x >> 32);
    if (cw) cw[e] = (uint32_t)x;
  }
}

// every cross proxy in ascending proxy order, keyed by its compute slot
__global__ void k_cross_list(uint32_t nreg, const uint32_t *__restrict__ xbase, const uint32_t *__restrict__ xcnt,
                             const uint32_t *__restrict__ xstart, const uint32_t *__restrict__ proxy_slot,
                             uint32_t *__restrict__ key, uint32_t *__restrict__ val)
{
  for (uint32_t k = blockIdx.y; k < nreg; k += gridDim.y)
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < xcnt[k]; i += gridDim.x * kBlock) {
      const uint32_t p = xbase[k] + i;
      key[xstart[k] + i] = proxy_slot[p];
      val[xstart[k] + i] = p;
    }
}

template <typename T>
grd_status exclusive_scan_u32(DevMem &mem, hipStream_t st, const T *in, T *out, uint64_t n)
{
  size_t tb = 0;
  GRD_HIP(rocprim::exclusive_scan(nullptr, tb, in, out, T(0), (size_t)n, rocprim::plus<T>(), st));
  void *tmp = nullptr;
  GRD_TRY(mem.alloc(&tmp, tb));
  GRD_HIP(rocprim::exclusive_scan(tmp, tb, in, out, T(0), (size_t)n, rocprim::plus<T>(), st));
  return GRD_OK;
}

// ptr[0..R] for `count` keys sorted ascending, each below R
grd_status rowptr_sorted(DevMem &mem, hipStream_t st, uint64_t count, const uint32_t *key, uint32_t R, uint32_t *ptr)
{
  uint32_t *start = nullptr, *len = nullptr;
  GRD_TRY(mem.alloc_n(&start, std::max<uint32_t>(R, 1)));
  GRD_TRY(mem.alloc_n(&len, (uint64_t)R + 1));
  hipLaunchKernelGGL(k_fill32, dim3(grid_for((uint64_t)R + 1)), dim3(kBlock), 0, st, len, (uint64_t)R + 1, 0u);
  if (count) {
    hipLaunchKernelGGL(k_run_start, dim3(grid_for(count)), dim3(kBlock), 0, st, count, key, start);
    hipLaunchKernelGGL(k_run_len, dim3(grid_for(count)), dim3(kBlock), 0, st, count, key, start, len);
  }
  GRD_HIP(hipGetLastError());
  return exclusive_scan_u32(mem, st, len, ptr, (uint64_t)R + 1);
}

template <typename K, typename V>
grd_status sort_pairs(DevMem &mem, hipStream_t st, K *&keys, V *&vals, uint64_t n, int end_bit)
{
  K *k2 = nullptr;
  V *v2 = nullptr;
  GRD_TRY(mem.alloc_n(&k2, std::max<uint64_t>(n, 1)));
  GRD_TRY(mem.alloc_n(&v2, std::max<uint64_t>(n, 1)));
  rocprim::double_buffer<K> kb(keys, k2);
  rocprim::double_buffer<V> vb(vals, v2);
  size_t tb = 0;
  GRD_HIP(rocprim::radix_sort_pairs(nullptr, tb, kb, vb, (size_t)n, 0, end_bit, st));
  void *tmp = nullptr;
  GRD_TRY(mem.alloc(&tmp, tb));
  GRD_HIP(rocprim::radix_sort_pairs(tmp, tb, kb, vb, (size_t)n, 0, end_bit, st));
  keys = kb.current();
  vals = vb.current();
  return GRD_OK;
}

int bits_of(uint64_t x)
{
  int b = 1;
  while (b < 64 && (x >> b)) b++;
  return b;
}

template <typename T>
grd_status download(hipStream_t st, std::vector<T> &dst, const T *src, uint64_t n)
{
  dst.resize(n);
  if (n) GRD_HIP(hipMemcpyAsync(dst.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, st));
  return GRD_OK;
}

template <typename T>
grd_status upload_small(DevMem &mem, hipStream_t st, T **dst, const std::vector<T> &src)
{
  GRD_TRY(mem.alloc_n(dst, std::max<size_t>(src.size(), 1)));
  if (!src.empty()) GRD_HIP(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, st));
  return GRD_OK;
}

grd_status build_dev_impl(grd_ctx *ctx, uint32_t n, const uint32_t *h_rp, const uint32_t *h_col, const uint32_t *h_w,
                          int C, int M, const uint32_t *mirror_in, grd_graph *g)
{
  hipStream_t st = ctx->stream;
  const uint64_t m = h_rp[n];
  const uint32_t NW = (n + 63) / 64;
  DevMem tmp;  // scratch of the build, freed on return
  // inputs
  uint32_t *rp = nullptr, *col = nullptr, *w = nullptr, *mirror = nullptr;
  GRD_TRY(tmp.alloc_n(&rp, (uint64_t)n + 1));
  GRD_TRY(tmp.alloc_n(&col, std::max<uint64_t>(m, 1)));
  GRD_HIP(hipMemcpyAsync(rp, h_rp, ((uint64_t)n + 1) * 4, hipMemcpyHostToDevice, st));
  if (m) GRD_HIP(hipMemcpyAsync(col, h_col, m * 4, hipMemcpyHostToDevice, st));
  if (h_w) {
    GRD_TRY(tmp.alloc_n(&w, std::max<uint64_t>(m, 1)));
    if (m) GRD_HIP(hipMemcpyAsync(w, h_w, m * 4, hipMemcpyHostToDevice, st));
  }
  if (mirror_in && M > 1) {
    GRD_TRY(tmp.alloc_n(&mirror, n));
    GRD_HIP(hipMemcpyAsync(mirror, mirror_in, (uint64_t)n * 4, hipMemcpyHostToDevice, st));
  }
  // vertices
  uint32_t *deg = nullptr, *err = nullptr, *maxdeg = nullptr;
  uint8_t *mir = nullptr;
  unsigned long long *sinks = nullptr;
  GRD_TRY(tmp.alloc_n(&deg, n));
  GRD_TRY(tmp.alloc_n(&mir, n));
  GRD_TRY(tmp.alloc_n(&err, 4));
  GRD_TRY(tmp.alloc_n(&maxdeg, 4));
  GRD_TRY(tmp.alloc_n(&sinks, 2));
  hipLaunchKernelGGL(k_fill32, dim3(1), dim3(kBlock), 0, st, err, 4, 0u);
  hipLaunchKernelGGL(k_fill32, dim3(1), dim3(kBlock), 0, st, maxdeg, 4, 0u);
  hipLaunchKernelGGL(k_fill64, dim3(1), dim3(kBlock), 0, st, (uint64_t *)sinks, 2, 0ull);
  hipLaunchKernelGGL(k_vertex_init, dim3(grid_for(n)), dim3(kBlock), 0, st, n, rp, mirror, (uint32_t)M, deg, mir, err,
                     sinks, maxdeg);
  // source of every arc: max-scan of the row heads
  uint32_t *src = nullptr, *head = nullptr;
  GRD_TRY(tmp.alloc_n(&src, std::max<uint64_t>(m, 1)));
  GRD_TRY(tmp.alloc_n(&head, std::max<uint64_t>(m, 1)));
  hipLaunchKernelGGL(k_fill32, dim3(grid_for(m)), dim3(kBlock), 0, st, head, m, 0u);
  hipLaunchKernelGGL(k_heads, dim3(grid_for(n)), dim3(kBlock), 0, st, n, rp, head);
  if (m) {
    size_t tb = 0;
    GRD_HIP(rocprim::inclusive_scan(nullptr, tb, head, src, (size_t)m, MaxOp(), st));
    void *t = nullptr;
    GRD_TRY(tmp.alloc(&t, tb));
    GRD_HIP(rocprim::inclusive_scan(t, tb, head, src, (size_t)m, MaxOp(), st));
  }
  // cross-mirror bitmaps and the column check
  uint64_t *crossw = nullptr;
  const uint64_t nwords = (uint64_t)M * NW;
  GRD_TRY(tmp.alloc_n(&crossw, nwords));
  hipLaunchKernelGGL(k_fill64, dim3(grid_for(nwords)), dim3(kBlock), 0, st, crossw, nwords, 0ull);
  hipLaunchKernelGGL(k_cross, dim3(grid_for(m)), dim3(kBlock), 0, st, m, n, src, col, mir, NW, crossw, err);
  GRD_HIP(hipGetLastError());
  uint32_t h_err = 0, h_maxdeg = 0;
  unsigned long long h_sinks = 0;
  GRD_HIP(hipMemcpyAsync(&h_err, err, 4, hipMemcpyDeviceToHost, st));
  GRD_HIP(hipMemcpyAsync(&h_maxdeg, maxdeg, 4, hipMemcpyDeviceToHost, st));
  GRD_HIP(hipMemcpyAsync(&h_sinks, sinks, 8, hipMemcpyDeviceToHost, st));
  GRD_HIP(hipStreamSynchronize(st));
  if (h_err & 1u) return fail(GRD_ERR_ARG, "graph: row_ptr is not monotone");
  if (h_err & 2u) return fail(GRD_ERR_ARG, "graph: a mirror partition id is >= M");
  if (h_err & 4u) return fail(GRD_ERR_ARG, "graph: column index out of range");
  g->n_sinks = h_sinks;
  g->max_deg = h_maxdeg;
  // group sizes
  const uint32_t NP = (uint32_t)(C * M);
  uint32_t *ccnt = nullptr, *xcnt = nullptr, *wcount = nullptr, *wpos = nullptr;
  GRD_TRY(tmp.alloc_n(&ccnt, NP));
  GRD_TRY(tmp.alloc_n(&xcnt, NP));
  GRD_TRY(tmp.alloc_n(&wcount, nwords + 1));
  GRD_TRY(tmp.alloc_n(&wpos, nwords + 1));
  hipLaunchKernelGGL(k_fill32, dim3(1), dim3(kBlock), 0, st, ccnt, NP, 0u);
  hipLaunchKernelGGL(k_fill32, dim3(1), dim3(kBlock), 0, st, xcnt, NP, 0u);
  hipLaunchKernelGGL(k_fill32, dim3(1), dim3(kBlock), 0, st, wcount + nwords, 1, 0u);
  hipLaunchKernelGGL(k_own_count, dim3(grid_for(n)), dim3(kBlock), 0, st, n, mir, (uint32_t)C, (uint32_t)M, ccnt);
  hipLaunchKernelGGL(k_cross_count, dim3(grid_for(nwords)), dim3(kBlock), 0, st, nwords, NW, crossw, (uint32_t)C,
                     (uint32_t)M, xcnt, wcount);
  GRD_TRY(exclusive_scan_u32(tmp, st, wcount, wpos, nwords + 1));
  GRD_TRY(download(st, g->ccnt, ccnt, NP));
  GRD_TRY(download(st, g->xcnt, xcnt, NP));
  GRD_HIP(hipStreamSynchronize(st));
  // bases (host, as in layout.cpp)
  g->cbase.assign(NP, 0);
  uint64_t cur = 0;
  for (int c = 0; c < C; c++)
    for (int mm = 0; mm < M; mm++) {
      g->cbase[(size_t)c * M + mm] = (uint32_t)cur;
      cur += pad64(g->ccnt[(size_t)c * M + mm]);
    }
  if (cur >= 0xFFFFFF00ull) return fail(GRD_ERR_ARG, "graph: compute slots exceed u32");
  g->S = (uint32_t)std::max<uint64_t>(cur, 64);
  g->obase.assign(NP, 0);
  g->ocnt.assign(NP, 0);
  g->xbase.assign(NP, 0);
  cur = 0;
  uint64_t X = 0;
  for (int mm = 0; mm < M; mm++)
    for (int c = 0; c < C; c++) {
      const size_t k = (size_t)mm * C + c;
      g->ocnt[k] = g->ccnt[(size_t)c * M + mm];
      g->obase[k] = (uint32_t)cur;
      cur += pad64(g->ocnt[k]);
      g->xbase[k] = (uint32_t)cur;
      cur += pad64(g->xcnt[k]);
      X += g->xcnt[k];
    }
  if (cur >= 0xFFFFFF00ull) return fail(GRD_ERR_ARG, "graph: memory proxies exceed u32");
  g->P = (uint32_t)std::max<uint64_t>(cur, 64);
  g->X = X;
  const uint32_t S = g->S, P = g->P;
  // pair / region starts in sorted order
  std::vector<uint32_t> gstart(NP, 0), xstart(NP, 0);
  for (uint32_t k = 1; k < NP; k++) gstart[k] = gstart[k - 1] + g->ccnt[k - 1];
  for (uint32_t k = 1; k < NP; k++) xstart[k] = xstart[k - 1] + g->xcnt[k - 1];
  uint32_t *d_gstart = nullptr, *d_xstart = nullptr, *d_cbase = nullptr, *d_obase = nullptr, *d_xbase = nullptr;
  GRD_TRY(upload_small(tmp, st, &d_gstart, gstart));
  GRD_TRY(upload_small(tmp, st, &d_xstart, xstart));
  GRD_TRY(upload_small(tmp, st, &d_cbase, g->cbase));
  GRD_TRY(upload_small(tmp, st, &d_obase, g->obase));
  GRD_TRY(upload_small(tmp, st, &d_xbase, g->xbase));
  // own slots: stable sort by (pair, degree descending), gid ascending in ties
  uint64_t *okey = nullptr;
  uint32_t *oval = nullptr, *gid_slot = nullptr, *own_proxy = nullptr;
  GRD_TRY(tmp.alloc_n(&okey, n));
  GRD_TRY(tmp.alloc_n(&oval, n));
  GRD_TRY(tmp.alloc_n(&gid_slot, n));
  GRD_TRY(tmp.alloc_n(&own_proxy, n));
  hipLaunchKernelGGL(k_own_keys, dim3(grid_for(n)), dim3(kBlock), 0, st, n, deg, mir, (uint32_t)C, (uint32_t)M, okey,
                     oval);
  GRD_TRY(sort_pairs(tmp, st, okey, oval, n, 32 + bits_of(NP)));
  hipLaunchKernelGGL(k_own_assign, dim3(grid_for(n)), dim3(kBlock), 0, st, n, okey, oval, d_gstart, d_cbase, d_obase,
                     (uint32_t)C, (uint32_t)M, gid_slot, own_proxy);
  // cross proxies: sort by (region, compute slot)
  uint32_t *xp = nullptr, *proxy_gid = nullptr;
  GRD_TRY(tmp.alloc_n(&xp, std::max<uint64_t>((uint64_t)M * n, 1)));
  GRD_TRY(tmp.alloc_n(&proxy_gid, P));
  hipLaunchKernelGGL(k_fill32, dim3(grid_for(P)), dim3(kBlock), 0, st, proxy_gid, (uint64_t)P, kNoneU);
  if (X) {
    uint64_t *xkey = nullptr;
    uint32_t *xval = nullptr;
    GRD_TRY(tmp.alloc_n(&xkey, X));
    GRD_TRY(tmp.alloc_n(&xval, X));
    hipLaunchKernelGGL(k_cross_keys, dim3(grid_for(nwords)), dim3(kBlock), 0, st, nwords, NW, crossw, wpos, gid_slot,
                       (uint32_t)C, xkey, xval);
    GRD_TRY(sort_pairs(tmp, st, xkey, xval, X, 32 + bits_of(NP)));
    hipLaunchKernelGGL(k_cross_assign, dim3(grid_for(X)), dim3(kBlock), 0, st, X, n, xkey, xval, d_xstart, d_xbase,
                       (uint32_t)C, xp, proxy_gid);
  }
  // per-slot and per-proxy arrays
  DevMem &gm = g->mem;
  uint32_t *prp_cnt = nullptr;
  GRD_TRY(gm.alloc_n(&g->slot_gid, S));
  GRD_TRY(gm.alloc_n(&g->slot_deg, S));
  GRD_TRY(gm.alloc_n(&g->proxy_slot, P));
  GRD_TRY(tmp.alloc_n(&prp_cnt, (uint64_t)P + 1));
  hipLaunchKernelGGL(k_fill32, dim3(grid_for(S)), dim3(kBlock), 0, st, g->slot_gid, (uint64_t)S, kNoneU);
  hipLaunchKernelGGL(k_fill32, dim3(grid_for(S)), dim3(kBlock), 0, st, g->slot_deg, (uint64_t)S, 0u);
  hipLaunchKernelGGL(k_fill32, dim3(grid_for(P + 1)), dim3(kBlock), 0, st, prp_cnt, (uint64_t)P + 1, 0u);
  hipLaunchKernelGGL(k_slots, dim3(grid_for(n)), dim3(kBlock), 0, st, n, gid_slot, own_proxy, deg, g->slot_gid,
                     g->slot_deg, proxy_gid, prp_cnt);
  hipLaunchKernelGGL(k_proxy_slot, dim3(grid_for(P)), dim3(kBlock), 0, st, P, proxy_gid, gid_slot, g->proxy_slot);
  {
    std::vector<uint32_t> word_dst(S / 64, kNone);
    for (int c = 0; c < C; c++)
      for (int mm = 0; mm < M; mm++) {
        const uint32_t sb = g->cbase[(size_t)c * M + mm] / 64;
        const uint32_t nwd = (uint32_t)(pad64(g->ccnt[(size_t)c * M + mm]) / 64);
        const uint32_t pb = g->obase[(size_t)mm * C + c] / 64;
        for (uint32_t k = 0; k < nwd; k++) word_dst[sb + k] = pb + k;
      }
    GRD_TRY(upload_small(gm, st, &g->word_dst, word_dst));
    GRD_HIP(hipStreamSynchronize(st));  // word_dst leaves scope
  }
  // push CSR over proxies
  GRD_TRY(gm.alloc_n(&g->rp, (uint64_t)P + 1));
  GRD_TRY(exclusive_scan_u32(tmp, st, prp_cnt, g->rp, (uint64_t)P + 1));
  GRD_TRY(gm.alloc_n(&g->col, std::max<uint64_t>(m, 1)));
  if (h_w) GRD_TRY(gm.alloc_n(&g->w, std::max<uint64_t>(m, 1)));
  uint64_t *pval = nullptr;
  GRD_TRY(tmp.alloc_n(&pval, std::max<uint64_t>(m, 1)));
  hipLaunchKernelGGL(k_push_csr, dim3(grid_for(m)), dim3(kBlock), 0, st, m, n, src, rp, col, w, mir, own_proxy, xp,
                     g->rp, g->col, g->w, pval);
  // pull CSC: stable sort of the push arcs (ascending source proxy) by target
  uint32_t *ckey = nullptr;
  GRD_TRY(tmp.alloc_n(&ckey, std::max<uint64_t>(m, 1)));
  if (m) GRD_HIP(hipMemcpyAsync(ckey, g->col, m * 4, hipMemcpyDeviceToDevice, st));
  if (m) GRD_TRY(sort_pairs(tmp, st, ckey, pval, m, bits_of(P)));
  GRD_TRY(gm.alloc_n(&g->crp, (uint64_t)P + 1));
  GRD_TRY(rowptr_sorted(tmp, st, m, ckey, P, g->crp));
  GRD_TRY(gm.alloc_n(&g->ccol, std::max<uint64_t>(m, 1)));
  if (h_w) GRD_TRY(gm.alloc_n(&g->cw, std::max<uint64_t>(m, 1)));
  hipLaunchKernelGGL(k_csc_unpack, dim3(grid_for(m)), dim3(kBlock), 0, st, m, pval, g->ccol, g->cw);
  // cross proxies of every slot, in proxy order
  GRD_TRY(gm.alloc_n(&g->xptr, (uint64_t)S + 1));
  GRD_TRY(gm.alloc_n(&g->xlist, std::max<uint64_t>(X, 1)));
  if (X) {
    uint32_t *lkey = nullptr, *lval = nullptr;
    GRD_TRY(tmp.alloc_n(&lkey, X));
    GRD_TRY(tmp.alloc_n(&lval, X));
    hipLaunchKernelGGL(k_cross_list, dim3(64, std::min<uint32_t>(NP, 65535)), dim3(kBlock), 0, st, NP, d_xbase, xcnt,
                       d_xstart, g->proxy_slot, lkey, lval);
    GRD_TRY(sort_pairs(tmp, st, lkey, lval, X, bits_of(S)));
    GRD_HIP(hipMemcpyAsync(g->xlist, lval, X * 4, hipMemcpyDeviceToDevice, st));
    GRD_TRY(rowptr_sorted(tmp, st, X, lkey, S, g->xptr));
  } else {
    GRD_TRY(rowptr_sorted(tmp, st, 0, nullptr, S, g->xptr));
  }
  GRD_HIP(hipGetLastError());
  // host copies
  std::vector<uint32_t> crp_h;
  GRD_TRY(download(st, g->h_deg, deg, n));
  GRD_TRY(download(st, g->h_gid_slot, gid_slot, n));
  GRD_TRY(download(st, g->h_proxy_gid, proxy_gid, P));
  GRD_TRY(download(st, crp_h, g->crp, (uint64_t)P + 1));
  GRD_HIP(hipStreamSynchronize(st));
  return build_row_blocks(g, crp_h);
}

}  // namespace

grd_status build_graph_device(grd_ctx *ctx, uint32_t n, const uint32_t *rp, const uint32_t *col, const uint32_t *w,
                              int C, int M, const uint32_t *mirror, grd_graph **out)
{
  if (n == 0 || C < 1 || M < 1 || C > 64 || M > 64)
    return fail(GRD_ERR_ARG, "graph: need n >= 1 and 1 <= C, M <= 64");
  if (rp[n] >= 0xFFFFFFF0ull) return fail(GRD_ERR_ARG, "graph: more than 2^32-16 arcs is beyond the u32 offset limit");
  GRD_TRY(check_row_ptr(n, rp));  // the device passes index by row_ptr
  grd_graph *g = new (std::nothrow) grd_graph();
  if (!g) return fail(GRD_ERR_OOM, "graph: host allocation failed");
  g->ctx = ctx;
  g->n = n;
  g->m = rp[n];
  g->C = C;
  g->M = M;
  g->weighted = (w != nullptr);
  g->avg_degree = (double)g->m / (double)n;
  grd_status st;
  try {
    st = build_dev_impl(ctx, n, rp, col, w, C, M, mirror, g);
  } catch (const std::bad_alloc &) {
    st = fail(GRD_ERR_OOM, "graph: host allocation failed");
  }
  if (st != GRD_OK) {
    (void)hipStreamSynchronize(ctx->stream);
    delete g;
    return st;
  }
  *out = g;
  return GRD_OK;
}

}  // namespace grd
