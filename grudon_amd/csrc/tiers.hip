// tiers.hip -- the tiled PageRank pull: every gather served from LDS.
//
// The pull sums, for every proxy row r of a memory partition, the values of
// its in-arcs' source proxies.  The sources are the partition's own proxies,
// numbered in out-degree order, and a random 4-byte gather from a table of
// tens of MB is served by L2 at best and by a whole cache line from the
// Infinity Cache / HBM at worst.  Here the partition's proxy range is cut
// into tiers of 2^tier_bits (at most 2^15) consecutive proxies, and the arcs
// of the CSC are regrouped by (memory partition, tier, row):
//
//   pair q       = one (row, tier) with at least one arc; its arcs are the
//                  contiguous run of the row's CSC arcs whose sources lie in
//                  the tier (rows keep their sources ascending)
//   tcol[e]      = the source's 16-bit offset within its tier
//   tile         = a run of row blocks of one (partition, tier) group
//
// A 512-thread workgroup takes a tile, stages the tier's window of mprop
// (<= 128 KB) into LDS with coalesced loads, then its two 256-thread halves
// walk the tile's row blocks like the plain pull (8 coalesced column loads
// and 8 LDS gathers per thread, a group of lanes per row, FP32 Kahan per lane
// merged by TwoSum).  Each pair yields one compensated partial; a row with a
// single pair writes its value straight into the reduce buffer, the others
// are merged per row in tier order (k_prp_tier_merge), so the result does not
// depend on scheduling.  A pair longer than one block is cut into segments
// merged in order by a fix-up pass.
//
// HBM sees only streams: 2 bytes of column per arc, the pair records, and the
// partials of rows with more than one pair.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <vector>

#include "grd_device.hpp"
#include "grd_internal.hpp"

namespace grd {

namespace {

constexpr int kTierBlock = 512;  // threads per tile workgroup: two halves of kBlock
constexpr uint32_t kSole = 0x80000000u;

inline uint32_t tgrid(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(65535, (n + kBlock - 1) / kBlock)); }

#define TIER_LOOP(i, n) for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < (n); i += (uint64_t)gridDim.x * kBlock)

struct MaxU32 {
  __host__ __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};

// memory partition of a proxy row: pb[0..M] are the partitions' first proxies
__device__ __forceinline__ uint32_t part_of(const uint32_t *__restrict__ pb, uint32_t M, uint32_t r)
{
  uint32_t m = 0;
  while (m + 1 < M && pb[m + 1] <= r) m++;
  return m;
}

__global__ void k_t_fill(uint32_t *p, uint64_t n, uint32_t v) { TIER_LOOP(i, n) p[i] = v; }

// head[crp[r]] = r for every row with arcs (positions are distinct)
__global__ void k_t_heads(uint32_t P, const uint32_t *__restrict__ crp, uint32_t *__restrict__ head)
{
  TIER_LOOP(r, P)
  if (crp[r] < crp[r + 1]) head[crp[r]] = (uint32_t)r;
}

// flag[e] = 1 when arc e starts a pair: the row's first arc, or a source in
// another tier than the previous arc's
__global__ void k_t_flags(uint64_t m, const uint32_t *__restrict__ arow, const uint32_t *__restrict__ crp,
                          const uint32_t *__restrict__ ccol, const uint32_t *__restrict__ pb, uint32_t M, uint32_t tb,
                          uint32_t *__restrict__ flag)
{
  TIER_LOOP(e, m)
  {
    const uint32_t r = arow[e];
    const uint32_t base = pb[part_of(pb, M, r)];
    const uint32_t t = (ccol[e] - base) >> tb;
    bool start = e == crp[r];
    if (!start) start = t != ((ccol[e - 1] - base) >> tb);  // e > crp[r] >= 0 here
    flag[e] = start ? 1u : 0u;
  }
}

// pair records in (row, tier) order
__global__ void k_t_pairs(uint64_t m, const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pid,
                          const uint32_t *__restrict__ arow, const uint32_t *__restrict__ ccol,
                          const uint32_t *__restrict__ pb, uint32_t M, uint32_t tb, uint32_t ntier_max,
                          uint32_t *__restrict__ prow, uint32_t *__restrict__ pe0, uint32_t *__restrict__ pkey,
                          uint32_t *__restrict__ pval)
{
  TIER_LOOP(e, m)
  if (flag[e]) {
    const uint32_t p = pid[e], r = arow[e];
    const uint32_t mp = part_of(pb, M, r);
    prow[p] = r;
    pe0[p] = (uint32_t)e;
    pkey[p] = mp * ntier_max + ((ccol[e] - pb[mp]) >> tb);
    pval[p] = p;
  }
}

__global__ void k_t_rptr(uint32_t P, const uint32_t *__restrict__ crp, const uint32_t *__restrict__ pid,
                         uint32_t *__restrict__ trptr)
{
  TIER_LOOP(r, (uint64_t)P + 1) trptr[r] = pid[crp[r]];
}

// per pair in tier order: length, row record, and the inverse permutation
__global__ void k_t_order(uint64_t np, const uint32_t *__restrict__ perm, const uint32_t *__restrict__ prow,
                          const uint32_t *__restrict__ pe0, const uint32_t *__restrict__ crp,
                          const uint32_t *__restrict__ trptr, uint32_t *__restrict__ tlen, uint32_t *__restrict__ trow,
                          uint32_t *__restrict__ tinv)
{
  TIER_LOOP(i, np)
  {
    const uint32_t p = perm[i], r = prow[p];
    uint32_t end = crp[r + 1];
    if (p + 1 < np && pe0[p + 1] < end) end = pe0[p + 1];
    tlen[i] = end - pe0[p];
    trow[i] = r | ((trptr[r + 1] - trptr[r] == 1) ? kSole : 0u);
    tinv[p] = (uint32_t)i;
  }
}

// every arc to its place in tier order, as an offset within its tier
__global__ void k_t_copy(uint64_t m, const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pid,
                         const uint32_t *__restrict__ pe0, const uint32_t *__restrict__ prow,
                         const uint32_t *__restrict__ tinv, const uint32_t *__restrict__ tptr,
                         const uint32_t *__restrict__ ccol, const uint32_t *__restrict__ pb, uint32_t M, uint32_t tb,
                         uint16_t *__restrict__ tcol)
{
  TIER_LOOP(e, m)
  {
    const uint32_t p = pid[e] + flag[e] - 1;  // the pair holding arc e
    const uint32_t base = pb[part_of(pb, M, prow[p])];
    const uint32_t off = ccol[e] - base;
    tcol[tptr[tinv[p]] + ((uint32_t)e - pe0[p])] = (uint16_t)(off & ((1u << tb) - 1u));
  }
}

template <typename T>
grd_status t_scan(DevMem &mem, hipStream_t st, const T *in, T *out, uint64_t n)
{
  size_t tb = 0;
  GRD_HIP(rocprim::exclusive_scan(nullptr, tb, in, out, T(0), (size_t)n, rocprim::plus<T>(), st));
  void *tmp = nullptr;
  GRD_TRY(mem.alloc(&tmp, std::max<size_t>(tb, 1)));
  GRD_HIP(rocprim::exclusive_scan(tmp, tb, in, out, T(0), (size_t)n, rocprim::plus<T>(), st));
  return GRD_OK;
}

template <typename T>
grd_status t_upload(DevMem &mem, hipStream_t st, T **dst, const std::vector<T> &v)
{
  GRD_TRY(mem.alloc_n(dst, std::max<size_t>(v.size(), 1)));
  if (!v.empty()) GRD_HIP(hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
  return GRD_OK;
}

// ---- the tiled pull ----

__global__ void __launch_bounds__(kTierBlock) k_prp_pull_tiled(const TierTile *__restrict__ tiles,
                                                               const PullBlock *__restrict__ blocks,
                                                               const uint32_t *__restrict__ tptr,
                                                               const uint32_t *__restrict__ trow,
                                                               const uint16_t *__restrict__ tcol,
                                                               const float *__restrict__ mprop,
                                                               float *__restrict__ out, float2 *__restrict__ part,
                                                               float2 *__restrict__ seg)
{
  __shared__ float s_win[1u << kTierBitsMax];
  __shared__ float s_v[2][kPullNnz];
  __shared__ float s_ws[2][kWavesPerBlock], s_wc[2][kWavesPerBlock];
  const TierTile T = tiles[blockIdx.x];
  for (uint32_t i = threadIdx.x; i < T.wlen; i += kTierBlock) s_win[i] = mprop[T.wbase + i];
  __syncthreads();
  const uint32_t half = threadIdx.x / kBlock, lt = threadIdx.x % kBlock;
  const uint32_t rounds = (T.nb + 1) / 2;
  for (uint32_t rd = 0; rd < rounds; rd++) {
    const uint32_t bi = 2 * rd + half;
    const bool valid = bi < T.nb;
    const PullBlock d = valid ? blocks[T.b0 + bi] : PullBlock{0, 0, 0, 0};
    const uint32_t ne = d.e1 - d.e0;
    if (valid && d.nrows == 0) {
      // one segment of a long pair: a compensated sum of its arcs
      Kahan k;
      for (uint32_t e = d.e0 + lt; e < d.e1; e += kBlock) k.add(s_win[tcol[e]]);
      kahan_wave_merge(k, kWave);
      if (lane_id() == 0) {
        s_ws[half][lt / kWave] = k.s;
        s_wc[half][lt / kWave] = k.c;
      }
    } else if (valid) {
      constexpr int kPer = kPullNnz / kBlock;
      uint32_t c[kPer];
#pragma unroll
      for (int k = 0; k < kPer; k++) {
        const uint32_t i = lt + k * kBlock;
        c[k] = i < ne ? tcol[d.e0 + i] : 0u;
      }
#pragma unroll
      for (int k = 0; k < kPer; k++) {
        const uint32_t i = lt + k * kBlock;
        if (i < ne) s_v[half][i] = s_win[c[k]];
      }
    }
    __syncthreads();
    if (valid && d.nrows == 0) {
      if (lt == 0) {
        Kahan t;
        t.s = s_ws[half][0];
        t.c = s_wc[half][0];
        for (int j = 1; j < kWavesPerBlock; j++) t.merge(s_ws[half][j], s_wc[half][j]);
        seg[T.b0 + bi] = make_float2(t.s, t.c);
      }
    } else if (valid) {
      // the pairs of the block, a group of G lanes each; every group runs
      // the same number of rounds so that the group shuffles stay in step
      int G = 64;
      while (G > 1 && (uint32_t)G * d.nrows > (uint32_t)kBlock) G >>= 1;
      const uint32_t sub = lt % G, ngroups = kBlock / G;
      const uint32_t nr = (d.nrows + ngroups - 1) / ngroups;
      for (uint32_t k = 0; k < nr; k++) {
        const uint32_t ql = k * ngroups + lt / G;
        const uint32_t q = d.r0 + ql;
        Kahan acc;
        if (ql < d.nrows) {
          const uint32_t b = tptr[q] - d.e0, e = tptr[q + 1] - d.e0;
          for (uint32_t x = b + sub; x < e; x += G) acc.add(s_v[half][x]);
        }
        kahan_wave_merge(acc, G);
        if (sub == 0 && ql < d.nrows) {
          const uint32_t rw = trow[q];
          if (rw & kSole) out[rw & ~kSole] = acc.value();
          else part[q] = make_float2(acc.s, acc.c);
        }
      }
    }
    __syncthreads();
  }
}

// long pairs: their segments merged in order
__global__ void __launch_bounds__(kBlock) k_prp_tier_fixup(const PullLong *__restrict__ longs, uint32_t nlong,
                                                           const float2 *__restrict__ seg,
                                                           const uint32_t *__restrict__ trow, float *__restrict__ out,
                                                           float2 *__restrict__ part)
{
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nlong; i += gridDim.x * kBlock) {
    const PullLong L = longs[i];
    Kahan t;
    for (uint32_t j = 0; j < L.nseg; j++) {
      const float2 p = seg[L.first + j];
      t.merge(p.x, p.y);
    }
    const uint32_t rw = trow[L.row];
    if (rw & kSole) out[rw & ~kSole] = t.value();
    else part[L.row] = make_float2(t.s, t.c);
  }
}

// rows with no pair (0) or several (their partials in tier order)
__global__ void __launch_bounds__(kBlock) k_prp_tier_merge(uint32_t P, const uint32_t *__restrict__ trptr,
                                                           const uint32_t *__restrict__ tinv,
                                                           const float2 *__restrict__ part, float *__restrict__ out)
{
  for (uint32_t r = blockIdx.x * kBlock + threadIdx.x; r < P; r += gridDim.x * kBlock) {
    const uint32_t a = trptr[r], b = trptr[r + 1];
    if (b - a == 1) continue;
    Kahan t;
    for (uint32_t x = a; x < b; x++) {
      const float2 p = part[tinv[x]];
      t.merge(p.x, p.y);
    }
    out[r] = t.value();
  }
}

}  // namespace

// Regroups the CSC by (partition, tier, row) on the device; the cut into row
// blocks and tiles, a sequential pass over the pair lengths, runs on the host.
grd_status build_tiers(grd_graph *g)
{
  if (g->tiers_built) return GRD_OK;
  grd_ctx *ctx = g->ctx;
  hipStream_t st = ctx->stream;
  const uint32_t P = g->P, M = (uint32_t)g->M, C = (uint32_t)g->C;
  const uint64_t m = g->m;
  const uint32_t tb = ctx->tier_bits ? ctx->tier_bits : kTierBitsMax;
  const uint32_t tile_arcs = ctx->tile_arcs ? ctx->tile_arcs : kTierTileArcs;
  if (tb < 4 || tb > kTierBitsMax) return fail(GRD_ERR_ARG, "tiled pull: tier bits must be 4..15");
  if (P >= kSole) return fail(GRD_ERR_ARG, "tiled pull: more than 2^31 proxies");
  // the partitions' proxy ranges
  std::vector<uint32_t> pb(M + 1);
  for (uint32_t k = 0; k < M; k++) pb[k] = g->obase[(size_t)k * C];
  pb[M] = P;
  uint32_t ntier_max = 1;
  for (uint32_t k = 0; k < M; k++) ntier_max = std::max<uint32_t>(ntier_max, ((pb[k + 1] - pb[k]) >> tb) + 1);
  if ((uint64_t)ntier_max * M >= (1ull << 31)) return fail(GRD_ERR_ARG, "tiled pull: too many tiers");
  DevMem tmp;
  uint32_t *d_pb = nullptr;
  GRD_TRY(t_upload(tmp, st, &d_pb, pb));
  // row of every arc: max-scan of the row heads
  uint32_t *head = nullptr, *arow = nullptr, *flag = nullptr, *pid = nullptr;
  GRD_TRY(tmp.alloc_n(&head, std::max<uint64_t>(m, 1)));
  GRD_TRY(tmp.alloc_n(&arow, std::max<uint64_t>(m, 1)));
  GRD_TRY(tmp.alloc_n(&flag, m + 1));
  GRD_TRY(tmp.alloc_n(&pid, m + 1));
  hipLaunchKernelGGL(k_t_fill, dim3(tgrid(m)), dim3(kBlock), 0, st, head, m, 0u);
  hipLaunchKernelGGL(k_t_heads, dim3(tgrid(P)), dim3(kBlock), 0, st, P, g->crp, head);
  if (m) {
    size_t tbytes = 0;
    GRD_HIP(rocprim::inclusive_scan(nullptr, tbytes, head, arow, (size_t)m, MaxU32(), st));
    void *t = nullptr;
    GRD_TRY(tmp.alloc(&t, std::max<size_t>(tbytes, 1)));
    GRD_HIP(rocprim::inclusive_scan(t, tbytes, head, arow, (size_t)m, MaxU32(), st));
  }
  hipLaunchKernelGGL(k_t_fill, dim3(1), dim3(kBlock), 0, st, flag + m, 1, 0u);
  hipLaunchKernelGGL(k_t_flags, dim3(tgrid(m)), dim3(kBlock), 0, st, m, arow, g->crp, g->ccol, d_pb, M, tb, flag);
  GRD_HIP(hipGetLastError());
  GRD_TRY(t_scan(tmp, st, flag, pid, m + 1));
  uint32_t np32 = 0;
  GRD_HIP(hipMemcpyAsync(&np32, pid + m, 4, hipMemcpyDeviceToHost, st));
  GRD_HIP(hipStreamSynchronize(st));
  const uint64_t np = np32;
  // pair records, sorted stably by (partition, tier)
  uint32_t *prow = nullptr, *pe0 = nullptr, *pkey = nullptr, *pval = nullptr;
  GRD_TRY(tmp.alloc_n(&prow, std::max<uint64_t>(np, 1)));
  GRD_TRY(tmp.alloc_n(&pe0, std::max<uint64_t>(np, 1)));
  GRD_TRY(tmp.alloc_n(&pkey, std::max<uint64_t>(np, 1)));
  GRD_TRY(tmp.alloc_n(&pval, std::max<uint64_t>(np, 1)));
  hipLaunchKernelGGL(k_t_pairs, dim3(tgrid(m)), dim3(kBlock), 0, st, m, flag, pid, arow, g->ccol, d_pb, M, tb,
                     ntier_max, prow, pe0, pkey, pval);
  DevMem &gm = g->tmem;
  GRD_TRY(gm.alloc_n(&g->trptr, (uint64_t)P + 1));
  hipLaunchKernelGGL(k_t_rptr, dim3(tgrid((uint64_t)P + 1)), dim3(kBlock), 0, st, P, g->crp, pid, g->trptr);
  GRD_HIP(hipGetLastError());
  if (np) {
    uint32_t *k2 = nullptr, *v2 = nullptr;
    GRD_TRY(tmp.alloc_n(&k2, np));
    GRD_TRY(tmp.alloc_n(&v2, np));
    rocprim::double_buffer<uint32_t> kb(pkey, k2), vb(pval, v2);
    int end_bit = 1;
    while (end_bit < 32 && ((uint64_t)ntier_max * M >> end_bit)) end_bit++;
    size_t tbytes = 0;
    GRD_HIP(rocprim::radix_sort_pairs(nullptr, tbytes, kb, vb, (size_t)np, 0, end_bit, st));
    void *t = nullptr;
    GRD_TRY(tmp.alloc(&t, std::max<size_t>(tbytes, 1)));
    GRD_HIP(rocprim::radix_sort_pairs(t, tbytes, kb, vb, (size_t)np, 0, end_bit, st));
    pkey = kb.current();
    pval = vb.current();
  }
  uint32_t *tlen = nullptr;
  GRD_TRY(tmp.alloc_n(&tlen, np + 1));
  GRD_TRY(gm.alloc_n(&g->trow, std::max<uint64_t>(np, 1)));
  GRD_TRY(gm.alloc_n(&g->tinv, std::max<uint64_t>(np, 1)));
  GRD_TRY(gm.alloc_n(&g->tptr, np + 1));
  GRD_TRY(gm.alloc_n(&g->tcol, std::max<uint64_t>(m, 1)));
  hipLaunchKernelGGL(k_t_fill, dim3(1), dim3(kBlock), 0, st, tlen + np, 1, 0u);
  hipLaunchKernelGGL(k_t_order, dim3(tgrid(np)), dim3(kBlock), 0, st, np, pval, prow, pe0, g->crp, g->trptr, tlen,
                     g->trow, g->tinv);
  GRD_HIP(hipGetLastError());
  GRD_TRY(t_scan(tmp, st, tlen, g->tptr, np + 1));
  hipLaunchKernelGGL(k_t_copy, dim3(tgrid(m)), dim3(kBlock), 0, st, m, flag, pid, pe0, prow, g->tinv, g->tptr, g->ccol,
                     d_pb, M, tb, g->tcol);
  GRD_HIP(hipGetLastError());
  // host: row blocks within each (partition, tier) group, then tiles
  std::vector<uint32_t> tptr(np + 1), key(np);
  GRD_HIP(hipMemcpyAsync(tptr.data(), g->tptr, (np + 1) * 4, hipMemcpyDeviceToHost, st));
  if (np) GRD_HIP(hipMemcpyAsync(key.data(), pkey, np * 4, hipMemcpyDeviceToHost, st));
  GRD_HIP(hipStreamSynchronize(st));
  std::vector<PullBlock> blks;
  std::vector<PullLong> longs;
  std::vector<TierTile> tiles;
  uint64_t q = 0;
  while (q < np) {
    uint64_t g_end = q;
    while (g_end < np && key[g_end] == key[q]) g_end++;
    const uint32_t mp = key[q] / ntier_max, t = key[q] % ntier_max;
    const uint32_t wbase = pb[mp] + (t << tb);
    const uint32_t wlen = std::min<uint32_t>(1u << tb, pb[mp + 1] - wbase);
    TierTile cur{(uint32_t)blks.size(), 0, wbase, wlen};
    uint64_t cur_arcs = 0;
    auto close_tile = [&]() {
      cur.nb = (uint32_t)blks.size() - cur.b0;
      if (cur.nb) tiles.push_back(cur);
      cur = TierTile{(uint32_t)blks.size(), 0, wbase, wlen};
      cur_arcs = 0;
    };
    while (q < g_end) {
      const uint32_t len = tptr[q + 1] - tptr[q];
      uint64_t add_arcs = 0;
      const size_t before = blks.size();
      if (len > kPullNnz) {
        PullLong L{(uint32_t)q, (uint32_t)blks.size(), 0, 0};
        for (uint32_t e = tptr[q]; e < tptr[q + 1]; e += kPullNnz) {
          blks.push_back(PullBlock{(uint32_t)q, 0, e, std::min(tptr[q + 1], e + kPullNnz)});
          L.nseg++;
        }
        longs.push_back(L);
        add_arcs = len;
        q++;
      } else {
        const uint64_t q0 = q;
        const uint32_t e0 = tptr[q];
        while (q < g_end && q - q0 < kPullMaxRows && tptr[q + 1] - tptr[q] <= kPullNnz &&
               tptr[q + 1] - e0 <= kPullNnz)
          q++;
        blks.push_back(PullBlock{(uint32_t)q0, (uint32_t)(q - q0), e0, tptr[q]});
        add_arcs = tptr[q] - e0;
      }
      cur_arcs += add_arcs;
      (void)before;
      if (cur_arcs >= tile_arcs) close_tile();
    }
    close_tile();
  }
  if (blks.size() >= 0x7FFFFFFFull || tiles.size() >= 0x7FFFFFFFull) return fail(GRD_ERR_ARG, "tiled pull: too many blocks");
  g->n_tblk = (uint32_t)blks.size();
  g->n_tlong = (uint32_t)longs.size();
  g->n_ttile = (uint32_t)tiles.size();
  GRD_TRY(t_upload(gm, st, &g->tblk, blks));
  GRD_TRY(t_upload(gm, st, &g->tlong, longs));
  GRD_TRY(t_upload(gm, st, &g->ttile, tiles));
  GRD_HIP(hipStreamSynchronize(st));  // the host vectors and temporaries go out of scope
  g->n_pairs = np;
  g->tier_bits = tb;
  g->tile_arcs = tile_arcs;
  g->tiers_built = true;
  return GRD_OK;
}

grd_status prp_pull_tiled(grd_state *s)
{
  grd_graph *g = s->g;
  hipStream_t st = g->ctx->stream;
  float *out = (float *)s->valR;
  if (g->n_ttile)
    hipLaunchKernelGGL(k_prp_pull_tiled, dim3(g->n_ttile), dim3(kTierBlock), 0, st, g->ttile, g->tblk, g->tptr,
                       g->trow, g->tcol, (const float *)s->mprop, out, s->tpart, s->tseg);
  if (g->n_tlong)
    hipLaunchKernelGGL(k_prp_tier_fixup, dim3(tgrid(g->n_tlong)), dim3(kBlock), 0, st, g->tlong, g->n_tlong, s->tseg,
                       g->trow, out, s->tpart);
  hipLaunchKernelGGL(k_prp_tier_merge, dim3(tgrid(g->P)), dim3(kBlock), 0, st, g->P, g->trptr, g->tinv, s->tpart,
                     out);
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

}  // namespace grd
