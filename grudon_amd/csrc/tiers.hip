// tiers.hip -- the source-blocked pulls: gathers served from LDS or L2.
//
// The pull sums, for every proxy row r of a memory partition, the values of
// its in-arcs' source proxies.  The sources are the partition's own proxies,
// numbered in out-degree order, and a random 4-byte gather from a table of
// tens of MB is served by L2 at best and by a whole cache line from the
// Infinity Cache / HBM at worst.  Three reorganisations of the CSC, built
// once per graph when its first state is created (grd_set_pull_tiling):
//
// mode 3, the hot / cold split (the default, for every pull): a row's arcs
// from its partition's 2^b hottest sources (b <= 15) are a prefix of the row,
// because rows keep their sources ascending.  The hot prefixes, with 16-bit
// offsets into the partition's window, are reduced from a 128 KB LDS window
// into one 4-byte partial per row (k_pull_hot, under an accumulation policy:
// FP32 Kahan for pull PageRank, the BSP operator for CC / SSSP / reference
// PageRank); the cold rest goes through the row-block kernel of pull.hip,
// which merges the row's partial first.  Arc weights (weighted SSSP) follow
// their arcs, one byte each when every weight is below 256.
//
// modes 1 and 2, the (partition, tier, row) regrouping (pull PageRank only):
// the partition's proxy range is cut into tiers of 2^tier_bits consecutive
// proxies, and the arcs of the CSC are regrouped by (memory partition, tier,
// row):
//
//   pair q       = one (row, tier) with at least one arc; its arcs are the
//                  contiguous run of the row's CSC arcs whose sources lie in
//                  the tier
//   tcol[e]      = the source's 16-bit offset within its tier (mode 1), or
//   tcol32[e]    = the source proxy (mode 2)
//   tile         = a run of row blocks of one (partition, tier) group
//
// Mode 1 (LDS tiles, tiers <= 2^15): a 1024-thread workgroup takes a tile and
// stages the tier's window of mprop (<= 128 KB) into LDS; then its four
// quarters walk the tile's row blocks four at a time (8 coalesced column
// loads and 8 LDS gathers per thread, a group of lanes per pair, FP32 Kahan
// per lane merged by TwoSum).  Mode 2 (L2 segments, tiers <= 2^24): the
// row-block kernel over the regrouped CSC, blocks ordered by segment so that
// a segment's values stay in L2.  Each pair yields one compensated partial; a
// row with a single pair writes its value straight into the reduce buffer,
// the others are merged per row in tier order (k_prp_tier_merge), so the
// result does not depend on scheduling.  A pair longer than one block is cut
// into segments merged in order by a fix-up pass.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <vector>

#include "grd_device.hpp"
#include "grd_internal.hpp"
#include "grd_ops.hpp"

namespace grd {

namespace {

constexpr int kTierBlock = 1024;                  // threads per tile workgroup
constexpr int kTierQuarters = kTierBlock / kBlock;  // row blocks per round, one per 256 threads
constexpr uint32_t kTNnz = kPullNnz;              // arcs per row block (8 per thread)
constexpr uint32_t kTRows = kPullMaxRows;         // pairs per row block (at most 4 per thread)
constexpr uint32_t kWideMerge = 32;               // rows with more pairs are merged by a wave
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
__global__ void k_w_max(uint64_t n, const uint32_t *__restrict__ w, uint32_t *__restrict__ out)
{
  uint32_t mx = 0;
  TIER_LOOP(i, n) mx = max(mx, w[i]);
  for (int o = kWave / 2; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, o, kWave));
  if (lane_id() == 0 && mx) atomicMax(out, mx);
}

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

// per pair in tier order: length, destination and the inverse permutation.
// The destination of a row's only pair is the row (its value goes straight
// to the reduce buffer); a pair of a row with several is written to its
// (row, tier)-order index, so that each row's partials are contiguous.
__global__ void k_t_order(uint64_t np, const uint32_t *__restrict__ perm, const uint32_t *__restrict__ prow,
                          const uint32_t *__restrict__ pe0, const uint32_t *__restrict__ crp,
                          const uint32_t *__restrict__ trptr, uint32_t *__restrict__ tlen, uint2 *__restrict__ trec,
                          uint32_t *__restrict__ tinv)
{
  TIER_LOOP(i, np)
  {
    const uint32_t p = perm[i], r = prow[p];
    uint32_t end = crp[r + 1];
    if (p + 1 < np && pe0[p + 1] < end) end = pe0[p + 1];
    tlen[i] = end - pe0[p];
    trec[i].y = (trptr[r + 1] - trptr[r] == 1) ? (r | kSole) : p;
    tinv[p] = (uint32_t)i;
  }
}

__global__ void k_t_starts(uint64_t np, const uint32_t *__restrict__ tptr, uint2 *__restrict__ trec)
{
  TIER_LOOP(i, np + 1) trec[i].x = tptr[i];
}

// every arc to its place in tier order: as a 16-bit offset within its tier
// (tcol, mode 1) or as its source proxy (tcol32, mode 2)
__global__ void k_t_copy(uint64_t m, const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pid,
                         const uint32_t *__restrict__ pe0, const uint32_t *__restrict__ prow,
                         const uint32_t *__restrict__ tinv, const uint32_t *__restrict__ tptr,
                         const uint32_t *__restrict__ ccol, const uint32_t *__restrict__ pb, uint32_t M, uint32_t tb,
                         uint16_t *__restrict__ tcol, uint32_t *__restrict__ tcol32)
{
  TIER_LOOP(e, m)
  {
    const uint32_t p = pid[e] + flag[e] - 1;  // the pair holding arc e
    const uint32_t at = tptr[tinv[p]] + ((uint32_t)e - pe0[p]);
    if (tcol32) {
      tcol32[at] = ccol[e];
    } else {
      const uint32_t base = pb[part_of(pb, M, prow[p])];
      tcol[at] = (uint16_t)((ccol[e] - base) & ((1u << tb) - 1u));
    }
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

// One 1024-thread workgroup per tile.  The tier window is staged with 16-byte
// loads, all issued before the first LDS store.  Then the four 256-thread
// quarters take four row blocks per round (at most kTNnz arcs and kTRows
// pairs each): every thread issues its 8 column loads and the records of the
// up to 4 pairs it will sum, gathers its arcs' values from the window into
// the quarter's slice of s_v, and after a barrier groups of G lanes sum the
// pairs.  LDS: 128 KB window + 4 x 8 KB slices = 160 KB.
__global__ void __launch_bounds__(kTierBlock) k_prp_pull_tiled(const TierTile *__restrict__ tiles,
                                                               const PullBlock *__restrict__ blocks,
                                                               const uint2 *__restrict__ trec,
                                                               const uint16_t *__restrict__ tcol,
                                                               const float *__restrict__ mprop,
                                                               float *__restrict__ out, float2 *__restrict__ part,
                                                               float2 *__restrict__ seg)
{
  __shared__ float4 s_win4[(1u << kTierBitsMax) / 4];
  __shared__ float s_v[kTierQuarters][kTNnz];
  const TierTile T = tiles[blockIdx.x];
  {
    // wlen is a multiple of 16 and wbase of 64: whole, aligned float4s
    const float4 *__restrict__ src = reinterpret_cast<const float4 *>(mprop + T.wbase);
    const uint32_t n4 = T.wlen >> 2;
    constexpr int kStage = (1u << kTierBitsMax) / 4 / kTierBlock;
    float4 x[kStage];
#pragma unroll
    for (int k = 0; k < kStage; k++) {
      const uint32_t i = threadIdx.x + k * kTierBlock;
      x[k] = src[i < n4 ? i : 0];  // unconditional loads keep x in registers
    }
#pragma unroll
    for (int k = 0; k < kStage; k++) {
      const uint32_t i = threadIdx.x + k * kTierBlock;
      if (i < n4) s_win4[i] = x[k];
    }
  }
  __syncthreads();
  const float *s_win = reinterpret_cast<const float *>(s_win4);
  const uint32_t qt = threadIdx.x / kBlock, lt = threadIdx.x % kBlock;
  float *sv = s_v[qt];
  const uint32_t *starts = reinterpret_cast<const uint32_t *>(trec);
  constexpr int kPer = kTNnz / kBlock, kRnd = kTRows / kBlock;
  const uint32_t rounds = (T.nb + kTierQuarters - 1) / kTierQuarters;
  for (uint32_t rd = 0; rd < rounds; rd++) {
    const uint32_t bi = rd * kTierQuarters + qt;
    const bool valid = bi < T.nb;
    const PullBlock d = valid ? blocks[T.b0 + bi] : PullBlock{0, 0, 0, 0};
    const uint32_t ne = d.e1 - d.e0;
    uint32_t G = 64;
    while (G > 1 && G * d.nrows > (uint32_t)kBlock) G >>= 1;
    const uint32_t sub = lt % G, ngroups = kBlock / G;
    const uint32_t nr = (d.nrows + ngroups - 1) / ngroups;
    uint32_t ps[kRnd], pe[kRnd], pd[kRnd];
    // the block's column loads, all in flight before the first gather
    uint32_t c[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint32_t i = lt + k * kBlock;
      c[k] = (valid && i < ne) ? tcol[d.e0 + i] : 0u;
    }
    if (valid && d.nrows == 0) {
      // one segment (<= kTNnz arcs) of a long pair: a compensated sum, each
      // thread over its arcs in order, the wave sums parked in the slice
      Kahan k;
#pragma unroll
      for (int j = 0; j < kPer; j++)
        if (lt + j * kBlock < ne) k.add(s_win[c[j]]);
      kahan_wave_merge(k, kWave);
      if (lane_id() == 0) {
        sv[2 * (lt / kWave)] = k.s;
        sv[2 * (lt / kWave) + 1] = k.c;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kRnd; k++) {
        const uint32_t ql = k * ngroups + lt / G;
        const bool ok = valid && (uint32_t)k < nr && ql < d.nrows;
        const uint2 r = ok ? trec[d.r0 + ql] : make_uint2(0u, 0u);
        ps[k] = r.x;
        pd[k] = r.y;
        pe[k] = ok ? starts[2 * (d.r0 + ql + 1)] : 0u;
      }
#pragma unroll
      for (int k = 0; k < kPer; k++) {
        const uint32_t i = lt + k * kBlock;
        if (valid && i < ne) sv[i] = s_win[c[k]];
      }
    }
    __syncthreads();
    if (valid && d.nrows == 0) {
      if (lt == 0) {
        Kahan t;
        t.s = sv[0];
        t.c = sv[1];
        for (int j = 1; j < kBlock / kWave; j++) t.merge(sv[2 * j], sv[2 * j + 1]);
        seg[T.b0 + bi] = make_float2(t.s, t.c);
      }
    } else if (valid) {
#pragma unroll
      for (int k = 0; k < kRnd; k++) {
        if ((uint32_t)k < nr) {  // uniform over the quarter
          const uint32_t ql = k * ngroups + lt / G;
          Kahan acc;
          if (ql < d.nrows)
            for (uint32_t x = ps[k] - d.e0 + sub; x < pe[k] - d.e0; x += G) acc.add(sv[x]);
          kahan_wave_merge(acc, (int)G);
          if (sub == 0 && ql < d.nrows) {
            if (pd[k] & kSole) out[pd[k] & ~kSole] = acc.value();
            else part[pd[k]] = make_float2(acc.s, acc.c);
          }
        }
      }
    }
    __syncthreads();
  }
}

// ---- the segmented pull (mode 2) ----
//
// Tiers of 2^19 sources by default: a segment's values (2 MB) stay in L2
// while its row blocks run, the blocks being ordered by segment.  The
// blocks are the row-block pull's (at most kPullNnz arcs and kPullMaxRows
// pairs, 8 column loads then 8 gathers per thread, a group of lanes per
// pair); the hottest kPullHot sources of the block's segment are staged in
// LDS.
__global__ void __launch_bounds__(kBlock) k_prp_pull_seg(const PullBlock *__restrict__ blocks,
                                                         const uint32_t *__restrict__ hot_base,
                                                         const uint2 *__restrict__ trec,
                                                         const uint32_t *__restrict__ tcol32,
                                                         const float *__restrict__ mprop, uint32_t P,
                                                         float *__restrict__ out, float2 *__restrict__ part,
                                                         float2 *__restrict__ seg)
{
  __shared__ float s_v[kPullNnz];
  __shared__ float s_hot[kPullHot];
  __shared__ float s_ws[kWavesPerBlock], s_wc[kWavesPerBlock];
  const uint32_t bid = blockIdx.x;
  const PullBlock d = blocks[bid];
  if (d.nrows == 0) {  // one segment (<= kPullNnz arcs) of a long pair
    constexpr int kPer = kPullNnz / kBlock;
    const uint32_t ne = d.e1 - d.e0;
    uint32_t src[kPer];
    float v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint32_t i = threadIdx.x + k * kBlock;
      src[k] = i < ne ? tcol32[d.e0 + i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kPer; k++) v[k] = threadIdx.x + k * kBlock < ne ? mprop[src[k]] : 0.0f;
    Kahan k;
#pragma unroll
    for (int j = 0; j < kPer; j++)
      if (threadIdx.x + j * kBlock < ne) k.add(v[j]);
    kahan_wave_merge(k, kWave);
    if (lane_id() == 0) {
      s_ws[wave_in_block()] = k.s;
      s_wc[wave_in_block()] = k.c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      Kahan t;
      t.s = s_ws[0];
      t.c = s_wc[0];
      for (int j = 1; j < kWavesPerBlock; j++) t.merge(s_ws[j], s_wc[j]);
      seg[bid] = make_float2(t.s, t.c);
    }
    return;
  }
  const uint32_t ne = d.e1 - d.e0;
  {
    constexpr int kPer = kPullNnz / kBlock;
    uint32_t src[kPer];
    float v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint32_t i = threadIdx.x + k * kBlock;
      src[k] = i < ne ? tcol32[d.e0 + i] : 0u;
    }
    const uint32_t hb = hot_base[bid];
    const uint32_t hn = min(kPullHot, P - hb);
    for (uint32_t i = threadIdx.x; i < hn; i += kBlock) s_hot[i] = mprop[hb + i];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint32_t i = threadIdx.x + k * kBlock;
      const uint32_t h = src[k] - hb;
      v[k] = i < ne ? (h < hn ? s_hot[h] : mprop[src[k]]) : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint32_t i = threadIdx.x + k * kBlock;
      if (i < ne) s_v[i] = v[k];
    }
  }
  __syncthreads();
  int G = 64;
  while (G > 1 && (uint32_t)G * d.nrows > (uint32_t)kBlock) G >>= 1;
  const uint32_t sub = threadIdx.x % G, ngroups = kBlock / G;
  const uint32_t rounds = (d.nrows + ngroups - 1) / ngroups;
  const uint32_t *starts = reinterpret_cast<const uint32_t *>(trec);
  for (uint32_t k = 0; k < rounds; k++) {
    const uint32_t ql = k * ngroups + threadIdx.x / G;
    const uint32_t q = d.r0 + ql;
    Kahan acc;
    uint32_t dst = 0;
    if (ql < d.nrows) {
      const uint2 r = trec[q];
      dst = r.y;
      const uint32_t b = r.x - d.e0, e = starts[2 * (q + 1)] - d.e0;
      for (uint32_t x = b + sub; x < e; x += G) acc.add(s_v[x]);
    }
    kahan_wave_merge(acc, G);
    if (sub == 0 && ql < d.nrows) {
      if (dst & kSole) out[dst & ~kSole] = acc.value();
      else part[dst] = make_float2(acc.s, acc.c);
    }
  }
}

// long pairs: a wave each, lane j taking segments j, j + 64, ... and the
// lanes merged by a fixed butterfly, so the order is fixed
__global__ void __launch_bounds__(kBlock) k_prp_tier_fixup(const PullLong *__restrict__ longs, uint32_t nlong,
                                                           const float2 *__restrict__ seg, const uint2 *__restrict__ trec,
                                                           float *__restrict__ out, float2 *__restrict__ part)
{
  const uint32_t lane = lane_id();
  for (uint32_t i = blockIdx.x * kWavesPerBlock + wave_in_block(); i < nlong; i += gridDim.x * kWavesPerBlock) {
    const PullLong L = longs[i];
    Kahan t;
    for (uint32_t j = lane; j < L.nseg; j += kWave) {
      const float2 p = seg[L.first + j];
      t.merge(p.x, p.y);
    }
    kahan_wave_merge(t, kWave);
    if (lane == 0) {
      const uint32_t dst = trec[L.row].y;
      if (dst & kSole) out[dst & ~kSole] = t.value();
      else part[dst] = make_float2(t.s, t.c);
    }
  }
}

// rows with no pair (0) or 2..kWideMerge pairs: a thread each, over the
// row's partials (contiguous, in tier order)
__global__ void __launch_bounds__(kBlock) k_prp_tier_merge(uint32_t P, const uint32_t *__restrict__ trptr,
                                                           const float2 *__restrict__ part, float *__restrict__ out)
{
  for (uint32_t r = blockIdx.x * kBlock + threadIdx.x; r < P; r += gridDim.x * kBlock) {
    const uint32_t a = trptr[r], b = trptr[r + 1];
    if (b - a == 1 || b - a > kWideMerge) continue;
    Kahan t;
    uint32_t x = a;
    for (; x + 4 <= b; x += 4) {
      const float2 p0 = part[x], p1 = part[x + 1], p2 = part[x + 2], p3 = part[x + 3];
      t.merge(p0.x, p0.y);
      t.merge(p1.x, p1.y);
      t.merge(p2.x, p2.y);
      t.merge(p3.x, p3.y);
    }
    for (; x < b; x++) {
      const float2 p = part[x];
      t.merge(p.x, p.y);
    }
    out[r] = t.value();
  }
}

// rows with more than kWideMerge pairs (the hubs): a wave each, as the
// long pairs
__global__ void __launch_bounds__(kBlock) k_prp_tier_merge_wide(const uint32_t *__restrict__ rows, uint32_t nrows,
                                                                const uint32_t *__restrict__ trptr,
                                                                const float2 *__restrict__ part, float *__restrict__ out)
{
  const uint32_t lane = lane_id();
  for (uint32_t i = blockIdx.x * kWavesPerBlock + wave_in_block(); i < nrows; i += gridDim.x * kWavesPerBlock) {
    const uint32_t r = rows[i], a = trptr[r], b = trptr[r + 1];
    Kahan t;
    for (uint32_t x = a + lane; x < b; x += kWave) {
      const float2 p = part[x];
      t.merge(p.x, p.y);
    }
    kahan_wave_merge(t, kWave);
    if (lane == 0) out[r] = t.value();
  }
}


// ---- mode 3: the hot / cold split ----
//
// A row's arcs from its memory partition's 2^b hottest sources (the first
// proxies of the partition, in out-degree order; b <= 15) are a prefix of
// the row, since rows keep their sources ascending.  1024-thread workgroups,
// each with one partition's window of those sources in LDS, sum the
// prefixes into one Kahan partial per row, written in row order.  The rest
// of each row goes to the row-block kernel over a CSC of the remaining arcs,
// which merges the row's partial first.

// hot and cold arc counts of every row: the first source at or beyond the
// partition's window end, by binary search over the sorted row
__global__ void k_h_split(uint32_t P, const uint32_t *__restrict__ crp, const uint32_t *__restrict__ ccol,
                          const uint32_t *__restrict__ pb, uint32_t M, uint32_t win, uint32_t *__restrict__ hlen,
                          uint32_t *__restrict__ clen)
{
  TIER_LOOP(r, P)
  {
    const uint32_t mp = part_of(pb, M, (uint32_t)r);
    const uint32_t lim = pb[mp] + min(win, pb[mp + 1] - pb[mp]);
    const uint32_t b = crp[r], e = crp[r + 1];
    uint32_t lo = b, hi = e;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (ccol[mid] < lim) lo = mid + 1;
      else hi = mid;
    }
    hlen[r] = lo - b;
    clen[r] = e - lo;
  }
}

// a wave per row: its hot arcs as 16-bit window offsets, its cold arcs as
// source proxies, and their weights as W (a byte when every weight fits)
template <class W>
__global__ void k_h_copy(uint32_t P, const uint32_t *__restrict__ crp, const uint32_t *__restrict__ ccol,
                         const uint32_t *__restrict__ cw, const uint32_t *__restrict__ hrp,
                         const uint32_t *__restrict__ corp, const uint32_t *__restrict__ pb, uint32_t M,
                         uint16_t *__restrict__ hcol, uint32_t *__restrict__ cocol, W *__restrict__ hw,
                         W *__restrict__ cow)
{
  const uint32_t lane = lane_id();
  const uint64_t stride = (uint64_t)gridDim.x * kWavesPerBlock;
  for (uint64_t r = (uint64_t)blockIdx.x * kWavesPerBlock + wave_in_block(); r < P; r += stride) {
    const uint32_t base = pb[part_of(pb, M, (uint32_t)r)];
    const uint32_t e0 = crp[r], n = crp[r + 1] - e0;
    const uint32_t h0 = hrp[r], nh = hrp[r + 1] - h0, c0 = corp[r];
    for (uint32_t i = lane; i < n; i += kWave) {
      const uint32_t v = ccol[e0 + i];
      if (i < nh) {
        hcol[h0 + i] = (uint16_t)(v - base);
        if (cw) hw[h0 + i] = (W)cw[e0 + i];
      } else {
        cocol[c0 + (i - nh)] = v;
        if (cw) cow[c0 + (i - nh)] = (W)cw[e0 + i];
      }
    }
  }
}

// Wave blocks of the hot prefixes: at most kWNnz arcs and kWRows rows each
// (a row with more hot arcs is cut into segments of kWNnz arcs).  Workgroup
// w owns the blocks [T.b0, T.b0 + T.nb) of one partition (T = ranges[w]):
// it stages that partition's window once, then its 16 waves take the blocks
// one at a time from a counter in LDS and work on their own -- column loads
// and row bounds, gathers from the window into the wave's LDS slice, a group
// of G lanes per row -- with no workgroup barrier, so that the waves hide
// each other's latency.  The next block's loads are in flight while the
// current one is gathered and summed.  LDS: 128 KB window + 16 slices of 504
// values + the block counter.
constexpr uint32_t kWNnz = 504;   // arcs per wave block (at most 8 per lane; 8 slots short of 512 leave
                                  // LDS room for the block counter)
constexpr uint32_t kWRows = 256;  // rows per wave block (at most 4 per lane)
constexpr int kRsrcFlags = 0x00020000;  // raw buffer resource, word 3, on gfx9

__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// a raw buffer over [base, base + bytes): loads past the end return 0, so the
// lanes beyond a block's arcs or rows need no clamped index
__device__ __forceinline__ __amdgpu_buffer_rsrc_t block_buffer(const void *base, uint32_t bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, kRsrcFlags);
}

// group width for a block of n rows: the largest power of two G with G n <= 64
__device__ __forceinline__ uint32_t hot_group(uint32_t n)
{
  if (n <= 1) return 64;
  if (n > 64) return 1;
  return 64u >> (32 - __builtin_clz(n - 1));  // 64 >> ceil(log2 n)
}

// How the hot pass accumulates.  Pull PageRank: FP32 Kahan per lane, lanes
// merged by TwoSum, the compensated sum of each row rounded once to FP32.  The BSP pull (CC,
// SSSP, reference PageRank): Op::edge of the source's broadcast value and the
// arc weight, combined with the operator (min, or + in a fixed order).
// The hot part of a row is a few arcs on average (2.7 at RMAT-26, M = 4), so
// its compensated sum is kept as one FP32 value (the sum rounded once): the
// per-row partial is 4 bytes written and read instead of 8.
struct HotPR {
  using V = float;
  using Part = float;
  using Acc = Kahan;
  static constexpr bool kWeighted = false;
  __device__ static Acc zero() { return Kahan{}; }
  __device__ static V value(V v, uint32_t) { return v; }
  __device__ static void add(Acc &a, V x) { a.add(x); }
  __device__ static void merge_lanes(Acc &a, int width) { kahan_wave_merge(a, width); }
  __device__ static void add_part(Acc &a, Part p) { a.merge(p, 0.0f); }
  __device__ static Part part(const Acc &a) { return a.value(); }
};
template <class Op>
struct HotBsp {
  using V = typename Op::V;
  using Part = V;
  using Acc = V;
  static constexpr bool kWeighted = std::is_same<Op, OpWeighted>::value;
  __device__ static Acc zero() { return op_identity<Op>(); }
  __device__ static V value(V v, uint32_t w) { return Op::edge(v, w); }
  __device__ static void add(Acc &a, V x) { a = op_combine<Op>(a, x); }
  __device__ static void merge_lanes(Acc &a, int width)
  {
    for (int o = width / 2; o > 0; o >>= 1) a = op_combine<Op>(a, __shfl_xor(a, o, kWave));
  }
  __device__ static void add_part(Acc &a, Part p) { a = op_combine<Op>(a, p); }
  __device__ static Part part(const Acc &a) { return a; }
};

// kSkip (CC / SSSP): skip[b] = 1 marks a wave block whose rows are all done
// or without arcs (k_blocks_done; 2, some rows done, runs in full); it is
// taken as an empty segment (no loads, the identity written to its hseg slot)
// and its rows' partials are not written, which the cold pass never reads for
// such rows.
template <class Pol, class W = uint32_t, bool kSkip = false>
__global__ void __launch_bounds__(kTierBlock) k_pull_hot(const TierTile *__restrict__ ranges,
                                                         const PullBlock *__restrict__ blocks,
                                                         const uint32_t *__restrict__ hrp,
                                                         const uint16_t *__restrict__ hcol,
                                                         const W *__restrict__ hw,
                                                         const typename Pol::V *__restrict__ mprop,
                                                         typename Pol::Part *__restrict__ hpart,
                                                         typename Pol::Part *__restrict__ hseg,
                                                         const uint8_t *__restrict__ skip)
{
  using V = typename Pol::V;
  static_assert(sizeof(V) == 4, "the window holds 4-byte values");
  constexpr uint32_t kWaves = kTierBlock / kWave;
  constexpr int kPer = (kWNnz + kWave - 1) / kWave, kRows = kWRows / kWave, kStage = kHotWin / 4 / kTierBlock;
  __shared__ uint4 s_win4[kHotWin / 4];
  __shared__ V s_v[kWaves][kWNnz];
  __shared__ uint32_t s_next;  // the range's next block not yet taken by a wave
  const TierTile T = ranges[blockIdx.x];
  if (threadIdx.x == 0) s_next = T.b0 + kWaves;
  {
    // wlen is a multiple of 4 and wbase of 64: whole, aligned 16-byte units
    const uint4 *__restrict__ src = reinterpret_cast<const uint4 *>(mprop + T.wbase);
    const uint32_t n4 = T.wlen >> 2;
    uint4 x[kStage];
#pragma unroll
    for (int k = 0; k < kStage; k++) {
      const uint32_t i = threadIdx.x + k * kTierBlock;
      x[k] = src[i < n4 ? i : 0];  // all loads in flight, then the stores
    }
    // every slot is written (those past the window with a value no arc
    // reads), so that no store waits on a branch
#pragma unroll
    for (int k = 0; k < kStage; k++) s_win4[threadIdx.x + k * kTierBlock] = x[k];
  }
  __syncthreads();
  const V *s_win = reinterpret_cast<const V *>(s_win4);
  // the wave index made uniform, so that block descriptors are scalar loads
  const uint32_t w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = lane_id();
  V *sv = s_v[w];
  const uint32_t bend = T.b0 + T.nb;
  // the global loads of a block: its columns (and weights), and the lanes'
  // row bounds, each through a buffer over the block's own range
  auto issue = [&](const PullBlock &d, uint32_t (&c)[kPer], uint32_t (&wt)[kPer], uint32_t (&rb)[kRows],
                   uint32_t (&re)[kRows]) {
    const uint32_t ne = d.e1 - d.e0;
    const uint32_t G = hot_group(d.nrows), ngroups = kWave / G;
    const auto bc = block_buffer(hcol + d.e0, ne * 2);
#pragma unroll
    for (int k = 0; k < kPer; k++) c[k] = __builtin_amdgcn_raw_buffer_load_b16(bc, (lane + k * kWave) * 2, 0, 0);
    if constexpr (Pol::kWeighted) {
      const auto bw = block_buffer(hw + d.e0, ne * (uint32_t)sizeof(W));
#pragma unroll
      for (int k = 0; k < kPer; k++) {
        if constexpr (sizeof(W) == 1) wt[k] = __builtin_amdgcn_raw_buffer_load_b8(bw, lane + k * kWave, 0, 0);
        else wt[k] = __builtin_amdgcn_raw_buffer_load_b32(bw, (lane + k * kWave) * 4, 0, 0);
      }
    } else {
#pragma unroll
      for (int k = 0; k < kPer; k++) wt[k] = 1u;
    }
    const auto br = block_buffer(hrp + d.r0, (d.nrows + 1) * 4);
#pragma unroll
    for (int k = 0; k < kRows; k++) {
      const uint32_t rl = k * ngroups + lane / G;
      rb[k] = __builtin_amdgcn_raw_buffer_load_b32(br, rl * 4, 0, 0);
      re[k] = __builtin_amdgcn_raw_buffer_load_b32(br, rl * 4 + 4, 0, 0);
    }
  };
  // a wave takes its blocks from the range's counter, each two blocks ahead
  // of its use (blocks are independent, so the order does not change any
  // result); descriptors come through a vector load, so that waiting for one
  // does not wait for the LDS, and read back into scalars one block later.
  // A descriptor past the range reads as zeros: no rows, no arcs.
  auto take = [&]() {
    uint32_t x = 0;
    if (lane == 0) x = atomicAdd(&s_next, 1u);
    return __builtin_amdgcn_readfirstlane(x);
  };
  const auto bdesc = block_buffer(blocks, bend * (uint32_t)sizeof(PullBlock));
  auto load_desc = [&](uint32_t i) {
    return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(bdesc, i * sizeof(PullBlock), 0, 0));
  };
  auto to_block = [](const uint4 &q) {
    return PullBlock{(uint32_t)__builtin_amdgcn_readfirstlane(q.x), (uint32_t)__builtin_amdgcn_readfirstlane(q.y),
                     (uint32_t)__builtin_amdgcn_readfirstlane(q.z), (uint32_t)__builtin_amdgcn_readfirstlane(q.w)};
  };
  // the skip flags travel with the descriptors (a flag past the range reads 0)
  const auto bflag = block_buffer(skip, kSkip ? bend : 0u);
  auto load_flag = [&](uint32_t i) { return kSkip ? (uint32_t)__builtin_amdgcn_raw_buffer_load_b8(bflag, i, 0, 0) : 0u; };
  auto apply_flag = [](PullBlock x, uint32_t f) {
    if (kSkip && __builtin_amdgcn_readfirstlane(f) == 1u) x = PullBlock{x.r0, 0, x.e0, x.e0};
    return x;
  };
  uint32_t b = T.b0 + w;
  PullBlock d = apply_flag(to_block(load_desc(b)), load_flag(b));
  uint32_t b1 = take();
  uint4 q1 = load_desc(b1);
  uint32_t f1 = load_flag(b1);
  uint32_t c[kPer], wt[kPer], rb[kRows], re[kRows];
  issue(d, c, wt, rb, re);
  while (b < bend) {
    const uint32_t b2 = take();
    const uint4 q2 = load_desc(b2);
    const uint32_t f2 = load_flag(b2);
    const uint32_t ne = d.e1 - d.e0;
    V v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) v[k] = Pol::value(s_win[c[k]], wt[k]);
    const PullBlock dn = apply_flag(to_block(q1), f1);
    uint32_t cn[kPer], wtn[kPer], rbn[kRows], ren[kRows];
    issue(dn, cn, wtn, rbn, ren);
    if (d.nrows == 0) {
      // a segment of a long row: each lane over its arcs in order, then the
      // wave's lanes merged in a fixed order
      auto a = Pol::zero();
#pragma unroll
      for (int j = 0; j < kPer; j++)
        if (lane + j * kWave < ne) Pol::add(a, v[j]);
      Pol::merge_lanes(a, kWave);
      if (lane == 0) hseg[b] = Pol::part(a);
    } else {
#pragma unroll
      for (int k = 0; k < kPer; k++) {
        const uint32_t i = lane + k * kWave;
        if (i < ne) sv[i] = v[k];
      }
      wave_lds_sync();
      const uint32_t G = hot_group(d.nrows);
      const uint32_t sub = lane % G, ngroups = kWave / G;
      const uint32_t nr = (d.nrows + ngroups - 1) / ngroups;
#pragma unroll
      for (int k = 0; k < kRows; k++) {
        if ((uint32_t)k < nr) {  // uniform over the wave
          const uint32_t rl = k * ngroups + lane / G;
          auto a = Pol::zero();
          if (rl < d.nrows)
            for (uint32_t x = rb[k] - d.e0 + sub; x < re[k] - d.e0; x += G) Pol::add(a, sv[x]);
          Pol::merge_lanes(a, (int)G);
          if (sub == 0 && rl < d.nrows) hpart[d.r0 + rl] = Pol::part(a);
        }
      }
      wave_lds_sync();  // the slice is rewritten by the next block
    }
    b = b1;
    d = dn;
    b1 = b2;
    q1 = q2;
    f1 = f2;
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      c[k] = cn[k];
      wt[k] = wtn[k];
    }
#pragma unroll
    for (int k = 0; k < kRows; k++) {
      rb[k] = rbn[k];
      re[k] = ren[k];
    }
  }
}

// long hot prefixes: kFixLanes lanes each (a wave holds kWave / kFixLanes
// rows), lane j of a row taking segments j, j + kFixLanes, ... and the
// row's lanes merged in a fixed order; the loop is uniform over the wave
constexpr uint32_t kFixLanes = 8;
template <class Pol>
__global__ void __launch_bounds__(kBlock) k_hot_fixup(const PullLong *__restrict__ longs, uint32_t nlong,
                                                      const typename Pol::Part *__restrict__ hseg,
                                                      typename Pol::Part *__restrict__ hpart)
{
  constexpr uint32_t kRowsPerWave = kWave / kFixLanes;
  const uint32_t lane = lane_id(), sub = lane % kFixLanes;
  const uint32_t stride = gridDim.x * kWavesPerBlock * kRowsPerWave;
  for (uint32_t i0 = (blockIdx.x * kWavesPerBlock + wave_in_block()) * kRowsPerWave; i0 < nlong; i0 += stride) {
    const uint32_t i = i0 + lane / kFixLanes;
    auto a = Pol::zero();
    PullLong L{0, 0, 0, 0};
    if (i < nlong) {
      L = longs[i];
      for (uint32_t j = sub; j < L.nseg; j += kFixLanes) Pol::add_part(a, hseg[L.first + j]);
    }
    Pol::merge_lanes(a, (int)kFixLanes);
    if (i < nlong && sub == 0) hpart[L.row] = Pol::part(a);
  }
}

}  // namespace

// mode 3: split every row at its partition's window end on the device; the
// cut into wave blocks and workgroup ranges runs on the host
static grd_status build_hotcold(grd_graph *g)
{
  grd_ctx *ctx = g->ctx;
  hipStream_t st = ctx->stream;
  const uint32_t P = g->P, M = (uint32_t)g->M, C = (uint32_t)g->C;
  const uint32_t tb = ctx->tier_bits ? ctx->tier_bits : kTierBitsMax;
  if (tb < 4 || tb > kTierBitsMax) return fail(GRD_ERR_ARG, "hot / cold pull: window bits out of range");
  const uint32_t win = 1u << tb;  // <= kHotWin
  std::vector<uint32_t> pb(M + 1);
  for (uint32_t k = 0; k < M; k++) pb[k] = g->obase[(size_t)k * C];
  pb[M] = P;
  DevMem tmp;
  uint32_t *d_pb = nullptr, *hlen = nullptr, *clen = nullptr;
  GRD_TRY(t_upload(tmp, st, &d_pb, pb));
  GRD_TRY(tmp.alloc_n(&hlen, (uint64_t)P + 1));
  GRD_TRY(tmp.alloc_n(&clen, (uint64_t)P + 1));
  hipLaunchKernelGGL(k_t_fill, dim3(1), dim3(kBlock), 0, st, hlen + P, 1, 0u);
  hipLaunchKernelGGL(k_t_fill, dim3(1), dim3(kBlock), 0, st, clen + P, 1, 0u);
  if (P) hipLaunchKernelGGL(k_h_split, dim3(tgrid(P)), dim3(kBlock), 0, st, P, g->crp, g->ccol, d_pb, M, win, hlen, clen);
  GRD_HIP(hipGetLastError());
  DevMem &gm = g->tmem;
  GRD_TRY(gm.alloc_n(&g->hrp, (uint64_t)P + 1));
  GRD_TRY(gm.alloc_n(&g->corp, (uint64_t)P + 1));
  GRD_TRY(t_scan(tmp, st, hlen, g->hrp, (uint64_t)P + 1));
  GRD_TRY(t_scan(tmp, st, clen, g->corp, (uint64_t)P + 1));
  std::vector<uint32_t> hrp((size_t)P + 1), corp((size_t)P + 1);
  GRD_HIP(hipMemcpyAsync(hrp.data(), g->hrp, ((size_t)P + 1) * 4, hipMemcpyDeviceToHost, st));
  GRD_HIP(hipMemcpyAsync(corp.data(), g->corp, ((size_t)P + 1) * 4, hipMemcpyDeviceToHost, st));
  GRD_HIP(hipStreamSynchronize(st));
  // padded: a wave block without hot arcs still reads the column at its start
  GRD_TRY(gm.alloc_n(&g->tcol, (uint64_t)hrp[P] + kWave));
  GRD_TRY(gm.alloc_n(&g->cocol, std::max<uint64_t>(corp[P], 1)));
  // arc weights follow their arcs (the BSP pull of weighted SSSP), one byte
  // each when every weight is below 256
  bool w8 = false;
  if (g->cw) {
    uint32_t *d_wmax = nullptr, wmax = 0;
    GRD_TRY(tmp.alloc_n(&d_wmax, 1));
    hipLaunchKernelGGL(k_t_fill, dim3(1), dim3(kBlock), 0, st, d_wmax, 1, 0u);
    hipLaunchKernelGGL(k_w_max, dim3(tgrid(g->m)), dim3(kBlock), 0, st, g->m, g->cw, d_wmax);
    GRD_HIP(hipMemcpyAsync(&wmax, d_wmax, 4, hipMemcpyDeviceToHost, st));
    GRD_HIP(hipStreamSynchronize(st));
    w8 = wmax < 256;
    if (w8) {
      GRD_TRY(gm.alloc_n(&g->hw8, (uint64_t)hrp[P] + kWave));
      GRD_TRY(gm.alloc_n(&g->cow8, std::max<uint64_t>(corp[P], 1)));
    } else {
      GRD_TRY(gm.alloc_n(&g->hw, (uint64_t)hrp[P] + kWave));
      GRD_TRY(gm.alloc_n(&g->cow, std::max<uint64_t>(corp[P], 1)));
    }
  }
  if (P) {
    const uint64_t nb = std::min<uint64_t>(65535, ((uint64_t)P + kWavesPerBlock - 1) / kWavesPerBlock);
    if (w8)
      hipLaunchKernelGGL(k_h_copy<uint8_t>, dim3((uint32_t)nb), dim3(kBlock), 0, st, P, g->crp, g->ccol, g->cw,
                         g->hrp, g->corp, d_pb, M, g->tcol, g->cocol, g->hw8, g->cow8);
    else
      hipLaunchKernelGGL(k_h_copy<uint32_t>, dim3((uint32_t)nb), dim3(kBlock), 0, st, P, g->crp, g->ccol, g->cw,
                         g->hrp, g->corp, d_pb, M, g->tcol, g->cocol, g->hw, g->cow);
  }
  GRD_HIP(hipGetLastError());
  // hot: wave blocks per partition, then each partition's blocks split into
  // contiguous workgroup ranges of about equal cost, the number of ranges of
  // a partition in proportion to its cost
  std::vector<PullBlock> hblk;
  std::vector<PullLong> hlong;
  std::vector<uint32_t> pblk0(M + 1, 0);
  std::vector<uint64_t> bcost;
  for (uint32_t mp = 0; mp < M; mp++) {
    pblk0[mp] = (uint32_t)hblk.size();
    uint32_t r = pb[mp];
    while (r < pb[mp + 1]) {
      const uint32_t len = hrp[r + 1] - hrp[r];
      if (len > kWNnz) {
        PullLong L{r, (uint32_t)hblk.size(), 0, 0};
        for (uint32_t e = hrp[r]; e < hrp[r + 1]; e += kWNnz) {
          hblk.push_back(PullBlock{r, 0, e, std::min(hrp[r + 1], e + kWNnz)});
          L.nseg++;
        }
        hlong.push_back(L);
        r++;
        continue;
      }
      const uint32_t r0 = r, e0 = hrp[r];
      while (r < pb[mp + 1] && r - r0 < kWRows && hrp[r + 1] - hrp[r] <= kWNnz && hrp[r + 1] - e0 <= kWNnz) r++;
      hblk.push_back(PullBlock{r0, r - r0, e0, hrp[r]});
    }
  }
  pblk0[M] = (uint32_t)hblk.size();
  bcost.reserve(hblk.size());
  uint64_t total = 0;
  for (const PullBlock &d : hblk) {
    bcost.push_back(64 + (d.e1 - d.e0) + 4ull * d.nrows);  // a block's fixed cost: its loads' latency
    total += bcost.back();
  }
  std::vector<TierTile> ranges;
  const uint64_t nwg = 16ull * (uint64_t)std::max(1, ctx->num_cus);  // sixteen per CU, in turn: dynamic balance
  for (uint32_t mp = 0; mp < M; mp++) {
    const uint32_t b0 = pblk0[mp], b1 = pblk0[mp + 1];
    if (b0 == b1) continue;
    const uint32_t wbase = pb[mp], wlen = std::min<uint32_t>(win, pb[mp + 1] - pb[mp]);
    uint64_t cm = 0;
    for (uint32_t i = b0; i < b1; i++) cm += bcost[i];
    const uint64_t k = std::max<uint64_t>(1, std::min<uint64_t>(b1 - b0, cm * nwg / std::max<uint64_t>(total, 1)));
    uint64_t acc = 0;
    uint32_t start = b0, j = 1;
    for (uint32_t i = b0; i < b1 && j < k; i++) {
      acc += bcost[i];
      if (acc * k >= cm * j) {
        ranges.push_back(TierTile{start, i + 1 - start, wbase, wlen});
        start = i + 1;
        j++;
      }
    }
    if (start < b1) ranges.push_back(TierTile{start, b1 - start, wbase, wlen});
  }
  // cold: the row-block pull's cut over the remaining arcs, without the
  // kernel's small LDS window (a window base of P is empty): the next 2048
  // sources after the hot window draw too few arcs to pay for staging them
  std::vector<PullBlock> cblk;
  std::vector<PullLong> clong;
  std::vector<uint32_t> chot;
  {
    uint32_t r = 0;
    auto window_after = [&](uint32_t) { return P; };
    while (r < P) {
      const uint32_t len = corp[r + 1] - corp[r];
      if (len > kPullNnz) {
        PullLong L{r, (uint32_t)cblk.size(), 0, 0};
        for (uint32_t e = corp[r]; e < corp[r + 1]; e += kPullNnz) {
          cblk.push_back(PullBlock{r, 0, e, std::min(corp[r + 1], e + kPullNnz)});
          chot.push_back(window_after(r));
          L.nseg++;
        }
        clong.push_back(L);
        r++;
        continue;
      }
      const uint32_t r0 = r, e0 = corp[r];
      while (r < P && r - r0 < kPullMaxRows && corp[r + 1] - corp[r] <= kPullNnz && corp[r + 1] - e0 <= kPullNnz) r++;
      cblk.push_back(PullBlock{r0, r - r0, e0, corp[r]});
      chot.push_back(window_after(r0));
    }
  }
  // the hot kernel reads its descriptors through a buffer of at most 2^31 bytes
  if (hblk.size() * sizeof(PullBlock) >= 0x7FFFFFFFull || cblk.size() >= 0x7FFFFFFFull)
    return fail(GRD_ERR_ARG, "hot / cold pull: too many row blocks");
  GRD_TRY(t_upload(gm, st, &g->tblk, hblk));
  GRD_TRY(t_upload(gm, st, &g->tlong, hlong));
  GRD_TRY(t_upload(gm, st, &g->ttile, ranges));
  GRD_TRY(t_upload(gm, st, &g->cblk, cblk));
  GRD_TRY(t_upload(gm, st, &g->cblk_hot, chot));
  GRD_TRY(t_upload(gm, st, &g->clong, clong));
  GRD_HIP(hipStreamSynchronize(st));  // the host vectors and temporaries go out of scope
  g->n_tblk = (uint32_t)hblk.size();
  g->n_tlong = (uint32_t)hlong.size();
  g->n_ttile = (uint32_t)ranges.size();
  g->n_cblk = (uint32_t)cblk.size();
  g->n_clong = (uint32_t)clong.size();
  g->n_pairs = P;
  g->tier_mode = 3;
  g->tier_bits = tb;
  g->tile_arcs = 0;
  g->tiers_built = true;
  return GRD_OK;
}

// Regroups the CSC by (partition, tier, row) on the device; the cut into row
// blocks and tiles, a sequential pass over the pair lengths, runs on the host.
grd_status build_tiers(grd_graph *g)
{
  if (g->tiers_built) return GRD_OK;
  if (g->ctx->pull_mode == 3) return build_hotcold(g);
  grd_ctx *ctx = g->ctx;
  hipStream_t st = ctx->stream;
  const uint32_t P = g->P, M = (uint32_t)g->M, C = (uint32_t)g->C;
  const uint64_t m = g->m;
  const int mode = ctx->pull_mode;
  if (mode != 1 && mode != 2) return fail(GRD_ERR_ARG, "tiled pull: mode 1 (LDS tiles), 2 (L2 segments) or 3 (hot / cold)");
  const uint32_t tb = ctx->tier_bits ? ctx->tier_bits : (mode == 1 ? kTierBitsMax : kSegBits);
  const uint32_t tile_arcs = ctx->tile_arcs ? ctx->tile_arcs : kTierTileArcs;
  if (tb < 4 || tb > (mode == 1 ? kTierBitsMax : kSegBitsMax)) return fail(GRD_ERR_ARG, "tiled pull: tier bits out of range");
  // row blocks: a wave's (mode 1) or a workgroup's (mode 2)
  const uint32_t blk_nnz = mode == 1 ? kTNnz : kPullNnz, blk_rows = mode == 1 ? kTRows : kPullMaxRows;
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
  uint32_t *tlen = nullptr, *tinv = nullptr, *d_tptr = nullptr;
  GRD_TRY(tmp.alloc_n(&tlen, np + 1));
  GRD_TRY(tmp.alloc_n(&tinv, std::max<uint64_t>(np, 1)));
  GRD_TRY(tmp.alloc_n(&d_tptr, np + 1));
  GRD_TRY(gm.alloc_n(&g->trec, np + 1));
  if (mode == 1) GRD_TRY(gm.alloc_n(&g->tcol, std::max<uint64_t>(m, 1)));
  else GRD_TRY(gm.alloc_n(&g->tcol32, std::max<uint64_t>(m, 1)));
  hipLaunchKernelGGL(k_t_fill, dim3(1), dim3(kBlock), 0, st, tlen + np, 1, 0u);
  hipLaunchKernelGGL(k_t_order, dim3(tgrid(np)), dim3(kBlock), 0, st, np, pval, prow, pe0, g->crp, g->trptr, tlen,
                     g->trec, tinv);
  GRD_HIP(hipGetLastError());
  GRD_TRY(t_scan(tmp, st, tlen, d_tptr, np + 1));
  hipLaunchKernelGGL(k_t_starts, dim3(tgrid(np + 1)), dim3(kBlock), 0, st, np, d_tptr, g->trec);
  hipLaunchKernelGGL(k_t_copy, dim3(tgrid(m)), dim3(kBlock), 0, st, m, flag, pid, pe0, prow, tinv, d_tptr, g->ccol,
                     d_pb, M, tb, g->tcol, g->tcol32);
  GRD_HIP(hipGetLastError());
  // host: row blocks within each (partition, tier) group, then tiles
  std::vector<uint32_t> tptr(np + 1), key(np);
  GRD_HIP(hipMemcpyAsync(tptr.data(), d_tptr, (np + 1) * 4, hipMemcpyDeviceToHost, st));
  if (np) GRD_HIP(hipMemcpyAsync(key.data(), pkey, np * 4, hipMemcpyDeviceToHost, st));
  GRD_HIP(hipStreamSynchronize(st));
  std::vector<PullBlock> blks;
  std::vector<PullLong> longs;
  std::vector<TierTile> tiles;
  std::vector<uint32_t> hot;  // mode 2: the window base of each block's tier
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
      if (len > blk_nnz) {
        PullLong L{(uint32_t)q, (uint32_t)blks.size(), 0, 0};
        for (uint32_t e = tptr[q]; e < tptr[q + 1]; e += blk_nnz) {
          blks.push_back(PullBlock{(uint32_t)q, 0, e, std::min(tptr[q + 1], e + blk_nnz)});
          L.nseg++;
        }
        longs.push_back(L);
        add_arcs = len;
        q++;
      } else {
        const uint64_t q0 = q;
        const uint32_t e0 = tptr[q];
        while (q < g_end && q - q0 < blk_rows && tptr[q + 1] - tptr[q] <= blk_nnz && tptr[q + 1] - e0 <= blk_nnz)
          q++;
        blks.push_back(PullBlock{(uint32_t)q0, (uint32_t)(q - q0), e0, tptr[q]});
        add_arcs = tptr[q] - e0;
      }
      cur_arcs += add_arcs;
      while (hot.size() < blks.size()) hot.push_back(wbase);
      (void)before;
      if (mode == 1 && cur_arcs >= tile_arcs) close_tile();
    }
    if (mode == 1) close_tile();
  }
  if (blks.size() >= 0x7FFFFFFFull || tiles.size() >= 0x7FFFFFFFull) return fail(GRD_ERR_ARG, "tiled pull: too many blocks");
  // rows whose partials a wave merges
  std::vector<uint32_t> trptr((size_t)P + 1), wide;
  GRD_HIP(hipMemcpyAsync(trptr.data(), g->trptr, ((size_t)P + 1) * 4, hipMemcpyDeviceToHost, st));
  GRD_HIP(hipStreamSynchronize(st));
  for (uint32_t r = 0; r < P; r++)
    if (trptr[r + 1] - trptr[r] > kWideMerge) wide.push_back(r);
  g->n_twide = (uint32_t)wide.size();
  GRD_TRY(t_upload(gm, st, &g->twide, wide));
  g->n_tblk = (uint32_t)blks.size();
  g->n_tlong = (uint32_t)longs.size();
  g->n_ttile = (uint32_t)tiles.size();
  GRD_TRY(t_upload(gm, st, &g->tblk, blks));
  GRD_TRY(t_upload(gm, st, &g->tlong, longs));
  GRD_TRY(t_upload(gm, st, &g->ttile, tiles));
  if (mode == 2) GRD_TRY(t_upload(gm, st, &g->tblk_hot, hot));
  GRD_HIP(hipStreamSynchronize(st));  // the host vectors and temporaries go out of scope
  g->n_pairs = np;
  g->tier_mode = mode;
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
  auto wave_grid = [](uint64_t n) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(65535, (n + kWavesPerBlock - 1) / kWavesPerBlock));
  };
  if (g->tier_mode == 3) {
    // hot prefixes from LDS into per-row partials, then the cold rest
    float *hpart = reinterpret_cast<float *>(s->tpart), *hseg = reinterpret_cast<float *>(s->tseg);
    if (g->n_ttile)
      hipLaunchKernelGGL(k_pull_hot<HotPR>, dim3(g->n_ttile), dim3(kTierBlock), 0, st, g->ttile, g->tblk, g->hrp,
                         g->tcol, (const uint32_t *)nullptr, (const float *)s->mprop, hpart, hseg,
                         (const uint8_t *)nullptr);
    if (g->n_tlong)
      hipLaunchKernelGGL(k_hot_fixup<HotPR>, dim3(wave_grid((g->n_tlong + 7) / 8)), dim3(kBlock), 0, st, g->tlong, g->n_tlong,
                         (const float *)hseg, hpart);
    GRD_HIP(hipGetLastError());
    GRD_TRY(prp_pull_rows(st, g->cblk, g->n_cblk, g->cblk_hot, g->corp, g->cocol, (const float *)s->mprop, g->P, out,
                          s->cseg, hpart));
    return prp_pull_rows_fixup(st, g->clong, g->n_clong, s->cseg, out, hpart);
  }
  if (g->tier_mode == 1 && g->n_ttile)
    hipLaunchKernelGGL(k_prp_pull_tiled, dim3(g->n_ttile), dim3(kTierBlock), 0, st, g->ttile, g->tblk, g->trec,
                       g->tcol, (const float *)s->mprop, out, s->tpart, s->tseg);
  if (g->tier_mode == 2 && g->n_tblk)
    hipLaunchKernelGGL(k_prp_pull_seg, dim3(g->n_tblk), dim3(kBlock), 0, st, g->tblk, g->tblk_hot, g->trec, g->tcol32,
                       (const float *)s->mprop, g->P, out, s->tpart, s->tseg);
  if (g->n_tlong)
    hipLaunchKernelGGL(k_prp_tier_fixup, dim3(wave_grid(g->n_tlong)), dim3(kBlock), 0, st, g->tlong, g->n_tlong,
                       s->tseg, g->trec, out, s->tpart);
  hipLaunchKernelGGL(k_prp_tier_merge, dim3(tgrid(g->P)), dim3(kBlock), 0, st, g->P, g->trptr, s->tpart, out);
  if (g->n_twide)
    hipLaunchKernelGGL(k_prp_tier_merge_wide, dim3(wave_grid(g->n_twide)), dim3(kBlock), 0, st, g->twide, g->n_twide,
                       g->trptr, s->tpart, out);
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

// the hot pass of the BSP pull over a graph split by mode 3: one partial per
// row (the operator's identity for a row without hot arcs) into s->tpart
template <class Op>
grd_status bsp_pull_hot(grd_state *s)
{
  grd_graph *g = s->g;
  hipStream_t st = g->ctx->stream;
  using V = typename Op::V;
  V *hpart = reinterpret_cast<V *>(s->tpart), *hseg = reinterpret_cast<V *>(s->tseg);
  if (g->n_ttile) {
    const uint32_t *hw = std::is_same<Op, OpWeighted>::value ? g->hw : nullptr;
    const bool w8 = std::is_same<Op, OpWeighted>::value && g->hw8;
    if (s->hskip_on && w8)
      hipLaunchKernelGGL((k_pull_hot<HotBsp<Op>, uint8_t, true>), dim3(g->n_ttile), dim3(kTierBlock), 0, st, g->ttile,
                         g->tblk, g->hrp, g->tcol, (const uint8_t *)g->hw8, (const V *)s->mprop, hpart, hseg,
                         (const uint8_t *)s->hskip);
    else if (s->hskip_on)
      hipLaunchKernelGGL((k_pull_hot<HotBsp<Op>, uint32_t, true>), dim3(g->n_ttile), dim3(kTierBlock), 0, st, g->ttile,
                         g->tblk, g->hrp, g->tcol, hw, (const V *)s->mprop, hpart, hseg, (const uint8_t *)s->hskip);
    else if (w8)
      hipLaunchKernelGGL((k_pull_hot<HotBsp<Op>, uint8_t>), dim3(g->n_ttile), dim3(kTierBlock), 0, st, g->ttile,
                         g->tblk, g->hrp, g->tcol, (const uint8_t *)g->hw8, (const V *)s->mprop, hpart, hseg,
                         (const uint8_t *)nullptr);
    else
      hipLaunchKernelGGL((k_pull_hot<HotBsp<Op>, uint32_t>), dim3(g->n_ttile), dim3(kTierBlock), 0, st, g->ttile,
                         g->tblk, g->hrp, g->tcol, hw, (const V *)s->mprop, hpart, hseg, (const uint8_t *)nullptr);
  }
  if (g->n_tlong) {
    const uint64_t waves = (g->n_tlong + kWave / kFixLanes - 1) / (kWave / kFixLanes);
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(65535, (waves + kWavesPerBlock - 1) / kWavesPerBlock));
    hipLaunchKernelGGL((k_hot_fixup<HotBsp<Op>>), dim3(grid), dim3(kBlock), 0, st, g->tlong, g->n_tlong, hseg, hpart);
  }
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}
template grd_status bsp_pull_hot<OpCC>(grd_state *);
template grd_status bsp_pull_hot<OpHop>(grd_state *);
template grd_status bsp_pull_hot<OpWeighted>(grd_state *);
template grd_status bsp_pull_hot<OpPR>(grd_state *);

}  // namespace grd
