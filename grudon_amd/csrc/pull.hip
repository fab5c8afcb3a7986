// pull.hip -- edge-balanced pull over the memory partitions' CSC.
//
// The CSC rows (one per memory proxy: own vertices and cross mirrors) are cut
// at graph-build time into row blocks holding at most kPullNnz arcs
// (adaptive CSR row blocks).  A 256-thread workgroup takes one block:
//   1. stage: every thread walks the block's arcs with stride 256 (coalesced
//      column reads), gathers the source value and stores the per-arc
//      contribution in LDS;
//   2. reduce: the block's rows are assigned to groups of G lanes, G being the
//      largest power of two with G * rows <= 256, and each group reduces its
//      rows from LDS.
// A row with more than kPullNnz arcs is split into segments of kPullNnz arcs,
// one workgroup each; segment partials are merged in segment order by a
// fix-up pass, so results do not depend on scheduling.
//
// PageRank pull sums in FP32 with Kahan compensation (per lane, merged with
// TwoSum); the BSP pull (CC / SSSP / reference PageRank) combines the
// contributions of the sources active this iteration with min or +.
#include <algorithm>

#include "grd_device.hpp"
#include "grd_internal.hpp"
#include "grd_ops.hpp"

namespace grd {

// per-block LDS bitmap of dirty rows (rows of one block span <= 17 words)
constexpr int kDirtyWords = kPullMaxRows / 64 + 2;

// workgroups of a wave-per-row fix-up pass
static inline uint32_t fixup_grid(uint32_t nlong)
{
  return std::max<uint32_t>(1, std::min<uint32_t>(65535, (nlong + kWavesPerBlock - 1) / kWavesPerBlock));
}

__device__ __forceinline__ int group_size(uint32_t nrows)
{
  int g = 64;
  while (g > 1 && (uint32_t)g * nrows > (uint32_t)kBlock) g >>= 1;
  return g;
}

// ---- PageRank pull: Kahan FP32 ----

__global__ void __launch_bounds__(kBlock) k_prp_pull_rb(const PullBlock *__restrict__ blocks,
                                                        const uint32_t *__restrict__ hot_base,
                                                        const uint32_t *__restrict__ crp,
                                                        const uint32_t *__restrict__ ccol,
                                                        const float *__restrict__ mprop, uint32_t P,
                                                        float *__restrict__ out, float2 *__restrict__ partial,
                                                        const float *__restrict__ pre)
{
  __shared__ float s_v[kPullNnz];
  __shared__ float s_hot[kPullHot];
  __shared__ float s_ws[kWavesPerBlock], s_wc[kWavesPerBlock];
  const uint32_t bid = blockIdx.x;
  const PullBlock d = blocks[bid];
  if (d.nrows == 0) {  // segment (<= kPullNnz arcs) of a long row
    // as below: all column loads, then all gathers (hot sources from LDS),
    // then the compensated sum of each thread's arcs in order
    constexpr int kPer = kPullNnz / kBlock;
    const uint32_t ne = d.e1 - d.e0;
    uint32_t src[kPer];
    float v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint32_t i = threadIdx.x + k * kBlock;
      src[k] = i < ne ? ccol[d.e0 + i] : 0u;
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
      partial[bid] = make_float2(t.s, t.c);
    }
    return;
  }
  const uint32_t ne = d.e1 - d.e0;
  {
    // all column loads first, then all gathers: 8 independent requests in
    // flight per thread for each stage.  The values of the partition's
    // kPullHot highest-degree sources (the first own proxies) are staged in
    // LDS and gathered from there.
    constexpr int kPer = kPullNnz / kBlock;
    uint32_t src[kPer];
    float v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      uint32_t i = threadIdx.x + k * kBlock;
      src[k] = i < ne ? ccol[d.e0 + i] : 0u;
    }
    const uint32_t hb = hot_base[bid];
    const uint32_t hn = min(kPullHot, P - hb);
    for (uint32_t i = threadIdx.x; i < hn; i += kBlock) s_hot[i] = mprop[hb + i];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      uint32_t i = threadIdx.x + k * kBlock;
      const uint32_t h = src[k] - hb;
      v[k] = i < ne ? (h < hn ? s_hot[h] : mprop[src[k]]) : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      uint32_t i = threadIdx.x + k * kBlock;
      if (i < ne) s_v[i] = v[k];
    }
  }
  __syncthreads();
  const int G = group_size(d.nrows);
  const int sub = threadIdx.x % G;
  const uint32_t ngroups = kBlock / G;
  // every group runs the same number of rounds so that the group shuffles
  // stay in step
  const uint32_t rounds = (d.nrows + ngroups - 1) / ngroups;
  for (uint32_t k = 0; k < rounds; k++) {
    uint32_t rl = k * ngroups + threadIdx.x / G;
    Kahan acc;
    uint32_t r = d.r0 + rl;
    if (rl < d.nrows) {
      uint32_t b = crp[r] - d.e0, e = crp[r + 1] - d.e0;
      for (uint32_t x = b + sub; x < e; x += G) acc.add(s_v[x]);
    }
    kahan_wave_merge(acc, G);
    if (sub == 0 && rl < d.nrows) {
      if (pre) {  // the row's partial from the hot pass comes first
        Kahan t;
        t.s = pre[r];
        t.merge(acc.s, acc.c);
        out[r] = t.value();
      } else {
        out[r] = acc.value();
      }
    }
  }
}

// a wave per long row: lane j merges segments j, j + 64, ..., then the lanes
// are merged in a fixed order after the row's hot partial
__global__ void __launch_bounds__(kBlock) k_prp_pull_fixup(const PullLong *__restrict__ longs, uint32_t nlong,
                                                           const float2 *__restrict__ partial, float *__restrict__ out,
                                                           const float *__restrict__ pre)
{
  const uint32_t lane = lane_id();
  for (uint32_t i = blockIdx.x * kWavesPerBlock + wave_in_block(); i < nlong; i += gridDim.x * kWavesPerBlock) {
    const PullLong L = longs[i];
    Kahan a;
    for (uint32_t j = lane; j < L.nseg; j += kWave) {
      const float2 p = partial[L.first + j];
      a.merge(p.x, p.y);
    }
    kahan_wave_merge(a, kWave);
    if (lane == 0) {
      Kahan t;
      if (pre) t.s = pre[L.row];
      t.merge(a.s, a.c);
      out[L.row] = t.value();
    }
  }
}

// ---- BSP pull: min / + over the active sources ----
//
// The broadcast receive of a pull superstep writes every own proxy: active
// ones with their value, inactive ones with the combine identity (Op::edge
// keeps the identity), so one gather per arc gives the arc's contribution.

// Rows of a CC / SSSP pull that cannot improve (the broadcast's rows-done
// bitmap, k_rows_done): a row block whose rows with arcs are all done is left
// out whole (flag 1); one with some done rows (flag 2) gathers only the arcs
// of its other rows.  This pass sets the flags and counts the arcs and rows of
// the blocks left out (one atomic per workgroup), so that the pull reads one
// flag per block.
__global__ void __launch_bounds__(kBlock) k_blocks_done(uint32_t nblk, const PullBlock *__restrict__ blocks,
                                                        const uint64_t *__restrict__ done,
                                                        const uint64_t *__restrict__ hasarc,
                                                        uint8_t *__restrict__ bskip, Counters *cnt, int cold)
{
  unsigned long long arcs = 0, rows = 0;
  for (uint32_t b = blockIdx.x * kBlock + threadIdx.x; b < nblk; b += gridDim.x * kBlock) {
    const PullBlock d = blocks[b];
    bool all, any = false;
    if (d.nrows == 0) {
      all = (done[d.r0 >> 6] >> (d.r0 & 63)) & 1ull;
    } else {
      all = true;
      const uint32_t r1 = d.r0 + d.nrows;
      for (uint32_t w = d.r0 >> 6; (all || !any) && w <= (r1 - 1) >> 6; w++) {
        const uint32_t lo = max(d.r0, w * 64) - w * 64, hi = min(r1 - w * 64, 64u);
        const uint64_t mask = (hi - lo == 64) ? ~0ull : (((1ull << (hi - lo)) - 1ull) << lo);
        const uint64_t dn = done[w], ha = hasarc[w];
        all = all && ((dn | ~ha) & mask) == mask;
        any = any || (dn & ha & mask) != 0;
      }
    }
    bskip[b] = all ? 1 : any ? 2 : 0;
    if (all) {
      arcs += d.e1 - d.e0;
      rows += d.nrows;
    }
  }
  arcs = block_sum(arcs);
  rows = block_sum(rows);
  if (threadIdx.x == 0) {
    if (arcs) atomicAdd(cold ? &cnt->pull_skipped : &cnt->pull_skipped_hot, arcs);
    if (rows && cold) atomicAdd(&cnt->pull_skipped_rows, rows);
  }
}

// the smallest candidate a pull superstep can produce: the smallest value
// broadcast plus the smallest Op::edge increment, saturating
__device__ __forceinline__ uint32_t pull_bound(const Counters *cnt, uint32_t inc)
{
  const uint32_t f = cnt->fmin;
  return f + inc < f ? 0xFFFFFFFFu : f + inc;
}

// Measurement only (time_kernel): the arcs, in the row-block pull's CSC (rp:
// the cold CSC, or the whole one), of the rows that gathered none: done rows
// (done may be null) and rows whose hot partial already reached the bound
// (pre may be null); a thread per row, one atomic per workgroup.
__global__ void __launch_bounds__(kBlock) k_done_arcs(uint32_t nwP, const uint32_t *__restrict__ rp,
                                                      const uint64_t *__restrict__ done,
                                                      const uint32_t *__restrict__ pre, uint32_t inc, Counters *cnt)
{
  const uint32_t bnd = pre ? pull_bound(cnt, inc) : 0u;
  unsigned long long arcs = 0;
  const uint64_t nrow = (uint64_t)nwP * 64;
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < nrow; r += (uint64_t)gridDim.x * kBlock) {
    const bool none = (done && ((done[r >> 6] >> (r & 63)) & 1ull)) || (pre && pre[r] <= bnd);
    if (none) arcs += rp[r + 1] - rp[r];
  }
  arcs = block_sum(arcs);
  if (threadIdx.x == 0 && arcs) atomicAdd(&cnt->pull_done_arcs, arcs);
}

// bit r = row r has an arc (rp[r + 1] > rp[r]); a wave covers one word
__global__ void __launch_bounds__(kBlock) k_rows_with_arcs(uint32_t nwP, const uint32_t *__restrict__ rp,
                                                           uint64_t *__restrict__ bits)
{
  const uint64_t nrow = (uint64_t)nwP * 64;
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < nrow; r += (uint64_t)gridDim.x * kBlock) {
    const uint64_t m = __ballot(rp[r + 1] > rp[r]);
    if (lane_id() == 0) bits[r >> 6] = m;
  }
}

// apply the pulled value to the proxy's running update; true = dirty
template <class Op>
__device__ __forceinline__ bool pull_apply(typename Op::V *upd, uint32_t r, typename Op::V acc)
{
  if constexpr (Op::kMin) {
    if (acc < upd[r]) {
      upd[r] = acc;
      return true;
    }
    return false;
  } else {
    // contributions of active sources are strictly positive (pr.cpp:80)
    if (acc > 0.0f) {
      upd[r] = upd[r] + acc;
      return true;
    }
    return false;
  }
}

// W: the arc weights' storage type (a byte when every weight fits)
template <class Op, class W = uint32_t>
__global__ void __launch_bounds__(kBlock) k_bsp_pull_rb(const PullBlock *__restrict__ blocks,
                                                        const uint32_t *__restrict__ hot_base,
                                                        const uint32_t *__restrict__ crp,
                                                        const uint32_t *__restrict__ ccol,
                                                        const W *__restrict__ cw,
                                                        const typename Op::V *__restrict__ mprop, uint32_t P,
                                                        typename Op::V *__restrict__ upd, uint64_t *__restrict__ dirty,
                                                        typename Op::V *__restrict__ partial,
                                                        const typename Op::V *__restrict__ pre,
                                                        const uint8_t *__restrict__ bskip,
                                                        const uint64_t *__restrict__ done,
                                                        const uint64_t *__restrict__ hasarc,
                                                        const Counters *__restrict__ bcnt, uint32_t inc)
{
  using V = typename Op::V;
  __shared__ V s_v[kPullNnz];
  __shared__ V s_hot[kPullHot];
  __shared__ unsigned long long s_dirty[kDirtyWords];
  __shared__ uint8_t s_live[kPullNnz];  // a partially done block: 1 = the arc's row may improve
  const uint32_t bid = blockIdx.x;
  const PullBlock d = blocks[bid];
  const uint32_t flag = bskip ? bskip[bid] : 0u;  // k_blocks_done: 1 = left out, 2 = some rows done
  if (flag == 1) return;
  // with bcnt (a CC / SSSP pull over the hot / cold split): a row whose hot
  // partial is already at or below the bound takes no cold candidate below it
  // (every cold candidate is at least the bound), so it gathers nothing and
  // its result is the hot partial
  const uint32_t bnd = bcnt ? pull_bound(bcnt, inc) : 0u;
  if (d.nrows == 0) {  // segment (<= kPullNnz arcs) of a long row
    // column ids (and weights), then the gathers (hot sources from LDS),
    // then each thread combines its arcs in order
    constexpr int kPer = kPullNnz / kBlock;
    const uint32_t ne = d.e1 - d.e0;
    uint32_t src[kPer], wt[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint32_t i = threadIdx.x + k * kBlock;
      src[k] = i < ne ? ccol[d.e0 + i] : 0u;
      wt[k] = (cw && i < ne) ? (uint32_t)cw[d.e0 + i] : 1u;
    }
    if (bcnt && pre[d.r0] <= bnd) {  // uniform over the block
      if (threadIdx.x == 0) partial[bid] = op_identity<Op>();
      return;
    }
    const uint32_t hb = hot_base[bid];
    const uint32_t hn = min(kPullHot, P - hb);
    for (uint32_t i = threadIdx.x; i < hn; i += kBlock) s_hot[i] = mprop[hb + i];
    __syncthreads();
    V acc = op_identity<Op>();
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      const uint32_t i = threadIdx.x + k * kBlock;
      const uint32_t h = src[k] - hb;
      if (i < ne) acc = op_combine<Op>(acc, Op::edge(h < hn ? s_hot[h] : mprop[src[k]], wt[k]));
    }
    if constexpr (Op::kMin) acc = block_min(acc);
    else acc = block_sum(acc);
    if (threadIdx.x == 0) partial[bid] = acc;
    return;
  }
  for (int i = threadIdx.x; i < kDirtyWords; i += kBlock) s_dirty[i] = 0ull;
  const uint32_t ne = d.e1 - d.e0;
  const int G = group_size(d.nrows);
  const int sub = threadIdx.x % G;
  const uint32_t ngroups = kBlock / G;
  const uint32_t rounds = (d.nrows + ngroups - 1) / ngroups;
  const uint32_t w0 = d.r0 >> 6;
  {
    // batched stages: column ids (and weights), then the gathers; the
    // partition's kPullHot highest-degree sources are read from LDS
    constexpr int kPer = kPullNnz / kBlock;
    uint32_t src[kPer], wt[kPer];
    V v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      uint32_t i = threadIdx.x + k * kBlock;
      src[k] = i < ne ? ccol[d.e0 + i] : 0u;
      wt[k] = (cw && i < ne) ? (uint32_t)cw[d.e0 + i] : 1u;
    }
    const bool marked = flag == 2 || bcnt;  // uniform over the block
    if (marked) {
      // each row's group marks the row's arcs to gather or not (every arc of
      // the block belongs to one of its rows): not for a done row, nor for
      // one whose hot partial already reached the bound
      for (uint32_t k = 0; k < rounds; k++) {
        const uint32_t rl = k * ngroups + threadIdx.x / G, r = d.r0 + rl;
        bool gat = false;
        uint32_t b = 0, e = 0;
        if (rl < d.nrows) {
          const uint64_t bit = 1ull << (r & 63);
          gat = !done || (!(done[r >> 6] & bit) && (hasarc[r >> 6] & bit));
          if (gat && bcnt) gat = pre[r] > bnd;
          b = crp[r] - d.e0;
          e = crp[r + 1] - d.e0;
        }
        for (uint32_t x = b + sub; x < e; x += G) s_live[x] = gat ? 1 : 0;
      }
    }
    const uint32_t hb = hot_base[bid];
    const uint32_t hn = min(kPullHot, P - hb);
    for (uint32_t i = threadIdx.x; i < hn; i += kBlock) s_hot[i] = mprop[hb + i];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      uint32_t i = threadIdx.x + k * kBlock;
      const uint32_t h = src[k] - hb;
      const bool on = i < ne && (!marked || s_live[i]);
      v[k] = on ? Op::edge(h < hn ? s_hot[h] : mprop[src[k]], wt[k]) : op_identity<Op>();
    }
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      uint32_t i = threadIdx.x + k * kBlock;
      if (i < ne) s_v[i] = v[k];
    }
  }
  __syncthreads();
  for (uint32_t k = 0; k < rounds; k++) {
    uint32_t rl = k * ngroups + threadIdx.x / G;
    V acc = op_identity<Op>();
    uint32_t r = d.r0 + rl;
    // a done row, or one without arcs, has nothing to apply (its hot partial
    // may not have been written)
    bool live = rl < d.nrows;
    if (done && live) {
      const uint64_t bit = 1ull << (r & 63);
      live = !(done[r >> 6] & bit) && (hasarc[r >> 6] & bit);
    }
    if (live) {
      uint32_t b = crp[r] - d.e0, e = crp[r + 1] - d.e0;
      for (uint32_t x = b + sub; x < e; x += G) acc = op_combine<Op>(acc, s_v[x]);
    }
    for (int o = G / 2; o > 0; o >>= 1) acc = op_combine<Op>(acc, __shfl_xor(acc, o, kWave));
    if (pre && sub == 0 && live) acc = op_combine<Op>(pre[r], acc);  // the row's hot partial first
    if (sub == 0 && live && pull_apply<Op>(upd, r, acc))
      atomicOr(&s_dirty[(r >> 6) - w0], 1ull << (r & 63));
  }
  __syncthreads();
  const uint32_t nw = ((d.r0 + d.nrows - 1) >> 6) - w0 + 1;
  for (uint32_t i = threadIdx.x; i < nw; i += kBlock)
    if (s_dirty[i]) atomicOr((unsigned long long *)(dirty + w0 + i), s_dirty[i]);
}

template <class Op>
__global__ void __launch_bounds__(kBlock) k_bsp_pull_fixup(const PullLong *__restrict__ longs, uint32_t nlong,
                                                           const typename Op::V *__restrict__ partial,
                                                           typename Op::V *__restrict__ upd, uint64_t *__restrict__ dirty,
                                                           const typename Op::V *__restrict__ pre,
                                                           const uint64_t *__restrict__ done)
{
  // a wave per long row, as k_prp_pull_fixup; a done row's segments were
  // left out (k_blocks_done flags every segment of it)
  const uint32_t lane = lane_id();
  for (uint32_t i = blockIdx.x * kWavesPerBlock + wave_in_block(); i < nlong; i += gridDim.x * kWavesPerBlock) {
    const PullLong L = longs[i];
    if (done && ((done[L.row >> 6] >> (L.row & 63)) & 1ull)) continue;
    typename Op::V acc = op_identity<Op>();
    for (uint32_t j = lane; j < L.nseg; j += kWave) acc = op_combine<Op>(acc, partial[L.first + j]);
    for (int o = kWave / 2; o > 0; o >>= 1) acc = op_combine<Op>(acc, __shfl_xor(acc, o, kWave));
    if (lane == 0) {
      if (pre) acc = op_combine<Op>(pre[L.row], acc);  // the row's hot partial first
      if (pull_apply<Op>(upd, L.row, acc))
        atomicOr((unsigned long long *)(dirty + (L.row >> 6)), 1ull << (L.row & 63));
    }
  }
}

// ------------------------------------------------------------------------

grd_status prp_pull_launch(grd_state *s)
{
  grd_graph *g = s->g;
  hipStream_t st = g->ctx->stream;
  return prp_pull_rows(st, g->pblk, g->n_pblk, g->pblk_hot, g->crp, g->ccol, (const float *)s->mprop, g->P,
                       (float *)s->valR, (float2 *)s->partial, nullptr);
}

grd_status prp_pull_fixup(grd_state *s)
{
  grd_graph *g = s->g;
  return prp_pull_rows_fixup(g->ctx->stream, g->plong, g->n_plong, (const float2 *)s->partial, (float *)s->valR,
                             nullptr);
}

grd_status prp_pull_rows(hipStream_t st, const PullBlock *blocks, uint32_t nblk, const uint32_t *hot_base,
                         const uint32_t *rptr, const uint32_t *col, const float *mprop, uint32_t P, float *out,
                         float2 *partial, const float *pre)
{
  if (nblk)
    hipLaunchKernelGGL(k_prp_pull_rb, dim3(nblk), dim3(kBlock), 0, st, blocks, hot_base, rptr, col, mprop, P, out,
                       partial, pre);
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

grd_status prp_pull_rows_fixup(hipStream_t st, const PullLong *longs, uint32_t nlong, const float2 *partial,
                               float *out, const float *pre)
{
  if (!nlong) return GRD_OK;
  hipLaunchKernelGGL(k_prp_pull_fixup, dim3(fixup_grid(nlong)), dim3(kBlock), 0, st, longs, nlong,
                     partial, out, pre);
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

// A state over a graph split by the hot / cold pull (mode 3) runs the hot
// pass first and then the row-block kernel over the cold arcs, which combines
// each row's hot partial with its own; otherwise the row-block kernel over
// the whole CSC.
// CC / SSSP after a broadcast that sent the rows-done bitmap: flag the row
// blocks (of the cold CSC, or of the whole CSC) that the pull leaves out
template <class Op>
static grd_status blocks_done(grd_state *s, bool cold, const uint8_t **flags)
{
  *flags = nullptr;
  s->hskip_on = false;
  if (!Op::kMin || !s->bound_sent) return GRD_OK;
  s->skip_pending = true;  // the rows that gathered nothing come off with the counters
  if (!s->rdone_sent) return GRD_OK;
  grd_graph *g = s->g;
  hipStream_t st = g->ctx->stream;
  const uint32_t nwP = g->P / 64;
  // rows with arcs in the whole CSC: in the hot / cold split a row without
  // cold arcs still combines its hot partial in the cold pass
  if (!g->rarc) {
    GRD_TRY(g->mem.alloc_n(&g->rarc, std::max<uint32_t>(nwP, 1)));
    if (nwP)
      hipLaunchKernelGGL(k_rows_with_arcs, dim3(std::min<uint32_t>((nwP + kWavesPerBlock - 1) / kWavesPerBlock,
                                                                   (uint32_t)g->ctx->num_cus * 8)),
                         dim3(kBlock), 0, st, nwP, g->crp, g->rarc);
  }
  const uint32_t cap = (uint32_t)g->ctx->num_cus * 8;
  const uint32_t nblk = cold ? g->n_cblk : g->n_pblk;
  if (nblk)
    hipLaunchKernelGGL(k_blocks_done, dim3(std::min<uint32_t>((nblk + kBlock - 1) / kBlock, cap)), dim3(kBlock), 0, st,
                       nblk, cold ? g->cblk : g->pblk, (const uint64_t *)s->Rdone, (const uint64_t *)g->rarc, s->bskip,
                       s->cnt, 1);
  // the hot pass's wave blocks (their rows are counted once, with the cold pass's)
  if (cold && g->n_tblk && s->hskip) {
    hipLaunchKernelGGL(k_blocks_done, dim3(std::min<uint32_t>((g->n_tblk + kBlock - 1) / kBlock, cap)), dim3(kBlock), 0,
                       st, g->n_tblk, g->tblk, (const uint64_t *)s->Rdone, (const uint64_t *)g->rarc, s->hskip,
                       s->cnt, 0);
    s->hskip_on = true;
  }
  GRD_HIP(hipGetLastError());
  s->skip_pending = true;
  *flags = s->bskip;
  return GRD_OK;
}

template <class Op>
grd_status bsp_pull_launch(grd_state *s, const uint32_t *cw)
{
  grd_graph *g = s->g;
  using V = typename Op::V;
  const uint8_t *skip = nullptr;
  if (bsp_hot_cold(s)) {
    GRD_TRY(blocks_done<Op>(s, true, &skip));
    // the hot partials against the bound (CC / SSSP)
    const Counters *bcnt = (Op::kMin && s->bound_sent) ? s->cnt : nullptr;
    GRD_TRY(bsp_pull_hot<Op>(s));
    if (g->n_cblk && cw && g->cow8)
      hipLaunchKernelGGL((k_bsp_pull_rb<Op, uint8_t>), dim3(g->n_cblk), dim3(kBlock), 0, g->ctx->stream, g->cblk,
                         g->cblk_hot, g->corp, g->cocol, (const uint8_t *)g->cow8, (const V *)s->mprop, g->P,
                         (V *)s->mupd, s->Du, (V *)s->cseg, (const V *)s->tpart, skip,
                         skip ? (const uint64_t *)s->Rdone : nullptr, skip ? (const uint64_t *)g->rarc : nullptr,
                         bcnt, s->bound_inc);
    else if (g->n_cblk)
      hipLaunchKernelGGL((k_bsp_pull_rb<Op, uint32_t>), dim3(g->n_cblk), dim3(kBlock), 0, g->ctx->stream, g->cblk,
                         g->cblk_hot, g->corp, g->cocol, cw ? g->cow : nullptr, (const V *)s->mprop, g->P,
                         (V *)s->mupd, s->Du, (V *)s->cseg, (const V *)s->tpart, skip,
                         skip ? (const uint64_t *)s->Rdone : nullptr, skip ? (const uint64_t *)g->rarc : nullptr,
                         bcnt, s->bound_inc);
  } else {
    GRD_TRY(blocks_done<Op>(s, false, &skip));
    hipLaunchKernelGGL((k_bsp_pull_rb<Op, uint32_t>), dim3(g->n_pblk), dim3(kBlock), 0, g->ctx->stream, g->pblk,
                       g->pblk_hot, g->crp, g->ccol, cw, (const V *)s->mprop, g->P, (V *)s->mupd, s->Du,
                       (V *)s->partial, (const V *)nullptr, skip, skip ? (const uint64_t *)s->Rdone : nullptr,
                       skip ? (const uint64_t *)g->rarc : nullptr, (const Counters *)nullptr, 0u);
  }
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

template <class Op>
grd_status bsp_pull_fixup(grd_state *s)
{
  grd_graph *g = s->g;
  using V = typename Op::V;
  const bool hc = bsp_hot_cold(s);
  const uint32_t nlong = hc ? g->n_clong : g->n_plong;
  if (!nlong) return GRD_OK;
  hipLaunchKernelGGL(k_bsp_pull_fixup<Op>, dim3(fixup_grid(nlong)), dim3(kBlock), 0, g->ctx->stream,
                     hc ? g->clong : g->plong, nlong, hc ? (const V *)s->cseg : (const V *)s->partial, (V *)s->mupd,
                     s->Du, hc ? (const V *)s->tpart : (const V *)nullptr,
                     (Op::kMin && s->rdone_sent) ? (const uint64_t *)s->Rdone : nullptr);
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

grd_status pull_done_arcs(grd_state *s)
{
  grd_graph *g = s->g;
  const uint32_t nwP = g->P / 64;
  if (nwP)
    hipLaunchKernelGGL(k_done_arcs, dim3(std::min<uint32_t>((nwP + kWavesPerBlock - 1) / kWavesPerBlock,
                                                          (uint32_t)g->ctx->num_cus * 8)),
                       dim3(kBlock), 0, g->ctx->stream, nwP, bsp_hot_cold(s) ? g->corp : g->crp,
                       s->rdone_sent ? (const uint64_t *)s->Rdone : nullptr,
                       bsp_hot_cold(s) ? (const uint32_t *)s->tpart : nullptr, s->bound_inc, s->cnt);
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

template grd_status bsp_pull_launch<OpCC>(grd_state *, const uint32_t *);
template grd_status bsp_pull_launch<OpHop>(grd_state *, const uint32_t *);
template grd_status bsp_pull_launch<OpWeighted>(grd_state *, const uint32_t *);
template grd_status bsp_pull_launch<OpPR>(grd_state *, const uint32_t *);
template grd_status bsp_pull_fixup<OpCC>(grd_state *);
template grd_status bsp_pull_fixup<OpHop>(grd_state *);
template grd_status bsp_pull_fixup<OpWeighted>(grd_state *);
template grd_status bsp_pull_fixup<OpPR>(grd_state *);

}  // namespace grd
