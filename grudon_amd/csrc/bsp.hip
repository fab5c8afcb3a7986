// bsp.hip -- the BSP superstep of the per-partition operator path on gfx950.
//
// One launch covers every partition: compute slots and memory proxies of all
// partitions are laid out back to back (grd_internal.hpp), so each kernel
// below is a single grid over the concatenation.  Wave64 throughout; the
// bitvector protocol works on 64-bit words, one wave per word (lane = bit).
//
// Reference loops replaced (src/GraphAlgorithm.cpp):
//   k_bcast_send / k_bcast_recv   compute pack + memory unpack       :199-343
//   k_push / k_push_heavy / (pull.hip)  gen_updates (pr/cc/sssp.cpp)  :348-356
//   k_reduce_send / k_reduce_recv memory pack + compute aggregate     :377-581
//   k_update                      update_frontier + apply_updates +
//                                 termination popcount               :643-662
//   k_popc_blocks/k_scan_blocks/k_word_rank  DynamicBitSet::getOffsets ranks
#include <algorithm>
#include <cstring>

#include "grd_device.hpp"
#include "grd_internal.hpp"
#include "grd_ops.hpp"

namespace grd {

// ------------------------------------------------------------------------
// word popcount ranks (exclusive prefix of set bits per 64-bit word)
// ------------------------------------------------------------------------

constexpr int kWordsPerScanBlock = 1024;  // 256 threads x 4 words

__global__ void __launch_bounds__(kBlock) k_popc_blocks(const uint64_t *__restrict__ w, uint32_t nw, uint32_t *blk)
{
  uint32_t base = blockIdx.x * kWordsPerScanBlock + threadIdx.x * 4;
  uint32_t c = 0;
#pragma unroll
  for (int i = 0; i < 4; i++)
    if (base + i < nw) c += __popcll(w[base + i]);
  c = block_sum(c);
  if (threadIdx.x == 0) blk[blockIdx.x] = c;
}

// single block of 1024 threads: exclusive scan of nb partials, total -> *total
__global__ void __launch_bounds__(1024) k_scan_blocks(uint32_t *blk, uint32_t nb, uint32_t *total)
{
  __shared__ uint32_t s_w[16];
  uint32_t per = (nb + 1023) / 1024;
  uint32_t b0 = threadIdx.x * per;
  uint32_t local = 0;
  for (uint32_t i = 0; i < per; i++)
    if (b0 + i < nb) local += blk[b0 + i];
  uint32_t inc = wave_incl_scan(local);
  if (lane_id() == kWave - 1) s_w[threadIdx.x / kWave] = inc;
  __syncthreads();
  uint32_t wb = 0, tot = 0;
  for (int i = 0; i < 16; i++) {
    if (i < (int)(threadIdx.x / kWave)) wb += s_w[i];
    tot += s_w[i];
  }
  uint32_t run = wb + inc - local;
  for (uint32_t i = 0; i < per; i++)
    if (b0 + i < nb) {
      uint32_t x = blk[b0 + i];
      blk[b0 + i] = run;
      run += x;
    }
  if (threadIdx.x == 0) *total = tot;
}

__global__ void __launch_bounds__(kBlock) k_word_rank(const uint64_t *__restrict__ w, uint32_t nw,
                                                      const uint32_t *__restrict__ blk, uint32_t *__restrict__ wrank)
{
  uint32_t base = blockIdx.x * kWordsPerScanBlock + threadIdx.x * 4;
  uint32_t c[4], s = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    c[i] = (base + i < nw) ? (uint32_t)__popcll(w[base + i]) : 0u;
    s += c[i];
  }
  uint32_t tot;
  uint32_t ex = block_excl_scan(s, &tot) + blk[blockIdx.x];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    if (base + i < nw) wrank[base + i] = ex;
    ex += c[i];
  }
}

static grd_status word_ranks(hipStream_t st, const uint64_t *w, uint32_t nw, uint32_t *blk, uint32_t *wrank,
                             uint32_t *total)
{
  uint32_t nb = (nw + kWordsPerScanBlock - 1) / kWordsPerScanBlock;
  if (nb == 0) nb = 1;
  hipLaunchKernelGGL(k_popc_blocks, dim3(nb), dim3(kBlock), 0, st, w, nw, blk);
  hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, blk, nb, total);
  hipLaunchKernelGGL(k_word_rank, dim3(nb), dim3(kBlock), 0, st, w, nw, blk, wrank);
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

// Grid of the grid-stride word loops (init): up to 64 workgroups per CU.
static inline uint32_t wave_grid(const grd_ctx *ctx, uint64_t nwords)
{
  uint64_t want = (nwords + kWavesPerBlock - 1) / kWavesPerBlock;
  uint64_t cap = (uint64_t)ctx->num_cus * 64;
  return (uint32_t)std::max<uint64_t>(1, std::min(want, cap));
}

// Word-group passes (exchange, update, bucket selection): each wave owns
// kGroupWords consecutive bitmap words.  Lane j < kGroupWords loads word j of
// the group (one coalesced access), the lanes finish their zero words
// themselves, and the wave then walks only the nonzero words with one lane per
// bit.  A sparse superstep therefore costs one pass over the bitmap, and a
// dense one a few dependent loads per word.
constexpr uint32_t kGroupWords = 8;

__device__ __forceinline__ uint32_t group_base()
{
  return (blockIdx.x * kWavesPerBlock + wave_in_block()) * kGroupWords;
}

__device__ __forceinline__ uint64_t shfl_u64(uint64_t x, int src)
{
  const uint32_t lo = __shfl((uint32_t)x, src, kWave), hi = __shfl((uint32_t)(x >> 32), src, kWave);
  return ((uint64_t)hi << 32) | lo;
}

static inline uint32_t group_grid(uint64_t nwords)
{
  const uint64_t per = (uint64_t)kGroupWords * kWavesPerBlock;
  return (uint32_t)std::max<uint64_t>(1, (nwords + per - 1) / per);
}

// Passes that end in same-address counter atomics (update, reduce pack,
// bucket selection) loop over the groups from a grid of this many
// workgroups per CU, so that few blocks issue those atomics.
constexpr uint32_t kLoopBlocksPerCU = 8;

static inline uint32_t loop_grid(const grd_ctx *ctx, uint64_t nwords)
{
  return std::min<uint32_t>(group_grid(nwords), (uint32_t)ctx->num_cus * kLoopBlocksPerCU);
}

// ------------------------------------------------------------------------
// init
// ------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_fill_u32(uint32_t *p, uint64_t n, uint32_t v)
{
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) p[i] = v;
}
__global__ void __launch_bounds__(kBlock) k_fill_u64(uint64_t *p, uint64_t n, uint64_t v)
{
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) p[i] = v;
}

// compute-side init: pr.cpp:25-50, cc.cpp:24-36, sssp.cpp:22-46
__global__ void __launch_bounds__(kBlock) k_init_compute(int algo, uint32_t nw, const uint32_t *__restrict__ slot_gid,
                                                         const uint32_t *__restrict__ deg, void *prop_, void *upd_,
                                                         float *pr, uint64_t *F, uint32_t source_slot)
{
  const int wpb = kWavesPerBlock;
  for (uint32_t w = blockIdx.x * wpb + wave_in_block(); w < nw; w += gridDim.x * wpb) {
    uint32_t s = w * 64 + lane_id();
    uint32_t gid = slot_gid[s];
    bool valid = gid != 0xFFFFFFFFu;
    bool f = false;
    if (algo == GRD_PR) {
      float *prop = (float *)prop_, *upd = (float *)upd_;
      uint32_t d = deg[s];
      if (valid && d != 0) {
        prop[s] = (float)(0.85 * (1.0 - 0.85) / (double)d);
        pr[s] = (float)(1.0 - 0.85);
        f = true;
      } else {
        prop[s] = 0.0f;
        pr[s] = 0.0f;
      }
      upd[s] = 0.0f;
    } else {
      uint32_t *prop = (uint32_t *)prop_, *upd = (uint32_t *)upd_;
      if (algo == GRD_CC) {
        prop[s] = valid ? gid : 0xFFFFFFFFu;
        upd[s] = 0xFFFFFFFFu;
        f = valid;
      } else {
        bool src = (s == source_slot);
        prop[s] = src ? 0u : 0xFFFFFFFFu;
        upd[s] = src ? 0u : 0xFFFFFFFFu;
        f = src;
      }
    }
    uint64_t fm = __ballot(f);
    if (lane_id() == 0) F[w] = fm;
  }
}

// ------------------------------------------------------------------------
// broadcast: compute -> memory (bitvector = frontier words of the pair)
// ------------------------------------------------------------------------

// With fmin (a CC / SSSP pull), also the smallest value sent: the grid then
// loops over the groups (loop_grid), so that one atomic per workgroup
// carries it.
template <typename V>
__global__ void __launch_bounds__(kBlock) k_bcast_send(uint32_t nw, const uint64_t *__restrict__ F,
                                                       const V *__restrict__ prop, const uint32_t *__restrict__ wrank,
                                                       int dense, uint64_t *__restrict__ bv, V *__restrict__ val,
                                                       unsigned int *fmin)
{
  static_assert(sizeof(V) == 4, "4-byte values on the wire");
  const int l = lane_id();
  uint32_t lo = 0xFFFFFFFFu;
  for (uint32_t base = group_base(); base < nw; base += gridDim.x * kWavesPerBlock * kGroupWords) {
    const bool ok = l < (int)kGroupWords && base + l < nw;
    const uint64_t mine = ok ? F[base + l] : 0ull;
    if (ok) bv[base + l] = mine;
    uint64_t todo = dense ? __ballot(ok) : __ballot(mine != 0);
    while (todo) {
      const int j = __ffsll((unsigned long long)todo) - 1;
      todo &= todo - 1;
      const uint64_t word = shfl_u64(mine, j);
      const uint32_t w = base + j, s = w * 64 + l;
      const bool on = (word >> l) & 1ull;
      if (dense || on) {
        const V x = prop[s];
        if (dense) val[s] = x;
        else val[wrank[w] + (uint32_t)__popcll(word & lanemask_lt())] = x;
        const uint32_t u = __builtin_bit_cast(uint32_t, x);
        if (on && u < lo) lo = u;
      }
    }
  }
  if (fmin) {  // uniform: a kernel argument
    lo = block_min(lo);
    if (threadIdx.x == 0 && lo != 0xFFFFFFFFu) atomicMin(fmin, lo);
  }
}

// Pull of CC / SSSP, compute side: bit r of Rdone is set when memory proxy r
// belongs to a vertex whose value is at or below b = the smallest value
// broadcast this superstep plus the smallest arc increment (0 for CC, 1 for
// hop counts, the smallest arc weight for weighted SSSP; saturating).  Every
// candidate of the superstep is Op::edge of a broadcast value over an arc,
// never below b, so no candidate can improve such a row and the pull may
// leave its arcs out.  A wave covers one word of rows.
__global__ void __launch_bounds__(kBlock) k_rows_done(uint32_t nwP, const uint32_t *__restrict__ proxy_slot,
                                                      const uint32_t *__restrict__ prop, const Counters *cnt,
                                                      uint32_t inc, uint64_t *__restrict__ Rdone)
{
  const uint32_t f = cnt->fmin, fmin = f + inc < f ? 0xFFFFFFFFu : f + inc;
  const uint64_t nrow = (uint64_t)nwP * 64;
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < nrow; r += (uint64_t)gridDim.x * kBlock) {
    const uint32_t sl = proxy_slot[r];
    const bool done = sl != kNone && prop[sl] <= fmin;
    const uint64_t m = __ballot(done);
    if (lane_id() == 0) Rdone[r >> 6] = m;
  }
}

// Receive of the broadcast.  For a push the active proxies get their value
// and the active list; for a pull (want_list == 0) every own proxy is written,
// inactive ones with the combine identity, so that the pull gathers one value
// per arc and needs no activity test.
template <typename V>
__global__ void __launch_bounds__(kBlock) k_bcast_recv(uint32_t nw, const uint64_t *__restrict__ bv,
                                                       const V *__restrict__ val, const uint32_t *__restrict__ wrank,
                                                       const uint32_t *__restrict__ word_dst, int dense, int want_list,
                                                       V ident, V *__restrict__ mprop, uint32_t *__restrict__ AL)
{
  const uint32_t base = group_base();
  const int l = lane_id();
  const bool ok = l < (int)kGroupWords && base + l < nw;
  const uint32_t pwl = ok ? word_dst[base + l] : 0xFFFFFFFFu;
  const uint64_t mine = (pwl != 0xFFFFFFFFu) ? bv[base + l] : 0ull;
  uint64_t todo = want_list ? __ballot(mine != 0) : __ballot(pwl != 0xFFFFFFFFu);
  while (todo) {
    const int j = __ffsll((unsigned long long)todo) - 1;
    todo &= todo - 1;
    const uint64_t word = shfl_u64(mine, j);
    const uint32_t pw = __shfl(pwl, j, kWave);
    const uint32_t w = base + j;
    const uint32_t p = pw * 64 + l;
    if ((word >> l) & 1ull) {
      uint32_t r = (dense && !want_list) ? 0u : wrank[w] + (uint32_t)__popcll(word & lanemask_lt());
      V x = dense ? val[w * 64 + l] : val[r];
      mprop[p] = x;
      if (want_list) AL[r] = p;
    } else if (!want_list) {
      mprop[p] = ident;
    }
  }
}

// ------------------------------------------------------------------------
// traversal: push over the active list (degree-binned, LDS-staged tiles)
// ------------------------------------------------------------------------

// A block stages 256 active vertices in LDS.  Vertices with out-degree below
// kHeavyDeg are processed together: the block scans their degrees and every thread
// walks the concatenated edge range (coalesced col reads), locating its
// source by binary search in the LDS prefix array.  Vertices with degree
// >= kHeavyDeg are appended to the heavy list as chunks of kHeavyDeg arcs and
// handled by k_push_heavy, one workgroup per chunk.
template <class Op>
__global__ void __launch_bounds__(kBlock) k_push(const uint32_t *__restrict__ AL, const uint32_t *__restrict__ nact,
                                                 const uint32_t *__restrict__ rp, const uint32_t *__restrict__ col,
                                                 const uint32_t *__restrict__ wt, const typename Op::V *__restrict__ mprop,
                                                 typename Op::V *upd, uint64_t *dirty, uint2 *heavy, Counters *cnt)
{
  using V = typename Op::V;
  __shared__ uint32_t s_off[kBlock + 1];
  __shared__ uint32_t s_beg[kBlock];
  __shared__ V s_val[kBlock];
  const uint32_t k = *nact;
  uint64_t my_edges = 0;
  for (uint32_t tile = blockIdx.x * kBlock; tile < k; tile += gridDim.x * kBlock) {
    uint32_t i = tile + threadIdx.x;
    uint32_t deg = 0, beg = 0;
    V v = op_identity<Op>();
    if (i < k) {
      uint32_t p = AL[i];
      beg = rp[p];
      deg = rp[p + 1] - beg;
      v = mprop[p];
      if (deg >= kHeavyDeg) {
        const uint32_t nch = (deg + kHeavyDeg - 1) / kHeavyDeg;
        unsigned long long slot = atomicAdd(&cnt->heavy, (unsigned long long)nch);
        for (uint32_t c = 0; c < nch; c++) heavy[slot + c] = make_uint2(p, c * kHeavyDeg);
        atomicAdd(&cnt->heavy_edges, (unsigned long long)deg);
        deg = 0;
      }
    }
    uint32_t total;
    uint32_t ex = block_excl_scan(deg, &total);
    s_off[threadIdx.x] = ex;
    s_beg[threadIdx.x] = beg;
    s_val[threadIdx.x] = v;
    if (threadIdx.x == 0) s_off[kBlock] = total;
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < total; j += kBlock) {
      // owner t: last index with s_off[t] <= j
      uint32_t lo = 0, hi = kBlock;
      while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (s_off[mid] <= j) lo = mid;
        else hi = mid;
      }
      uint32_t e = s_beg[lo] + (j - s_off[lo]);
      relax<Op>(upd, dirty, col[e], Op::edge(s_val[lo], wt ? wt[e] : 1u));
    }
    if (threadIdx.x == 0) my_edges += total;
    __syncthreads();
  }
  my_edges = block_sum(my_edges);
  if (threadIdx.x == 0 && my_edges) atomicAdd(&cnt->edges, (unsigned long long)my_edges);
}

template <class Op>
__global__ void __launch_bounds__(kBlock) k_push_heavy(const uint2 *__restrict__ heavy, const Counters *cnt,
                                                       const uint32_t *__restrict__ rp, const uint32_t *__restrict__ col,
                                                       const uint32_t *__restrict__ wt,
                                                       const typename Op::V *__restrict__ mprop, typename Op::V *upd,
                                                       uint64_t *dirty)
{
  const uint32_t nch = (uint32_t)cnt->heavy;
  for (uint32_t h = blockIdx.x; h < nch; h += gridDim.x) {
    const uint2 c = heavy[h];
    const uint32_t end = rp[c.x + 1];
    const uint32_t b = rp[c.x] + c.y, e = min(end, b + kHeavyDeg);
    const typename Op::V v = mprop[c.x];
    for (uint32_t x = b + threadIdx.x; x < e; x += kBlock) relax<Op>(upd, dirty, col[x], Op::edge(v, wt ? wt[x] : 1u));
  }
}

// ------------------------------------------------------------------------
// reduce: memory -> compute (bitvector = dirty words of the pair)
// ------------------------------------------------------------------------

template <typename V>
__global__ void __launch_bounds__(kBlock) k_reduce_send(uint32_t nw, uint64_t *__restrict__ Du, V *__restrict__ mupd,
                                                        const uint32_t *__restrict__ wrank, int dense, int clear_vals,
                                                        uint64_t *__restrict__ bv, V *__restrict__ val, Counters *cnt)
{
  const int l = lane_id();
  uint32_t ndirty = 0;
  // capped grid, wave-uniform loop over the word groups (fewer counter atomics)
  for (uint32_t base = group_base(); base < nw; base += gridDim.x * kWavesPerBlock * kGroupWords) {
    const bool ok = l < (int)kGroupWords && base + l < nw;
    const uint64_t mine = ok ? Du[base + l] : 0ull;
    if (ok) {
      bv[base + l] = mine;
      if (mine) Du[base + l] = 0;
    }
    ndirty += (uint32_t)__popcll(mine);
    uint64_t todo = dense ? __ballot(ok) : __ballot(mine != 0);
    while (todo) {
      const int j = __ffsll((unsigned long long)todo) - 1;
      todo &= todo - 1;
      const uint64_t word = shfl_u64(mine, j);
      const uint32_t w = base + j, p = w * 64 + l;
      const bool bit = (word >> l) & 1ull;
      if (dense) val[p] = mupd[p];
      else if (bit) val[wrank[w] + (uint32_t)__popcll(word & lanemask_lt())] = mupd[p];
      if (clear_vals && bit) mupd[p] = V(0);  // PropertyList::clear (DistributedGraph.hpp:140-166)
    }
  }
  uint32_t t = block_sum(ndirty);
  if (threadIdx.x == 0 && t) atomicAdd(&cnt->dirty, (unsigned long long)t);
}

template <class Op>
__global__ void __launch_bounds__(kBlock) k_reduce_recv(uint32_t nw, const uint64_t *__restrict__ bv,
                                                        const typename Op::V *__restrict__ val,
                                                        const uint32_t *__restrict__ wrank,
                                                        const uint32_t *__restrict__ proxy_slot, int dense,
                                                        typename Op::V *upd, uint64_t *Dc)
{
  using V = typename Op::V;
  const uint32_t base = group_base();
  const int l = lane_id();
  const bool ok = l < (int)kGroupWords && base + l < nw;
  const uint64_t mine = ok ? bv[base + l] : 0ull;
  const uint32_t myrank = (ok && !dense) ? wrank[base + l] : 0u;
  if (__ballot(mine != 0) == 0) return;
  // the group's words in every lane, then every load of the group (values,
  // slots, and for min-combines the current values) issued before any is
  // waited for; the relaxations and marks then follow word by word
  uint64_t word[kGroupWords];
  uint32_t s[kGroupWords];
  V x[kGroupWords];
#pragma unroll
  for (int j = 0; j < (int)kGroupWords; j++) {
    word[j] = shfl_u64(mine, j);
    const uint32_t rank = __shfl(myrank, j, kWave);
    s[j] = 0xFFFFFFFFu;
    x[j] = V{};
    if ((word[j] >> l) & 1ull) {
      const uint32_t p = (base + j) * 64 + l;
      x[j] = dense ? val[p] : val[rank + (uint32_t)__popcll(word[j] & lanemask_lt())];
      s[j] = proxy_slot[p];
    }
  }
  V cur[kGroupWords];
  if constexpr (Op::kMin) {
#pragma unroll
    for (int j = 0; j < (int)kGroupWords; j++)
      cur[j] = s[j] != 0xFFFFFFFFu ? __atomic_load_n(upd + s[j], __ATOMIC_RELAXED) : V{};
  }
#pragma unroll
  for (int j = 0; j < (int)kGroupWords; j++) {
    if (word[j] == 0) continue;  // uniform over the wave
    // aggregate(): minUpdate / addUpdate, then the dirty marks of the wave
    bool mark = false;
    if (s[j] != 0xFFFFFFFFu) {
      if constexpr (Op::kMin) mark = x[j] < cur[j] && x[j] < atomicMin(upd + s[j], x[j]);
      else mark = relax_nomark<Op, true>(upd, s[j], x[j]);
    }
    mark_wave(Dc, s[j], mark);
  }
}

// ------------------------------------------------------------------------
// update: update_frontier + apply_updates + frontier popcount
// ------------------------------------------------------------------------

template <int ALGO>  // GRD_PR, GRD_CC or GRD_SSSP
__global__ void __launch_bounds__(kBlock) k_update(uint32_t nw, uint64_t *__restrict__ Dc, void *prop_, void *upd_,
                                                   float *__restrict__ pr, const uint32_t *__restrict__ deg,
                                                   uint64_t *__restrict__ F, uint64_t *__restrict__ Pend,
                                                   unsigned long long hi, int use_delta, Counters *cnt)
{
  const int l = lane_id();
  uint64_t f_cnt = 0, f_edges = 0, improved = 0;
  uint32_t f_max = 0;
  // a capped grid walks the word groups (the loop is wave-uniform), so the
  // block's counter atomics below are issued by fewer blocks
  for (uint32_t base = group_base(); base < nw; base += gridDim.x * kWavesPerBlock * kGroupWords) {
    const bool ok = l < (int)kGroupWords && base + l < nw;
    const uint64_t mine = ok ? Dc[base + l] : 0ull;
    if (ok && !mine) F[base + l] = 0;  // nothing changed: empty next-frontier word
    uint64_t todo = __ballot(mine != 0);
    while (todo) {
      const int j = __ffsll((unsigned long long)todo) - 1;
      todo &= todo - 1;
      const uint64_t d = shfl_u64(mine, j);
      const uint32_t w = base + j, s = w * 64 + l;
      bool f = false, pend = false, imp = false;
      if ((d >> l) & 1ull) {
        if constexpr (ALGO == GRD_PR) {
          float *prop = (float *)prop_, *upd = (float *)upd_;
          float u = upd[s];
          if ((double)u > 1e-6) {  // pr.cpp:133-142 and :76-83 (prev is always 0 here)
            f = true;
            pr[s] = pr[s] + u;
            prop[s] = (float)(0.85 * (double)u / (double)deg[s]);
            upd[s] = 0.0f;
          }
        } else {
          uint32_t *prop = (uint32_t *)prop_, *upd = (uint32_t *)upd_;
          uint32_t u = upd[s];
          if (prop[s] > u) {  // cc.cpp:100-106, sssp.cpp:102-108
            prop[s] = u;
            imp = true;
            if (use_delta && (unsigned long long)u >= hi) pend = true;
            else f = true;
          }
        }
      }
      const uint64_t fm = __ballot(f);
      if (l == 0) {
        F[w] = fm;
        Dc[w] = 0;
        f_cnt += (uint64_t)__popcll(fm);
      }
      if (use_delta) {
        const uint64_t pm = __ballot(pend);
        if (l == 0 && (pm | fm)) Pend[w] = (Pend[w] | pm) & ~fm;
      }
      if (f) {
        uint32_t dg = deg[s];
        f_edges += dg;
        f_max = dg > f_max ? dg : f_max;
      }
      improved += imp;
    }
  }
  uint64_t a = block_sum(f_cnt);
  uint64_t b = block_sum(f_edges);
  uint64_t c = block_sum(improved);
  uint32_t mx = block_max(f_max);
  if (threadIdx.x == 0) {
    if (a) atomicAdd(&cnt->frontier, (unsigned long long)a);
    if (b) atomicAdd(&cnt->frontier_edges, (unsigned long long)b);
    if (c) atomicAdd(&cnt->improved, (unsigned long long)c);
    if (mx) atomicMax(&cnt->frontier_maxdeg, mx);
  }
}

// delta-stepping: smallest pending distance and pending count
__global__ void __launch_bounds__(kBlock) k_pend_min(uint32_t nw, const uint64_t *__restrict__ Pend,
                                                     const uint32_t *__restrict__ prop, Counters *cnt)
{
  const int l = lane_id();
  uint32_t mn = 0xFFFFFFFFu;
  uint64_t c = 0;
  // capped grid, wave-uniform loop over the word groups
  for (uint32_t base = group_base(); base < nw; base += gridDim.x * kWavesPerBlock * kGroupWords) {
    const bool ok = l < (int)kGroupWords && base + l < nw;
    const uint64_t mine = ok ? Pend[base + l] : 0ull;
    uint64_t todo = __ballot(mine != 0);
    while (todo) {
      const int j = __ffsll((unsigned long long)todo) - 1;
      todo &= todo - 1;
      const uint64_t pm = shfl_u64(mine, j);
      if ((pm >> l) & 1ull) {
        uint32_t v = prop[(base + j) * 64 + l];
        mn = v < mn ? v : mn;
        c++;
      }
    }
  }
  mn = block_min(mn);
  c = block_sum(c);
  if (threadIdx.x == 0 && c) {
    atomicMin(&cnt->pend_min, mn);
    atomicAdd(&cnt->pending, (unsigned long long)c);
  }
}

// delta-stepping: move pending vertices below the new bucket bound into F
__global__ void __launch_bounds__(kBlock) k_select(uint32_t nw, uint64_t *__restrict__ Pend,
                                                   const uint32_t *__restrict__ prop, const uint32_t *__restrict__ deg,
                                                   unsigned long long hi, uint64_t *__restrict__ F, Counters *cnt)
{
  const int l = lane_id();
  uint64_t f_cnt = 0, f_edges = 0;
  uint32_t f_max = 0;
  // capped grid, wave-uniform loop over the word groups
  for (uint32_t base = group_base(); base < nw; base += gridDim.x * kWavesPerBlock * kGroupWords) {
    const bool ok = l < (int)kGroupWords && base + l < nw;
    const uint64_t mine = ok ? Pend[base + l] : 0ull;
    if (ok && !mine) F[base + l] = 0;
    uint64_t todo = __ballot(mine != 0);
    while (todo) {
      const int j = __ffsll((unsigned long long)todo) - 1;
      todo &= todo - 1;
      const uint64_t pm = shfl_u64(mine, j);
      const uint32_t w = base + j, s = w * 64 + l;
      const bool f = ((pm >> l) & 1ull) && (unsigned long long)prop[s] < hi;
      const uint64_t fm = __ballot(f);
      if (l == 0) {
        F[w] = fm;
        if (fm) Pend[w] = pm & ~fm;
        f_cnt += (uint64_t)__popcll(fm);
      }
      if (f) {
        uint32_t dg = deg[s];
        f_edges += dg;
        f_max = dg > f_max ? dg : f_max;
      }
    }
  }
  uint64_t a = block_sum(f_cnt), b = block_sum(f_edges);
  uint32_t mx = block_max(f_max);
  if (threadIdx.x == 0) {
    if (a) atomicAdd(&cnt->frontier, (unsigned long long)a);
    if (b) atomicAdd(&cnt->frontier_edges, (unsigned long long)b);
    if (mx) atomicMax(&cnt->frontier_maxdeg, mx);
  }
}

__global__ void k_zero_counters(Counters *cnt)
{
  if (threadIdx.x == 0) {
    cnt->frontier = 0;
    cnt->frontier_edges = 0;
    cnt->dirty = 0;
    cnt->pending = 0;
    cnt->improved = 0;
    cnt->heavy = 0;
    cnt->heavy_edges = 0;
    cnt->edges = 0;
    cnt->frontier_maxdeg = 0;
    cnt->pend_min = 0xFFFFFFFFu;
    cnt->fmin = 0xFFFFFFFFu;
    cnt->pull_skipped = 0;
    cnt->pull_skipped_hot = 0;
    cnt->pull_skipped_rows = 0;
    cnt->pull_done_arcs = 0;
    cnt->frontier_uncovered = 0;
    cnt->ship_error = 0;
  }
  for (uint32_t c = threadIdx.x; c < (uint32_t)kMaxParts; c += blockDim.x) cnt->fpart[c] = 0;
}

// single-element aggregate() (API compatibility path)
template <class Op>
__global__ void k_aggregate_one(typename Op::V *upd, uint64_t *Dc, uint32_t slot, typename Op::V x)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) relax<Op, true>(upd, Dc, slot, x);
}

// results by gid
template <typename V>
__global__ void __launch_bounds__(kBlock) k_gather_result(uint32_t S, const uint32_t *__restrict__ slot_gid,
                                                          const V *__restrict__ src, V *__restrict__ out)
{
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < S; s += gridDim.x * kBlock) {
    uint32_t g = slot_gid[s];
    if (g != 0xFFFFFFFFu) out[g] = src[s];
  }
}

// ------------------------------------------------------------------------
// host orchestration
// ------------------------------------------------------------------------

grd_status op_zero_counters(grd_state *s)
{
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s->g->ctx->stream, s->cnt);
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

grd_status op_read_counters(grd_state *s)
{
  GRD_HIP(hipMemcpyAsync(s->hcnt, s->cnt, sizeof(Counters), hipMemcpyDeviceToHost, s->g->ctx->stream));
  GRD_HIP(hipStreamSynchronize(s->g->ctx->stream));
  return GRD_OK;
}

grd_status state_alloc(grd_state *s)
{
  grd_graph *g = s->g;
  const uint64_t S = g->S, P = g->P;
  const uint32_t nwS = (uint32_t)(S / 64), nwP = (uint32_t)(P / 64);
  const uint64_t V = 4;
  DevMem &m = s->mem;
  GRD_TRY(m.alloc(&s->prop, S * V));
  GRD_TRY(m.alloc(&s->upd, S * V));
  if (s->algo == GRD_PR || s->algo == GRD_PR_PULL) GRD_TRY(m.alloc_n(&s->pr, S));
  GRD_TRY(m.alloc_n(&s->F, nwS));
  GRD_TRY(m.alloc_n(&s->Dc, nwS));
  GRD_TRY(m.alloc_n(&s->Pend, nwS));
  GRD_TRY(m.alloc(&s->mprop, P * V));
  GRD_TRY(m.alloc(&s->mupd, P * V));
  GRD_TRY(m.alloc_n(&s->Du, nwP));
  if (s->algo == GRD_CC || s->algo == GRD_SSSP) GRD_TRY(m.alloc_n(&s->Rdone, std::max<uint32_t>(nwP, 1)));
  GRD_TRY(m.alloc_n(&s->bvB, nwS));
  GRD_TRY(m.alloc_n(&s->bvR, nwP));
  GRD_TRY(m.alloc(&s->valB, S * V));
  GRD_TRY(m.alloc(&s->valR, P * V));
  uint32_t nb = (std::max(nwS, nwP) + kWordsPerScanBlock - 1) / kWordsPerScanBlock + 1;
  GRD_TRY(m.alloc_n(&s->blk, nb));
  GRD_TRY(m.alloc_n(&s->wrankB, nwS));
  GRD_TRY(m.alloc_n(&s->wrankR, nwP));
  GRD_TRY(m.alloc_n(&s->totals, 64));
  GRD_TRY(m.alloc_n(&s->AL, S));
  // heavy chunks: sum over heavy vertices of ceil(deg / kHeavyDeg) <= 2 m / kHeavyDeg
  GRD_TRY(m.alloc_n(&s->heavy, 2 * (g->m / kHeavyDeg) + 64));
  GRD_TRY(m.alloc(&s->partial, (uint64_t)std::max<uint32_t>(1, g->n_pblk) * 8));
  if (s->algo == GRD_PR_PULL) {
    GRD_TRY(m.alloc_n(&s->red, 2ull * prp_grid(g)));
    if (!g->word_nvalid) {
      // a pair group's valid slots are a prefix of its padded range, so the
      // valid slots of a word are a prefix of the word
      std::vector<uint8_t> nv(std::max<uint32_t>(nwS, 1), 0);
      for (size_t k = 0; k < g->cbase.size(); k++) {
        const uint64_t b0 = g->cbase[k], b1 = b0 + g->ccnt[k];
        if (b0 % 64) return fail(GRD_ERR_STATE, "layout: a pair group does not start on a slot word");
        for (uint64_t w = b0 / 64; w * 64 < b1 && w < nwS; w++)
          nv[w] = (uint8_t)std::min<uint64_t>(64, b1 - std::max<uint64_t>(b0, w * 64));
      }
      GRD_TRY(g->mem.alloc_n(&g->word_nvalid, nv.size()));
      GRD_HIP(hipMemcpyAsync(g->word_nvalid, nv.data(), nv.size(), hipMemcpyHostToDevice, g->ctx->stream));
      GRD_HIP(hipStreamSynchronize(g->ctx->stream));
    }
  }
  if (s->tiled) {
    if (g->tier_mode == 3) {
      // hot / cold: one 4-byte partial per row and per hot segment
      GRD_TRY(m.alloc(reinterpret_cast<void **>(&s->tpart), std::max<uint64_t>(g->n_pairs, 1) * 4));
      GRD_TRY(m.alloc(reinterpret_cast<void **>(&s->tseg), (uint64_t)std::max<uint32_t>(g->n_tblk, 1) * 4));
      GRD_TRY(m.alloc_n(&s->cseg, std::max<uint32_t>(g->n_cblk, 1)));
    } else {
      GRD_TRY(m.alloc_n(&s->tpart, std::max<uint64_t>(g->n_pairs, 1)));
      GRD_TRY(m.alloc_n(&s->tseg, std::max<uint32_t>(g->n_tblk, 1)));
    }
  }
  if (s->Rdone) {
    // one flag per row block of the pull this state runs
    const uint32_t nblk = bsp_hot_cold(s) ? g->n_cblk : g->n_pblk;
    GRD_TRY(m.alloc_n(&s->bskip, std::max<uint32_t>(nblk, 1)));
    if (bsp_hot_cold(s)) GRD_TRY(m.alloc_n(&s->hskip, std::max<uint32_t>(g->n_tblk, 1)));
  }
  GRD_TRY(m.alloc_n(&s->cnt, 1));
  GRD_HIP(hipHostMalloc((void **)&s->hcnt, sizeof(Counters), hipHostMallocDefault));
  memset(s->hcnt, 0, sizeof(Counters));
  // snapshot list: every array whose content carries state between iterations
  s->snap = {{s->prop, nullptr}, {s->upd, nullptr}, {s->F, nullptr},     {s->Dc, nullptr},
             {s->Pend, nullptr}, {s->mprop, nullptr}, {s->mupd, nullptr},
             {s->Du, nullptr}};
  s->snap_bytes = {S * V, S * V, nwS * 8ull, nwS * 8ull, nwS * 8ull, P * V, P * V, nwP * 8ull};
  if (s->pr) {
    s->snap.push_back({s->pr, nullptr});
    s->snap_bytes.push_back(S * 4);
  }
  // the counters carry the pull PageRank's FP64 dangling mass of the next
  // update (ping-pong by iteration parity) and the last residual
  s->snap.push_back({s->cnt, nullptr});
  s->snap_bytes.push_back(sizeof(Counters));
  return GRD_OK;
}

static grd_status fill_u32(hipStream_t st, void *p, uint64_t n, uint32_t v)
{
  hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)std::min<uint64_t>(4096, (n + kBlock - 1) / kBlock + 1)), dim3(kBlock), 0,
                     st, (uint32_t *)p, n, v);
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}
static grd_status fill_u64(hipStream_t st, void *p, uint64_t n, uint64_t v)
{
  hipLaunchKernelGGL(k_fill_u64, dim3((uint32_t)std::min<uint64_t>(4096, (n + kBlock - 1) / kBlock + 1)), dim3(kBlock), 0,
                     st, (uint64_t *)p, n, v);
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

grd_status state_init(grd_state *s)
{
  grd_graph *g = s->g;
  hipStream_t st = g->ctx->stream;
  const uint32_t nwS = g->S / 64, nwP = g->P / 64;
  GRD_TRY(fill_u64(st, s->Dc, nwS, 0));
  GRD_TRY(fill_u64(st, s->Pend, nwS, 0));
  GRD_TRY(fill_u64(st, s->Du, nwP, 0));
  GRD_TRY(fill_u64(st, s->bvB, nwS, 0));
  GRD_TRY(fill_u64(st, s->bvR, nwP, 0));
  const uint32_t neutral = (s->algo == GRD_PR) ? 0u : 0xFFFFFFFFu;  // PR: 0.0f bit pattern
  GRD_TRY(fill_u32(st, s->mprop, g->P, neutral));
  GRD_TRY(fill_u32(st, s->mupd, g->P, neutral));
  GRD_TRY(fill_u32(st, s->valR, g->P, neutral));
  GRD_TRY(fill_u32(st, s->valB, g->S, neutral));
  uint32_t source_slot = 0xFFFFFFFFu;
  if (s->algo == GRD_SSSP) source_slot = g->h_gid_slot[s->prm.source];
  hipLaunchKernelGGL(k_init_compute, dim3(wave_grid(g->ctx, nwS)), dim3(kBlock), 0, st, s->algo, nwS, g->slot_gid,
                     g->slot_deg, s->prop, s->upd, s->pr, s->F, source_slot);
  GRD_HIP(hipGetLastError());
  GRD_TRY(op_zero_counters(s));
  s->hi = (s->algo == GRD_SSSP && s->prm.delta) ? (uint64_t)s->prm.delta : ~0ull;
  // initial frontier statistics for the first schedule decision: PR starts
  // from every vertex with out-edges, CC from every vertex, SSSP from the source
  uint64_t fr = 0, fe = 0;
  uint32_t mx = 0;
  if (s->algo == GRD_PR) {
    fr = g->n - g->n_sinks;
    fe = g->m;
    mx = g->max_deg;
  } else if (s->algo == GRD_CC) {
    fr = g->n;
    fe = g->m;
    mx = g->max_deg;
  } else {
    fr = 1;
    fe = mx = g->h_deg[s->prm.source];
  }
  s->last_frontier = fr;
  s->prev_frontier = fr;
  s->last_frontier_edges = fe;
  s->last_maxdeg = mx;
  s->last_dirty = 0;
  s->iteration = 0;
  s->bytes = 0;
  s->edges = 0;
  s->bcast_option = -1;
  s->sel_queued = false;
  s->xb_dense = s->xr_dense = -1;
  if (g->C > 1) {  // per-partition frontier sizes for the first NDPEngine vote
    GRD_TRY(op_frontier_parts(s));
    GRD_TRY(op_read_counters(s));
  }
  take_frontier_parts(s);
  return GRD_OK;
}

// translation-table words of every pair (for the exchange byte count)
static uint64_t bitvector_bytes(const grd_graph *g)
{
  uint64_t b = 0;
  for (int c = 0; c < g->C; c++)
    for (int m = 0; m < g->M; m++) {
      uint64_t T = (uint64_t)g->ccnt[(size_t)c * g->M + m] + g->xcnt[(size_t)m * g->C + c];
      b += 8 * ((T + 63) / 64);
    }
  return b;
}

__global__ void __launch_bounds__(kBlock) k_min_u32(uint64_t n, const uint32_t *__restrict__ x, uint32_t *out)
{
  uint32_t lo = 0xFFFFFFFFu;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
    lo = x[i] < lo ? x[i] : lo;
  lo = block_min(lo);
  if (threadIdx.x == 0 && lo != 0xFFFFFFFFu) atomicMin(out, lo);
}

// the smallest arc weight of a weighted graph (once per graph; 2^32 - 1
// when it has no arcs)
static grd_status graph_min_weight(grd_graph *g, uint32_t *out)
{
  if (!g->wmin_known) {
    hipStream_t st = g->ctx->stream;
    DevMem tmp;
    uint32_t *d = nullptr;
    GRD_TRY(tmp.alloc_n(&d, 1));
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(kBlock), 0, st, d, 1, 0xFFFFFFFFu);
    if (g->m && g->cw)
      hipLaunchKernelGGL(k_min_u32, dim3(std::min<uint64_t>((g->m + kBlock - 1) / kBlock, (uint64_t)g->ctx->num_cus * 8)),
                         dim3(kBlock), 0, st, g->m, g->cw, d);
    GRD_HIP(hipGetLastError());
    GRD_HIP(hipMemcpyAsync(&g->wmin, d, 4, hipMemcpyDeviceToHost, st));
    GRD_HIP(hipStreamSynchronize(st));
    g->wmin_known = true;
  }
  *out = g->wmin;
  return GRD_OK;
}

grd_status op_broadcast(grd_state *s, int option)
{
  grd_graph *g = s->g;
  hipStream_t st = g->ctx->stream;
  const uint32_t nwS = g->S / 64;
  const bool pull = option & 1, dense = option & 2;
  const bool need_rank = !dense || !pull;
  s->xb_dense = dense ? 1 : 0;
  s->xr_dense = -1;
  if (need_rank) GRD_TRY(word_ranks(st, s->F, nwS, s->blk, s->wrankB, s->totals + 0));
  const uint32_t grid = group_grid(nwS);
  // a pull of CC / SSSP also sends the smallest value broadcast and the rows
  // that cannot improve (the pull leaves out what cannot change unless the
  // caller asked for every row); the rows-done bitmap not in the first
  // superstep, where every vertex still holds its initial value and at most
  // the frontier's smallest is done
  const bool bound = pull && s->algo != GRD_PR && !s->prm.pull_every_row;
  const bool rows_done = bound && s->iteration > 0;
  hipLaunchKernelGGL(k_bcast_send<uint32_t>, dim3(bound ? loop_grid(g->ctx, nwS) : grid), dim3(kBlock), 0, st,
                     nwS, s->F, (const uint32_t *)s->prop, s->wrankB, (int)dense, s->bvB, (uint32_t *)s->valB,
                     bound ? &s->cnt->fmin : nullptr);
  const uint32_t ident = (s->algo == GRD_PR) ? 0u : 0xFFFFFFFFu;  // 0.0f or the min identity
  hipLaunchKernelGGL(k_bcast_recv<uint32_t>, dim3(grid), dim3(kBlock), 0, st, nwS, s->bvB, (const uint32_t *)s->valB,
                     s->wrankB, g->word_dst, (int)dense, (int)!pull, ident, (uint32_t *)s->mprop, s->AL);
  const uint32_t nwP = g->P / 64;
  uint32_t inc = 0;  // the smallest Op::edge increment over an arc
  if (bound && s->algo == GRD_SSSP) {
    if (!s->prm.weighted) inc = 1;
    else GRD_TRY(graph_min_weight(g, &inc));
  }
  if (rows_done && nwP)
    hipLaunchKernelGGL(k_rows_done, dim3(loop_grid(g->ctx, nwP)), dim3(kBlock), 0, st, nwP, g->proxy_slot,
                       (const uint32_t *)s->prop, s->cnt, inc, s->Rdone);
  GRD_HIP(hipGetLastError());
  s->rdone_sent = rows_done;
  s->bound_sent = bound;
  s->bound_inc = inc;
  // exchange bytes (GraphAlgorithm.cpp:251-265): bitvector of every pair +
  // packed values (sparse) or the pair's full own slice (dense), and the
  // rows-done bitmap of a CC / SSSP pull
  uint64_t vals = dense ? (uint64_t)g->n : s->last_frontier;
  s->bytes += bitvector_bytes(g) + vals * s->vbytes + (rows_done ? 8ull * nwP : 0ull);
  return GRD_OK;
}

template <class Op>
static grd_status traverse_t(grd_state *s, int option)
{
  grd_graph *g = s->g;
  grd_ctx *ctx = g->ctx;
  hipStream_t st = ctx->stream;
  using V = typename Op::V;
  const uint32_t *wt = std::is_same<Op, OpWeighted>::value ? g->w : nullptr;
  const uint32_t *cwt = std::is_same<Op, OpWeighted>::value ? g->cw : nullptr;
  const bool timed = s->prm.time_kernel != 0;
  if (option & 1) {
    if (timed) GRD_HIP(hipEventRecord(s->kev[0], st));
    GRD_TRY(bsp_pull_launch<Op>(s, cwt));
    if (timed) {
      GRD_HIP(hipEventRecord(s->kev[1], st));
      s->kern_pending = 1;
      // a column and a gathered value per arc, the weight as stored (one
      // byte in the hot / cold split when every weight fits), a value per
      // row, and the rows-done bitmap; the arcs and rows of the blocks left
      // out come off when the counters are read
      const uint64_t wb = !cwt ? 0 : (bsp_hot_cold(s) && g->cow8) ? 1 : 4;
      s->kern_pending_bytes = (8 + wb) * g->m + 8ull * g->P + (s->rdone_sent ? g->P / 8 : 0);
      s->skip_arc_bytes = 8 + wb;
    }
    GRD_TRY(bsp_pull_fixup<Op>(s));
    // exact gather count (outside the events; time_kernel 2 leaves it out)
    if (timed && s->skip_pending && s->prm.time_kernel != 2) GRD_TRY(pull_done_arcs(s));
    s->edges += g->m;  // every arc of the memory partitions, less those left out (counters)
  } else {
    uint32_t want = (uint32_t)std::min<uint64_t>((s->last_frontier + kBlock - 1) / kBlock, (uint64_t)ctx->num_cus * 8);
    uint32_t grid = std::max<uint32_t>(1, want);
    if (timed) GRD_HIP(hipEventRecord(s->kev[0], st));
    hipLaunchKernelGGL((k_push<Op>), dim3(grid), dim3(kBlock), 0, st, s->AL, s->totals + 0, g->rp, g->col, wt,
                       (const V *)s->mprop, (V *)s->mupd, s->Du, s->heavy, s->cnt);
    hipLaunchKernelGGL((k_push_heavy<Op>), dim3(ctx->num_cus * 4), dim3(kBlock), 0, st, s->heavy, s->cnt, g->rp, g->col,
                       wt, (const V *)s->mprop, (V *)s->mupd, s->Du);
    if (timed) {
      GRD_HIP(hipEventRecord(s->kev[1], st));
      s->kern_pending = 2;  // bytes from the edge counters after the iteration
    }
  }
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

grd_status op_traverse(grd_state *s, int option)
{
  switch (s->algo) {
  case GRD_PR: return traverse_t<OpPR>(s, option);
  case GRD_CC: return traverse_t<OpCC>(s, option);
  case GRD_SSSP: return s->prm.weighted ? traverse_t<OpWeighted>(s, option) : traverse_t<OpHop>(s, option);
  default: return fail(GRD_ERR_STATE, "traverse: algorithm has no BSP traversal");
  }
}

template <class Op>
static grd_status reduce_t(grd_state *s, int option)
{
  grd_graph *g = s->g;
  hipStream_t st = g->ctx->stream;
  using V = typename Op::V;
  const uint32_t nwP = g->P / 64;
  const bool dense = option & 2;
  s->xr_dense = dense ? 1 : 0;
  if (!dense) GRD_TRY(word_ranks(st, s->Du, nwP, s->blk, s->wrankR, s->totals + 1));
  const uint32_t grid = group_grid(nwP);
  const uint32_t grid_send = loop_grid(g->ctx, nwP);
  hipLaunchKernelGGL(k_reduce_send<V>, dim3(grid_send), dim3(kBlock), 0, st, nwP, s->Du, (V *)s->mupd, s->wrankR,
                     (int)dense, (int)(s->algo == GRD_PR), s->bvR, (V *)s->valR, s->cnt);
  hipLaunchKernelGGL((k_reduce_recv<Op>), dim3(grid), dim3(kBlock), 0, st, nwP, s->bvR, (const V *)s->valR, s->wrankR,
                     g->proxy_slot, (int)dense, (V *)s->upd, s->Dc);
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

grd_status op_reduce(grd_state *s, int option)
{
  switch (s->algo) {
  case GRD_PR: return reduce_t<OpPR>(s, option);
  case GRD_CC: return reduce_t<OpCC>(s, option);
  case GRD_SSSP: return reduce_t<OpHop>(s, option);  // combine = min for both SSSP forms
  default: return fail(GRD_ERR_STATE, "reduce: algorithm has no BSP exchange");
  }
}

grd_status op_update(grd_state *s)
{
  grd_graph *g = s->g;
  hipStream_t st = g->ctx->stream;
  const uint32_t nwS = g->S / 64;
  const uint32_t grid = loop_grid(g->ctx, nwS);
  const int use_delta = (s->algo == GRD_SSSP && s->prm.delta) ? 1 : 0;
  switch (s->algo) {
  case GRD_PR:
    hipLaunchKernelGGL(k_update<GRD_PR>, dim3(grid), dim3(kBlock), 0, st, nwS, s->Dc, s->prop, s->upd, s->pr,
                       g->slot_deg, s->F, s->Pend, (unsigned long long)s->hi, 0, s->cnt);
    break;
  case GRD_CC:
    hipLaunchKernelGGL(k_update<GRD_CC>, dim3(grid), dim3(kBlock), 0, st, nwS, s->Dc, s->prop, s->upd, s->pr,
                       g->slot_deg, s->F, s->Pend, (unsigned long long)s->hi, 0, s->cnt);
    break;
  case GRD_SSSP:
    hipLaunchKernelGGL(k_update<GRD_SSSP>, dim3(grid), dim3(kBlock), 0, st, nwS, s->Dc, s->prop, s->upd, s->pr,
                       g->slot_deg, s->F, s->Pend, (unsigned long long)s->hi, use_delta, s->cnt);
    break;
  default: return fail(GRD_ERR_STATE, "update: algorithm has no BSP update");
  }
  GRD_HIP(hipGetLastError());
  if (g->C > 1) GRD_TRY(op_frontier_parts(s));
  return GRD_OK;
}

void take_frontier_parts(grd_state *s)
{
  if (s->g->C == 1) {
    s->last_fpart[0] = s->last_frontier;
    s->last_uncovered = 0;
    return;
  }
  for (int c = 0; c < s->g->C; c++) s->last_fpart[c] = s->hcnt->fpart[c];
  s->last_uncovered = s->hcnt->frontier_uncovered;
}

// Delta-stepping bucket advance when the frontier came out empty.  Reads the
// counters (synchronizes) and launches k_select.  *any_pending = frontier
// non-empty afterwards.
grd_status op_select_bucket(grd_state *s, bool *any_pending)
{
  grd_graph *g = s->g;
  hipStream_t st = g->ctx->stream;
  const uint32_t nwS = g->S / 64;
  const uint32_t grid = loop_grid(g->ctx, nwS);
  *any_pending = false;
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, s->cnt);
  hipLaunchKernelGGL(k_pend_min, dim3(grid), dim3(kBlock), 0, st, nwS, s->Pend, (const uint32_t *)s->prop, s->cnt);
  GRD_HIP(hipGetLastError());
  GRD_TRY(op_read_counters(s));
  if (s->hcnt->pending == 0) return GRD_OK;
  uint64_t mn = s->hcnt->pend_min, d = s->prm.delta;
  s->hi = (mn / d + 1) * d;
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, st, s->cnt);
  hipLaunchKernelGGL(k_select, dim3(grid), dim3(kBlock), 0, st, nwS, s->Pend, (const uint32_t *)s->prop, g->slot_deg,
                     (unsigned long long)s->hi, s->F, s->cnt);
  GRD_HIP(hipGetLastError());
  if (g->C > 1) GRD_TRY(op_frontier_parts(s));
  GRD_TRY(op_read_counters(s));
  *any_pending = s->hcnt->frontier != 0;
  return GRD_OK;
}

grd_status op_aggregate_one(grd_state *s, uint32_t slot, const void *value)
{
  hipStream_t st = s->g->ctx->stream;
  if (s->algo == GRD_PR) {
    float x;
    memcpy(&x, value, 4);
    hipLaunchKernelGGL(k_aggregate_one<OpPR>, dim3(1), dim3(64), 0, st, (float *)s->upd, s->Dc, slot, x);
  } else {
    uint32_t x;
    memcpy(&x, value, 4);
    hipLaunchKernelGGL(k_aggregate_one<OpCC>, dim3(1), dim3(64), 0, st, (uint32_t *)s->upd, s->Dc, slot, x);
  }
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

grd_status op_download(grd_state *s, void *out)
{
  grd_graph *g = s->g;
  hipStream_t st = g->ctx->stream;
  void *dout = nullptr;
  GRD_HIP(hipMallocAsync(&dout, (uint64_t)g->n * 4, st));
  const void *src = (s->algo == GRD_PR || s->algo == GRD_PR_PULL) ? (const void *)s->pr : s->prop;
  hipLaunchKernelGGL(k_gather_result<uint32_t>, dim3(std::min<uint32_t>(4096, g->S / kBlock + 1)), dim3(kBlock), 0, st,
                     g->S, g->slot_gid, (const uint32_t *)src, (uint32_t *)dout);
  GRD_HIP(hipGetLastError());
  GRD_HIP(hipMemcpyAsync(out, dout, (uint64_t)g->n * 4, hipMemcpyDeviceToHost, st));
  GRD_HIP(hipFreeAsync(dout, st));
  GRD_HIP(hipStreamSynchronize(st));
  if (s->wide) {
    // 8-byte properties: widen in place from the top, the unreached 2^32 - 1
    // to 2^64 - 1 (numeric_limits<uint64_t>::max)
    const uint32_t *in = (const uint32_t *)out;
    uint64_t *o = (uint64_t *)out;
    for (uint64_t i = g->n; i-- > 0;) {
      const uint32_t x = in[i];
      o[i] = x == 0xFFFFFFFFu ? ~0ull : (uint64_t)x;
    }
  }
  return GRD_OK;
}

}  // namespace grd
