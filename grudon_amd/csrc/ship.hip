// ship.hip -- the NO_OFFLOAD edge-shipping schedule and the inputs of its
// NDPEngine vote.
//
// Reference: src/GraphAlgorithm.cpp:426-474 (memory side: for every vertex
// the broadcast updated, send [gid, gid of each out-neighbour] to its master)
// and :582-637 (compute side: receive the lists of its frontier and aggregate
// the source's property into every listed neighbour); the vote is
// include/offload_engine/NDPEngine.hpp:10-51 with the coverage vector of
// src/DistributedGraph.cpp:286-370.
//
// On one GPU the wire is one buffer of records in active-list order (the
// active list is grouped by (compute, memory) pair, so each pair's records
// are contiguous).  A record names the source and its out-neighbours by
// compute slot, the receiver's local id: the sender translates each proxy
// with one lookup of the proxy -> slot table (the table the offloaded reduce
// uses too), where gids would cost proxy -> gid on the way out and gid ->
// slot on the way in.  An entry is 4 bytes either way.
//   k_ship_len    record length (1 + out-degree) of every active proxy
//   (rocPRIM exclusive scan: record offsets)
//   k_ship_tiles  the record holding the first position of every tile
//   k_ship_send   memory side: writes the records
//   k_ship_recv   compute side: aggregate() of the source value
// Both wire passes are merge-path tiled: a tile of kShipTile positions stages
// the offsets of its records in LDS and each thread finds the record of its
// position by binary search there, so the work per thread is even whatever
// the degree spread, and the wire is read and written with unit stride.
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstring>
#include <type_traits>

#include "grd_device.hpp"
#include "grd_internal.hpp"
#include "grd_ops.hpp"

namespace grd {

// ------------------------------------------------------------------------
// coverage: a vertex is covered when all of its out-neighbours share its
// master, so that its compute partition can apply every update of its list
// ------------------------------------------------------------------------

// One thread per compute slot; a wave covers exactly one slot word.  The own
// proxy of slot w*64+l is word_dst[w]*64+l (the own region repeats the slot
// order), and it holds all of the vertex's out-arcs.
__global__ void __launch_bounds__(kBlock) k_uncovered(uint32_t S, uint32_t C, const uint32_t *__restrict__ slot_gid,
                                                      const uint32_t *__restrict__ word_dst,
                                                      const uint32_t *__restrict__ rp, const uint32_t *__restrict__ col,
                                                      const uint32_t *__restrict__ proxy_slot,
                                                      uint64_t *__restrict__ uncov)
{
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < S; s += gridDim.x * kBlock) {
    const uint32_t v = slot_gid[s];
    bool unc = false;
    if (v != kNone) {
      const uint32_t pw = word_dst[s >> 6];
      if (pw != kNone) {
        const uint32_t p = pw * 64 + (s & 63), mc = v % C;
        for (uint32_t e = rp[p]; e < rp[p + 1]; e++)
          if (slot_gid[proxy_slot[col[e]]] % C != mc) {
            unc = true;
            break;
          }
      }
    }
    const uint64_t m = __ballot(unc);
    if (lane_id() == 0) uncov[s >> 6] = m;
  }
}

// ------------------------------------------------------------------------
// per-partition frontier sizes and uncovered frontier vertices
// ------------------------------------------------------------------------

__global__ void __launch_bounds__(kBlock) k_frontier_parts(uint32_t nw, const uint64_t *__restrict__ F,
                                                           const uint8_t *__restrict__ word_c,
                                                           const uint64_t *__restrict__ uncov, uint32_t C,
                                                           Counters *cnt)
{
  __shared__ unsigned long long s_fc[kMaxParts];
  for (uint32_t i = threadIdx.x; i < (uint32_t)kMaxParts; i += kBlock) s_fc[i] = 0;
  __syncthreads();
  uint64_t unc = 0;
  for (uint32_t w = blockIdx.x * kBlock + threadIdx.x; w < nw; w += gridDim.x * kBlock) {
    const uint64_t f = F[w];
    if (f) {
      atomicAdd(&s_fc[word_c[w]], (unsigned long long)__popcll(f));
      unc += (uint64_t)__popcll(f & uncov[w]);
    }
  }
  __syncthreads();
  unc = block_sum(unc);
  if (threadIdx.x == 0 && unc) atomicAdd(&cnt->frontier_uncovered, (unsigned long long)unc);
  for (uint32_t c = threadIdx.x; c < C; c += kBlock)
    if (s_fc[c]) atomicAdd(&cnt->fpart[c], s_fc[c]);
}

// ------------------------------------------------------------------------
// the wire
// ------------------------------------------------------------------------

// Record lengths of the k active proxies the host sized the wire for, and
// len[k] = 0 so that the exclusive scan ends with the wire size.  If the
// device's active list holds another count, every length is zero and the
// error flag is raised; the later passes then find a wire size that differs
// from the host's and do nothing.
__global__ void __launch_bounds__(kBlock) k_ship_len(const uint32_t *__restrict__ AL, const uint32_t *__restrict__ nact,
                                                     uint32_t k, const uint32_t *__restrict__ rp,
                                                     uint32_t *__restrict__ len, Counters *cnt)
{
  const bool ok = *nact == k;
  if (!ok && blockIdx.x == 0 && threadIdx.x == 0) cnt->ship_error = 1;
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i <= k; i += gridDim.x * kBlock) {
    uint32_t L = 0;
    if (ok && i < k) {
      const uint32_t p = AL[i];
      L = rp[p + 1] - rp[p] + 1;
    }
    len[i] = L;
  }
}

// tile[t] = the record holding wire position t * kShipTile.  Every record has
// at least one position (its header), so each tile start is written once.
__global__ void __launch_bounds__(kBlock) k_ship_tiles(const uint32_t *__restrict__ off, uint32_t k, uint32_t total,
                                                       uint32_t *__restrict__ tile, Counters *cnt)
{
  if (off[k] != total) {
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt->ship_error = 1;
    return;
  }
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < k; i += gridDim.x * kBlock) {
    const uint32_t a = off[i], b = off[i + 1];
    for (uint64_t t = (a + (uint64_t)kShipTile - 1) / kShipTile; t * kShipTile < b; t++) tile[t] = i;
  }
}

// Records of tile t: [tile[t], last], at most kShipTile + 1 of them, since all
// records after the first start inside the tile.  Stages their offsets in
// LDS; returns the record count (0: skip the tile).
__device__ __forceinline__ uint32_t ship_tile_records(const uint32_t *__restrict__ off,
                                                      const uint32_t *__restrict__ tile, uint32_t k, uint32_t t,
                                                      uint32_t ntiles, uint32_t *s_off, uint32_t *first)
{
  const uint32_t a = tile[t], b = (t + 1 < ntiles) ? tile[t + 1] : k - 1;
  if (b < a || b >= k || b - a + 1 > kShipTile + 1) return 0;
  const uint32_t nr = b - a + 1;
  for (uint32_t i = threadIdx.x; i <= nr; i += kBlock) s_off[i] = off[a + i];
  *first = a;
  return nr;
}

// the record of wire position q among the staged records: last i with s_off[i] <= q
__device__ __forceinline__ uint32_t ship_owner(const uint32_t *s_off, uint32_t nr, uint32_t q)
{
  uint32_t lo = 0, hi = nr;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (s_off[mid] <= q) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ void __launch_bounds__(kBlock) k_ship_wmax(uint64_t m, const uint32_t *__restrict__ w, uint32_t *__restrict__ out)
{
  uint32_t mx = 0;
  for (uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x; e < m; e += (uint64_t)gridDim.x * kBlock) mx = max(mx, w[e]);
  for (int o = kWave / 2; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, o, kWave));
  if (lane_id() == 0 && mx) atomicMax(out, mx);
}

__global__ void __launch_bounds__(kBlock) k_ship_w8(uint64_t m, const uint32_t *__restrict__ w, uint8_t *__restrict__ w8)
{
  for (uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x; e < m; e += (uint64_t)gridDim.x * kBlock)
    w8[e] = (uint8_t)w[e];
}

__global__ void __launch_bounds__(kBlock) k_slot_col(uint64_t m, const uint32_t *__restrict__ col,
                                                     const uint32_t *__restrict__ proxy_slot, uint32_t *__restrict__ scol)
{
  for (uint64_t e = (uint64_t)blockIdx.x * kBlock + threadIdx.x; e < m; e += (uint64_t)gridDim.x * kBlock)
    scol[e] = proxy_slot[col[e]];
}

// Memory side (GraphAlgorithm.cpp:426-474): position 0 of a record is the
// source's slot, position r > 0 the slot of its r-th out-neighbour (with the
// arc weight in wbuf for weighted SSSP; the header's weight slot is 0).
// The out-neighbours' slots come from scol, the push CSR with every column
// translated to its compute slot once (a streaming read per shipped arc, not
// a gather from the proxy -> slot table).
// W: the weights' storage, a byte when every weight of the graph fits.
template <class W>
__global__ void __launch_bounds__(kBlock) k_ship_send(const uint32_t *__restrict__ AL, const uint32_t *__restrict__ off,
                                                      const uint32_t *__restrict__ tile, uint32_t k, uint32_t total,
                                                      uint32_t ntiles, const uint32_t *__restrict__ rp,
                                                      const uint32_t *__restrict__ scol, const W *__restrict__ wt,
                                                      const uint32_t *__restrict__ proxy_slot,
                                                      uint32_t *__restrict__ buf, W *__restrict__ wbuf)
{
  __shared__ uint32_t s_off[kShipTile + 2];
  __shared__ uint32_t s_beg[kShipTile + 1];
  __shared__ uint32_t s_src[kShipTile + 1];
  if (off[k] != total) return;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    uint32_t a = 0;
    const uint32_t nr = ship_tile_records(off, tile, k, t, ntiles, s_off, &a);
    for (uint32_t i = threadIdx.x; i < nr; i += kBlock) {
      const uint32_t p = AL[a + i];
      s_beg[i] = rp[p];
      s_src[i] = proxy_slot[p];
    }
    __syncthreads();
    const uint32_t q0 = t * kShipTile, q1 = min(total, q0 + kShipTile);
    if (nr)
      for (uint32_t q = q0 + threadIdx.x; q < q1; q += kBlock) {
        const uint32_t i = ship_owner(s_off, nr, q);
        const uint32_t r = q - s_off[i];
        uint32_t x;
        W wx = 0;
        if (r == 0) {
          x = s_src[i];
        } else {
          const uint32_t e = s_beg[i] + r - 1;
          x = scol[e];
          if (wbuf) wx = wt[e];
        }
        buf[q] = x;
        if (wbuf) wbuf[q] = wx;
      }
    __syncthreads();
  }
}

// Compute side (GraphAlgorithm.cpp:582-637): aggregate() of the source's
// property along each listed arc (the entries are already local ids, what
// getLocalNode returns there).  The reference
// applies addUpdate here whatever the algorithm (a TODO at :624); this applies
// the algorithm's own aggregate, which is what makes the schedule give the
// same result as the offloaded one.
template <class Op, class W>
__global__ void __launch_bounds__(kBlock) k_ship_recv(const uint32_t *__restrict__ off,
                                                      const uint32_t *__restrict__ tile, uint32_t k, uint32_t total,
                                                      uint32_t ntiles, const uint32_t *__restrict__ buf,
                                                      const W *__restrict__ wbuf,
                                                      const typename Op::V *__restrict__ prop, typename Op::V *upd,
                                                      uint64_t *Dc)
{
  using V = typename Op::V;
  __shared__ uint32_t s_off[kShipTile + 2];
  __shared__ V s_val[kShipTile + 1];
  if (off[k] != total) return;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    uint32_t a = 0;
    const uint32_t nr = ship_tile_records(off, tile, k, t, ntiles, s_off, &a);
    (void)a;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nr; i += kBlock) s_val[i] = prop[buf[s_off[i]]];
    __syncthreads();
    const uint32_t q0 = t * kShipTile, q1 = min(total, q0 + kShipTile);
    if (nr)
      for (uint32_t q = q0 + threadIdx.x; q < q1; q += kBlock) {
        const uint32_t i = ship_owner(s_off, nr, q);
        if (q == s_off[i]) continue;  // the record's header
        relax<Op, true>(upd, Dc, buf[q], Op::edge(s_val[i], wbuf ? (uint32_t)wbuf[q] : 1u));
      }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------

static uint32_t grid_for(uint64_t n, uint32_t cap = 4096)
{
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(cap, (n + kBlock - 1) / kBlock));
}

grd_status build_ship_tables(grd_graph *g)
{
  hipStream_t st = g->ctx->stream;
  const uint32_t nwS = g->S / 64;
  // compute partition of every slot word: pair (c, m) owns its own count
  // rounded up to whole words from cbase[c * M + m]
  std::vector<uint8_t> wc(std::max<uint32_t>(nwS, 1), 0);
  for (int c = 0; c < g->C; c++)
    for (int m = 0; m < g->M; m++) {
      const size_t pi = (size_t)c * g->M + m;
      const uint64_t w0 = g->cbase[pi] / 64, w1 = (g->cbase[pi] + (uint64_t)g->ccnt[pi] + 63) / 64;
      for (uint64_t w = w0; w < w1 && w < nwS; w++) wc[w] = (uint8_t)c;
    }
  GRD_TRY(g->mem.alloc_n(&g->word_c, wc.size()));
  GRD_HIP(hipMemcpyAsync(g->word_c, wc.data(), wc.size(), hipMemcpyHostToDevice, st));
  GRD_TRY(g->mem.alloc_n(&g->uncov, std::max<uint32_t>(nwS, 1)));
  g->n_uncovered = 0;
  if (g->C == 1) {
    GRD_HIP(hipMemsetAsync(g->uncov, 0, (uint64_t)std::max<uint32_t>(nwS, 1) * 8, st));
  } else {
    hipLaunchKernelGGL(k_uncovered, dim3(grid_for(g->S, 1u << 16)), dim3(kBlock), 0, st, g->S, (uint32_t)g->C,
                       g->slot_gid, g->word_dst, g->rp, g->col, g->proxy_slot, g->uncov);
    GRD_HIP(hipGetLastError());
    std::vector<uint64_t> h(nwS);
    GRD_HIP(hipMemcpyAsync(h.data(), g->uncov, (uint64_t)nwS * 8, hipMemcpyDeviceToHost, st));
    GRD_HIP(hipStreamSynchronize(st));
    for (uint64_t x : h) g->n_uncovered += (uint64_t)__builtin_popcountll(x);
  }
  GRD_HIP(hipStreamSynchronize(st));
  return GRD_OK;
}

grd_status op_frontier_parts(grd_state *s)
{
  grd_graph *g = s->g;
  const uint32_t nwS = g->S / 64;
  hipLaunchKernelGGL(k_frontier_parts, dim3(grid_for(nwS, 1024)), dim3(kBlock), 0, g->ctx->stream, nwS, s->F,
                     g->word_c, g->uncov, (uint32_t)g->C, s->cnt);
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

// wire buffers for k records over `total` positions; grown (never shrunk)
// outside any launch in flight
static grd_status ship_reserve(grd_state *s, uint64_t k, uint64_t total, bool weighted, uint32_t wbytes)
{
  if (k + 1 <= s->ship_kcap && total <= s->ship_cap && (!weighted || s->ship_w)) return GRD_OK;
  hipStream_t st = s->g->ctx->stream;
  GRD_HIP(hipStreamSynchronize(st));
  s->ship_mem.release();
  s->ship_off = s->ship_tile = s->ship_buf = nullptr;
  s->ship_w = nullptr;
  s->ship_tmp = nullptr;
  s->ship_tmp_bytes = 0;
  const uint64_t kc = std::max<uint64_t>(k + 1 + (k + 1) / 4, 1024), cap = std::max<uint64_t>(total + total / 4, 4096);
  const uint64_t tc = (cap + kShipTile - 1) / kShipTile + 1;
  // lengths and offsets share one allocation: [len | off]
  GRD_TRY(s->ship_mem.alloc_n(&s->ship_off, 2 * kc));
  GRD_TRY(s->ship_mem.alloc_n(&s->ship_tile, tc));
  GRD_TRY(s->ship_mem.alloc_n(&s->ship_buf, cap));
  if (weighted) GRD_TRY(s->ship_mem.alloc(&s->ship_w, cap * wbytes));
  s->ship_kcap = kc;
  s->ship_cap = cap;
  return GRD_OK;
}

template <class Op>
static grd_status ship_t(grd_state *s)
{
  using V = typename Op::V;
  grd_graph *g = s->g;
  grd_ctx *ctx = g->ctx;
  hipStream_t st = ctx->stream;
  const uint64_t k = s->last_frontier, E = s->last_frontier_edges;
  if (k == 0) return GRD_OK;
  const uint64_t total = k + E;
  if (total > 0xFFFFFFFFull - kShipTile) return fail(GRD_ERR_ARG, "edge shipping: the wire would exceed 2^32 entries");
  const bool weighted = std::is_same<Op, OpWeighted>::value;
  if (weighted && !g->ship_wbytes) {
    // the weights travel as bytes when every one of them is below 256
    uint32_t *d_max = nullptr, wmax = 0;
    GRD_TRY(g->mem.alloc_n(&d_max, 1));
    GRD_HIP(hipMemsetAsync(d_max, 0, 4, st));
    hipLaunchKernelGGL(k_ship_wmax, dim3(grid_for(g->m, 1u << 16)), dim3(kBlock), 0, st, g->m, g->w, d_max);
    GRD_HIP(hipMemcpyAsync(&wmax, d_max, 4, hipMemcpyDeviceToHost, st));
    GRD_HIP(hipStreamSynchronize(st));
    if (wmax < 256) {
      GRD_TRY(g->mem.alloc_n(&g->w8, std::max<uint64_t>(g->m, 1)));
      hipLaunchKernelGGL(k_ship_w8, dim3(grid_for(g->m, 1u << 16)), dim3(kBlock), 0, st, g->m, g->w, g->w8);
      GRD_HIP(hipGetLastError());
    }
    g->ship_wbytes = wmax < 256 ? 1 : 4;
  }
  const uint32_t wbytes = weighted ? (uint32_t)g->ship_wbytes : 4u;
  GRD_TRY(ship_reserve(s, k, total, weighted, wbytes));
  uint32_t *len = s->ship_off + s->ship_kcap, *off = s->ship_off;
  const uint32_t ntiles = (uint32_t)((total + kShipTile - 1) / kShipTile);
  hipLaunchKernelGGL(k_ship_len, dim3(grid_for(k + 1)), dim3(kBlock), 0, st, s->AL, s->totals + 0, (uint32_t)k, g->rp,
                     len, s->cnt);
  size_t tb = 0;
  GRD_HIP(rocprim::exclusive_scan(nullptr, tb, len, off, 0u, (size_t)k + 1, rocprim::plus<uint32_t>(), st));
  if (tb > s->ship_tmp_bytes) {  // scan storage of the wire's own allocation
    GRD_HIP(hipStreamSynchronize(st));
    GRD_TRY(s->ship_mem.alloc(&s->ship_tmp, tb));
    s->ship_tmp_bytes = tb;
  }
  GRD_HIP(rocprim::exclusive_scan(s->ship_tmp, tb, len, off, 0u, (size_t)k + 1, rocprim::plus<uint32_t>(), st));
  hipLaunchKernelGGL(k_ship_tiles, dim3(grid_for(k)), dim3(kBlock), 0, st, off, (uint32_t)k, (uint32_t)total,
                     s->ship_tile, s->cnt);
  const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(ntiles, (uint32_t)ctx->num_cus * 8));
  if (!g->scol) {  // the push CSR's columns as compute slots, once per graph
    GRD_TRY(g->mem.alloc_n(&g->scol, std::max<uint64_t>(g->m, 1)));
    hipLaunchKernelGGL(k_slot_col, dim3(grid_for(g->m, 1u << 16)), dim3(kBlock), 0, st, g->m, g->col, g->proxy_slot,
                       g->scol);
    GRD_HIP(hipGetLastError());
  }
  const bool timed = s->prm.time_kernel != 0;
  if (timed) GRD_HIP(hipEventRecord(s->kev[0], st));
  if (weighted && wbytes == 1) {
    hipLaunchKernelGGL(k_ship_send<uint8_t>, dim3(grid), dim3(kBlock), 0, st, s->AL, off, s->ship_tile, (uint32_t)k,
                       (uint32_t)total, ntiles, g->rp, g->scol, (const uint8_t *)g->w8, g->proxy_slot, s->ship_buf,
                       (uint8_t *)s->ship_w);
    hipLaunchKernelGGL((k_ship_recv<Op, uint8_t>), dim3(grid), dim3(kBlock), 0, st, off, s->ship_tile, (uint32_t)k,
                       (uint32_t)total, ntiles, s->ship_buf, (const uint8_t *)s->ship_w, (const V *)s->prop,
                       (V *)s->upd, s->Dc);
  } else {
    hipLaunchKernelGGL(k_ship_send<uint32_t>, dim3(grid), dim3(kBlock), 0, st, s->AL, off, s->ship_tile, (uint32_t)k,
                       (uint32_t)total, ntiles, g->rp, g->scol, weighted ? (const uint32_t *)g->w : nullptr,
                       g->proxy_slot, s->ship_buf, weighted ? (uint32_t *)s->ship_w : nullptr);
    hipLaunchKernelGGL((k_ship_recv<Op, uint32_t>), dim3(grid), dim3(kBlock), 0, st, off, s->ship_tile, (uint32_t)k,
                       (uint32_t)total, ntiles, s->ship_buf, weighted ? (const uint32_t *)s->ship_w : nullptr,
                       (const V *)s->prop, (V *)s->upd, s->Dc);
  }
  GRD_HIP(hipGetLastError());
  if (timed) {
    // per arc: its slot (the slot column), wire write + read, the atomic
    // read-modify-write (+ weight read, write, read, at its stored width);
    // per record: active list, row bounds, offsets (twice), slot lookup,
    // header write + read, property
    GRD_HIP(hipEventRecord(s->kev[1], st));
    s->kern_pending = 1;
    s->kern_pending_bytes = (20 + (weighted ? 3ull * wbytes : 0ull)) * E + 36 * k;
  }
  s->edges += E;
  // wire bytes (MPI.hpp:44 counts count * sizeof(GNode) per send): one id
  // per record header and per arc, plus the weights of a weighted SSSP
  s->bytes += 4 * total + (weighted ? 4 * E : 0);
  return GRD_OK;
}

grd_status op_ship(grd_state *s)
{
  switch (s->algo) {
  case GRD_PR: return ship_t<OpPR>(s);
  case GRD_CC: return ship_t<OpCC>(s);
  case GRD_SSSP: return s->prm.weighted ? ship_t<OpWeighted>(s) : ship_t<OpHop>(s);
  default: return fail(GRD_ERR_STATE, "edge shipping: algorithm has no BSP traversal");
  }
}

}  // namespace grd
