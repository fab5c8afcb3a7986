// layout.cpp -- partitioned graph construction (host) and upload to HBM.
//
// Restates the reference's DistributedGraph constructor
// (src/DistributedGraph.cpp:13-785) for all C compute and M memory partitions
// at once, with dense arrays instead of its per-rank unordered_maps:
//   master(g) = g % C                        Partitioner.hpp:32, :176
//   mirror(g) = g % M, or mirror[g] from a partition file when M > 1
//                                            Partitioner.hpp:33, :73-105, :177
//   cross mirrors of m = {v : u -> v, mirror(u) = m, mirror(v) != m}
//                                            DistributedGraph.cpp:345-349, :367-371
//   memory lids: own then cross (ascending gid)
//                                            DistributedGraph.cpp:517-525, :579-588
//   translation tables: compute side own-on-m then cross-on-m, memory side
//   own then cross per master, both ascending  DistributedGraph.cpp:644-660, :701-734
// The resulting slots/proxies are stored in the padded, pair-grouped order
// described in grd_internal.hpp: own vertices of a (master, mirror) pair by
// out-degree rank, cross mirrors by compute slot.  grd_graph_table still
// returns the reference's ascending tables.  The pull CSC rows are sorted by
// source.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <functional>
#include <new>
#include <thread>

#include "grd_internal.hpp"

namespace grd {

int host_threads(const grd_ctx *ctx)
{
  int t = ctx && ctx->host_threads > 0 ? ctx->host_threads : (int)std::thread::hardware_concurrency();
  if (t < 1) t = 1;
  if (ctx == nullptr || ctx->host_threads <= 0) t = std::min(t, 16);
  return t;
}

void parallel_for(int nthreads, uint64_t n, uint64_t grain, const std::function<void(uint64_t, uint64_t)> &fn)
{
  if (n == 0) return;
  if (grain == 0) grain = 1;
  if (nthreads <= 1 || n <= grain) {
    fn(0, n);
    return;
  }
  std::atomic<uint64_t> next{0};
  auto body = [&]() {
    for (;;) {
      uint64_t b = next.fetch_add(grain);
      if (b >= n) break;
      fn(b, std::min(n, b + grain));
    }
  };
  std::vector<std::thread> th;
  int nt = (int)std::min<uint64_t>((uint64_t)nthreads, (n + grain - 1) / grain);
  for (int t = 1; t < nt; t++) th.emplace_back(body);
  body();
  for (auto &x : th) x.join();
}

grd_status DevMem::alloc(void **p, uint64_t nbytes)
{
  *p = nullptr;
  uint64_t b = nbytes ? nbytes : 64;
  hipError_t e = hipMalloc(p, b);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(GRD_ERR_OOM, "hipMalloc(" + std::to_string(b) + " bytes) failed: " + hipGetErrorString(e));
  }
  ptrs.push_back(*p);
  bytes += b;
  return GRD_OK;
}

void DevMem::release()
{
  for (void *p : ptrs) (void)hipFree(p);
  ptrs.clear();
  bytes = 0;
}

template <typename T>
static grd_status upload(grd_graph *g, T **dst, const std::vector<T> &src)
{
  GRD_TRY(g->mem.alloc_n(dst, src.size()));
  if (!src.empty()) GRD_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return GRD_OK;
}

// Cut the CSC rows into pull row blocks: consecutive rows while the block
// holds at most kPullNnz arcs and kPullMaxRows rows; a row with more than
// kPullNnz arcs becomes segments of kPullNnz arcs (one block each) plus a
// PullLong entry that merges them in order.
grd_status build_row_blocks(grd_graph *g, const std::vector<uint32_t> &crp)
{
  std::vector<PullBlock> blks;
  std::vector<PullLong> longs;
  const uint32_t P = g->P;
  uint32_t r = 0;
  while (r < P) {
    uint32_t deg = crp[r + 1] - crp[r];
    if (deg > kPullNnz) {
      PullLong L{r, (uint32_t)blks.size(), 0, 0};
      for (uint32_t e = crp[r]; e < crp[r + 1]; e += kPullNnz) {
        blks.push_back(PullBlock{r, 0, e, std::min(crp[r + 1], e + kPullNnz)});
        L.nseg++;
      }
      longs.push_back(L);
      r++;
      continue;
    }
    const uint32_t r0 = r, e0 = crp[r];
    while (r < P && r - r0 < kPullMaxRows && crp[r + 1] - crp[r] <= kPullNnz && crp[r + 1] - e0 <= kPullNnz) r++;
    blks.push_back(PullBlock{r0, r - r0, e0, crp[r]});
  }
  if (blks.size() >= 0x7FFFFFFFull) return fail(GRD_ERR_ARG, "graph: too many pull row blocks");
  // hot window of each block: the first own proxies (highest out-degree) of
  // the memory partition its first row belongs to
  std::vector<uint32_t> hot(blks.size());
  {
    const int C = g->C, M = g->M;
    int m = 0;
    for (size_t i = 0; i < blks.size(); i++) {
      while (m + 1 < M && g->obase[(size_t)(m + 1) * C] <= blks[i].r0) m++;
      hot[i] = g->obase[(size_t)m * C];
    }
  }
  GRD_TRY(upload(g, &g->pblk_hot, hot));
  g->n_pblk = (uint32_t)blks.size();
  g->n_plong = (uint32_t)longs.size();
  if (longs.empty()) longs.push_back(PullLong{0, 0, 0, 0});
  GRD_TRY(upload(g, &g->pblk, blks));
  return upload(g, &g->plong, longs);
}

static grd_status build_impl(grd_ctx *ctx, uint32_t n, const uint32_t *rp, const uint32_t *col, const uint32_t *w,
                             int C, int M, const uint32_t *mirror_in, grd_graph **out)
{
  if (n == 0 || C < 1 || M < 1 || C > 64 || M > 64)
    return fail(GRD_ERR_ARG, "graph: need n >= 1 and 1 <= C, M <= 64");
  const uint64_t m = rp[n];
  if (m >= 0xFFFFFFF0ull) return fail(GRD_ERR_ARG, "graph: more than 2^32-16 arcs is beyond the u32 offset limit");
  GRD_TRY(check_row_ptr(n, rp));
  const int T = host_threads(ctx);
  const uint32_t NW = (n + 63) / 64;

  grd_graph *g = new grd_graph();
  g->ctx = ctx;
  g->n = n;
  g->m = m;
  g->C = C;
  g->M = M;
  g->weighted = (w != nullptr);
  g->avg_degree = (double)m / (double)n;

  // master / mirror (Partitioner.hpp)
  std::vector<uint8_t> mir(n);
  for (uint32_t v = 0; v < n; v++) {
    uint32_t mm = (M == 1 || mirror_in == nullptr) ? v % (uint32_t)M : mirror_in[v];
    if (mm >= (uint32_t)M) {
      delete g;
      return fail(GRD_ERR_ARG, "graph: mirror partition id " + std::to_string(mm) + " >= M");
    }
    mir[v] = (uint8_t)mm;
  }
  for (uint32_t e = 0; e < m; e++)
    if (col[e] >= n) {
      delete g;
      return fail(GRD_ERR_ARG, "graph: column index out of range");
    }
  g->h_deg.resize(n);
  for (uint32_t v = 0; v < n; v++) {
    const uint32_t d = rp[v + 1] - rp[v];
    g->h_deg[v] = d;
    g->n_sinks += (d == 0);
    g->max_deg = std::max(g->max_deg, d);
  }

  // cross-node mirrors: one bitmap per memory partition
  std::vector<uint64_t> crossw((uint64_t)M * NW, 0);
  parallel_for(T, n, 4096, [&](uint64_t b, uint64_t e) {
    for (uint64_t u = b; u < e; u++) {
      uint32_t mu = mir[u];
      for (uint32_t x = rp[u]; x < rp[u + 1]; x++) {
        uint32_t v = col[x];
        if (mir[v] != mu) {
          uint64_t *wd = &crossw[(uint64_t)mu * NW + (v >> 6)];
          uint64_t bit = 1ull << (v & 63);
          if (!(__atomic_load_n(wd, __ATOMIC_RELAXED) & bit)) __atomic_fetch_or(wd, bit, __ATOMIC_RELAXED);
        }
      }
    }
  });

  // group sizes
  g->ccnt.assign((size_t)C * M, 0);
  g->xcnt.assign((size_t)M * C, 0);
  for (uint32_t v = 0; v < n; v++) g->ccnt[(size_t)(v % C) * M + mir[v]]++;
  for (int mm = 0; mm < M; mm++)
    for (uint32_t wd = 0; wd < NW; wd++) {
      uint64_t x = crossw[(uint64_t)mm * NW + wd];
      while (x) {
        uint32_t v = wd * 64 + (uint32_t)__builtin_ctzll(x);
        x &= x - 1;
        g->xcnt[(size_t)mm * C + (v % C)]++;
      }
    }
  // bases
  g->cbase.assign((size_t)C * M, 0);
  uint64_t cur = 0;
  for (int c = 0; c < C; c++)
    for (int mm = 0; mm < M; mm++) {
      g->cbase[(size_t)c * M + mm] = (uint32_t)cur;
      cur += pad64(g->ccnt[(size_t)c * M + mm]);
    }
  if (cur >= 0xFFFFFF00ull) {
    delete g;
    return fail(GRD_ERR_ARG, "graph: compute slots exceed u32");
  }
  g->S = (uint32_t)std::max<uint64_t>(cur, 64);
  g->obase.assign((size_t)M * C, 0);
  g->ocnt.assign((size_t)M * C, 0);
  g->xbase.assign((size_t)M * C, 0);
  cur = 0;
  uint64_t X = 0;
  for (int mm = 0; mm < M; mm++)
    for (int c = 0; c < C; c++) {
      size_t k = (size_t)mm * C + c;
      g->ocnt[k] = g->ccnt[(size_t)c * M + mm];
      g->obase[k] = (uint32_t)cur;
      cur += pad64(g->ocnt[k]);
      g->xbase[k] = (uint32_t)cur;
      cur += pad64(g->xcnt[k]);
      X += g->xcnt[k];
    }
  if (cur >= 0xFFFFFF00ull) {
    delete g;
    return fail(GRD_ERR_ARG, "graph: memory proxies exceed u32");
  }
  g->P = (uint32_t)std::max<uint64_t>(cur, 64);
  g->X = X;

  // own slot / own proxy per gid
  // (position in a group: out-degree rank, highest first, ties by gid)
  std::vector<uint32_t> own_proxy(n);
  g->h_gid_slot.assign(n, kNone);
  {
    const size_t NG = (size_t)C * M;
    std::vector<uint64_t> gstart(NG + 1, 0);
    for (size_t k = 0; k < NG; k++) gstart[k + 1] = gstart[k] + g->ccnt[k];
    // key = (~deg << 32) | gid sorts by degree descending, then gid ascending
    std::vector<uint64_t> key(n);
    {
      std::vector<uint64_t> ctr(gstart.begin(), gstart.end() - 1);
      for (uint32_t v = 0; v < n; v++)
        key[ctr[(size_t)(v % C) * M + mir[v]]++] = ((uint64_t)(~g->h_deg[v]) << 32) | v;
    }
    parallel_for(T, NG, 1, [&](uint64_t b, uint64_t e) {
      for (uint64_t k = b; k < e; k++) std::sort(key.begin() + gstart[k], key.begin() + gstart[k + 1]);
    });
    for (size_t k = 0; k < NG; k++) {
      const size_t c = k / M, mm = k % M;
      for (uint64_t i = gstart[k]; i < gstart[k + 1]; i++) {
        const uint32_t v = (uint32_t)key[i], r = (uint32_t)(i - gstart[k]);
        g->h_gid_slot[v] = g->cbase[k] + r;
        own_proxy[v] = g->obase[mm * C + c] + r;
      }
    }
  }
  // cross proxy per (m, gid).  Within a (m, c) cross region the proxies are
  // ordered by their vertex's compute slot, so that the compute side, which
  // walks its slots in order, reads each partition's cross partials in
  // ascending order.
  std::vector<uint32_t> xp((uint64_t)M * n, kNone);
  g->h_proxy_gid.assign(g->P, kNone);
  for (uint32_t v = 0; v < n; v++) g->h_proxy_gid[own_proxy[v]] = v;
  parallel_for(T, (uint64_t)M * C, 1, [&](uint64_t b, uint64_t e) {
    for (uint64_t k = b; k < e; k++) {
      const int mm = (int)(k / C);
      const uint32_t c = (uint32_t)(k % C);
      std::vector<uint64_t> key;  // slot << 32 | gid
      key.reserve(g->xcnt[k]);
      for (uint32_t wd = 0; wd < NW; wd++) {
        uint64_t x = crossw[(uint64_t)mm * NW + wd];
        while (x) {
          const uint32_t v = wd * 64 + (uint32_t)__builtin_ctzll(x);
          x &= x - 1;
          if (v % C == c) key.push_back(((uint64_t)g->h_gid_slot[v] << 32) | v);
        }
      }
      std::sort(key.begin(), key.end());
      for (size_t i = 0; i < key.size(); i++) {
        const uint32_t v = (uint32_t)key[i], p = g->xbase[k] + (uint32_t)i;
        xp[(uint64_t)mm * n + v] = p;
        g->h_proxy_gid[p] = v;
      }
    }
  });
  std::vector<uint64_t>().swap(crossw);

  // per-slot and per-proxy arrays
  std::vector<uint32_t> slot_gid(g->S, kNone), slot_deg(g->S, 0);
  for (uint32_t v = 0; v < n; v++) {
    slot_gid[g->h_gid_slot[v]] = v;
    slot_deg[g->h_gid_slot[v]] = g->h_deg[v];
  }
  std::vector<uint32_t> proxy_slot(g->P, kNone);
  for (uint32_t p = 0; p < g->P; p++)
    if (g->h_proxy_gid[p] != kNone) proxy_slot[p] = g->h_gid_slot[g->h_proxy_gid[p]];
  std::vector<uint32_t> word_dst(g->S / 64, kNone);
  for (int c = 0; c < C; c++)
    for (int mm = 0; mm < M; mm++) {
      uint32_t sb = g->cbase[(size_t)c * M + mm] / 64;
      uint32_t nwd = (uint32_t)(pad64(g->ccnt[(size_t)c * M + mm]) / 64);
      uint32_t pb = g->obase[(size_t)mm * C + c] / 64;
      for (uint32_t k = 0; k < nwd; k++) word_dst[sb + k] = pb + k;
    }

  // push CSR over proxies
  std::vector<uint32_t> prp((uint64_t)g->P + 1, 0);
  for (uint32_t v = 0; v < n; v++) prp[own_proxy[v] + 1] = g->h_deg[v];
  for (uint32_t p = 0; p < g->P; p++) prp[p + 1] += prp[p];
  std::vector<uint32_t> pcol(m), pw(w ? m : 0);
  parallel_for(T, n, 2048, [&](uint64_t b, uint64_t e) {
    for (uint64_t u = b; u < e; u++) {
      uint32_t mm = mir[u];
      uint32_t o = prp[own_proxy[u]];
      for (uint32_t x = rp[u]; x < rp[u + 1]; x++, o++) {
        uint32_t v = col[x];
        pcol[o] = (mir[v] == mm) ? own_proxy[v] : xp[(uint64_t)mm * n + v];
        if (w) pw[o] = w[x];
      }
    }
  });
  std::vector<uint32_t>().swap(xp);

  // pull CSC over proxies: count, prefix, fill (rows sorted by source)
  std::vector<uint32_t> crp((uint64_t)g->P + 1, 0);
  {
    std::vector<uint32_t> indeg(g->P, 0);
    for (uint64_t x = 0; x < m; x++) indeg[pcol[x]]++;
    for (uint32_t p = 0; p < g->P; p++) crp[p + 1] = crp[p] + indeg[p];
  }
  std::vector<uint32_t> ccol(m), cwv(w ? m : 0);
  {
    // sources are visited in ascending proxy order, so every CSC row comes
    // out sorted by source (the graph has no duplicate arcs)
    std::vector<uint32_t> cursor(crp.begin(), crp.end() - 1);
    for (uint32_t p = 0; p < g->P; p++)
      for (uint32_t x = prp[p]; x < prp[p + 1]; x++) {
        uint32_t o = cursor[pcol[x]]++;
        ccol[o] = p;
        if (w) cwv[o] = pw[x];
      }
  }

  // cross proxies per compute slot (PR pull aggregation), ascending proxy order
  std::vector<uint32_t> xptr((uint64_t)g->S + 1, 0), xlist(X);
  for (uint32_t p = 0; p < g->P; p++) {
    uint32_t v = g->h_proxy_gid[p];
    if (v != kNone && own_proxy[v] != p) xptr[g->h_gid_slot[v] + 1]++;
  }
  for (uint32_t s = 0; s < g->S; s++) xptr[s + 1] += xptr[s];
  {
    std::vector<uint32_t> cursor(xptr.begin(), xptr.end() - 1);
    for (uint32_t p = 0; p < g->P; p++) {
      uint32_t v = g->h_proxy_gid[p];
      if (v != kNone && own_proxy[v] != p) xlist[cursor[g->h_gid_slot[v]]++] = p;
    }
  }

  // upload
  grd_status st = GRD_OK;
  if ((st = upload(g, &g->slot_gid, slot_gid)) != GRD_OK ||
      (st = upload(g, &g->slot_deg, slot_deg)) != GRD_OK ||
      (st = upload(g, &g->word_dst, word_dst)) != GRD_OK ||
      (st = upload(g, &g->proxy_slot, proxy_slot)) != GRD_OK ||
      (st = upload(g, &g->rp, prp)) != GRD_OK ||
      (st = upload(g, &g->col, pcol)) != GRD_OK ||
      (st = upload(g, &g->crp, crp)) != GRD_OK ||
      (st = upload(g, &g->ccol, ccol)) != GRD_OK ||
      (st = upload(g, &g->xptr, xptr)) != GRD_OK ||
      (st = upload(g, &g->xlist, xlist)) != GRD_OK) {
    delete g;
    return st;
  }
  if (w) {
    if ((st = upload(g, &g->w, pw)) != GRD_OK || (st = upload(g, &g->cw, cwv)) != GRD_OK) {
      delete g;
      return st;
    }
  }
  if ((st = build_row_blocks(g, crp)) != GRD_OK) {
    delete g;
    return st;
  }
  *out = g;
  return GRD_OK;
}

grd_status check_row_ptr(uint32_t n, const uint32_t *rp)
{
  if (rp[0] != 0) return fail(GRD_ERR_ARG, "graph: row_ptr[0] must be 0");
  for (uint32_t v = 0; v < n; v++)
    if (rp[v + 1] < rp[v]) return fail(GRD_ERR_ARG, "graph: row_ptr is not monotone");
  return GRD_OK;
}

grd_status build_graph(grd_ctx *ctx, uint32_t n, const uint32_t *rp, const uint32_t *col, const uint32_t *w, int C,
                       int M, const uint32_t *mirror, grd_graph **out)
{
  grd_status st = GRD_OK;
  if (ctx->layout_on_device) {
    st = build_graph_device(ctx, n, rp, col, w, C, M, mirror, out);
  } else {
    try {
      st = build_impl(ctx, n, rp, col, w, C, M, mirror, out);
    } catch (const std::bad_alloc &) {
      st = fail(GRD_ERR_OOM, "graph: host allocation failed");
    }
  }
  if (st != GRD_OK) return st;
  // both constructions end with the tables of the edge-shipping schedule
  try {
    st = build_ship_tables(*out);
  } catch (const std::bad_alloc &) {
    st = fail(GRD_ERR_OOM, "graph: host allocation failed");
  }
  if (st != GRD_OK) {
    (void)hipStreamSynchronize(ctx->stream);
    delete *out;
    *out = nullptr;
  }
  return st;
}

}  // namespace grd
