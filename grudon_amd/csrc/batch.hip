// batch.hip -- many graphs of uneven size processed by one persistent
// launch per phase, with the iterations captured in a hipGraph.
//
// The graphs' CSCs are concatenated (global vertex id = graph base + local
// id) and cut into row blocks of at most `tile` arcs, each tagged with its
// graph; rows longer than `tile` arcs become segments merged in order by a
// fix-up pass.  A persistent grid of `grid_per_cu` x CUs workgroups walks the
// block list; blocks of graphs that have already converged are skipped, so a
// replay of the captured iterations does no work for finished graphs.
//
// PageRank (per graph, normalised pull, FP32 Kahan contributions, FP64
// dangling mass and L1 residual, converged when the L1 change < tolerance)
// and CC (per graph, synchronous min-label propagation over in-edges, the
// fixed point of src/graph_algorithms/cc.cpp) share the block structure.
//
// Tunables: tile (arcs per block = long-row threshold), grid_per_cu
// (persistent workgroups per CU) and K (iterations per captured graph);
// grd_batch_autotune times candidate settings on the batch itself.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "grd_device.hpp"
#include "grd_internal.hpp"

namespace grd {

struct BatchBlock {
  uint32_t r0, nrows, e0, e1, graph, pad0, pad1, pad2;
};
struct BatchLong {
  uint32_t row, first, nseg, graph;
};
struct BatchTile {
  uint32_t v0, v1, graph, pad;
};

}  // namespace grd

struct grd_batch {
  grd_ctx *ctx = nullptr;
  uint32_t count = 0;
  uint64_t N = 0, E = 0;
  std::vector<uint64_t> vbase;  // [count+1]
  std::vector<uint32_t> h_crp;  // [N+1] concatenated CSC row pointers (global arc ids)
  std::vector<uint32_t> h_deg;  // [N] out-degree
  grd::DevMem mem;
  uint32_t *crp = nullptr, *ccol = nullptr, *deg = nullptr, *gid_local = nullptr;
  uint64_t *vb = nullptr;  // [count+1]
  // tunables
  uint32_t tile = 2048, grid_per_cu = 8, K = 4;
  // blocks for the current tile
  grd::DevMem blk_mem;
  grd::BatchBlock *blocks = nullptr;
  grd::BatchLong *longs = nullptr;
  grd::BatchTile *tiles = nullptr;
  uint32_t *gtile = nullptr;  // [count+1] first vertex tile of each graph
  double *tred = nullptr;     // [2 * ntile] (residual, dangling mass) of each vertex tile
  uint32_t nblk = 0, nlong = 0, ntile = 0;
  // per-graph state
  grd::DevMem st_mem;
  float *r = nullptr, *contrib = nullptr, *acc = nullptr;
  float2 *partial = nullptr;
  uint32_t *lab[2] = {nullptr, nullptr};
  uint32_t *active = nullptr, *iters = nullptr, *changed = nullptr, *alive = nullptr;
  double *residual = nullptr, *dcur = nullptr;
  uint32_t *h_alive = nullptr;  // pinned
  int last_algo = -1;
  std::vector<uint32_t> last_iters;
};

namespace grd {

// ------------------------------------------------------------------------
// PageRank
// ------------------------------------------------------------------------

template <int NNZ>
__global__ void __launch_bounds__(kBlock) k_bpr_pull(const BatchBlock *__restrict__ blocks, uint32_t nblk,
                                                     const uint32_t *__restrict__ crp, const uint32_t *__restrict__ ccol,
                                                     const float *__restrict__ contrib, float *__restrict__ acc,
                                                     float2 *__restrict__ partial, const uint32_t *__restrict__ active)
{
  __shared__ float s_v[NNZ];
  __shared__ float s_ws[kWavesPerBlock], s_wc[kWavesPerBlock];
  for (uint32_t b = blockIdx.x; b < nblk; b += gridDim.x) {
    const BatchBlock d = blocks[b];
    if (!active[d.graph]) continue;  // uniform over the workgroup
    if (d.nrows == 0) {
      Kahan k;
      for (uint32_t e = d.e0 + threadIdx.x; e < d.e1; e += kBlock) k.add(contrib[ccol[e]]);
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
        partial[b] = make_float2(t.s, t.c);
      }
      __syncthreads();
      continue;
    }
    const uint32_t ne = d.e1 - d.e0;
    constexpr int kPer = NNZ / kBlock;
    uint32_t src[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      uint32_t i = threadIdx.x + k * kBlock;
      src[k] = i < ne ? ccol[d.e0 + i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      uint32_t i = threadIdx.x + k * kBlock;
      if (i < ne) s_v[i] = contrib[src[k]];
    }
    __syncthreads();
    int G = 64;
    while (G > 1 && (uint32_t)G * d.nrows > (uint32_t)kBlock) G >>= 1;
    const int sub = threadIdx.x % G;
    const uint32_t ngroups = kBlock / G;
    const uint32_t rounds = (d.nrows + ngroups - 1) / ngroups;
    for (uint32_t k = 0; k < rounds; k++) {
      uint32_t rl = k * ngroups + threadIdx.x / G;
      Kahan a;
      uint32_t r = d.r0 + rl;
      if (rl < d.nrows) {
        uint32_t b0 = crp[r] - d.e0, e0 = crp[r + 1] - d.e0;
        for (uint32_t x = b0 + sub; x < e0; x += G) a.add(s_v[x]);
      }
      kahan_wave_merge(a, G);
      if (sub == 0 && rl < d.nrows) acc[r] = a.value();
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kBlock) k_bpr_fixup(const BatchLong *__restrict__ longs, uint32_t nlong,
                                                      const float2 *__restrict__ partial, float *__restrict__ acc,
                                                      const uint32_t *__restrict__ active)
{
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nlong; i += gridDim.x * kBlock) {
    BatchLong L = longs[i];
    if (!active[L.graph]) continue;
    Kahan t;
    for (uint32_t j = 0; j < L.nseg; j++) {
      float2 p = partial[L.first + j];
      t.merge(p.x, p.y);
    }
    acc[L.row] = t.value();
  }
}

// vertex tiles lie within one graph: residual and dangling mass are reduced
// per tile into the tile's own FP64 slot (no atomics), summed per graph in
// tile order by k_bpr_finish, so convergence does not depend on scheduling
__global__ void __launch_bounds__(kBlock) k_bpr_update(const BatchTile *__restrict__ tiles, uint32_t ntile,
                                                       const uint64_t *__restrict__ vb, const uint32_t *__restrict__ deg,
                                                       const float *__restrict__ acc, float *__restrict__ r,
                                                       float *__restrict__ contrib, const uint32_t *__restrict__ active,
                                                       const double *__restrict__ dcur, double *__restrict__ tred,
                                                       float damping)
{
  for (uint32_t t = blockIdx.x; t < ntile; t += gridDim.x) {
    const BatchTile T = tiles[t];
    if (!active[T.graph]) continue;
    const double n = (double)(vb[T.graph + 1] - vb[T.graph]);
    const float base = (float)((1.0 - (double)damping) / n + (double)damping * dcur[T.graph] / n);
    double res = 0.0, dang = 0.0;
    for (uint32_t v = T.v0 + threadIdx.x; v < T.v1; v += kBlock) {
      float rn = base + damping * acc[v];
      res += fabs((double)rn - (double)r[v]);
      uint32_t dg = deg[v];
      if (dg == 0) dang += (double)rn;
      r[v] = rn;
      contrib[v] = dg ? rn / (float)dg : 0.0f;
    }
    res = block_sum(res);
    dang = block_sum(dang);
    if (threadIdx.x == 0) {
      tred[2 * t] = res;
      tred[2 * t + 1] = dang;
    }
  }
}

__global__ void k_bpr_finish(uint32_t count, uint32_t *active, uint32_t *iters, const uint32_t *__restrict__ gtile,
                             const double *__restrict__ tred, double *residual, double *dcur, double tolerance,
                             uint32_t max_iter, uint32_t *alive)
{
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < count; g += gridDim.x * blockDim.x) {
    if (!active[g]) continue;
    double res = 0.0, dang = 0.0;
    for (uint32_t t = gtile[g]; t < gtile[g + 1]; t++) {
      res += tred[2 * t];
      dang += tred[2 * t + 1];
    }
    uint32_t it = ++iters[g];
    residual[g] = res;
    if (res < tolerance || it >= max_iter) active[g] = 0;
    else atomicAdd(alive, 1u);
    dcur[g] = dang;
  }
}

__global__ void __launch_bounds__(kBlock) k_bpr_init(uint64_t N, const uint64_t *__restrict__ vb, uint32_t count,
                                                     const uint32_t *__restrict__ gidx, const uint32_t *__restrict__ deg,
                                                     float *r, float *contrib)
{
  for (uint64_t v = blockIdx.x * (uint64_t)kBlock + threadIdx.x; v < N; v += (uint64_t)gridDim.x * kBlock) {
    uint32_t g = gidx[v];
    float r0 = (float)(1.0 / (double)(vb[g + 1] - vb[g]));
    r[v] = r0;
    contrib[v] = deg[v] ? r0 / (float)deg[v] : 0.0f;
  }
}

// ------------------------------------------------------------------------
// CC (synchronous min-label propagation over in-edges)
// ------------------------------------------------------------------------

template <int NNZ>
__global__ void __launch_bounds__(kBlock) k_bcc_pull(const BatchBlock *__restrict__ blocks, uint32_t nblk,
                                                     const uint32_t *__restrict__ crp, const uint32_t *__restrict__ ccol,
                                                     const uint32_t *__restrict__ lin, uint32_t *__restrict__ lout,
                                                     uint32_t *__restrict__ partial, const uint32_t *__restrict__ active,
                                                     uint32_t *changed)
{
  __shared__ uint32_t s_v[NNZ];
  __shared__ uint32_t s_changed;
  for (uint32_t b = blockIdx.x; b < nblk; b += gridDim.x) {
    const BatchBlock d = blocks[b];
    if (!active[d.graph]) continue;
    if (d.nrows == 0) {
      uint32_t m = 0xFFFFFFFFu;
      for (uint32_t e = d.e0 + threadIdx.x; e < d.e1; e += kBlock) m = min(m, lin[ccol[e]]);
      m = block_min(m);
      if (threadIdx.x == 0) partial[b] = m;
      __syncthreads();
      continue;
    }
    if (threadIdx.x == 0) s_changed = 0;
    const uint32_t ne = d.e1 - d.e0;
    constexpr int kPer = NNZ / kBlock;
    uint32_t src[kPer];
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      uint32_t i = threadIdx.x + k * kBlock;
      src[k] = i < ne ? ccol[d.e0 + i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kPer; k++) {
      uint32_t i = threadIdx.x + k * kBlock;
      if (i < ne) s_v[i] = lin[src[k]];
    }
    __syncthreads();
    int G = 64;
    while (G > 1 && (uint32_t)G * d.nrows > (uint32_t)kBlock) G >>= 1;
    const int sub = threadIdx.x % G;
    const uint32_t ngroups = kBlock / G;
    const uint32_t rounds = (d.nrows + ngroups - 1) / ngroups;
    for (uint32_t k = 0; k < rounds; k++) {
      uint32_t rl = k * ngroups + threadIdx.x / G;
      uint32_t m = 0xFFFFFFFFu;
      uint32_t r = d.r0 + rl;
      if (rl < d.nrows) {
        uint32_t b0 = crp[r] - d.e0, e0 = crp[r + 1] - d.e0;
        for (uint32_t x = b0 + sub; x < e0; x += G) m = min(m, s_v[x]);
      }
      for (int o = G / 2; o > 0; o >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, o, kWave));
      if (sub == 0 && rl < d.nrows) {
        uint32_t old = lin[r];
        uint32_t nv = min(old, m);
        lout[r] = nv;
        if (nv < old) s_changed = 1;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_changed) changed[d.graph] = 1;
  }
}

__global__ void __launch_bounds__(kBlock) k_bcc_fixup(const BatchLong *__restrict__ longs, uint32_t nlong,
                                                      const uint32_t *__restrict__ partial,
                                                      const uint32_t *__restrict__ lin, uint32_t *__restrict__ lout,
                                                      const uint32_t *__restrict__ active, uint32_t *changed)
{
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < nlong; i += gridDim.x * kBlock) {
    BatchLong L = longs[i];
    if (!active[L.graph]) continue;
    uint32_t m = 0xFFFFFFFFu;
    for (uint32_t j = 0; j < L.nseg; j++) m = min(m, partial[L.first + j]);
    uint32_t old = lin[L.row], nv = min(old, m);
    lout[L.row] = nv;
    if (nv < old) changed[L.graph] = 1;
  }
}

__global__ void k_bcc_finish(uint32_t count, uint32_t *active, uint32_t *iters, uint32_t *changed, uint32_t max_iter,
                             uint32_t *alive)
{
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < count; g += gridDim.x * blockDim.x) {
    if (!active[g]) continue;
    uint32_t it = ++iters[g];
    if (!changed[g] || it >= max_iter) active[g] = 0;
    else atomicAdd(alive, 1u);
    changed[g] = 0;
  }
}

__global__ void __launch_bounds__(kBlock) k_bcc_init(uint64_t N, const uint32_t *__restrict__ local, uint32_t *l0,
                                                     uint32_t *l1)
{
  for (uint64_t v = blockIdx.x * (uint64_t)kBlock + threadIdx.x; v < N; v += (uint64_t)gridDim.x * kBlock) {
    l0[v] = local[v];
    l1[v] = local[v];
  }
}

__global__ void k_bfill_u32(uint32_t *p, uint32_t n, uint32_t v)
{
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_bfill_f64(double *p, uint32_t n, double v)
{
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// ------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------

static grd_status batch_build_blocks(grd_batch *b)
{
  std::vector<BatchBlock> blks;
  std::vector<BatchLong> longs;
  std::vector<BatchTile> tiles;
  std::vector<uint32_t> gtile(b->count + 1, 0);
  const uint32_t T = b->tile;
  for (uint32_t g = 0; g < b->count; g++) {
    gtile[g] = (uint32_t)tiles.size();
    const uint64_t v0 = b->vbase[g], v1 = b->vbase[g + 1];
    uint64_t r = v0;
    while (r < v1) {
      uint32_t deg = b->h_crp[r + 1] - b->h_crp[r];
      if (deg > T) {
        BatchLong L{(uint32_t)r, (uint32_t)blks.size(), 0, g};
        for (uint32_t e = b->h_crp[r]; e < b->h_crp[r + 1]; e += T) {
          blks.push_back(BatchBlock{(uint32_t)r, 0, e, std::min(b->h_crp[r + 1], e + T), g, 0, 0, 0});
          L.nseg++;
        }
        longs.push_back(L);
        r++;
        continue;
      }
      const uint64_t rs = r;
      const uint32_t e0 = b->h_crp[r];
      while (r < v1 && r - rs < kPullMaxRows && b->h_crp[r + 1] - b->h_crp[r] <= T && b->h_crp[r + 1] - e0 <= T) r++;
      blks.push_back(BatchBlock{(uint32_t)rs, (uint32_t)(r - rs), e0, b->h_crp[r], g, 0, 0, 0});
    }
    for (uint64_t v = v0; v < v1; v += 4096)
      tiles.push_back(BatchTile{(uint32_t)v, (uint32_t)std::min<uint64_t>(v1, v + 4096), g, 0});
  }
  gtile[b->count] = (uint32_t)tiles.size();
  b->blk_mem.release();
  b->nblk = (uint32_t)blks.size();
  b->nlong = (uint32_t)longs.size();
  b->ntile = (uint32_t)tiles.size();
  if (longs.empty()) longs.push_back(BatchLong{0, 0, 0, 0});
  GRD_TRY(b->blk_mem.alloc_n(&b->blocks, std::max<size_t>(1, blks.size())));
  GRD_TRY(b->blk_mem.alloc_n(&b->longs, longs.size()));
  GRD_TRY(b->blk_mem.alloc_n(&b->tiles, std::max<size_t>(1, tiles.size())));
  GRD_TRY(b->blk_mem.alloc(reinterpret_cast<void **>(&b->partial), (uint64_t)std::max<uint32_t>(1, b->nblk) * 8));
  GRD_TRY(b->blk_mem.alloc_n(&b->gtile, gtile.size()));
  GRD_TRY(b->blk_mem.alloc_n(&b->tred, 2ull * std::max<size_t>(1, tiles.size())));
  GRD_HIP(hipMemcpy(b->gtile, gtile.data(), gtile.size() * 4, hipMemcpyHostToDevice));
  GRD_HIP(hipMemcpy(b->blocks, blks.data(), blks.size() * sizeof(BatchBlock), hipMemcpyHostToDevice));
  GRD_HIP(hipMemcpy(b->longs, longs.data(), longs.size() * sizeof(BatchLong), hipMemcpyHostToDevice));
  GRD_HIP(hipMemcpy(b->tiles, tiles.data(), tiles.size() * sizeof(BatchTile), hipMemcpyHostToDevice));
  return GRD_OK;
}

static uint32_t persistent_grid(const grd_batch *b) { return (uint32_t)b->ctx->num_cus * b->grid_per_cu; }

static void launch_bpr_pull(grd_batch *b, hipStream_t st)
{
  const uint32_t grid = persistent_grid(b);
  switch (b->tile) {
  case 1024:
    hipLaunchKernelGGL(k_bpr_pull<1024>, dim3(grid), dim3(kBlock), 0, st, b->blocks, b->nblk, b->crp, b->ccol,
                       b->contrib, b->acc, b->partial, b->active);
    break;
  case 4096:
    hipLaunchKernelGGL(k_bpr_pull<4096>, dim3(grid), dim3(kBlock), 0, st, b->blocks, b->nblk, b->crp, b->ccol,
                       b->contrib, b->acc, b->partial, b->active);
    break;
  default:
    hipLaunchKernelGGL(k_bpr_pull<2048>, dim3(grid), dim3(kBlock), 0, st, b->blocks, b->nblk, b->crp, b->ccol,
                       b->contrib, b->acc, b->partial, b->active);
  }
}

static void launch_bcc_pull(grd_batch *b, hipStream_t st, int parity)
{
  const uint32_t grid = persistent_grid(b);
  uint32_t *lin = b->lab[parity], *lout = b->lab[parity ^ 1];
  uint32_t *part = reinterpret_cast<uint32_t *>(b->partial);
  switch (b->tile) {
  case 1024:
    hipLaunchKernelGGL(k_bcc_pull<1024>, dim3(grid), dim3(kBlock), 0, st, b->blocks, b->nblk, b->crp, b->ccol, lin,
                       lout, part, b->active, b->changed);
    break;
  case 4096:
    hipLaunchKernelGGL(k_bcc_pull<4096>, dim3(grid), dim3(kBlock), 0, st, b->blocks, b->nblk, b->crp, b->ccol, lin,
                       lout, part, b->active, b->changed);
    break;
  default:
    hipLaunchKernelGGL(k_bcc_pull<2048>, dim3(grid), dim3(kBlock), 0, st, b->blocks, b->nblk, b->crp, b->ccol, lin,
                       lout, part, b->active, b->changed);
  }
  hipLaunchKernelGGL(k_bcc_fixup, dim3(std::max<uint32_t>(1, (b->nlong + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                     b->longs, b->nlong, part, lin, lout, b->active, b->changed);
}

// capture K iterations of `algo` into a graph; CC captures pairs of
// iterations so that the label buffers alternate inside one replay
static grd_status capture(grd_batch *b, int algo, double damping, double tol, uint32_t max_iter, hipGraphExec_t *exec)
{
  hipStream_t st = b->ctx->stream;
  hipGraph_t graph = nullptr;
  GRD_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const uint32_t reps = algo == GRD_CC ? std::max<uint32_t>(2, b->K + (b->K & 1)) : b->K;
  for (uint32_t i = 0; i < reps; i++) {
    if (algo == GRD_PR_PULL) {
      launch_bpr_pull(b, st);
      hipLaunchKernelGGL(k_bpr_fixup, dim3(std::max<uint32_t>(1, (b->nlong + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                         st, b->longs, b->nlong, b->partial, b->acc, b->active);
      hipLaunchKernelGGL(k_bpr_update, dim3(std::min<uint32_t>(b->ntile, persistent_grid(b))), dim3(kBlock), 0, st,
                         b->tiles, b->ntile, b->vb, b->deg, b->acc, b->r, b->contrib, b->active, b->dcur, b->tred,
                         (float)damping);
      hipLaunchKernelGGL(k_bpr_finish, dim3((b->count + 255) / 256), dim3(256), 0, st, b->count, b->active, b->iters,
                         b->gtile, b->tred, b->residual, b->dcur, tol, max_iter, b->alive);
    } else {
      launch_bcc_pull(b, st, (int)(i & 1));
      hipLaunchKernelGGL(k_bcc_finish, dim3((b->count + 255) / 256), dim3(256), 0, st, b->count, b->active, b->iters,
                         b->changed, max_iter, b->alive);
    }
  }
  hipError_t e = hipStreamEndCapture(st, &graph);
  if (e != hipSuccess) return fail(GRD_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(GRD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  return GRD_OK;
}

static grd_status batch_run(grd_batch *b, int algo, double damping, double tol, uint32_t max_iter, uint32_t *iters_out,
                            double *ms_out)
{
  hipStream_t st = b->ctx->stream;
  const uint32_t cnt = b->count;
  // per-graph state (fill kernels: nothing memset-like inside the graph)
  hipLaunchKernelGGL(k_bfill_u32, dim3((cnt + 255) / 256), dim3(256), 0, st, b->active, cnt, 1u);
  hipLaunchKernelGGL(k_bfill_u32, dim3((cnt + 255) / 256), dim3(256), 0, st, b->iters, cnt, 0u);
  hipLaunchKernelGGL(k_bfill_u32, dim3((cnt + 255) / 256), dim3(256), 0, st, b->changed, cnt, 0u);
  hipLaunchKernelGGL(k_bfill_f64, dim3((cnt + 255) / 256), dim3(256), 0, st, b->residual, cnt, 0.0);
  const uint32_t ig = (uint32_t)std::min<uint64_t>(4096, b->N / kBlock + 1);
  if (algo == GRD_PR_PULL) {
    hipLaunchKernelGGL(k_bpr_init, dim3(ig), dim3(kBlock), 0, st, b->N, b->vb, cnt, b->changed + cnt, b->deg, b->r,
                       b->contrib);
    // D0 per graph: sinks * (1/n) in FP64
    std::vector<double> d0(cnt, 0.0);
    for (uint32_t g = 0; g < cnt; g++) {
      uint64_t sinks = 0;
      for (uint64_t v = b->vbase[g]; v < b->vbase[g + 1]; v++) sinks += b->h_deg[v] == 0;
      float r0 = (float)(1.0 / (double)(b->vbase[g + 1] - b->vbase[g]));
      d0[g] = (double)sinks * (double)r0;
    }
    GRD_HIP(hipMemcpyAsync(b->dcur, d0.data(), cnt * sizeof(double), hipMemcpyHostToDevice, st));
  } else {
    hipLaunchKernelGGL(k_bcc_init, dim3(ig), dim3(kBlock), 0, st, b->N, b->gid_local, b->lab[0], b->lab[1]);
  }
  GRD_HIP(hipGetLastError());
  GRD_HIP(hipStreamSynchronize(st));
  hipGraphExec_t exec = nullptr;
  GRD_TRY(capture(b, algo, damping, tol, max_iter, &exec));
  GRD_HIP(hipEventRecord(b->ctx->ev[0], st));
  uint32_t launched = 0;
  const uint32_t per = algo == GRD_CC ? std::max<uint32_t>(2, b->K + (b->K & 1)) : b->K;
  grd_status rc = GRD_OK;
  while (launched < max_iter) {
    hipLaunchKernelGGL(k_bfill_u32, dim3(1), dim3(64), 0, st, b->alive, 1u, 0u);
    hipError_t e = hipGraphLaunch(exec, st);
    if (e != hipSuccess) {
      rc = fail(GRD_ERR_HIP, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
      break;
    }
    launched += per;
    e = hipMemcpyAsync(b->h_alive, b->alive, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      rc = fail(GRD_ERR_HIP, std::string("batch readback: ") + hipGetErrorString(e));
      break;
    }
    if (*b->h_alive == 0) break;
  }
  hipEventRecord(b->ctx->ev[1], st);
  hipEventSynchronize(b->ctx->ev[1]);
  float ms = 0.0f;
  hipEventElapsedTime(&ms, b->ctx->ev[0], b->ctx->ev[1]);
  (void)hipGraphExecDestroy(exec);
  GRD_TRY(rc);
  b->last_iters.assign(cnt, 0);
  GRD_HIP(hipMemcpy(b->last_iters.data(), b->iters, cnt * 4, hipMemcpyDeviceToHost));
  if (iters_out) memcpy(iters_out, b->last_iters.data(), cnt * 4);
  if (ms_out) *ms_out = ms;
  b->last_algo = algo;
  return GRD_OK;
}

}  // namespace grd

using namespace grd;

extern "C" {

grd_status grd_batch_create(grd_ctx *ctx, uint32_t count, const uint32_t *n, const uint64_t *m,
                            const uint32_t *const *row_ptr, const uint32_t *const *col, grd_batch **out)
{
  if (!ctx || !n || !m || !row_ptr || !col || !out || count == 0) return fail(GRD_ERR_ARG, "grd_batch_create: bad argument");
  GRD_HIP(hipSetDevice(ctx->device));
  try {
    grd_batch *b = new grd_batch();
    b->ctx = ctx;
    b->count = count;
    b->vbase.assign(count + 1, 0);
    for (uint32_t g = 0; g < count; g++) {
      if (n[g] == 0 || row_ptr[g][n[g]] != m[g]) {
        delete b;
        return fail(GRD_ERR_ARG, "grd_batch_create: graph " + std::to_string(g) + " is empty or inconsistent");
      }
      b->vbase[g + 1] = b->vbase[g] + n[g];
      b->E += m[g];
    }
    b->N = b->vbase[count];
    if (b->N >= 0xFFFFFFF0ull || b->E >= 0xFFFFFFF0ull) {
      delete b;
      return fail(GRD_ERR_ARG, "grd_batch_create: batch exceeds the u32 vertex / arc limit");
    }
    // concatenated CSC (in-edges, sources ascending), out-degrees, graph ids
    b->h_crp.assign(b->N + 1, 0);
    b->h_deg.assign(b->N, 0);
    std::vector<uint32_t> gidx(b->N), local(b->N);
    for (uint32_t g = 0; g < count; g++) {
      const uint64_t vb = b->vbase[g];
      for (uint32_t u = 0; u < n[g]; u++) {
        b->h_deg[vb + u] = row_ptr[g][u + 1] - row_ptr[g][u];
        gidx[vb + u] = g;
        local[vb + u] = u;
        for (uint32_t x = row_ptr[g][u]; x < row_ptr[g][u + 1]; x++) {
          if (col[g][x] >= n[g]) {
            delete b;
            return fail(GRD_ERR_ARG, "grd_batch_create: column out of range in graph " + std::to_string(g));
          }
          b->h_crp[vb + col[g][x] + 1]++;
        }
      }
    }
    for (uint64_t v = 0; v < b->N; v++) b->h_crp[v + 1] += b->h_crp[v];
    std::vector<uint32_t> hcol(b->E), cur(b->h_crp.begin(), b->h_crp.end() - 1);
    for (uint32_t g = 0; g < count; g++) {
      const uint64_t vb = b->vbase[g];
      for (uint32_t u = 0; u < n[g]; u++)
        for (uint32_t x = row_ptr[g][u]; x < row_ptr[g][u + 1]; x++)
          hcol[cur[vb + col[g][x]]++] = (uint32_t)(vb + u);
    }
    grd_status st = GRD_OK;
    auto up = [&](uint32_t **dst, const std::vector<uint32_t> &src) -> grd_status {
      GRD_TRY(b->mem.alloc_n(dst, std::max<size_t>(1, src.size())));
      if (!src.empty()) GRD_HIP(hipMemcpy(*dst, src.data(), src.size() * 4, hipMemcpyHostToDevice));
      return GRD_OK;
    };
    if ((st = up(&b->crp, b->h_crp)) != GRD_OK || (st = up(&b->ccol, hcol)) != GRD_OK ||
        (st = up(&b->deg, b->h_deg)) != GRD_OK || (st = up(&b->gid_local, local)) != GRD_OK) {
      delete b;
      return st;
    }
    // per-graph and per-vertex state; the graph index of each vertex rides
    // behind `changed` (changed[count + v])
    DevMem &sm = b->st_mem;
    if ((st = b->mem.alloc_n(&b->vb, count + 1)) != GRD_OK || (st = sm.alloc_n(&b->r, b->N)) != GRD_OK ||
        (st = sm.alloc_n(&b->contrib, b->N)) != GRD_OK || (st = sm.alloc_n(&b->acc, b->N)) != GRD_OK ||
        (st = sm.alloc_n(&b->lab[0], b->N)) != GRD_OK || (st = sm.alloc_n(&b->lab[1], b->N)) != GRD_OK ||
        (st = sm.alloc_n(&b->active, count)) != GRD_OK || (st = sm.alloc_n(&b->iters, count)) != GRD_OK ||
        (st = sm.alloc_n(&b->changed, count + b->N)) != GRD_OK || (st = sm.alloc_n(&b->alive, 1)) != GRD_OK ||
        (st = sm.alloc_n(&b->residual, count)) != GRD_OK || (st = sm.alloc_n(&b->dcur, count)) != GRD_OK) {
      delete b;
      return st;
    }
    GRD_HIP(hipMemcpy(b->vb, b->vbase.data(), (count + 1) * 8, hipMemcpyHostToDevice));
    GRD_HIP(hipMemcpy(b->changed + count, gidx.data(), b->N * 4, hipMemcpyHostToDevice));
    GRD_HIP(hipHostMalloc((void **)&b->h_alive, 4, hipHostMallocDefault));
    if ((st = batch_build_blocks(b)) != GRD_OK) {
      (void)hipHostFree(b->h_alive);
      delete b;
      return st;
    }
    *out = b;
    return GRD_OK;
  } catch (const std::bad_alloc &) {
    return fail(GRD_ERR_OOM, "grd_batch_create: host allocation failed");
  }
}

grd_status grd_batch_free(grd_batch *b)
{
  if (!b) return GRD_OK;
  (void)hipStreamSynchronize(b->ctx->stream);
  if (b->h_alive) (void)hipHostFree(b->h_alive);
  delete b;
  return GRD_OK;
}

grd_status grd_batch_tune(grd_batch *b, uint32_t tile, uint32_t grid_per_cu, uint32_t K)
{
  if (!b) return fail(GRD_ERR_ARG, "grd_batch_tune: NULL batch");
  if (tile && tile != 1024 && tile != 2048 && tile != 4096)
    return fail(GRD_ERR_ARG, "grd_batch_tune: tile must be 1024, 2048 or 4096 arcs");
  if (grid_per_cu > 64 || K > 64) return fail(GRD_ERR_ARG, "grd_batch_tune: grid_per_cu and K must be <= 64");
  if (grid_per_cu) b->grid_per_cu = grid_per_cu;
  if (K) b->K = K;
  if (tile && tile != b->tile) {
    b->tile = tile;
    GRD_TRY(batch_build_blocks(b));
  }
  return GRD_OK;
}

grd_status grd_batch_autotune(grd_batch *b, uint32_t *chosen3)
{
  if (!b) return fail(GRD_ERR_ARG, "grd_batch_autotune: NULL batch");
  const uint32_t tiles[] = {1024, 2048, 4096}, grids[] = {4, 8, 16}, Ks[] = {2, 4, 8};
  double best = 1e30;
  uint32_t bt = b->tile, bg = b->grid_per_cu, bk = b->K;
  for (uint32_t t : tiles)
    for (uint32_t gpc : grids)
      for (uint32_t k : Ks) {
        GRD_TRY(grd_batch_tune(b, t, gpc, k));
        // a fixed amount of work per candidate: 16 PageRank iterations
        double ms = 0.0;
        GRD_TRY(batch_run(b, GRD_PR_PULL, 0.85, 0.0, 16, nullptr, &ms));
        if (ms < best) {
          best = ms;
          bt = t;
          bg = gpc;
          bk = k;
        }
      }
  GRD_TRY(grd_batch_tune(b, bt, bg, bk));
  if (chosen3) {
    chosen3[0] = bt;
    chosen3[1] = bg;
    chosen3[2] = bk;
  }
  return GRD_OK;
}

grd_status grd_batch_pagerank(grd_batch *b, double damping, double tolerance, uint32_t max_iter, uint32_t *iters_out,
                              double *ms_out)
{
  if (!b || max_iter == 0) return fail(GRD_ERR_ARG, "grd_batch_pagerank: bad argument");
  GRD_HIP(hipSetDevice(b->ctx->device));
  return batch_run(b, GRD_PR_PULL, damping, tolerance, max_iter, iters_out, ms_out);
}

grd_status grd_batch_cc(grd_batch *b, uint32_t max_iter, uint32_t *iters_out, double *ms_out)
{
  if (!b || max_iter == 0) return fail(GRD_ERR_ARG, "grd_batch_cc: bad argument");
  GRD_HIP(hipSetDevice(b->ctx->device));
  return batch_run(b, GRD_CC, 0.0, 0.0, max_iter, iters_out, ms_out);
}

grd_status grd_batch_download(grd_batch *b, uint32_t i, int algo, void *out)
{
  if (!b || !out || i >= b->count) return fail(GRD_ERR_ARG, "grd_batch_download: bad argument");
  if (algo != b->last_algo) return fail(GRD_ERR_STATE, "grd_batch_download: that algorithm has not run last");
  const uint64_t v0 = b->vbase[i], n = b->vbase[i + 1] - v0;
  if (algo == GRD_PR_PULL) {
    GRD_HIP(hipMemcpy(out, b->r + v0, n * 4, hipMemcpyDeviceToHost));
  } else if (algo == GRD_CC) {
    // the last superstep of graph i wrote buffer iters % 2 (and changed nothing)
    const uint32_t *src = b->lab[b->last_iters[i] & 1];
    GRD_HIP(hipMemcpy(out, src + v0, n * 4, hipMemcpyDeviceToHost));
  } else {
    return fail(GRD_ERR_ARG, "grd_batch_download: algo must be GRD_PR_PULL or GRD_CC");
  }
  return GRD_OK;
}

}  // extern "C"
