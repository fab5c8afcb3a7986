// prpull.hip -- normalised pull PageRank over the partitioned layout.
//
//   r'(v) = (1-d)/n + d * ( sum_{u->v} r(u)/deg(u) + D/n ),  D = sum_{deg(u)=0} r(u)
//
// Every vertex is active every iteration, so both exchange directions use the
// dense form: the compute partitions put r/deg of their vertices into the
// memory partitions' own proxies, each memory partition pulls over its local
// CSC (FP32, Kahan-compensated per row) and writes one partial sum per proxy
// (own and cross mirrors) into its reduce buffer, and the compute partitions
// combine the partial sums of each vertex's proxies (own first, then cross
// mirrors in ascending partition order, again compensated).  The dangling
// mass D and the L1 residual are accumulated in FP64.
#include <algorithm>

#include "grd_device.hpp"
#include "grd_internal.hpp"

namespace grd {

// dense receive: memory own proxies <- compute contributions
__global__ void __launch_bounds__(kBlock) k_prp_recv(uint32_t nw, const uint32_t *__restrict__ word_dst,
                                                     const float *__restrict__ valB, float *__restrict__ mprop)
{
  for (uint32_t w = blockIdx.x * kWavesPerBlock + wave_in_block(); w < nw; w += gridDim.x * kWavesPerBlock) {
    uint32_t pw = word_dst[w];
    if (pw == 0xFFFFFFFFu) continue;
    mprop[pw * 64 + lane_id()] = valB[w * 64 + lane_id()];
  }
}

// G lanes per CSC row, Kahan per lane, compensated merge across the group
template <int G>
__global__ void __launch_bounds__(kBlock) k_prp_pull(uint32_t P, const uint32_t *__restrict__ crp,
                                                     const uint32_t *__restrict__ ccol, const float *__restrict__ mprop,
                                                     float *__restrict__ out, uint32_t H)
{
  constexpr int RPW = kWave / G;
  const int sub = lane_id() % G;
  const uint32_t nwaves = gridDim.x * kWavesPerBlock;
  for (uint32_t wv = blockIdx.x * kWavesPerBlock + wave_in_block(); (uint64_t)wv * RPW < P; wv += nwaves) {
    uint32_t r = wv * RPW + lane_id() / G;
    uint32_t b = 0, e = 0;
    bool heavy = false;
    if (r < P) {
      b = crp[r];
      e = crp[r + 1];
      if (e - b >= H) {
        heavy = true;
        e = b;
      }
    }
    Kahan k;
    for (uint32_t x = b + sub; x < e; x += G) k.add(mprop[ccol[x]]);
    kahan_wave_merge(k, G);
    if (sub == 0 && r < P && !heavy) out[r] = k.value();
  }
}

__global__ void __launch_bounds__(kBlock) k_prp_pull_heavy(const uint32_t *__restrict__ rows, uint32_t nrows,
                                                           const uint32_t *__restrict__ crp,
                                                           const uint32_t *__restrict__ ccol,
                                                           const float *__restrict__ mprop, float *__restrict__ out)
{
  __shared__ float s_s[kWavesPerBlock], s_c[kWavesPerBlock];
  for (uint32_t i = blockIdx.x; i < nrows; i += gridDim.x) {
    uint32_t r = rows[i];
    uint32_t b = crp[r], e = crp[r + 1];
    Kahan k;
    for (uint32_t x = b + threadIdx.x; x < e; x += kBlock) k.add(mprop[ccol[x]]);
    kahan_wave_merge(k, kWave);
    if (lane_id() == 0) {
      s_s[wave_in_block()] = k.s;
      s_c[wave_in_block()] = k.c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      Kahan t;
      t.s = s_s[0];
      t.c = s_c[0];
      for (int j = 1; j < kWavesPerBlock; j++) t.merge(s_s[j], s_c[j]);
      out[r] = t.value();
    }
    __syncthreads();
  }
}

// compute update: combine proxy partials, new rank, FP64 residual and
// dangling mass, and the next contributions (dense put source)
__global__ void __launch_bounds__(kBlock) k_prp_update(uint32_t nw, const uint32_t *__restrict__ slot_gid,
                                                       const uint32_t *__restrict__ deg,
                                                       const uint32_t *__restrict__ word_dst,
                                                       const uint32_t *__restrict__ xptr,
                                                       const uint32_t *__restrict__ xlist,
                                                       const float *__restrict__ valR, float *__restrict__ r,
                                                       float *__restrict__ valB, float damping, double inv_n,
                                                       int parity, Counters *cnt)
{
  const double Dcur = cnt->dangling[parity];
  const float base = (float)((1.0 - (double)damping) * inv_n + (double)damping * Dcur * inv_n);
  double res = 0.0, dang = 0.0;
  for (uint32_t w = blockIdx.x * kWavesPerBlock + wave_in_block(); w < nw; w += gridDim.x * kWavesPerBlock) {
    uint32_t s = w * 64 + lane_id();
    if (slot_gid[s] == 0xFFFFFFFFu) continue;
    uint32_t pw = word_dst[w];
    Kahan k;
    k.add(valR[pw * 64 + lane_id()]);
    for (uint32_t x = xptr[s]; x < xptr[s + 1]; x++) k.add(valR[xlist[x]]);
    float rn = base + damping * k.value();
    float ro = r[s];
    res += fabs((double)rn - (double)ro);
    uint32_t dg = deg[s];
    if (dg == 0) dang += (double)rn;
    r[s] = rn;
    valB[s] = dg ? rn / (float)dg : 0.0f;
  }
  res = block_sum(res);
  dang = block_sum(dang);
  if (threadIdx.x == 0) {
    if (res != 0.0) atomicAdd(&cnt->residual, res);
    if (dang != 0.0) atomicAdd(&cnt->dangling[parity ^ 1], dang);
  }
}

__global__ void __launch_bounds__(kBlock) k_prp_init(uint32_t S, const uint32_t *__restrict__ slot_gid,
                                                     const uint32_t *__restrict__ deg, float *__restrict__ r,
                                                     float *__restrict__ valB, float r0)
{
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < S; s += gridDim.x * kBlock) {
    bool valid = slot_gid[s] != 0xFFFFFFFFu;
    uint32_t dg = deg[s];
    r[s] = valid ? r0 : 0.0f;
    valB[s] = (valid && dg) ? r0 / (float)dg : 0.0f;
  }
}

__global__ void k_prp_counters(Counters *cnt, int parity, double d0)
{
  if (threadIdx.x == 0) {
    cnt->residual = 0.0;
    if (d0 >= 0.0) cnt->dangling[parity] = d0;
    cnt->dangling[parity ^ 1] = 0.0;
  }
}

grd_status prp_init(grd_state *s)
{
  grd_graph *g = s->g;
  hipStream_t st = g->ctx->stream;
  const float r0 = (float)(1.0 / (double)g->n);
  hipLaunchKernelGGL(k_prp_init, dim3(std::min<uint32_t>(4096, g->S / kBlock + 1)), dim3(kBlock), 0, st, g->S,
                     g->slot_gid, g->slot_deg, s->pr, (float *)s->valB, r0);
  uint64_t sinks = 0;
  for (uint32_t v = 0; v < g->n; v++) sinks += (g->h_deg[v] == 0);
  // D0 = sum over sinks of the FP32 initial rank, accumulated in FP64
  double d0 = (double)sinks * (double)r0;
  hipLaunchKernelGGL(k_prp_counters, dim3(1), dim3(64), 0, st, s->cnt, 0, d0);
  GRD_HIP(hipGetLastError());
  s->iteration = 0;
  s->bytes = 0;
  s->edges = 0;
  s->last_frontier = g->n;
  s->last_frontier_edges = g->m;
  return GRD_OK;
}

grd_status prp_iteration(grd_state *s)
{
  grd_graph *g = s->g;
  grd_ctx *ctx = g->ctx;
  hipStream_t st = ctx->stream;
  const uint32_t nwS = g->S / 64;
  const int parity = (int)(s->iteration & 1);
  hipLaunchKernelGGL(k_prp_counters, dim3(1), dim3(64), 0, st, s->cnt, parity, -1.0);
  uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((nwS + 3) / 4, (uint64_t)ctx->num_cus * 16));
  hipLaunchKernelGGL(k_prp_recv, dim3(grid), dim3(kBlock), 0, st, nwS, g->word_dst, (const float *)s->valB,
                     (float *)s->mprop);
  const uint32_t rows_per_block = kBlock / 8;
  uint32_t pgrid = std::max<uint32_t>(1, std::min<uint32_t>((g->P + rows_per_block - 1) / rows_per_block,
                                                            ctx->num_cus * 16));
  hipLaunchKernelGGL(k_prp_pull<8>, dim3(pgrid), dim3(kBlock), 0, st, g->P, g->crp, g->ccol, (const float *)s->mprop,
                     (float *)s->valR, g->heavy_in_threshold);
  if (g->n_heavy_in)
    hipLaunchKernelGGL(k_prp_pull_heavy, dim3(std::min<uint32_t>(g->n_heavy_in, ctx->num_cus * 8)), dim3(kBlock), 0, st,
                       g->heavy_in, g->n_heavy_in, g->crp, g->ccol, (const float *)s->mprop, (float *)s->valR);
  hipLaunchKernelGGL(k_prp_update, dim3(grid), dim3(kBlock), 0, st, nwS, g->slot_gid, g->slot_deg, g->word_dst,
                     g->xptr, g->xlist, (const float *)s->valR, s->pr, (float *)s->valB, (float)s->prm.damping,
                     1.0 / (double)g->n, parity, s->cnt);
  GRD_HIP(hipGetLastError());
  s->bytes += (uint64_t)g->n * 4 + ((uint64_t)g->n + g->X) * 4;
  s->edges += g->m;
  return GRD_OK;
}

}  // namespace grd
