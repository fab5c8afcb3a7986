// prpull.hip -- normalised pull PageRank over the partitioned layout.
//
//   r'(v) = (1-d)/n + d * ( sum_{u->v} r(u)/deg(u) + D/n ),  D = sum_{deg(u)=0} r(u)
//
// Every vertex is active every iteration, so both exchange directions use the
// dense form: the compute partitions put r/deg of their vertices into the
// memory partitions' own proxies (the update kernel stores them there
// directly: the own-proxy words of a (master, mirror) pair map one to one onto
// the pair's slot words), each memory partition pulls over its local
// CSC (FP32, Kahan-compensated per row) and writes one partial sum per proxy
// (own and cross mirrors) into its reduce buffer, and the compute partitions
// combine the partial sums of each vertex's proxies (own first, then cross
// mirrors in ascending partition order, again compensated).  The dangling
// mass D and the L1 residual are accumulated in FP64, per workgroup and then
// in a fixed order by k_prp_finish.
#include <algorithm>

#include "grd_device.hpp"
#include "grd_internal.hpp"

namespace grd {

grd_status prp_pull_launch(grd_state *s);
grd_status prp_pull_fixup(grd_state *s);

// compute update: combine proxy partials, new rank, FP64 residual and
// dangling mass, and the next contributions put into the own proxies
__global__ void __launch_bounds__(kBlock) k_prp_update(uint32_t nw, const uint8_t *__restrict__ word_nvalid,
                                                       const uint32_t *__restrict__ deg,
                                                       const uint32_t *__restrict__ word_dst,
                                                       const uint32_t *__restrict__ xptr,
                                                       const uint32_t *__restrict__ xlist,
                                                       const float *__restrict__ valR, float *__restrict__ r,
                                                       float *__restrict__ mprop, float damping, double inv_n,
                                                       int parity, const Counters *cnt, double *__restrict__ red)
{
  const double Dcur = cnt->dangling[parity];
  const float base = (float)((1.0 - (double)damping) * inv_n + (double)damping * Dcur * inv_n);
  const uint32_t lane = lane_id(), stride = gridDim.x * kWavesPerBlock;
  double res = 0.0, dang = 0.0;
  // two slot words per wave and pass; the loads that depend on nothing else
  // (degree, old rank, cross-list bounds, the word's proxy word and valid
  // count) are all issued before any is waited for, then the cross lists,
  // then the partials.  A word's valid slots are a prefix of it.
  constexpr int kW = 2;
  for (uint32_t w0 = blockIdx.x * kWavesPerBlock + wave_in_block(); w0 < nw; w0 += kW * stride) {
    uint32_t s[kW], pw[kW], x0[kW], nx[kW], dg[kW];
    float ro[kW];
    bool ok[kW];
#pragma unroll
    for (int j = 0; j < kW; j++) {
      const uint32_t w = w0 + j * stride;
      ok[j] = w < nw;
      s[j] = (ok[j] ? w : w0) * 64 + lane;
      const uint32_t nvalid = word_nvalid[ok[j] ? w : w0];
      pw[j] = word_dst[ok[j] ? w : w0];
      x0[j] = xptr[s[j]];
      nx[j] = xptr[s[j] + 1];
      dg[j] = deg[s[j]];
      ro[j] = r[s[j]];
      ok[j] = ok[j] && lane < nvalid && pw[j] != 0xFFFFFFFFu;
    }
    // the own partial and up to three cross-mirror partials (every vertex
    // at M <= 4) loaded together, then summed in the same order as the loop
    // that takes any further ones
    uint32_t i[kW][3];
    float own[kW], v[kW][3];
#pragma unroll
    for (int j = 0; j < kW; j++) {
      nx[j] = ok[j] ? nx[j] - x0[j] : 0u;
#pragma unroll
      for (int q = 0; q < 3; q++) i[j][q] = nx[j] > (uint32_t)q ? xlist[x0[j] + q] : 0u;
      own[j] = ok[j] ? valR[pw[j] * 64 + lane] : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < kW; j++)
#pragma unroll
      for (int q = 0; q < 3; q++) v[j][q] = nx[j] > (uint32_t)q ? valR[i[j][q]] : 0.0f;
#pragma unroll
    for (int j = 0; j < kW; j++) {
      if (!ok[j]) continue;
      Kahan k;
      k.add(own[j]);
#pragma unroll
      for (int q = 0; q < 3; q++)
        if (nx[j] > (uint32_t)q) k.add(v[j][q]);
      for (uint32_t x = x0[j] + 3; x < x0[j] + nx[j]; x++) k.add(valR[xlist[x]]);
      const float rn = base + damping * k.value();
      res += fabs((double)rn - (double)ro[j]);
      if (dg[j] == 0) dang += (double)rn;
      r[s[j]] = rn;
      mprop[pw[j] * 64 + lane] = dg[j] ? rn / (float)dg[j] : 0.0f;
    }
  }
  res = block_sum(res);
  dang = block_sum(dang);
  if (threadIdx.x == 0) {
    red[2 * blockIdx.x] = res;
    red[2 * blockIdx.x + 1] = dang;
  }
}

// Sums the update's per-workgroup (residual, dangling) pairs in a fixed order,
// so the convergence test does not depend on scheduling.
__global__ void __launch_bounds__(kBlock) k_prp_finish(const double *__restrict__ red, uint32_t nblk, int parity,
                                                       Counters *cnt)
{
  double res = 0.0, dang = 0.0;
  for (uint32_t b = threadIdx.x; b < nblk; b += kBlock) {
    res += red[2 * b];
    dang += red[2 * b + 1];
  }
  res = block_sum(res);
  dang = block_sum(dang);
  if (threadIdx.x == 0) {
    cnt->residual = res;
    cnt->dangling[parity ^ 1] = dang;
  }
}

__global__ void __launch_bounds__(kBlock) k_prp_init(uint32_t S, const uint32_t *__restrict__ slot_gid,
                                                     const uint32_t *__restrict__ deg,
                                                     const uint32_t *__restrict__ word_dst, float *__restrict__ r,
                                                     float *__restrict__ mprop, float r0)
{
  for (uint32_t s = blockIdx.x * kBlock + threadIdx.x; s < S; s += gridDim.x * kBlock) {
    bool valid = slot_gid[s] != 0xFFFFFFFFu;
    uint32_t dg = deg[s];
    r[s] = valid ? r0 : 0.0f;
    uint32_t pw = word_dst[s >> 6];
    if (pw != 0xFFFFFFFFu) mprop[pw * 64 + (s & 63)] = (valid && dg) ? r0 / (float)dg : 0.0f;
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
                     g->slot_gid, g->slot_deg, g->word_dst, s->pr, (float *)s->mprop, r0);
  // D0 = sum over sinks of the FP32 initial rank, accumulated in FP64
  double d0 = (double)g->n_sinks * (double)r0;
  hipLaunchKernelGGL(k_prp_counters, dim3(1), dim3(64), 0, st, s->cnt, 0, d0);
  GRD_HIP(hipGetLastError());
  s->iteration = 0;
  s->bytes = 0;
  s->edges = 0;
  s->last_frontier = g->n;
  s->last_frontier_edges = g->m;
  return GRD_OK;
}

// workgroups of the update: eight per CU (every wave slot of the CU), each
// wave striding over slot words; one (residual, dangling) pair per workgroup
// for k_prp_finish
uint32_t prp_grid(const grd_graph *g)
{
  const uint32_t nwS = g->S / 64;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((nwS + 3) / 4, (uint64_t)g->ctx->num_cus * 8));
}

// broadcast step: the dense put of r/deg into the memory partitions' own
// proxies was stored by the previous update (or init); only the byte count
// of the exchange is left
grd_status prp_broadcast(grd_state *s)
{
  s->bytes += (uint64_t)s->g->n * 4;
  return GRD_OK;
}

// traverse step: the pull writes each proxy's partial sum straight into the
// dense reduce buffer (the memory partition's outgoing message)
grd_status prp_traverse(grd_state *s)
{
  grd_graph *g = s->g;
  hipStream_t st = g->ctx->stream;
  const bool timed = s->prm.time_kernel != 0;
  if (timed) GRD_HIP(hipEventRecord(s->kev[0], st));
  if (s->tiled) {
    // tiled pull, long-pair fix-up and the per-row merge: the whole pull
    GRD_TRY(prp_pull_tiled(s));
  } else {
    GRD_TRY(prp_pull_launch(s));
  }
  if (timed) {
    GRD_HIP(hipEventRecord(s->kev[1], st));
    s->kern_pending = 1;
    s->kern_pending_bytes = 8 * g->m + 8ull * g->P;
  }
  if (!s->tiled) GRD_TRY(prp_pull_fixup(s));
  s->edges += g->m;
  s->bytes += ((uint64_t)g->n + g->X) * 4;
  return GRD_OK;
}

// update step: aggregate the proxies of every vertex, new ranks, residual
grd_status prp_update(grd_state *s)
{
  grd_graph *g = s->g;
  const int parity = (int)(s->iteration & 1);
  const uint32_t grid = prp_grid(g);
  hipLaunchKernelGGL(k_prp_update, dim3(grid), dim3(kBlock), 0, g->ctx->stream, g->S / 64, g->word_nvalid, g->slot_deg,
                     g->word_dst, g->xptr, g->xlist, (const float *)s->valR, s->pr, (float *)s->mprop,
                     (float)s->prm.damping, 1.0 / (double)g->n, parity, s->cnt, s->red);
  hipLaunchKernelGGL(k_prp_finish, dim3(1), dim3(kBlock), 0, g->ctx->stream, (const double *)s->red, grid, parity,
                     s->cnt);
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

grd_status prp_iteration(grd_state *s)
{
  GRD_TRY(prp_broadcast(s));
  GRD_TRY(prp_traverse(s));
  return prp_update(s);
}

}  // namespace grd
