// mlp.hip -- schedule selector: MLP 16 -> 64 -> 64 -> 4 (ReLU), trained and
// evaluated by these kernels on the GPU.
//
// Parameter layout (5,508 values, FP32 masters during training, BF16 for
// inference):
//   W1[64][16]  b1[64]  W2[64][64]  b2[64]  W3[4][64]  b3[4]
// Output o indexes the schedule option (bit 0 pull, bit 1 dense sync).  The
// network is trained as a classifier of the fastest option with a
// cost-weighted softmax cross-entropy: sample weight = clamp((t_max -
// t_min) / t_min, 0.1, 10), so iterations where the choice matters weigh
// more.  One wave evaluates one sample, lane j owning hidden unit j; the
// forward pass of training reads BF16-rounded weights so that training and
// inference see the same network.  The selector only picks the schedule of
// an iteration: every option computes the same result.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "grd_device.hpp"
#include "grd_internal.hpp"

namespace grd {

constexpr int kIn = 16, kH = 64, kOut = 4;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH, oB3 = oW3 + kOut * kH;
constexpr int kParams = oB3 + kOut;  // 5508

// ---- staging of the allocation-free inference (grd_selector_infer_noalloc) ----
//
// Features and logits of up to kInferMax samples, in a file-scope device
// array: every GPU the library runs on holds its own copy (320 KB) until the
// process ends.  In host code `scratch` names a host-side stand-in, not the
// device copy: its device address is taken once per device by
// mlp_scratch_init (from grd_create, with that device current) and only that
// address is used.  Calls on one GPU share the array: calls that overlap on
// one GPU race.
constexpr uint32_t kInferMax = 4096;
static __device__ __align__(256) float scratch[kInferMax * (kIn + kOut)];
constexpr int kMaxDevices = 64;
static float *g_scratch_ptr[kMaxDevices] = {};

grd_status mlp_scratch_init(int device)
{
  if (device < 0 || device >= kMaxDevices) return fail(GRD_ERR_ARG, "selector scratch: device id beyond 63");
  if (g_scratch_ptr[device]) return GRD_OK;
  float *ptr = nullptr;
  GRD_HIP(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) return fail(GRD_ERR_HIP, "selector scratch: no device address");
  g_scratch_ptr[device] = ptr;
  return GRD_OK;
}

__host__ __device__ __forceinline__ float bf16_to_f32(uint16_t h)
{
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
// round to nearest even
__host__ __device__ __forceinline__ uint16_t f32_to_bf16(float f)
{
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)(u >> 16);  // inf / nan
  uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}

// forward of one sample by one wave; returns logits in lanes 0..3 (value
// broadcast to all lanes through the out array).  Weights are read through
// `wf` (FP32) or `wb` (BF16) depending on which pointer is non-null.
struct Fwd {
  float h1, h2;      // this lane's hidden units (post-ReLU)
  float logit[kOut];  // all lanes hold all logits
};

template <bool BF16>
__device__ __forceinline__ float wload(const float *wf, const uint16_t *wb, int i)
{
  if constexpr (BF16) return bf16_to_f32(wb[i]);
  else return wf[i];
}

template <bool BF16>
__device__ __forceinline__ Fwd forward_wave(const float *wf, const uint16_t *wb, const float *x, float *s_h1)
{
  const int j = lane_id();
  Fwd r;
  float a = wload<BF16>(wf, wb, oB1 + j);
#pragma unroll
  for (int i = 0; i < kIn; i++) a += wload<BF16>(wf, wb, oW1 + j * kIn + i) * x[i];
  r.h1 = a > 0.0f ? a : 0.0f;
  s_h1[j] = r.h1;
  __syncthreads();  // callers keep whole blocks in step
  float b = wload<BF16>(wf, wb, oB2 + j);
  for (int k = 0; k < kH; k++) b += wload<BF16>(wf, wb, oW2 + j * kH + k) * s_h1[k];
  r.h2 = b > 0.0f ? b : 0.0f;
#pragma unroll
  for (int o = 0; o < kOut; o++) r.logit[o] = wave_sum(wload<BF16>(wf, wb, oW3 + o * kH + j) * r.h2);
#pragma unroll
  for (int o = 0; o < kOut; o++) r.logit[o] += wload<BF16>(wf, wb, oB3 + o);
  return r;
}

// Forward + backward over samples [0, n), gradients accumulated into grad
// (atomic adds), loss into *loss.  Each wave handles samples w, w + W, ...
template <bool BF16>
__global__ void __launch_bounds__(kBlock) k_mlp_grad(const float *__restrict__ wf, const uint16_t *__restrict__ wb,
                                                     const float *__restrict__ feat, const uint32_t *__restrict__ label,
                                                     const float *__restrict__ sw, const uint32_t *__restrict__ idx,
                                                     uint32_t n, float *grad, float *loss)
{
  __shared__ float s_h1[kWavesPerBlock][kH];
  __shared__ float s_d2[kWavesPerBlock][kH];
  const int j = lane_id(), wv = wave_in_block();
  const uint32_t nwaves = gridDim.x * kWavesPerBlock;
  // per-lane gradient accumulators: row j of W2, column j of W3, row j of W1
  float gW2[kH];
  float gW1[kIn];
  float gW3[kOut];
  float gb1 = 0.0f, gb2 = 0.0f, gb3 = 0.0f;
#pragma unroll
  for (int k = 0; k < kH; k++) gW2[k] = 0.0f;
#pragma unroll
  for (int i = 0; i < kIn; i++) gW1[i] = 0.0f;
#pragma unroll
  for (int o = 0; o < kOut; o++) gW3[o] = 0.0f;
  float lsum = 0.0f;
  const uint32_t iters = (n + nwaves - 1) / nwaves;
  for (uint32_t it = 0; it < iters; it++) {
    uint32_t t = it * nwaves + blockIdx.x * kWavesPerBlock + wv;
    bool active = t < n;
    uint32_t si = active ? (idx ? idx[t] : t) : 0u;
    float x[kIn];
#pragma unroll
    for (int i = 0; i < kIn; i++) x[i] = active ? feat[(uint64_t)si * kIn + i] : 0.0f;
    // forward (wave-local LDS row for h1)
    float a = wload<BF16>(wf, wb, oB1 + j);
#pragma unroll
    for (int i = 0; i < kIn; i++) a += wload<BF16>(wf, wb, oW1 + j * kIn + i) * x[i];
    float h1 = a > 0.0f ? a : 0.0f;
    s_h1[wv][j] = h1;
    __syncthreads();
    float b = wload<BF16>(wf, wb, oB2 + j);
    for (int k = 0; k < kH; k++) b += wload<BF16>(wf, wb, oW2 + j * kH + k) * s_h1[wv][k];
    float h2 = b > 0.0f ? b : 0.0f;
    float lg[kOut];
#pragma unroll
    for (int o = 0; o < kOut; o++) lg[o] = wave_sum(wload<BF16>(wf, wb, oW3 + o * kH + j) * h2) + wload<BF16>(wf, wb, oB3 + o);
    // softmax cross-entropy
    float mx = lg[0];
#pragma unroll
    for (int o = 1; o < kOut; o++) mx = lg[o] > mx ? lg[o] : mx;
    float den = 0.0f, p[kOut];
#pragma unroll
    for (int o = 0; o < kOut; o++) {
      p[o] = expf(lg[o] - mx);
      den += p[o];
    }
    const uint32_t y = active ? label[si] : 0u;
    const float wgt = active ? (sw ? sw[si] : 1.0f) : 0.0f;
#pragma unroll
    for (int o = 0; o < kOut; o++) p[o] /= den;
    if (j == 0) lsum += wgt * (-logf(fmaxf(p[y], 1e-30f)));
    float dl[kOut];
#pragma unroll
    for (int o = 0; o < kOut; o++) dl[o] = wgt * (p[o] - (o == (int)y ? 1.0f : 0.0f));
    // backward
    float d2 = 0.0f;
#pragma unroll
    for (int o = 0; o < kOut; o++) {
      gW3[o] += dl[o] * h2;  // W3[o][j]
      d2 += wload<BF16>(wf, wb, oW3 + o * kH + j) * dl[o];
    }
    if (j < kOut) gb3 += dl[j];
    d2 = h2 > 0.0f ? d2 : 0.0f;
    gb2 += d2;
#pragma unroll
    for (int k = 0; k < kH; k++) gW2[k] += d2 * s_h1[wv][k];  // W2[j][k]
    s_d2[wv][j] = d2;
    __syncthreads();
    float d1 = 0.0f;
    for (int k = 0; k < kH; k++) d1 += wload<BF16>(wf, wb, oW2 + k * kH + j) * s_d2[wv][k];
    d1 = h1 > 0.0f ? d1 : 0.0f;
    gb1 += d1;
#pragma unroll
    for (int i = 0; i < kIn; i++) gW1[i] += d1 * x[i];
    __syncthreads();
  }
  // flush (one atomic per parameter per wave)
#pragma unroll
  for (int k = 0; k < kH; k++) atomicAdd(grad + oW2 + j * kH + k, gW2[k]);
#pragma unroll
  for (int i = 0; i < kIn; i++) atomicAdd(grad + oW1 + j * kIn + i, gW1[i]);
#pragma unroll
  for (int o = 0; o < kOut; o++) atomicAdd(grad + oW3 + o * kH + j, gW3[o]);
  atomicAdd(grad + oB1 + j, gb1);
  atomicAdd(grad + oB2 + j, gb2);
  if (j < kOut) atomicAdd(grad + oB3 + j, gb3);
  if (j == 0) atomicAdd(loss, lsum);
}

// Adam step on FP32 masters; refreshes the BF16 copy and clears the gradient
__global__ void __launch_bounds__(kBlock) k_mlp_adam(float *__restrict__ w, float *__restrict__ m1,
                                                     float *__restrict__ m2, float *__restrict__ g,
                                                     uint16_t *__restrict__ wb, float scale, float lr, float b1,
                                                     float b2, float c1, float c2)
{
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < kParams; i += gridDim.x * kBlock) {
    float gi = g[i] * scale;
    m1[i] = b1 * m1[i] + (1.0f - b1) * gi;
    m2[i] = b2 * m2[i] + (1.0f - b2) * gi * gi;
    float mh = m1[i] / c1, vh = m2[i] / c2;
    w[i] -= lr * mh / (sqrtf(vh) + 1e-8f);
    wb[i] = f32_to_bf16(w[i]);
    g[i] = 0.0f;
  }
}

// inference: one wave per sample, BF16 weights; logits[n*4]
__global__ void __launch_bounds__(kBlock) k_mlp_infer(const uint16_t *__restrict__ wb, const float *__restrict__ feat,
                                                      uint32_t n, float *__restrict__ logits)
{
  __shared__ float s_h1[kWavesPerBlock][kH];
  const uint32_t nwaves = gridDim.x * kWavesPerBlock;
  const uint32_t iters = (n + nwaves - 1) / nwaves;
  for (uint32_t it = 0; it < iters; it++) {
    uint32_t t = it * nwaves + blockIdx.x * kWavesPerBlock + wave_in_block();
    bool active = t < n;
    float x[kIn];
#pragma unroll
    for (int i = 0; i < kIn; i++) x[i] = active ? feat[(uint64_t)t * kIn + i] : 0.0f;
    Fwd r = forward_wave<true>(nullptr, wb, x, s_h1[wave_in_block()]);
    if (active && lane_id() < kOut) {
      float v = r.logit[0];
#pragma unroll
      for (int o = 1; o < kOut; o++)
        if (lane_id() == o) v = r.logit[o];
      logits[(uint64_t)t * kOut + lane_id()] = v;
    }
    __syncthreads();
  }
}

// per-iteration selection: one wave; writes 4 logits + argmax to out[0..4]
__global__ void k_mlp_select(const uint16_t *__restrict__ wb, const float *__restrict__ feat, float *__restrict__ out)
{
  __shared__ float s_h1[kWavesPerBlock][kH];
  float x[kIn];
#pragma unroll
  for (int i = 0; i < kIn; i++) x[i] = feat[i];
  Fwd r = forward_wave<true>(nullptr, wb, x, s_h1[wave_in_block()]);
  if (threadIdx.x == 0) {
    int best = 0;
    for (int o = 0; o < kOut; o++) {
      out[o] = r.logit[o];
      if (r.logit[o] > r.logit[best]) best = o;
    }
    out[4] = (float)best;
  }
}

// Graph-level inputs of the features (the per-superstep ones come from the
// counters); the same formulas as make_features on the host.
struct SelInputs {
  double n, m, P, X, avg_deg, prev_frontier;
  uint32_t M, C, algo, pad;
};

// the logits and the argmax go into the counters (cnt->sel), so that the
// superstep's counter read-back brings the decision along
__global__ void k_mlp_select_state(const uint16_t *__restrict__ wb, Counters *cnt, SelInputs in)
{
  __shared__ float s_x[kIn];
  __shared__ float s_h1[kWavesPerBlock][kH];
  if (threadIdx.x == 0) {
    const double fr = (double)cnt->frontier, fe = (double)cnt->frontier_edges;
    const double m = in.m > 1.0 ? in.m : 1.0;
    s_x[0] = (float)(log2(1.0 + fr) / 32.0);
    s_x[1] = (float)(fr / in.n);
    s_x[2] = (float)(fe / m);
    s_x[3] = (float)(log2(1.0 + fe) / 40.0);
    s_x[4] = (float)((double)cnt->dirty / in.P);
    s_x[5] = (float)((double)cnt->frontier_maxdeg / (fe + 1.0));
    s_x[6] = (float)fmin(4.0, (fr > 0 ? fe / fr : 0.0) / fmax(1e-9, in.avg_deg * 4.0));
    s_x[7] = (float)(log2(in.n) / 32.0);
    s_x[8] = (float)fmin(4.0, in.avg_deg / 64.0);
    s_x[9] = (float)(in.M / 16.0);
    s_x[10] = (float)(in.C / 16.0);
    s_x[11] = (float)fmin(4.0, in.X / in.n / 4.0);
    const double ratio = in.prev_frontier > 0 ? fr / in.prev_frontier : 1.0;
    s_x[12] = (float)(fmax(-1.0, fmin(1.0, log2(fmax(ratio, 1e-9)) / 16.0)) * 0.5 + 0.5);
    s_x[13] = in.algo == GRD_PR ? 1.0f : 0.0f;
    s_x[14] = in.algo == GRD_CC ? 1.0f : 0.0f;
    s_x[15] = in.algo == GRD_SSSP ? 1.0f : 0.0f;
  }
  __syncthreads();
  float x[kIn];
#pragma unroll
  for (int i = 0; i < kIn; i++) x[i] = s_x[i];
  Fwd r = forward_wave<true>(nullptr, wb, x, s_h1[wave_in_block()]);
  if (threadIdx.x == 0) {
    int best = 0;
    for (int o = 0; o < kOut; o++) {
      cnt->sel[o] = r.logit[o];
      if (r.logit[o] > r.logit[best]) best = o;
    }
    cnt->sel[4] = (float)best;
  }
}

// ------------------------------------------------------------------------

static grd_status sel_buffers(grd_ctx *ctx)
{
  if (!ctx->sel_out) {
    GRD_HIP(hipMalloc((void **)&ctx->sel_out, 32 * sizeof(float)));
    GRD_HIP(hipHostMalloc((void **)&ctx->sel_out_host, 32 * sizeof(float), hipHostMallocDefault));
  }
  return GRD_OK;
}

static void decode_logits(const float *o, int *option, float *margin)
{
  int best = (int)o[4];
  float second = -1e30f;
  for (int k = 0; k < kOut; k++)
    if (k != best) second = std::max(second, o[k]);
  *option = best;
  *margin = o[best] - second;
}

grd_status mlp_select_queue(grd_state *s, uint64_t prev_frontier)
{
  grd_graph *g = s->g;
  grd_ctx *ctx = g->ctx;
  SelInputs in{(double)g->n, (double)g->m, (double)g->P, (double)g->X, g->avg_degree, (double)prev_frontier,
               (uint32_t)g->M, (uint32_t)g->C, (uint32_t)s->algo, 0u};
  hipLaunchKernelGGL(k_mlp_select_state, dim3(1), dim3(kWave), 0, ctx->stream, ctx->sel_w, s->cnt, in);
  GRD_HIP(hipGetLastError());
  return GRD_OK;
}

void mlp_select_read(const grd_state *s, int *option, float *margin) { decode_logits(s->hcnt->sel, option, margin); }

grd_status mlp_select(grd_ctx *ctx, const float *feat16, int *option, float *margin)
{
  GRD_TRY(sel_buffers(ctx));
  memcpy(ctx->sel_out_host + 8, feat16, 16 * sizeof(float));
  GRD_HIP(hipMemcpyAsync(ctx->sel_out + 8, ctx->sel_out_host + 8, 16 * sizeof(float), hipMemcpyHostToDevice,
                         ctx->stream));
  hipLaunchKernelGGL(k_mlp_select, dim3(1), dim3(kWave), 0, ctx->stream, ctx->sel_w, ctx->sel_out + 8, ctx->sel_out);
  GRD_HIP(hipGetLastError());
  GRD_HIP(hipMemcpyAsync(ctx->sel_out_host, ctx->sel_out, 8 * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  GRD_HIP(hipStreamSynchronize(ctx->stream));
  decode_logits(ctx->sel_out_host, option, margin);
  return GRD_OK;
}

}  // namespace grd

using namespace grd;

extern "C" {

grd_status grd_selector_load(grd_ctx *ctx, const uint16_t *w, uint64_t count, float min_margin)
{
  if (!ctx || !w) return fail(GRD_ERR_ARG, "grd_selector_load: NULL argument");
  if (count != (uint64_t)kParams)
    return fail(GRD_ERR_ARG, "grd_selector_load: expected " + std::to_string(kParams) + " BF16 weights");
  GRD_HIP(hipSetDevice(ctx->device));
  if (!ctx->sel_w) GRD_HIP(hipMalloc((void **)&ctx->sel_w, kParams * sizeof(uint16_t)));
  GRD_HIP(hipMemcpy(ctx->sel_w, w, kParams * sizeof(uint16_t), hipMemcpyHostToDevice));
  ctx->sel_loaded = true;
  ctx->sel_margin = min_margin;
  return GRD_OK;
}

grd_status grd_selector_train(grd_ctx *ctx, const float *feat, const uint32_t *label, const float *cost, uint64_t n,
                              uint32_t epochs, uint32_t batch, float lr, uint64_t seed, uint16_t *out_bf16,
                              float *final_loss)
{
  if (!ctx || !feat || !label || n == 0 || n > (1ull << 31) || batch == 0)
    return fail(GRD_ERR_ARG, "grd_selector_train: bad argument");
  for (uint64_t i = 0; i < n; i++)
    if (label[i] >= (uint32_t)kOut) return fail(GRD_ERR_ARG, "grd_selector_train: label out of range");
  GRD_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  // sample weights from the per-option costs
  std::vector<float> sw(n, 1.0f);
  if (cost)
    for (uint64_t i = 0; i < n; i++) {
      float mn = cost[i * 4], mx = cost[i * 4];
      for (int o = 1; o < 4; o++) {
        mn = std::min(mn, cost[i * 4 + o]);
        mx = std::max(mx, cost[i * 4 + o]);
      }
      float r = mn > 0 ? (mx - mn) / mn : 1.0f;
      sw[i] = std::min(10.0f, std::max(0.1f, r));
    }
  // He initialisation from a seeded SplitMix stream
  std::vector<float> w0(kParams, 0.0f);
  uint64_t z = seed ? seed : 0x9E3779B97F4A7C15ull;
  auto uni = [&]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
  };
  auto init = [&](int off, int cnt, int fan_in) {
    double a = std::sqrt(6.0 / fan_in);
    for (int i = 0; i < cnt; i++) w0[off + i] = (float)((uni() * 2.0 - 1.0) * a);
  };
  init(oW1, kH * kIn, kIn);
  init(oW2, kH * kH, kH);
  init(oW3, kOut * kH, kH);
  DevMem mem;
  float *dw, *dm1, *dm2, *dg, *dfeat, *dsw, *dloss;
  uint32_t *dlab, *didx;
  uint16_t *dwb;
  GRD_TRY(mem.alloc_n(&dw, kParams));
  GRD_TRY(mem.alloc_n(&dm1, kParams));
  GRD_TRY(mem.alloc_n(&dm2, kParams));
  GRD_TRY(mem.alloc_n(&dg, kParams));
  GRD_TRY(mem.alloc_n(&dwb, kParams));
  GRD_TRY(mem.alloc_n(&dfeat, n * kIn));
  GRD_TRY(mem.alloc_n(&dsw, n));
  GRD_TRY(mem.alloc_n(&dlab, n));
  GRD_TRY(mem.alloc_n(&didx, n));
  GRD_TRY(mem.alloc_n(&dloss, 1));
  std::vector<uint16_t> wb0(kParams);
  for (int i = 0; i < kParams; i++) wb0[i] = f32_to_bf16(w0[i]);
  GRD_HIP(hipMemcpy(dw, w0.data(), kParams * 4, hipMemcpyHostToDevice));
  GRD_HIP(hipMemcpy(dwb, wb0.data(), kParams * 2, hipMemcpyHostToDevice));
  GRD_HIP(hipMemsetAsync(dm1, 0, kParams * 4, st));
  GRD_HIP(hipMemsetAsync(dm2, 0, kParams * 4, st));
  GRD_HIP(hipMemsetAsync(dg, 0, kParams * 4, st));
  GRD_HIP(hipMemcpy(dfeat, feat, n * kIn * 4, hipMemcpyHostToDevice));
  GRD_HIP(hipMemcpy(dsw, sw.data(), n * 4, hipMemcpyHostToDevice));
  GRD_HIP(hipMemcpy(dlab, label, n * 4, hipMemcpyHostToDevice));
  std::vector<uint32_t> perm(n);
  for (uint64_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
  const float b1 = 0.9f, b2 = 0.999f;
  uint64_t step = 0;
  float last = 0.0f;
  for (uint32_t ep = 0; ep < epochs; ep++) {
    for (uint64_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[(uint64_t)(uni() * (double)(i + 1)) % (i + 1)]);
    GRD_HIP(hipMemcpyAsync(didx, perm.data(), n * 4, hipMemcpyHostToDevice, st));
    GRD_HIP(hipMemsetAsync(dloss, 0, 4, st));
    for (uint64_t b0 = 0; b0 < n; b0 += batch) {
      uint32_t nb = (uint32_t)std::min<uint64_t>(batch, n - b0);
      uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(64, (nb + 15) / 16));
      hipLaunchKernelGGL(k_mlp_grad<true>, dim3(grid), dim3(kBlock), 0, st, dw, dwb, dfeat, dlab, dsw, didx + b0, nb, dg,
                         dloss);
      step++;
      float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
      hipLaunchKernelGGL(k_mlp_adam, dim3((kParams + kBlock - 1) / kBlock), dim3(kBlock), 0, st, dw, dm1, dm2, dg, dwb,
                         1.0f / (float)nb, lr, b1, b2, c1, c2);
    }
    GRD_HIP(hipGetLastError());
    float ls = 0.0f;
    GRD_HIP(hipMemcpyAsync(&ls, dloss, 4, hipMemcpyDeviceToHost, st));
    GRD_HIP(hipStreamSynchronize(st));
    last = ls / (float)n;
  }
  std::vector<uint16_t> wb(kParams);
  GRD_HIP(hipMemcpyAsync(wb.data(), dwb, kParams * 2, hipMemcpyDeviceToHost, st));
  GRD_HIP(hipStreamSynchronize(st));
  if (out_bf16) memcpy(out_bf16, wb.data(), kParams * 2);
  if (final_loss) *final_loss = last;
  return grd_selector_load(ctx, wb.data(), kParams, ctx->sel_margin);
}

grd_status grd_selector_infer(grd_ctx *ctx, const float *feat, uint64_t n, float *logits)
{
  if (!ctx || !feat || !logits || n == 0) return fail(GRD_ERR_ARG, "grd_selector_infer: bad argument");
  if (!ctx->sel_loaded) return fail(GRD_ERR_STATE, "grd_selector_infer: no weights loaded");
  GRD_HIP(hipSetDevice(ctx->device));
  DevMem mem;
  float *df, *dl;
  GRD_TRY(mem.alloc_n(&df, n * kIn));
  GRD_TRY(mem.alloc_n(&dl, n * kOut));
  GRD_HIP(hipMemcpy(df, feat, n * kIn * 4, hipMemcpyHostToDevice));
  uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(1024, (n + 3) / 4));
  hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->sel_w, df, (uint32_t)n, dl);
  GRD_HIP(hipGetLastError());
  GRD_HIP(hipMemcpyAsync(logits, dl, n * kOut * 4, hipMemcpyDeviceToHost, ctx->stream));
  GRD_HIP(hipStreamSynchronize(ctx->stream));
  return GRD_OK;
}

grd_status grd_selector_infer_noalloc(grd_ctx *ctx, const float *feat, uint64_t n, float *logits)
{
  if (!ctx || !feat || !logits || n == 0) return fail(GRD_ERR_ARG, "grd_selector_infer_noalloc: bad argument");
  if (n > kInferMax)
    return fail(GRD_ERR_ARG, "grd_selector_infer_noalloc: at most " + std::to_string(kInferMax) + " samples per call");
  if (!ctx->sel_loaded) return fail(GRD_ERR_STATE, "grd_selector_infer_noalloc: no weights loaded");
  if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_scratch_ptr[ctx->device])
    return fail(GRD_ERR_STATE, "grd_selector_infer_noalloc: scratch not set up for this device");
  GRD_HIP(hipSetDevice(ctx->device));
  // staging in the device's scratch copy: features, then logits; every
  // element read is written first in this call
  float *df = g_scratch_ptr[ctx->device];
  float *dl = df + (size_t)kInferMax * kIn;
  GRD_HIP(hipMemcpyAsync(df, feat, n * kIn * 4, hipMemcpyHostToDevice, ctx->stream));
  uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(1024, (n + 3) / 4));
  hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(kBlock), 0, ctx->stream, ctx->sel_w, df, (uint32_t)n, dl);
  GRD_HIP(hipGetLastError());
  GRD_HIP(hipMemcpyAsync(logits, dl, n * kOut * 4, hipMemcpyDeviceToHost, ctx->stream));
  GRD_HIP(hipStreamSynchronize(ctx->stream));
  return GRD_OK;
}

grd_status grd_selector_grad(grd_ctx *ctx, const float *w, const float *feat, const uint32_t *label, const float *cost,
                             uint64_t n, float *loss, float *grad)
{
  if (!ctx || !w || !feat || !label || !loss || !grad || n == 0) return fail(GRD_ERR_ARG, "grd_selector_grad: bad argument");
  for (uint64_t i = 0; i < n; i++)
    if (label[i] >= (uint32_t)kOut) return fail(GRD_ERR_ARG, "grd_selector_grad: label out of range");
  GRD_HIP(hipSetDevice(ctx->device));
  DevMem mem;
  float *dw, *df, *dsw = nullptr, *dg, *dloss;
  uint32_t *dl;
  GRD_TRY(mem.alloc_n(&dw, kParams));
  GRD_TRY(mem.alloc_n(&df, n * kIn));
  GRD_TRY(mem.alloc_n(&dg, kParams));
  GRD_TRY(mem.alloc_n(&dloss, 1));
  GRD_TRY(mem.alloc_n(&dl, n));
  GRD_HIP(hipMemcpy(dw, w, kParams * 4, hipMemcpyHostToDevice));
  GRD_HIP(hipMemcpy(df, feat, n * kIn * 4, hipMemcpyHostToDevice));
  GRD_HIP(hipMemcpy(dl, label, n * 4, hipMemcpyHostToDevice));
  if (cost) {
    GRD_TRY(mem.alloc_n(&dsw, n));
    GRD_HIP(hipMemcpy(dsw, cost, n * 4, hipMemcpyHostToDevice));
  }
  GRD_HIP(hipMemsetAsync(dg, 0, kParams * 4, ctx->stream));
  GRD_HIP(hipMemsetAsync(dloss, 0, 4, ctx->stream));
  uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(64, (n + 15) / 16));
  hipLaunchKernelGGL(k_mlp_grad<false>, dim3(grid), dim3(kBlock), 0, ctx->stream, dw, nullptr, df, dl, dsw, nullptr,
                     (uint32_t)n, dg, dloss);
  GRD_HIP(hipGetLastError());
  GRD_HIP(hipMemcpyAsync(grad, dg, kParams * 4, hipMemcpyDeviceToHost, ctx->stream));
  GRD_HIP(hipMemcpyAsync(loss, dloss, 4, hipMemcpyDeviceToHost, ctx->stream));
  GRD_HIP(hipStreamSynchronize(ctx->stream));
  return GRD_OK;
}

}  // extern "C"
