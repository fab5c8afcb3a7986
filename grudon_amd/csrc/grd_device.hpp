// grd_device.hpp -- wave64 device helpers shared by the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace grd {

constexpr int kWave = 64;
constexpr int kBlock = 256;  // 4 waves
constexpr int kWavesPerBlock = kBlock / kWave;

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }
__device__ __forceinline__ int wave_in_block() { return (int)(threadIdx.x / kWave); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T t = __shfl_xor(v, o, kWave);
    v = t > v ? t : v;
  }
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_min(T v)
{
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T t = __shfl_xor(v, o, kWave);
    v = t < v ? t : v;
  }
  return v;
}

// Inclusive scan across the wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t t = __shfl_up(v, o, kWave);
    if (l >= o) v += t;
  }
  return v;
}

// Exclusive scan across a 256-thread block; returns the exclusive prefix of
// this thread and writes the block total to *total.  Uses 8 words of LDS.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *total)
{
  __shared__ uint32_t s_w[kWavesPerBlock];
  uint32_t inc = wave_incl_scan(v);
  if (lane_id() == kWave - 1) s_w[wave_in_block()] = inc;
  __syncthreads();
  uint32_t base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kWavesPerBlock; i++) {
    uint32_t x = s_w[i];
    if (i < wave_in_block()) base += x;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

template <typename T>
__device__ __forceinline__ T block_sum(T v)
{
  __shared__ T s_p[kWavesPerBlock];
  v = wave_sum(v);
  if (lane_id() == 0) s_p[wave_in_block()] = v;
  __syncthreads();
  T t = 0;
#pragma unroll
  for (int i = 0; i < kWavesPerBlock; i++) t += s_p[i];
  __syncthreads();
  return t;
}
template <typename T>
__device__ __forceinline__ T block_max(T v)
{
  __shared__ T s_p[kWavesPerBlock];
  v = wave_max(v);
  if (lane_id() == 0) s_p[wave_in_block()] = v;
  __syncthreads();
  T t = s_p[0];
#pragma unroll
  for (int i = 1; i < kWavesPerBlock; i++) t = s_p[i] > t ? s_p[i] : t;
  __syncthreads();
  return t;
}
template <typename T>
__device__ __forceinline__ T block_min(T v)
{
  __shared__ T s_p[kWavesPerBlock];
  v = wave_min(v);
  if (lane_id() == 0) s_p[wave_in_block()] = v;
  __syncthreads();
  T t = s_p[0];
#pragma unroll
  for (int i = 1; i < kWavesPerBlock; i++) t = s_p[i] < t ? s_p[i] : t;
  __syncthreads();
  return t;
}

// Error-free transformation: s + e == a + b exactly (Knuth TwoSum).
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e)
{
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Kahan-compensated accumulator in FP32.
struct Kahan {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float x)
  {
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  // merge another compensated partial (s2 - c2 is its value)
  __device__ __forceinline__ void merge(float s2, float c2)
  {
    float hs, he;
    two_sum(s, s2, hs, he);
    c = c + c2 - he;
    s = hs;
  }
  __device__ __forceinline__ float value() const { return s - c; }
};

__device__ __forceinline__ void kahan_wave_merge(Kahan &k, int width)
{
  for (int o = width / 2; o > 0; o >>= 1) {
    float s2 = __shfl_xor(k.s, o, kWave);
    float c2 = __shfl_xor(k.c, o, kWave);
    k.merge(s2, c2);
  }
}

}  // namespace grd
