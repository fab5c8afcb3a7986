// grd_ops.hpp -- operator policies of the BSP traversal and their combine
// rules (shared by the push, pull and exchange kernels).
#pragma once

#include "grd_device.hpp"

namespace grd {

// CC: label propagation, atomicMin (cc.cpp:65-78, DistributedGraph.hpp:103-112)
struct OpCC {
  using V = uint32_t;
  static constexpr bool kMin = true;
  __device__ static V edge(V v, uint32_t) { return v; }
};
// SSSP hop count: dist + 1 (sssp.cpp:69); the unreached value stays unreached
struct OpHop {
  using V = uint32_t;
  static constexpr bool kMin = true;
  __device__ static V edge(V v, uint32_t) { return v == 0xFFFFFFFFu ? v : v + 1u; }
};
// SSSP weighted (extension): saturating dist + w (the unreached value stays
// unreached)
struct OpWeighted {
  using V = uint32_t;
  static constexpr bool kMin = true;
  __device__ static V edge(V v, uint32_t w)
  {
    uint32_t t = v + w;
    return t < v ? 0xFFFFFFFFu : t;
  }
};
// Reference PageRank: push the vertex property, atomicAdd (pr.cpp:95-120)
struct OpPR {
  using V = float;
  static constexpr bool kMin = false;
  __device__ static V edge(V v, uint32_t) { return v; }
};

template <class Op>
__device__ __forceinline__ typename Op::V op_identity()
{
  if constexpr (Op::kMin) return 0xFFFFFFFFu;
  else return 0.0f;
}

template <class Op>
__device__ __forceinline__ typename Op::V op_combine(typename Op::V a, typename Op::V b)
{
  if constexpr (Op::kMin) return a < b ? a : b;
  else return a + b;
}

__device__ __forceinline__ void set_bit(uint64_t *bits, uint32_t i)
{
  uint64_t b = 1ull << (i & 63);
  uint64_t *wd = bits + (i >> 6);
  if (!(__atomic_load_n(wd, __ATOMIC_RELAXED) & b)) atomicOr((unsigned long long *)wd, (unsigned long long)b);
}

// minUpdate / addUpdate, marking the dirty bit (DistributedGraph.hpp:103-130).
// minUpdate marks only on improvement.  addUpdate always marks: on the
// memory side a proxy's update is zero until its first add of the iteration
// (the previous reduce cleared it), so the first add (old == 0) marks; on the
// compute side residuals below the threshold persist, so every add marks.
template <class Op, bool kAlwaysMark = false>
__device__ __forceinline__ void relax(typename Op::V *upd, uint64_t *dirty, uint32_t dst, typename Op::V cand)
{
  if constexpr (Op::kMin) {
    uint32_t cur = __atomic_load_n(upd + dst, __ATOMIC_RELAXED);
    if (cand < cur) {
      uint32_t old = atomicMin(upd + dst, cand);
      if (cand < old) set_bit(dirty, dst);
    }
  } else {
    float old = atomicAdd(upd + dst, cand);
    if (kAlwaysMark || old == 0.0f) set_bit(dirty, dst);
  }
}

// relax without the dirty mark; returns whether the destination must be
// marked (same rules as relax)
template <class Op, bool kAlwaysMark = false>
__device__ __forceinline__ bool relax_nomark(typename Op::V *upd, uint32_t dst, typename Op::V cand)
{
  if constexpr (Op::kMin) {
    uint32_t cur = __atomic_load_n(upd + dst, __ATOMIC_RELAXED);
    if (cand < cur) return cand < atomicMin(upd + dst, cand);
    return false;
  } else {
    float old = atomicAdd(upd + dst, cand);
    return kAlwaysMark || old == 0.0f;
  }
}

// Sets bit i for every lane with pred.  Called by the whole wave: when all
// marked bits fall into one 64-bit word (a wave over a word of own proxies
// maps onto one word of slots) they are OR-ed across the wave and written by
// one atomic instead of up to 64 atomics on the same word.
__device__ __forceinline__ void mark_wave(uint64_t *bits, uint32_t i, bool pred)
{
  const uint64_t m = __ballot(pred);
  if (m == 0) return;
  const int leader = __ffsll((unsigned long long)m) - 1;
  const uint32_t wl = __shfl(i >> 6, leader, kWave);
  const bool same = !pred || (i >> 6) == wl;
  if (__ballot(!same) == 0) {
    uint32_t lo = 0, hi = 0;
    if (pred) {
      if ((i & 63) < 32) lo = 1u << (i & 31);
      else hi = 1u << (i & 31);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      lo |= __shfl_xor(lo, o, kWave);
      hi |= __shfl_xor(hi, o, kWave);
    }
    if (lane_id() == leader) atomicOr((unsigned long long *)(bits + wl), ((unsigned long long)hi << 32) | lo);
  } else if (pred) {
    set_bit(bits, i);
  }
}

// pull launchers (pull.hip)
template <class Op>
grd_status bsp_pull_launch(grd_state *s, const uint32_t *cw);
template <class Op>
grd_status bsp_pull_fixup(grd_state *s);
// CC / SSSP pull measured with time_kernel: counts the cold arcs of done rows
grd_status pull_done_arcs(grd_state *s);
// the hot pass of the hot / cold split for the BSP pull (tiers.hip)
template <class Op>
grd_status bsp_pull_hot(grd_state *s);

}  // namespace grd
