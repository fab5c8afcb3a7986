// Graph.hpp -- node roles and offload decisions of the operator API
// (reference include/Graph.hpp:9-24).  One process owns every partition, so
// GNode is a 32-bit vertex id and the graph type itself lives behind the
// C-ABI (grd_graph).
#ifndef GRUDON_GRAPH_HPP
#define GRUDON_GRAPH_HPP

#include <cstdint>
#include <vector>

using GNode = uint32_t;

enum NODE_TYPE
{
  COMPUTE_NODE,
  MEMORY_NODE,
  SWITCH_NODE
};

enum OFFLOAD_DECISION
{
  NDP_OFFLOAD,
  NO_OFFLOAD,
  INC_OFFLOAD,
};

// Host bitvector with the verbs the operator API applies to its
// bitCommVector entries (the reference uses galois::DynamicBitSet):
// resize / size / test / set / reset / count / getOffsets (set positions,
// ascending).  Bits live in 64-bit words, bit i in word i / 64.
class BitVector
{
 public:
  std::vector<uint64_t> bitvec;

  void resize(size_t n)
  {
    nbits_ = n;
    bitvec.assign((n + 63) / 64, 0ull);
  }
  size_t size() const { return nbits_; }
  // bytes of the word array, as the wire carries it
  size_t size_bytes() const { return bitvec.size() * sizeof(uint64_t); }
  bool test(size_t i) const { return (bitvec[i >> 6] >> (i & 63)) & 1ull; }
  void set(size_t i) { bitvec[i >> 6] |= 1ull << (i & 63); }
  void reset(size_t i) { bitvec[i >> 6] &= ~(1ull << (i & 63)); }
  void reset()
  {
    for (auto &w : bitvec) w = 0ull;
  }
  uint64_t count() const
  {
    uint64_t c = 0;
    for (uint64_t w : bitvec) c += (uint64_t)__builtin_popcountll(w);
    return c;
  }
  std::vector<GNode> getOffsets() const
  {
    std::vector<GNode> out;
    for (size_t w = 0; w < bitvec.size(); w++)
      for (uint64_t x = bitvec[w]; x; x &= x - 1) out.push_back((GNode)(w * 64 + (size_t)__builtin_ctzll(x)));
    return out;
  }

 private:
  size_t nbits_ = 0;
};

#endif  // GRUDON_GRAPH_HPP
