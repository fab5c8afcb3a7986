// sssp.hpp -- single-source shortest paths (reference
// include/graph_algorithms/sssp.hpp, src/graph_algorithms/sssp.cpp).
//
// The reference computes hop counts from gid 0 (dist + 1 per edge,
// sssp.cpp:69).  setSource / setWeighted / setDelta are extensions: any
// source, edge weights, and the delta-stepping bucket rule.
#ifndef GRUDON_SSSP_HPP
#define GRUDON_SSSP_HPP

#include "GraphAlgorithm.hpp"

template <typename VertexProperty>
class SSSP : public GraphAlgorithm<VertexProperty>
{
 public:
  SSSP(NODE_TYPE node_type, std::string algorithm_name, std::string graph_path, size_t num_compute, size_t num_memory,
       uint32_t node_id, Exchange &net, std::string partitioning_scheme_file = "");

  void init() override;
  void apply_updates() override;
  void gen_updates() override;
  void update_frontier() override;
  void aggregate(GNode &lid, const VertexProperty &buffer_val) override;
  bool termination_check() override;
  void printState() override;
  void verify() override;

  void setSource(uint32_t gid) { this->params.source = gid; }
  void setWeighted(bool w) { this->params.weighted = w ? 1u : 0u; }
  void setDelta(uint32_t delta) { this->params.delta = delta; }

  uint64_t violations = 0;  // set by verify()
};

#endif  // GRUDON_SSSP_HPP
