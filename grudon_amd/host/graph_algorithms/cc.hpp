// cc.hpp -- connected components by min-label propagation
// (reference include/graph_algorithms/cc.hpp, src/graph_algorithms/cc.cpp).
#ifndef GRUDON_CC_HPP
#define GRUDON_CC_HPP

#include "GraphAlgorithm.hpp"

template <typename VertexProperty>
class CC : public GraphAlgorithm<VertexProperty>
{
 public:
  CC(NODE_TYPE node_type, std::string algorithm_name, std::string graph_path, size_t num_compute, size_t num_memory,
     uint32_t node_id, Exchange &net, std::string partitioning_scheme_file = "");

  void init() override;
  void apply_updates() override;
  void gen_updates() override;
  void update_frontier() override;
  void aggregate(GNode &lid, const VertexProperty &buffer_val) override;
  bool termination_check() override;
  void printState() override;
  void verify() override;

  uint64_t violations = 0;  // set by verify()
};

#endif  // GRUDON_CC_HPP
