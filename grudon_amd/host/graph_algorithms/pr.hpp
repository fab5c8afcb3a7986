// pr.hpp -- PageRank operators (reference include/graph_algorithms/pr.hpp).
//
// PageRank<float> keeps the reference's delta-push semantics
// (src/graph_algorithms/pr.cpp): unnormalised, per-vertex activation above
// TOLERANCE, sinks start at 0.  PageRankPull<float> is the normalised pull
// form (FP32 Kahan contributions, FP64 dangling mass and residual).
#ifndef GRUDON_PR_HPP
#define GRUDON_PR_HPP

#include "GraphAlgorithm.hpp"

#define DAMPING_FACTOR 0.85
#define TOLERANCE 1e-6

template <typename VertexProperty>
class PageRank : public GraphAlgorithm<VertexProperty>
{
 public:
  PageRank(NODE_TYPE node_type, std::string algorithm_name, std::string graph_path, size_t num_compute,
           size_t num_memory, uint32_t node_id, Exchange &net, std::string partitioning_scheme_file = "");

  void init() override;
  void apply_updates() override;
  void gen_updates() override;
  void update_frontier() override;
  void aggregate(GNode &lid, const VertexProperty &buffer_val) override;
  bool termination_check() override;
  void printState() override;
  void verify() override;

 protected:
  PageRank(NODE_TYPE node_type, std::string algorithm_name, std::string graph_path, size_t num_compute,
           size_t num_memory, uint32_t node_id, Exchange &net, std::string partitioning_scheme_file, int grd_algo);
};

template <typename VertexProperty>
class PageRankPull : public PageRank<VertexProperty>
{
 public:
  PageRankPull(NODE_TYPE node_type, std::string algorithm_name, std::string graph_path, size_t num_compute,
               size_t num_memory, uint32_t node_id, Exchange &net, std::string partitioning_scheme_file = "");
  void init() override;
  void verify() override;
};

#endif  // GRUDON_PR_HPP
