// pr.cpp -- PageRank operators bound to the device path
// (reference src/graph_algorithms/pr.cpp:20-206).
#include "graph_algorithms/pr.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>

template <typename VP>
PageRank<VP>::PageRank(NODE_TYPE node_type, std::string algorithm_name, std::string graph_path, size_t num_compute,
                       size_t num_memory, uint32_t node_id, Exchange &net, std::string partitioning_scheme_file)
    : PageRank(node_type, algorithm_name, graph_path, num_compute, num_memory, node_id, net, partitioning_scheme_file,
               GRD_PR)
{
}

// PageRank<double> (pr.hpp:39) is instantiated, and bindDevice refuses it
// with the library's GRD_ERR_ARG message: FP64 vertex properties are not built

template <typename VP>
PageRank<VP>::PageRank(NODE_TYPE node_type, std::string algorithm_name, std::string graph_path, size_t num_compute,
                       size_t num_memory, uint32_t node_id, Exchange &net, std::string partitioning_scheme_file,
                       int grd_algo)
    : GraphAlgorithm<VP>(node_type, algorithm_name, graph_path, num_compute, num_memory, node_id, net,
                         partitioning_scheme_file)
{
  this->algorithm_name = grd_algo == GRD_PR ? "PageRank" : "PageRankPull";
  this->bindDevice(grd_algo);
}

// pr.cpp:20-61: prop = d(1-d)/deg, pr = 1-d for vertices with out-edges
// (frontier), sinks at 0
template <typename VP>
void PageRank<VP>::init()
{
  this->check(grd_init(this->state, &this->params), "init");
}

// pr.cpp:63-88 runs on the frontier right after update_frontier; on the
// device both are one fused pass (k_update), so this step has no work left
template <typename VP>
void PageRank<VP>::apply_updates()
{
}

// pr.cpp:90-121: push prop along the local out-edges of every updated vertex
template <typename VP>
void PageRank<VP>::gen_updates()
{
  this->check(grd_traverse(this->state, this->option), "gen_updates");
}

// pr.cpp:123-147 (+ apply_updates fused)
template <typename VP>
void PageRank<VP>::update_frontier()
{
  this->check(grd_update(this->state, &this->frontier_count), "update_frontier");
}

// pr.cpp:149-153: addUpdate on the vertex with local id lid of the compute
// partition the worker serves
template <typename VP>
void PageRank<VP>::aggregate(GNode &lid, const VP &buffer_val)
{
  this->check(grd_aggregate_one(this->state, (int)this->aggregate_partition, lid, &buffer_val), "aggregate");
}

template <typename VP>
bool PageRank<VP>::termination_check()
{
  return this->frontier_count == 0;
}

template <typename VP>
void PageRank<VP>::printState()
{
  std::vector<VP> r = this->result();
  for (size_t v = 0; v < r.size(); v++) std::printf("[Proc 0] Vertex/PR: %zu/%.9g\n", v, (double)r[v]);
}

// pr.cpp:178-206: top 20 values
template <typename VP>
void PageRank<VP>::verify()
{
  std::vector<VP> r = this->result();
  std::vector<uint32_t> idx(r.size());
  std::iota(idx.begin(), idx.end(), 0u);
  size_t k = std::min<size_t>(20, idx.size());
  std::partial_sort(idx.begin(), idx.begin() + k, idx.end(), [&](uint32_t a, uint32_t b) { return r[a] > r[b]; });
  std::printf("Top 20 PageRank Values:\n");
  for (size_t i = 0; i < k; i++) std::printf("Node: %u, PR: %.9g\n", idx[i], (double)r[idx[i]]);
}

template <typename VP>
PageRankPull<VP>::PageRankPull(NODE_TYPE node_type, std::string algorithm_name, std::string graph_path,
                               size_t num_compute, size_t num_memory, uint32_t node_id, Exchange &net,
                               std::string partitioning_scheme_file)
    : PageRank<VP>(node_type, algorithm_name, graph_path, num_compute, num_memory, node_id, net,
                   partitioning_scheme_file, GRD_PR_PULL)
{
}

template <typename VP>
void PageRankPull<VP>::init()
{
  this->check(grd_init(this->state, &this->params), "init");
}

// the normalised ranks sum to one; print the mass and the top 20
template <typename VP>
void PageRankPull<VP>::verify()
{
  std::vector<VP> r = this->result();
  double sum = 0.0;
  for (VP x : r) sum += (double)x;
  std::printf("PageRank mass: %.9f, last L1 change: %.3e\n", sum, this->stats().residual);
  PageRank<VP>::verify();
}

template class PageRank<float>;
template class PageRank<double>;
template class PageRankPull<float>;
