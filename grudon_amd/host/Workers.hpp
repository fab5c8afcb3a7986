// Workers.hpp -- role adapters of the operator API (reference
// include/Workers.hpp:12-115, src/Workers.cpp).
//
// UpdateWorker drives the compute partitions (update = update_frontier then
// apply_updates, aggregate = the algorithm's aggregate), TraverseWorker the
// memory partitions (traverse = gen_updates).  With every partition in one
// process a GraphAlgorithm owns one worker of each role, both sharing the
// resident DistributedGraph.
#ifndef GRUDON_WORKERS_HPP
#define GRUDON_WORKERS_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "DistributedGraph.hpp"
#include "Exchange.hpp"
#include "Graph.hpp"

template <typename T>
class GraphAlgorithm;

template <typename T>
class Worker
{
 public:
  Worker(DistributedGraph *graph, size_t num_compute, size_t num_memory, uint32_t node_id, NODE_TYPE node_type,
         Exchange &net)
      : distributed_graph(graph),
        net(net),
        num_compute(num_compute),
        num_memory(num_memory),
        node_id(node_id),
        node_type(node_type)
  {
    const HostGraph &h = graph->host();
    total_vertices = h.n;
    num_vertices = h.n;
    num_edges = h.num_edges();
    out_degrees.resize(h.n);
    for (uint32_t v = 0; v < h.n; v++) out_degrees[v] = h.row_ptr[v + 1] - h.row_ptr[v];
    coverage_vector.resize(h.n);
    if (grd_graph_coverage(graph->handle(), coverage_vector.data(), nullptr) != GRD_OK)
      throw std::runtime_error(std::string("coverage vector: ") + grd_last_error());
  }
  virtual ~Worker() = default;

  // master = gid % C, mirror = gid % M or the partition file (Partitioner.hpp)
  uint32_t getVertexComputePartition(const GNode &gid) { return (uint32_t)distributed_graph->getMasterPartition(gid); }
  uint32_t getVertexMemoryPartition(const GNode &gid) { return (uint32_t)distributed_graph->getMirrorPartition(gid); }

  DistributedGraph *distributed_graph;
  uint64_t num_vertices = 0;
  uint64_t total_vertices = 0;
  uint64_t num_edges = 0;
  std::vector<uint64_t> out_degrees;  // by gid
  // 1 when every out-neighbour shares the vertex's master (the vector the
  // NDPEngine vote reads, DistributedGraph.cpp:286-370), by gid
  std::vector<uint8_t> coverage_vector;
  Exchange &net;
  size_t num_compute;
  size_t num_memory;
  uint32_t node_id;
  NODE_TYPE node_type;

  virtual void update(GraphAlgorithm<T> &algorithm) = 0;
  virtual void traverse(GraphAlgorithm<T> &algorithm) = 0;
  virtual void aggregate(GraphAlgorithm<T> &algorithm, GNode &lid, const T &buffer_val) = 0;
};

template <typename T>
class UpdateWorker : public Worker<T>
{
 public:
  using Worker<T>::Worker;
  // Workers.cpp:77-82
  void update(GraphAlgorithm<T> &algorithm) override
  {
    algorithm.update_frontier();
    algorithm.apply_updates();
  }
  void aggregate(GraphAlgorithm<T> &algorithm, GNode &lid, const T &buffer_val) override
  {
    algorithm.aggregate(lid, buffer_val);
  }
  void traverse(GraphAlgorithm<T> &) override {}
};

template <typename T>
class TraverseWorker : public Worker<T>
{
 public:
  using Worker<T>::Worker;
  // Workers.cpp:109-113
  void traverse(GraphAlgorithm<T> &algorithm) override { algorithm.gen_updates(); }
  void update(GraphAlgorithm<T> &) override {}
  void aggregate(GraphAlgorithm<T> &, GNode &, const T &) override {}
};

#endif  // GRUDON_WORKERS_HPP
