// DistributedGraph.hpp -- the partitioned graph of the operator API
// (reference include/DistributedGraph.hpp:222-287, src/DistributedGraph.cpp).
//
// The reference builds one rank's share per process; here the whole C x M
// layout (compute slots, memory proxies, cross mirrors, translation tables,
// local CSR/CSC) is built once and kept resident in HBM by grd_graph_from_csr.
// The host keeps the global CSR for input/verification and answers the same
// queries: local <-> global ids of the compute partitions (lid = gid / C,
// DistributedGraph.cpp:170-191), master and mirror partition of a vertex,
// out-degrees and translation tables.
#ifndef GRUDON_DISTRIBUTEDGRAPH_HPP
#define GRUDON_DISTRIBUTEDGRAPH_HPP

#include <cstdint>
#include <string>
#include <vector>

#include "Exchange.hpp"
#include "Graph.hpp"

// Global CSR kept on the host.
struct HostGraph {
  uint32_t n = 0;
  std::vector<uint32_t> row_ptr, col, weights;  // weights empty when the graph has none
  uint64_t num_edges() const { return col.size(); }
};

// Load a graph: a .gr / .mtx file, or a seeded generator spec
//   rmat:SCALE[:EF[:SEED[:sym[:WSEED]]]]   (a/b/c/d .57/.19/.19/.05)
//   er:N:MEAN_DEGREE[:SEED[:WSEED]]
// Throws std::runtime_error with the library's message on failure.
HostGraph load_host_graph(Exchange &net, const std::string &spec);

// One memory-partition id per line (Partitioner.hpp:89-105).
std::vector<uint32_t> load_partition_file(const std::string &path, uint32_t n, size_t num_memory);

class DistributedGraph
{
 public:
  DistributedGraph(Exchange &net, const HostGraph &g, size_t num_compute, size_t num_memory,
                   const std::string &partitioning_scheme_file);
  ~DistributedGraph();
  DistributedGraph(const DistributedGraph &) = delete;
  DistributedGraph &operator=(const DistributedGraph &) = delete;

  grd_graph *handle() const { return graph_; }
  const HostGraph &host() const { return host_; }

  // compute-partition ids (master = gid % C, lid = gid / C)
  GNode getLocalNode(const GNode &gid) const { return gid / (GNode)num_compute_; }
  GNode getGlobalNode(const GNode &lid, uint32_t compute_partition = 0) const
  {
    return lid * (GNode)num_compute_ + compute_partition;
  }
  uint64_t getOutDegree(const GNode &gid) const { return host_.row_ptr[gid + 1] - host_.row_ptr[gid]; }
  uint64_t getMasterPartition(const GNode &gid) const { return gid % num_compute_; }
  uint64_t getMirrorPartition(const GNode &gid) const { return mirror_.empty() ? gid % num_memory_ : mirror_[gid]; }
  // translation table of (c, m) as gids in wire-index order
  std::vector<GNode> translationTable(uint32_t c, uint32_t m) const;
  uint64_t getNumVertices() const { return host_.n; }
  uint64_t numProxies() const;
  uint64_t numCrossMirrors() const;

 private:
  HostGraph host_;
  std::vector<uint32_t> mirror_;
  size_t num_compute_, num_memory_;
  grd_graph *graph_ = nullptr;
};

#endif  // GRUDON_DISTRIBUTEDGRAPH_HPP
