// Workers.hpp -- role adapters of the operator API (reference
// include/Workers.hpp:12-115, src/Workers.cpp).
//
// A Worker is one node of the reference's process model: node_id < C is
// compute partition node_id, node_id >= C memory partition node_id - C
// (src/main.cpp:92-93).  UpdateWorker drives the compute partitions (update =
// update_frontier then apply_updates, aggregate = the algorithm's aggregate),
// TraverseWorker the memory partitions (traverse = gen_updates),
// AggregateWorker the switch's in-network aggregation (aggregate only).
//
// The public members are the reference's, filled from the resident layout:
//   sTranslationTable[p] / rTranslationTable[p]  local id <-> wire index of the
//     pair with peer p (DistributedGraph.cpp:632-741).  Local ids are the
//     reference's: compute lid = gid / C; memory lids count the partition's
//     own vertices (ascending gid), then its cross mirrors (ascending gid).
//   bitCommVector[p]  sized like the pair's table (DistributedGraph.cpp:760-775);
//     loadExchange() fills it with the current superstep's message.
//   out_degrees / coverage_vector  by compute lid (DistributedGraph.cpp:286-370).
// One process owns every partition, so a GraphAlgorithm shares one resident
// DistributedGraph between its workers (the second constructor); the
// reference constructor builds and owns its own.
#ifndef GRUDON_WORKERS_HPP
#define GRUDON_WORKERS_HPP

#include <algorithm>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
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
  // reference constructor (include/Workers.hpp:16-23): loads and partitions
  // the graph itself
  Worker(std::string &graph_path, size_t &num_compute, size_t &num_memory, uint32_t &node_id, NODE_TYPE node_type,
         Exchange &net, std::string partitioning_scheme_file = "")
      : net(net), num_compute(num_compute), num_memory(num_memory), node_id(node_id), node_type(node_type)
  {
    owned_.reset(new DistributedGraph(net, load_host_graph(net, graph_path), num_compute, num_memory,
                                      partitioning_scheme_file));
    distributed_graph = owned_.get();
    fill_tables();
  }
  // shared resident graph (one process, every partition)
  Worker(DistributedGraph *graph, size_t num_compute, size_t num_memory, uint32_t node_id, NODE_TYPE node_type,
         Exchange &net)
      : distributed_graph(graph),
        net(net),
        num_compute(num_compute),
        num_memory(num_memory),
        node_id(node_id),
        node_type(node_type)
  {
    fill_tables();
  }
  virtual ~Worker() = default;

  // master = gid % C, mirror = gid % M or the partition file (Partitioner.hpp)
  uint32_t getVertexComputePartition(const GNode &gid) { return (uint32_t)distributed_graph->getMasterPartition(gid); }
  uint32_t getVertexMemoryPartition(const GNode &gid) { return (uint32_t)distributed_graph->getMirrorPartition(gid); }

  bool isCompute() const { return node_id < num_compute; }
  // compute partition c or memory partition m of this node
  uint32_t partition() const { return isCompute() ? node_id : node_id - (uint32_t)num_compute; }
  // global id of one of this node's local ids
  GNode globalId(GNode lid) const { return isCompute() ? lid * (GNode)num_compute + node_id : mem_gid_[lid]; }

  // bitCommVector[p] := the current superstep's message between this node
  // and peer p: dir 0 the broadcast compute -> memory, dir 1 the reduce
  // memory -> compute (bits at wire indices); values[p] receives the packed
  // properties in wire order, as propertyBuffers[p] holds them
  void loadExchange(GraphAlgorithm<T> &algorithm, int dir, std::vector<std::vector<T>> *values = nullptr)
  {
    const size_t peers = bitCommVector.size();
    if (values) values->assign(peers, {});
    for (size_t p = 0; p < peers; p++) {
      const uint32_t c = isCompute() ? node_id : (uint32_t)p;
      const uint32_t m = isCompute() ? (uint32_t)p : partition();
      std::vector<GNode> gids;
      std::vector<T> vals;
      algorithm.exchangeView(dir, c, m, gids, vals);
      bitCommVector[p].reset();
      std::vector<std::pair<GNode, T>> byidx;
      byidx.reserve(gids.size());
      for (size_t i = 0; i < gids.size(); i++) {
        const GNode idx = sTranslationTable[p].at(localId(gids[i]));
        bitCommVector[p].set(idx);
        byidx.emplace_back(idx, vals[i]);
      }
      if (values) {
        std::sort(byidx.begin(), byidx.end(),
                  [](const std::pair<GNode, T> &a, const std::pair<GNode, T> &b) { return a.first < b.first; });
        for (auto &x : byidx) (*values)[p].push_back(x.second);
      }
    }
  }

  std::vector<BitVector> bitCommVector;
  std::vector<std::unordered_map<GNode, GNode>> sTranslationTable;
  std::vector<std::unordered_map<GNode, GNode>> rTranslationTable;
  DistributedGraph *distributed_graph = nullptr;
  uint64_t num_vertices = 0;
  std::vector<uint8_t> coverage_vector;  // by compute lid: every out-neighbour shares the master
  std::vector<uint64_t> out_degrees;     // by compute lid
  Exchange &net;
  size_t num_compute;
  size_t num_memory;
  uint32_t node_id;
  uint64_t total_vertices = 0;
  uint64_t num_edges = 0;
  NODE_TYPE node_type;

  virtual void update(GraphAlgorithm<T> &algorithm) = 0;
  virtual void traverse(GraphAlgorithm<T> &algorithm) = 0;
  virtual void aggregate(GraphAlgorithm<T> &algorithm, GNode &lid, const T &buffer_val) = 0;

 private:
  std::unique_ptr<DistributedGraph> owned_;
  std::vector<GNode> mem_gid_;                  // memory node: memory lid -> gid
  std::unordered_map<GNode, GNode> mem_lid_;    // memory node: gid -> memory lid

  GNode localId(GNode gid) const { return isCompute() ? gid / (GNode)num_compute : mem_lid_.at(gid); }

  void fill_tables()
  {
    if (node_id >= num_compute + num_memory) throw std::runtime_error("Worker: node id beyond C + M");
    const HostGraph &h = distributed_graph->host();
    const uint32_t C = (uint32_t)num_compute;
    total_vertices = h.n;
    num_edges = h.num_edges();
    const size_t peers = isCompute() ? num_memory : num_compute;
    std::vector<std::vector<GNode>> tables(peers);
    for (size_t p = 0; p < peers; p++)
      tables[p] = isCompute() ? distributed_graph->translationTable(node_id, (uint32_t)p)
                              : distributed_graph->translationTable((uint32_t)p, partition());
    if (isCompute()) {
      num_vertices = h.n > node_id ? (h.n - node_id + C - 1) / C : 0;
      out_degrees.resize(num_vertices);
      std::vector<uint8_t> cov(h.n);
      if (grd_graph_coverage(distributed_graph->handle(), cov.data(), nullptr) != GRD_OK)
        throw std::runtime_error(std::string("coverage vector: ") + grd_last_error());
      coverage_vector.resize(num_vertices);
      for (uint64_t l = 0; l < num_vertices; l++) {
        const GNode g = (GNode)(l * C + node_id);
        out_degrees[l] = distributed_graph->getOutDegree(g);
        coverage_vector[l] = cov[g];
      }
    } else {
      // memory lids: own vertices (ascending gid), then cross mirrors
      // (ascending gid), DistributedGraph.cpp:517-525, 579-588
      const uint32_t m = partition();
      std::vector<GNode> cross;
      for (uint32_t g = 0; g < h.n; g++)
        if (distributed_graph->getMirrorPartition(g) == m) mem_gid_.push_back(g);
      for (auto &t : tables)
        for (GNode g : t)
          if (distributed_graph->getMirrorPartition(g) != m) cross.push_back(g);
      std::sort(cross.begin(), cross.end());
      cross.erase(std::unique(cross.begin(), cross.end()), cross.end());
      mem_gid_.insert(mem_gid_.end(), cross.begin(), cross.end());
      num_vertices = mem_gid_.size();
      mem_lid_.reserve(mem_gid_.size());
      for (size_t l = 0; l < mem_gid_.size(); l++) mem_lid_[mem_gid_[l]] = (GNode)l;
    }
    bitCommVector.assign(peers, BitVector());
    sTranslationTable.assign(peers, {});
    rTranslationTable.assign(peers, {});
    for (size_t p = 0; p < peers; p++) {
      bitCommVector[p].resize(tables[p].size());
      sTranslationTable[p].reserve(tables[p].size());
      rTranslationTable[p].reserve(tables[p].size());
      for (size_t i = 0; i < tables[p].size(); i++) {
        const GNode lid = localId(tables[p][i]);
        sTranslationTable[p][lid] = (GNode)i;
        rTranslationTable[p][(GNode)i] = lid;
      }
    }
  }
};

template <typename T>
class UpdateWorker : public Worker<T>
{
 public:
  UpdateWorker(std::string &graph_path, size_t &num_compute, size_t &num_memory, uint32_t &node_id, NODE_TYPE node_type,
               Exchange &net, std::string &partitioning_scheme_file)
      : Worker<T>(graph_path, num_compute, num_memory, node_id, node_type, net, partitioning_scheme_file)
  {
  }
  UpdateWorker(DistributedGraph *graph, size_t num_compute, size_t num_memory, uint32_t node_id, NODE_TYPE node_type,
               Exchange &net)
      : Worker<T>(graph, num_compute, num_memory, node_id, node_type, net)
  {
  }
  // Workers.cpp:77-82
  void update(GraphAlgorithm<T> &algorithm) override
  {
    algorithm.update_frontier();
    algorithm.apply_updates();
  }
  // Workers.cpp:71-75: lid is a local id of this compute partition
  void aggregate(GraphAlgorithm<T> &algorithm, GNode &lid, const T &buffer_val) override
  {
    algorithm.aggregateOn(this->node_id, lid, buffer_val);
  }
  void traverse(GraphAlgorithm<T> &) override {}
};

template <typename T>
class TraverseWorker : public Worker<T>
{
 public:
  TraverseWorker(std::string &graph_path, size_t &num_compute, size_t &num_memory, uint32_t &node_id,
                 NODE_TYPE node_type, Exchange &net, std::string &partitioning_scheme_file)
      : Worker<T>(graph_path, num_compute, num_memory, node_id, node_type, net, partitioning_scheme_file)
  {
  }
  TraverseWorker(DistributedGraph *graph, size_t num_compute, size_t num_memory, uint32_t node_id, NODE_TYPE node_type,
                 Exchange &net)
      : Worker<T>(graph, num_compute, num_memory, node_id, node_type, net)
  {
  }
  // Workers.cpp:109-113
  void traverse(GraphAlgorithm<T> &algorithm) override { algorithm.gen_updates(); }
  void update(GraphAlgorithm<T> &) override {}
  void aggregate(GraphAlgorithm<T> &, GNode &, const T &) override {}
};

// The switch node's worker (Workers.cpp:139-170): aggregate only.  Its
// node_id names the compute partition whose local ids it receives.
template <typename T>
class AggregateWorker : public Worker<T>
{
 public:
  AggregateWorker(std::string &graph_path, size_t &num_compute, size_t &num_memory, uint32_t &node_id,
                  NODE_TYPE node_type, Exchange &net)
      : Worker<T>(graph_path, num_compute, num_memory, node_id, node_type, net)
  {
  }
  AggregateWorker(DistributedGraph *graph, size_t num_compute, size_t num_memory, uint32_t node_id,
                  NODE_TYPE node_type, Exchange &net)
      : Worker<T>(graph, num_compute, num_memory, node_id, node_type, net)
  {
  }
  void aggregate(GraphAlgorithm<T> &algorithm, GNode &lid, const T &buffer_val) override
  {
    algorithm.aggregateOn(this->node_id, lid, buffer_val);
  }
  void update(GraphAlgorithm<T> &) override {}
  void traverse(GraphAlgorithm<T> &) override {}
};

#endif  // GRUDON_WORKERS_HPP
