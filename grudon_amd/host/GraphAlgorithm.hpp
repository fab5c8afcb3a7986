// GraphAlgorithm.hpp -- the operator API (reference include/GraphAlgorithm.hpp:24-86).
//
// Same constructor arguments and pure virtuals as the reference; the
// MPICore& argument becomes the on-device Exchange&.  The device state is
// created by the concrete algorithm's constructor (bindDevice), which knows
// the operator and the property width.  Instantiated for float, double,
// uint64_t and uint32_t like the reference (GraphAlgorithm.hpp:88-92); 8-byte
// properties are accepted by CC and SSSP and refused by PageRank (the C-ABI's
// grd_state_create_typed says why).  run() is the BSP
// superstep loop of src/GraphAlgorithm.cpp:145-701: per iteration the
// schedule decision (the offload vote of :187-195: push/pull x sparse/dense
// for the offloaded superstep, or NO_OFFLOAD edge shipping), the compute ->
// memory broadcast, gen_updates on the memory partitions (or the shipping of
// the frontier's edge lists), the memory -> compute reduce with aggregate()
// applied in bulk on the device, update_frontier + apply_updates, and the
// termination check.
#ifndef GRUDON_GRAPHALGORITHM_HPP
#define GRUDON_GRAPHALGORITHM_HPP

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "DistributedGraph.hpp"
#include "Exchange.hpp"
#include "Graph.hpp"
#include "Workers.hpp"

#define MAX_ITERATIONS 1000

template <typename VertexProperty>
class GraphAlgorithm
{
 public:
  GraphAlgorithm(NODE_TYPE node_type, std::string algorithm_name, std::string graph_path, size_t num_compute,
                 size_t num_memory, uint32_t node_id, Exchange &net, std::string &partitioning_scheme_file);
  virtual ~GraphAlgorithm();

  virtual void init() = 0;
  virtual void apply_updates() = 0;
  virtual void gen_updates() = 0;
  virtual void update_frontier() = 0;
  virtual void aggregate(GNode &lid, const VertexProperty &buffer_val) = 0;
  virtual bool termination_check() = 0;
  virtual void printState() = 0;
  virtual void verify() = 0;

  void run();

  // extensions of this build (defaults reproduce the reference)
  void setSchedule(grd_sched sched, int fixed_option = 0);
  void setMaxIterations(uint32_t max_iter) { max_iterations = max_iter; }
  const grd_run_stats &stats() const { return last_stats; }
  std::vector<VertexProperty> result();  // by gid
  const DistributedGraph &graph() const { return *dgraph; }
  // read-back of the current superstep's bitCommVector + propertyBuffers of
  // pair (c, m): dir 0 = broadcast c -> m, dir 1 = reduce m -> c; the
  // vertices whose bit is set, ascending gid, with the values packed for them
  void exchangeView(int dir, uint32_t c, uint32_t m, std::vector<GNode> &gids, std::vector<VertexProperty> &values);
  // aggregate(lid, value) for a local id of compute partition c (the
  // reference's process for partition c calls aggregate with its own lids)
  void aggregateOn(uint32_t c, GNode &lid, const VertexProperty &buffer_val)
  {
    aggregate_partition = c;
    aggregate(lid, buffer_val);
  }
  Worker<VertexProperty> &computeWorker() { return *worker; }
  Worker<VertexProperty> &memoryWorker() { return *traverser; }

 protected:
  std::string algorithm_name;
  std::string graph_path;
  NODE_TYPE node_type;
  Exchange &net;
  uint32_t num_compute;
  uint32_t num_memory;
  uint32_t num_workers;
  uint32_t worker_completion_count = 0;

  std::unique_ptr<DistributedGraph> dgraph;
  std::unique_ptr<Worker<VertexProperty>> worker;     // compute partitions
  std::unique_ptr<Worker<VertexProperty>> traverser;  // memory partitions

  grd_state *state = nullptr;
  grd_params params{};
  uint32_t aggregate_partition = 0;  // compute partition of the lids aggregate() receives
  int option = GRD_OPT_PUSH_SPARSE;  // schedule of the current superstep
  uint64_t frontier_count = 0;
  uint32_t max_iterations = MAX_ITERATIONS;
  grd_run_stats last_stats{};

  void check(grd_status st, const char *what);
  // creates the device state for operator grd_algo with sizeof(VertexProperty)
  void bindDevice(int grd_algo);

  friend class UpdateWorker<VertexProperty>;
  friend class TraverseWorker<VertexProperty>;
  friend class AggregateWorker<VertexProperty>;
  friend class Worker<VertexProperty>;
};

#endif  // GRUDON_GRAPHALGORITHM_HPP
